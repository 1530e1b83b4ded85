// Internal declarations shared by the translation units of libmycel_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mycel_hip.h"
#include "ke_sym.hpp"

namespace mycel {

// ------------------------------------------------------------------------------------------------
// Errors: internal code throws mycel::Error; every C entry point converts it into a status code and
// a thread-local message (mycel_last_error).
// ------------------------------------------------------------------------------------------------
struct Error : std::runtime_error {
    int32_t status;
    Error(int32_t s, const std::string &msg) : std::runtime_error(msg), status(s) {}
};

void set_last_error(const std::string &msg);

#define MYCEL_HIP(call)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            throw ::mycel::Error(e_ == hipErrorOutOfMemory ? MYCEL_E_OUT_OF_MEMORY : MYCEL_E_DEVICE, \
                                 std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

#define MYCEL_LAUNCH_CHECK() MYCEL_HIP(hipGetLastError())

// ------------------------------------------------------------------------------------------------
// Device memory owned by a handle. No copies, moves only.
// ------------------------------------------------------------------------------------------------
template <typename T>
class DevBuf {
   public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { resize(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void resize(size_t n) {
        if (n == n_) return;
        release();
        if (n) MYCEL_HIP(hipMalloc(&p_, n * sizeof(T)));
        n_ = n;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *get() const { return p_; }
    size_t size() const { return n_; }
    void upload(const T *h, size_t n, hipStream_t s, size_t off = 0) {
        MYCEL_HIP(hipMemcpyAsync(p_ + off, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t n, hipStream_t s, size_t off = 0) const {
        MYCEL_HIP(hipMemcpyAsync(h, p_ + off, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }

   private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

// ------------------------------------------------------------------------------------------------
// Structured grid as the kernels see it. Element data live on a padded (nx+2)(ny+2)(nz+2) grid
// with a zero halo so that the node-centric gather never branches on the box boundary.
// ------------------------------------------------------------------------------------------------
struct Dims {
    int nx, ny, nz;
    int nx1, ny1, nz1;  // node counts per axis
    int n_nodes, n_dofs, n_elem;
    int px, py, pz;     // padded element grid
    int n_pad;          // px*py*pz
};

inline Dims make_dims(int nx, int ny, int nz) {
    Dims d;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.nx1 = nx + 1; d.ny1 = ny + 1; d.nz1 = nz + 1;
    d.n_nodes = d.nx1 * d.ny1 * d.nz1;
    d.n_dofs = 3 * d.n_nodes;
    d.n_elem = nx * ny * nz;
    d.px = nx + 2; d.py = ny + 2; d.pz = nz + 2;
    d.n_pad = d.px * d.py * d.pz;
    return d;
}

// DOF flag bits stored per DOF on the device.
enum : uint8_t { kFlagConstrained = 1, kFlagFixed = 2 };

// Apply modes of the K*u kernels (FemOperator.ApplyK / ApplyA).
enum ApplyMode : int { kApplyK = 0, kApplyAProjected = 1, kApplyAGeneral = 2 };

// Reduction partial-sum geometry: one FP64 partial per block, summed in a fixed order.
constexpr int kBlock = 256;

// Host-side Hex8 element (Fem/Hex8Element.cs).
void hex8_ke(double h, double nu, double *ke576);
void hex8_bc(double h, double *bc144);
void constitutive(double e, double nu, double *d36);

struct PcgWork;
struct Gmg;
struct OrdWork;
struct SolveBroker;

// Several right-hand sides in one launch: grid dimension y is the RHS index; vectors of RHS k start at
// k * vstride, block partials at k * pstride, and the solver-state word that stops a launch at k * sstride
// (in ints). count = 1 with zero strides is a single RHS.
struct Batch {
    int count = 1;
    long long vstride = 0;
    int pstride = 0;
    int sstride = 0;
};

}  // namespace mycel

// ------------------------------------------------------------------------------------------------
// The operator handle: FeModel + FemOperator state, device-resident.
// ------------------------------------------------------------------------------------------------
struct mycel_op {
    int device = 0;
    hipStream_t stream = nullptr;
    int ku_form = MYCEL_MATVEC_AUTO;  // K*u form of the launches without dot partials (mycel_op_set_matvec_form)
    int ku_kz = 0;                    // its node layers per thread or tile
    int ku_tuned[2] = {0, 0};         // autotuned tile depth of the automatic form's launches without dot partials
                                      // (FP64, FP32; 0 until the first such launch)
    mycel::Dims d{};
    double h = 1, e0 = 1, nu = 0.3, emin = 0, penal = 3;

    // Host model description (FeModel): kept for the counters and the getters.
    std::vector<uint8_t> domain;         // box elements
    std::vector<int32_t> active;         // ascending box element indices
    std::vector<uint8_t> active_nodes;   // box nodes
    std::vector<uint8_t> dof_flags;      // box DOFs (kFlag*)
    std::vector<int32_t> fixed;          // FixedDofs as given
    int active_fixed = 0;
    int n_active_nodes = 0;
    double ke[576];
    double diag_ke[24];
    mycel::KeSym<double> ks64;  // Ke as 14 symmetry-class values (kernel arguments)
    mycel::KeSym<float> ks32;
    mycel::KeMono<double> km64;  // Ke in the monomial basis (element-centric K*u)
    mycel::KeMono<float> km32;

    // Device state.
    mycel::DevBuf<double> d_ke;        // [0,576) Ke, [576,600) DiagKe, [600,744) Bc(1), [744,780) D(1,nu)
    mycel::DevBuf<float> d_ke32;       // FP32 Ke
    mycel::DevBuf<double> d_ep;        // padded E_e (0 for inactive / halo)
    mycel::DevBuf<float> d_ep32;       // FP32 copy of the same
    mycel::DevBuf<uint8_t> d_flags;    // per DOF
    mycel::DevBuf<uint8_t> d_domain;   // per box element (active mask)
    mycel::DevBuf<uint8_t> d_active_nodes;  // per box node (FeModel.ActiveNodes)
    mycel::DevBuf<int32_t> d_active;   // active list
    mycel::DevBuf<double> d_density;   // box densities (ElementDensity)
    mycel::DevBuf<double> d_diag;      // DiagA
    mycel::DevBuf<double> d_inv_diag;  // Jacobi 1/DiagA
    mycel::DevBuf<float> d_inv_diag32;
    mycel::DevBuf<double> d_x, d_y;    // staging for the host-pointer entry points
    mycel::DevBuf<double> d_part;      // block partials for one-off reductions
    mycel::PcgWork *pcg = nullptr;
    mycel::Gmg *gmg = nullptr;         // multigrid hierarchy (built on first use)
    mycel::OrdWork *ord = nullptr;     // workspace of the reference-order solver (built on first use)
    uint64_t epoch = 0;                // bumped whenever the element moduli change
    double island_ratio = 0.0;         // multigrid island deflation: solid threshold E >= ratio E0 (0: off)
    mycel::SolveBroker *broker = nullptr;  // job server: multigrid solves run in its packed launches (not owned)
};

namespace mycel {

// ---- fem.hip ---------------------------------------------------------------------------------
void op_upload_model(mycel_op *op);
void op_update_e(mycel_op *op, const double *d_density_box, double penal, hipStream_t s);
void launch_apply_k(mycel_op *op, const double *d_x, double *d_y, int mode, hipStream_t s,
                    double *d_dot_partial = nullptr, const int *d_stop = nullptr, const Batch &bt = Batch{});
// Blocks (= dot partials) of the K*u launch that writes x.y partials; batched: a multi-RHS launch.
int apply_k_blocks(const mycel_op *op);
// Node layers per tile of the automatic element-centric K*u form on this operator's grid.
int ku_element_layers(const mycel_op *op);
// Autotuner. Among launch variants whose results are bit-identical (the tile depth of the element-centric K*u
// march in launches without dot partials: each node's sum is formed in the same order at every depth), the fastest
// on this device for this grid, timed with HIP events on s the first time a (family, device, grid) key is seen in
// the process and reused afterwards. cand[0] is the rule's choice: returned when MYCEL_AUTOTUNE=0 or while s is
// being captured. launch(c) enqueues one launch of variant c on s.
enum TuneFamily : int { kTuneKu64 = 0, kTuneKu32 = 1, kTuneSmooth = 2 };
int autotune(int family, int device, const Dims &d, const int *cand, int n, hipStream_t s,
             const std::function<void(int)> &launch);
// The tile depths the autotuner tries for an element-centric march: the size rule's first, then the others.
int tile_depth_candidates(int rule_kz, int *out);
void launch_apply_k_f32(mycel_op *op, const float *d_x, float *d_y, int mode, hipStream_t s,
                        double *d_dot_partial = nullptr, const int *d_stop = nullptr);
void launch_von_mises(mycel_op *op, const double *d_u, double *d_vm, hipStream_t s);
void launch_sensitivities(mycel_op *op, const double *d_u, const double *d_rho, double *d_dc, double penal,
                          hipStream_t s);
void launch_avg_e2n(mycel_op *op, const double *d_elem, double *d_nodal, hipStream_t s);
// SampleNodalScalar at n_points points (float32 xyz triples); d_fallback: n_points ints of scratch.
void launch_sample_nodal(mycel_op *op, const double *d_nodal, int n_points, const float *d_pts, const float *origin,
                         double *d_out, int *d_fallback, hipStream_t s);

// Deterministic FP64 reductions (fixed block partition, fixed tree) on the op's geometry.
double reduce_dot(mycel_op *op, const double *d_a, const double *d_b, int n, hipStream_t s);
double reduce_dot_masked(mycel_op *op, const double *d_a, const double *d_b, int n, const uint8_t *d_flags,
                         hipStream_t s);

// ---- pcg.hip ---------------------------------------------------------------------------------
struct PcgOutcome {
    mycel_pcg_result r;
};
int32_t pcg_solve_device(mycel_op *op, const double *d_b, double *d_x, double tol, int max_iter, int precision,
                         const volatile int32_t *cancel, mycel_pcg_progress_cb cb, void *user,
                         mycel_pcg_result *res);
void pcg_free(PcgWork *w);
// R independent systems A x_k = b_k (b, x: R box-DOF vectors back to back) solved in lockstep, each with its
// own state and stopping test; MYCEL_PRECISION_MIXED_GMG only. res: R results.
int32_t pcg_solve_batch(mycel_op *op, const double *d_B, double *d_X, int R, double tol, int max_iter, int precision,
                        mycel_pcg_result *res);
// Packed multigrid solves: a broker thread advances every multigrid solve handed to it (op->broker set, no
// cancel flag or progress callback) in launches shared by all of them, joining and leaving between batches.
SolveBroker *broker_create(int device, int slots);  // slots: members of its pack (<= 16)
void broker_destroy(SolveBroker *b);
mycel_pack_stats broker_stats(SolveBroker *b);
// count (<= 16) independent multigrid systems on distinct operators, solved in packed launches by the calling
// thread; each result is bit-identical to the system's multigrid solve alone.
void pcg_solve_packed(mycel_op *const *ops, const double *const *d_b, double *const *d_x, int count, double tol,
                      int max_iter, mycel_pcg_result *res);

// ---- ordered.hip: the reference's evaluation order (MYCEL_PRECISION_FP64_ORDERED) ---------------
// K*u with the reference's per-DOF colour order; d_halt: a PCG halt flag that turns the launch into a no-op.
void launch_apply_k_ordered(mycel_op *op, const double *d_x, double *d_y, int mode, hipStream_t s,
                            const int *d_halt = nullptr);
int32_t pcg_solve_ordered(mycel_op *op, const double *d_b, double *d_x, double tol, int max_iter,
                          const volatile int32_t *cancel, mycel_pcg_progress_cb cb, void *user,
                          mycel_pcg_result *res);
// Sequential sum over i ascending of a_i * b_i (b null: a_i), skipping DOFs d_flags marks constrained.
double seq_sum(mycel_op *op, const double *d_a, const double *d_b, const uint8_t *d_flags, int n, hipStream_t s);
double oc_update_ordered(mycel_op *op, const double *d_rho, const double *d_dc, const double *d_s,
                         const uint8_t *d_passive, double *d_rho_new, int n, double vf, double move, double rho_min,
                         double lambda_guess, int max_bisect, double vol_tol, hipStream_t s, int *evals);
void simp_sums_ordered(mycel_op *op, int na, const double *d_rho_bar, const double *d_grey_src, double *out3,
                       hipStream_t s);
void ord_free(OrdWork *w);

// ---- gmg.hip ---------------------------------------------------------------------------------
// Builds the hierarchy on first use and refreshes its moduli, constraints and coarsest inverse when
// op->epoch changed (enqueued on s, no host synchronisation).
void gmg_prepare(mycel_op *op, hipStream_t s);
// Whether the grid has the multigrid hierarchy (2x coarsening per axis down to a dense coarsest level); the
// solvers fall back to the FP64 Jacobi iteration on grids that do not.
bool gmg_supported(const mycel_op *op);
// z = M(r) (one FP32 V-cycle) plus per-block partials of z.r and z.ap, gmg_partials(op) of each.
// fine_ready: the finest level's b, x = d = D^-1 b / theta were written by the caller (gmg_fine_view).
// n_rhs: right-hand sides in one pass (grid dimension y; vectors of RHS k at k * n_dofs of each level;
// partials of RHS k at k * gmg_partials(op)); at most what gmg_reserve provided. halt: the outer iteration's
// halt flag of RHS k at halt[k * halt_stride], or null; a halted RHS skips the coarse-level tail.
void gmg_enqueue_vcycle(mycel_op *op, const double *r64, const double *ap64, double *z64, double *part_zr,
                        double *part_zap, hipStream_t s, bool fine_ready = false, int n_rhs = 1,
                        const int *halt = nullptr, int halt_stride = 0);
// Sizes the level vectors for n_rhs right-hand sides (after gmg_prepare); bumps gmg_generation when they move.
void gmg_reserve(mycel_op *op, int n_rhs);
uint64_t gmg_generation(const mycel_op *op);
// The finest level's V-cycle inputs, for a caller that fuses their initialisation into its own kernel.
struct GmgFineView {
    float *b, *x, *d;
    const float *inv;   // D^-1 of the finest level: 0 on constrained DOFs (the mask)
    const float *binv;  // the smoother's nodal block inverses (6 per node: xx, yy, zz, xy, xz, yz)
    float theta_inv;
};
GmgFineView gmg_fine_view(mycel_op *op);
int gmg_partials(const mycel_op *op);
void gmg_free(Gmg *g);
// Packed V-cycles: several operators' V-cycles (different grids) in one launch per phase, grid dimension y the
// member, each member's arguments in a device table; each member's results are bit-identical to its V-cycle alone.
struct GmgPack;
struct GmgPackMember {
    mycel_op *op;
    const double *r64, *ap64;        // the outer iteration's residual and A p (FP64)
    double *z64, *part_zr, *part_zap;  // z and the z.r, z.Ap partials (gmg_partials(op) each)
    const int *halt;                 // the member's halt flag (its V-cycle is skipped once set)
};
// Whether op's V-cycle can join a pack: a prepared hierarchy of at least two levels, no island deflation and the
// default kernel set (element-centric finest level, slab-split coarse gathers, fused coarse residual-restriction).
bool gmg_packable(const mycel_op *op);
// A pack of `slots` member slots: grid dimension y of every launch; its launch list (one launch per step, kernel
// and tile depth any member has brought, with the largest grid and LDS seen) only grows.
GmgPack *gmg_pack_create(int slots);
void gmg_pack_free(GmgPack *p);
// Rewrites the device tables for the members in m[0 .. slots) (m[q].op null: slot q is empty, its entries zero),
// growing the launch list when a member brings a new step or a larger grid; synchronises s, then uploads on s.
void gmg_pack_build(GmgPack *p, const GmgPackMember *m, int count, hipStream_t s);
// One V-cycle of every member (each member's finest-level b, x, d written by the caller, as fine_ready).
void gmg_pack_enqueue(const GmgPack *p, hipStream_t s);
int gmg_pack_launches(const GmgPack *p);
// Bumped whenever the launch list, a grid, an LDS size or the table buffer changes (a captured graph is stale).
uint64_t gmg_pack_shape(const GmgPack *p);

// ---- capi.hip (internal forms of the entry points, used by the job server) -------------------
mycel_op *op_create(const mycel_grid &g, const uint8_t *domain, const int32_t *fixed, int n_fixed, double e0,
                    double nu, double emin_ratio, double penal, const double *density);
void op_destroy(mycel_op *op);
void simp_run(mycel_op *op, const double *f, int n_cases, const mycel_simp_options &o, const uint8_t *passive_mask,
              const int32_t *patch_nodes, int n_patch, const float *init_density, int init_len,
              const double *warm_u, mycel_snapshot_cb cb, void *user, mycel_simp_result *result,
              mycel_snapshot *snaps, int snaps_cap, double *rho_out, double *rho_bar_out, double *u_out);
void static_analysis_run(mycel_op *op, const double *f, double *u, double tol, int max_iter, bool throw_on_nonconv,
                         int precision, double *compliance, double *max_disp, double *reactions, double *von_mises,
                         mycel_pcg_result *result);

}  // namespace mycel
