// The reference's evaluation order on the GPU: MYCEL_PRECISION_FP64_ORDERED (Jacobi-PCG, SIMP) and the
// MYCEL_MATVEC_REFERENCE_ORDER form of K*u.
//
// Floating-point addition is not associative, so a CG trajectory depends on the order in which its sums are
// formed; over hundreds of iterations and tens of SIMP iterations those last-bit differences grow into
// visible ones. These kernels form every sum of the path in the order the reference forms it, without
// contracting a*b+c, so their results are bit-identical to the reference restatement:
//   - K*u (Fem/FemOperator.cs:196-243,379-422): y = 0, then the element colours 0..7 one after another, each
//     element adding its 24 rows; a row is the E-scaled gather ue = x * E dotted with the Ke row as four
//     lanes (columns j, j+4, ..., j+20, summed left to right) combined as (l0 + l1) + (l2 + l3). Elements of
//     one colour share no node, so per DOF this is the ordered sum over its <= 8 incident elements, one per
//     colour: one thread per node forms exactly that.
//   - PCG dot products (Fem/PcgSolver.cs:414-527): 8,192-DOF blocks, four lane accumulators per block over
//     the DOFs j, j+4, ..., then (l0 + l1) + (l2 + l3), the block's scalar tail, and the block partials summed
//     in block order. One thread per (block, lane) runs its lane's chain; one thread combines.
//   - ||Pb||^2 (:128-137), the compliance F.u (TopologyOptimizer.cs:256-260), the OC volumes
//     (OcOptimizer.cs:30-34,143-174) and the SIMP metrics (:323-338): plain sequential sums, formed by one
//     thread from LDS (terms that are exactly zero do not change a sequential sum and are skipped in
//     64-term groups).
// The vector updates are element-wise (no order to keep) and run on the whole grid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>

#include "device_common.hpp"
#include "mycel_internal.hpp"
#include "mycel_opt.hpp"

namespace mycel {

namespace {

constexpr int kPcgBlock = 8192;  // Fem/PcgSolver.cs:18
constexpr int kBatch = 16;       // iterations enqueued between two state reads
__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

__device__ __constant__ int kCornerOf[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // (ax + 2 ay + 4 az) -> Hex8 corner

// ---- K*u --------------------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256) void k_ord_apply_k(Dims g, const double *__restrict__ ke,
                                                     const double *__restrict__ ep, const double *__restrict__ x,
                                                     const uint8_t *__restrict__ flags, double *__restrict__ y,
                                                     const int *__restrict__ halt) {
#pragma clang fp contract(off)
    if (halt && *halt) return;
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= g.n_nodes) return;
    const int i = n % g.nx1, t = n / g.nx1, j = t % g.ny1, k = t / g.ny1;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 8; c++) {  // colour c = (ix & 1) | (iy & 1) << 1 | (iz & 1) << 2 (FeModel.cs:96-107)
        const int ax = (i ^ c) & 1, ay = (j ^ (c >> 1)) & 1, az = (k ^ (c >> 2)) & 1;
        const int ex = i - ax, ey = j - ay, ez = k - az;
        const double e = ep[(ex + 1) + g.px * ((ey + 1) + g.py * (ez + 1))];
        if (e == 0.0) continue;  // no active element of this colour at this node (inactive or outside)
        const int a = kCornerOf[ax + 2 * ay + 4 * az];
        const int n0 = ex + g.nx1 * (ey + g.ny1 * ez);
        double ue[24];
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int nd = n0 + corner_x(b) + g.nx1 * (corner_y(b) + g.ny1 * corner_z(b));
#pragma unroll
            for (int d = 0; d < 3; d++) {
                double xv = x[3 * nd + d];
                if (MODE == kApplyAGeneral && (flags[3 * nd + d] & kFlagConstrained)) xv = 0.0;  // v = P x
                ue[3 * b + d] = xv * e;
            }
        }
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double *row = ke + (3 * a + d) * 24;
            double l[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                double s = row[q] * ue[q] + row[4 + q] * ue[4 + q];
                s = s + row[8 + q] * ue[8 + q];
                s = s + row[12 + q] * ue[12 + q];
                s = s + row[16 + q] * ue[16 + q];
                s = s + row[20 + q] * ue[20 + q];
                l[q] = s;
            }
            acc[d] += (l[0] + l[1]) + (l[2] + l[3]);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int dof = 3 * n + d;
        double out = acc[d];
        if (MODE == kApplyAProjected && (flags[dof] & kFlagFixed)) out = x[dof];      // FemOperator.cs:259-273
        if (MODE == kApplyAGeneral && (flags[dof] & kFlagConstrained)) out = x[dof];  // :274-293
        y[dof] = out;
    }
}

// ---- PCG -------------------------------------------------------------------------------------
struct OrdState {
    double bnorm, tol, rho0, alpha, beta, rel, prev_rel, value, scratch;
    int iter, stag, max_iter, halt, message, pad;
};

// Lane chains of the blocked dot product: thread (block b, lane j) sums a[i] * b[i] over i = 8192 b + j,
// + 4, ... for the block's count / 4 quadruples, left to right from 0.
__global__ __launch_bounds__(256) void k_ord_lanes(const double *__restrict__ a, const double *__restrict__ b, int n,
                                                   double *__restrict__ lanes, const int *__restrict__ halt) {
#pragma clang fp contract(off)
    if (halt && *halt) return;
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int blk = t >> 2, j = t & 3;
    const int nb = cdiv(n, kPcgBlock);
    if (blk >= nb) return;
    const int s = blk * kPcgBlock, cnt = min(kPcgBlock, n - s), vc = cnt / 4;
    const double *pa = a + s + j, *pb = b + s + j;
    double acc = 0.0;
    int i = 0;
    for (; i + 8 <= vc; i += 8) {
        double va[8], vb[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            va[q] = pa[4 * (i + q)];
            vb[q] = pb[4 * (i + q)];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) acc = acc + va[q] * vb[q];
    }
    for (; i < vc; i++) acc = acc + pa[4 * i] * pb[4 * i];
    lanes[t] = acc;
}

// The block sums (lanes combined as (l0 + l1) + (l2 + l3), then the block's tail) added in block order
// (ReducePartialSums, PcgSolver.cs:414-422). The tail products are re-formed from the stored vectors, which
// hold exactly the values the reference's tail loop forms.
__device__ double ord_reduce(const double *__restrict__ lanes, const double *__restrict__ a,
                             const double *__restrict__ b, int n) {
#pragma clang fp contract(off)
    const int nb = cdiv(n, kPcgBlock);
    double total = 0.0;
    for (int blk = 0; blk < nb; blk++) {
        const int s = blk * kPcgBlock, cnt = min(kPcgBlock, n - s);
        double sum = (lanes[4 * blk] + lanes[4 * blk + 1]) + (lanes[4 * blk + 2] + lanes[4 * blk + 3]);
        for (int k = (cnt / 4) * 4; k < cnt; k++) sum += a[s + k] * b[s + k];
        total += sum;
    }
    return total;
}

__device__ __forceinline__ void ord_halt(OrdState *st, int message, double value) {
    st->message = message;
    st->value = value;
    st->halt = 1;
}

// Start of an iteration (PcgSolver.cs:235-271): iter++, pAp, the curvature and step tests, alpha.
__global__ void k_ord_alpha(OrdState *__restrict__ st, const double *__restrict__ lanes, const double *__restrict__ p,
                            const double *__restrict__ ap, int n) {
#pragma clang fp contract(off)
    if (st->halt) return;
    st->iter += 1;
    const double pap = ord_reduce(lanes, p, ap, n);
    if (!(pap > 1e-30) || !isfinite(pap)) {
        ord_halt(st, MYCEL_PCG_CURVATURE_BREAKDOWN, pap);
        return;
    }
    const double alpha = st->rho0 / pap;
    if (!isfinite(alpha) || alpha <= 0.0) {
        ord_halt(st, MYCEL_PCG_STEP_BREAKDOWN, alpha);
        return;
    }
    st->alpha = alpha;
}

__global__ __launch_bounds__(256) void k_ord_update_xr(int n, const OrdState *__restrict__ st, double *__restrict__ x,
                                                       double *__restrict__ r, const double *__restrict__ p,
                                                       const double *__restrict__ ap) {
#pragma clang fp contract(off)
    if (st->halt) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double alpha = st->alpha;
    x[i] = x[i] + alpha * p[i];
    r[i] = r[i] - alpha * ap[i];
}

// ||r||, convergence and stagnation (PcgSolver.cs:276-313).
__global__ void k_ord_check(OrdState *__restrict__ st, const double *__restrict__ lanes, const double *__restrict__ r,
                            int n) {
#pragma clang fp contract(off)
    if (st->halt) return;
    const double rr = ord_reduce(lanes, r, r, n);
    const double rel = sqrt(rr) / st->bnorm;
    st->rel = rel;
    if (rel <= st->tol) {
        ord_halt(st, MYCEL_PCG_CONVERGED, 0.0);
        return;
    }
    if (fabs(rel - st->prev_rel) < 1e-15 * rel) {
        st->stag += 1;
        if (st->stag > 50) {
            ord_halt(st, MYCEL_PCG_STAGNATION, 0.0);
            return;
        }
    } else {
        st->stag = 0;
    }
    st->prev_rel = rel;
}

// z = r * invDiag (JacobiPreconditioner.cs:29-35); st = null: no halt test (set-up).
__global__ __launch_bounds__(256) void k_ord_z(int n, const OrdState *__restrict__ st, const double *__restrict__ r,
                                               const double *__restrict__ inv, double *__restrict__ z) {
#pragma clang fp contract(off)
    if (st && st->halt) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) z[i] = r[i] * inv[i];
}

// rho1 = r.z, beta and its test, rho0 = rho1, and the iteration limit (PcgSolver.cs:315-369).
__global__ void k_ord_beta(OrdState *__restrict__ st, const double *__restrict__ lanes, const double *__restrict__ r,
                           const double *__restrict__ z, int n) {
#pragma clang fp contract(off)
    if (st->halt) return;
    const double rho1 = ord_reduce(lanes, r, z, n);
    const double beta = rho1 / st->rho0;
    if (!isfinite(beta) || beta < 0.0) {
        ord_halt(st, MYCEL_PCG_DIRECTION_BREAKDOWN, beta);
        return;
    }
    st->beta = beta;
    st->rho0 = rho1;
    if (st->iter >= st->max_iter) ord_halt(st, MYCEL_PCG_MAX_ITERATIONS, 0.0);
}

__global__ __launch_bounds__(256) void k_ord_update_p(int n, const OrdState *__restrict__ st,
                                                      const double *__restrict__ z, double *__restrict__ p) {
#pragma clang fp contract(off)
    if (st->halt) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = z[i] + st->beta * p[i];
}

// Set-up kernels: x[c] = 0 and the non-zero test (PcgSolver.cs:113-125); r = P b or P (b - A x).
__global__ __launch_bounds__(256) void k_ord_project_x(int n, const uint8_t *__restrict__ flags, double *__restrict__ x,
                                                       int *__restrict__ nonzero) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (flags[i] & kFlagConstrained) x[i] = 0.0;
    else if (x[i] != 0.0) atomicOr(nonzero, 1);
}

__global__ __launch_bounds__(256) void k_ord_init_r(int n, const uint8_t *__restrict__ flags, const double *__restrict__ b,
                                                    const double *__restrict__ ax, double *__restrict__ r) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    r[i] = (flags[i] & kFlagConstrained) ? 0.0 : (ax ? b[i] - ax[i] : b[i]);
}

__global__ __launch_bounds__(256) void k_ord_fill(int n, double *__restrict__ v, double value) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = value;
}

__global__ void k_ord_clear_int(int *__restrict__ v) { *v = 0; }

// The blocked dot of the lanes already formed, into st->scratch (set-up only).
__global__ void k_ord_reduce_to(OrdState *__restrict__ st, const double *__restrict__ lanes,
                                const double *__restrict__ a, const double *__restrict__ b, int n) {
    st->scratch = ord_reduce(lanes, a, b, n);
}

__global__ void k_ord_start(OrdState *__restrict__ st, const double *__restrict__ lanes, const double *__restrict__ r,
                            const double *__restrict__ z, int n, double bnorm, double tol, double rel, int max_iter) {
    st->rho0 = ord_reduce(lanes, r, z, n);
    st->bnorm = bnorm;
    st->tol = tol;
    st->rel = rel;
    st->prev_rel = rel;
    st->alpha = st->beta = st->value = 0.0;
    st->iter = 0;
    st->stag = 0;
    st->max_iter = max_iter;
    st->halt = 0;
    st->message = MYCEL_PCG_MAX_ITERATIONS;
}

// ---- sequential sums -------------------------------------------------------------------------
// out = sum over i ascending of term_i, term_i = a_i * b_i (b null: a_i), 0 where flags mark the DOF
// constrained; one thread adds, 4,096 terms per round staged in LDS by the workgroup; a 64-term group whose
// terms are all zero (+-0) is skipped, which leaves a sequential sum unchanged.
__global__ __launch_bounds__(1024) void k_ord_seq_sum(const double *__restrict__ a, const double *__restrict__ b,
                                                      const uint8_t *__restrict__ flags, int n,
                                                      double *__restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double buf[4096];
    double sum = 0.0;
    for (int base = 0; base < n; base += 4096) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = base + threadIdx.x + 1024 * q;
            double t = 0.0;
            if (i < n) {
                t = b ? a[i] * b[i] : a[i];
                if (flags && (flags[i] & kFlagConstrained)) t = 0.0;
            }
            buf[threadIdx.x + 1024 * q] = t;
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            for (int gq = 0; gq < 64; gq++) {
                const unsigned long long m = __ballot(buf[64 * gq + threadIdx.x] != 0.0);
                if (m != 0ull && threadIdx.x == 0) {
#pragma unroll 16
                    for (int l = 0; l < 64; l++) sum += buf[64 * gq + l];
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sum;
}

// ---- OC and SIMP metrics ---------------------------------------------------------------------
// The volume term of element i at lambda (OcOptimizer.cs:143-172), or with final != 0 the new density
// (:103-129) written to rho_new (which may alias rho).
__global__ __launch_bounds__(256) void k_ord_oc(int n, const double *rho, const double *__restrict__ dc,
                                                const double *__restrict__ s, const uint8_t *__restrict__ passive,
                                                double lambda, double move, double rho_min, int final,
                                                double *out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double si = s[i];
    if (passive && passive[i]) {
        out[i] = final ? 1.0 : si * 1.0;
        return;
    }
    const double r = rho[i];
    const double lo = fmax(rho_min, r - move), hi = fmin(1.0, r + move);
    const double neg = -dc[i];
    double v;
    if (neg <= 0.0 || si <= 0.0) {
        v = lo;
    } else {
        const double bb = neg / (lambda * si);
        const double tr = r * sqrt(bb);
        v = tr < lo ? lo : (tr > hi ? hi : tr);
    }
    out[i] = final ? v : si * v;
}

// Grey-level terms rb (1 - rb) (TopologyOptimizer.cs:330-334).
__global__ __launch_bounds__(256) void k_ord_grey_terms(int n, const double *__restrict__ rb, double *__restrict__ t) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) t[i] = rb[i] * (1.0 - rb[i]);
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------------
struct OrdWork {
    int n = 0;
    DevBuf<double> r, z, p, ap, lanes, seq;
    DevBuf<OrdState> st;
    DevBuf<int> flag;
    OrdState *h_st = nullptr;
    double *h_seq = nullptr;
    ~OrdWork() {
        if (h_st) (void)hipHostFree(h_st);
        if (h_seq) (void)hipHostFree(h_seq);
    }
};

void ord_free(OrdWork *w) { delete w; }

static OrdWork *ord_work(mycel_op *op) {
    if (!op->ord) {
        auto *w = new OrdWork();
        try {
            const int n = op->d.n_dofs;
            w->n = n;
            w->r.resize(n);
            w->z.resize(n);
            w->p.resize(n);
            w->ap.resize(n);
            w->lanes.resize(4 * cdiv(n, kPcgBlock));
            w->seq.resize(2);
            w->st.resize(1);
            w->flag.resize(1);
            MYCEL_HIP(hipHostMalloc(reinterpret_cast<void **>(&w->h_st), sizeof(OrdState), hipHostMallocDefault));
            MYCEL_HIP(hipHostMalloc(reinterpret_cast<void **>(&w->h_seq), 2 * sizeof(double), hipHostMallocDefault));
        } catch (...) {
            delete w;
            throw;
        }
        op->ord = w;
    }
    return op->ord;
}

void launch_apply_k_ordered(mycel_op *op, const double *d_x, double *d_y, int mode, hipStream_t s,
                            const int *d_halt) {
    const dim3 grid(cdiv(op->d.n_nodes, 256));
    if (mode == kApplyAProjected)
        hipLaunchKernelGGL(k_ord_apply_k<kApplyAProjected>, grid, dim3(256), 0, s, op->d, op->d_ke.get(),
                           op->d_ep.get(), d_x, op->d_flags.get(), d_y, d_halt);
    else if (mode == kApplyAGeneral)
        hipLaunchKernelGGL(k_ord_apply_k<kApplyAGeneral>, grid, dim3(256), 0, s, op->d, op->d_ke.get(),
                           op->d_ep.get(), d_x, op->d_flags.get(), d_y, d_halt);
    else
        hipLaunchKernelGGL(k_ord_apply_k<kApplyK>, grid, dim3(256), 0, s, op->d, op->d_ke.get(), op->d_ep.get(),
                           d_x, op->d_flags.get(), d_y, d_halt);
    MYCEL_LAUNCH_CHECK();
}

// Sequential sum (see k_ord_seq_sum) into device memory.
static void seq_sum_to(const double *d_a, const double *d_b, const uint8_t *d_flags, int n, double *d_out,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_ord_seq_sum, dim3(1), dim3(1024), 0, s, d_a, d_b, d_flags, n, d_out);
    MYCEL_LAUNCH_CHECK();
}

double seq_sum(mycel_op *op, const double *d_a, const double *d_b, const uint8_t *d_flags, int n, hipStream_t s) {
    OrdWork *w = ord_work(op);
    seq_sum_to(d_a, d_b, d_flags, n, w->seq.get(), s);
    MYCEL_HIP(hipMemcpyAsync(w->h_seq, w->seq.get(), sizeof(double), hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipStreamSynchronize(s));
    return w->h_seq[0];
}

int32_t pcg_solve_ordered(mycel_op *op, const double *d_b, double *d_x, double tol, int max_iter,
                          const volatile int32_t *cancel, mycel_pcg_progress_cb cb, void *user,
                          mycel_pcg_result *res) {
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    *res = mycel_pcg_result{};
    if (op->active_fixed < 6)  // PcgSolver.cs:95-100
        throw Error(MYCEL_E_INVALID_OPERATION,
                    "Model has fewer than 6 unique fixed DOFs on active nodes (found " +
                        std::to_string(op->active_fixed) +
                        "). At least 6 fixed DOFs are required to prevent rigid body modes in 3D.");
    hipStream_t s = op->stream;
    OrdWork *w = ord_work(op);
    const int n = op->d.n_dofs;
    const dim3 gv(cdiv(n, 256)), b256(256);
    const int lane_blocks = cdiv(4 * cdiv(n, kPcgBlock), 256);
    OrdState *st = w->st.get();
    const int *halt = &st->halt;
    const uint8_t *flags = op->d_flags.get();

    hipLaunchKernelGGL(k_ord_clear_int, dim3(1), dim3(1), 0, s, w->flag.get());
    hipLaunchKernelGGL(k_ord_project_x, gv, b256, 0, s, n, flags, d_x, w->flag.get());
    seq_sum_to(d_b, d_b, flags, n, w->seq.get(), s);  // ||P b||^2, sequential (:128-137)
    MYCEL_LAUNCH_CHECK();
    int nonzero = 0;
    MYCEL_HIP(hipMemcpyAsync(&nonzero, w->flag.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipMemcpyAsync(w->h_seq, w->seq.get(), sizeof(double), hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipStreamSynchronize(s));
    const double bnorm_sq = w->h_seq[0];
    const double bnorm = std::sqrt(bnorm_sq);
    if (bnorm <= 1e-30) {  // :139-151
        hipLaunchKernelGGL(k_ord_fill, gv, b256, 0, s, n, d_x, 0.0);
        MYCEL_LAUNCH_CHECK();
        MYCEL_HIP(hipStreamSynchronize(s));
        res->converged = 1;
        res->message = MYCEL_PCG_RHS_ZERO;
        res->elapsed_ms = elapsed();
        return MYCEL_OK;
    }
    double rr;
    if (nonzero) {  // r = P (b - A x) (:175-184)
        launch_apply_k_ordered(op, d_x, w->ap.get(), kApplyAProjected, s, nullptr);
        hipLaunchKernelGGL(k_ord_init_r, gv, b256, 0, s, n, flags, d_b, w->ap.get(), w->r.get());
        hipLaunchKernelGGL(k_ord_lanes, dim3(lane_blocks), b256, 0, s, w->r.get(), w->r.get(), n, w->lanes.get(),
                           nullptr);
        hipLaunchKernelGGL(k_ord_reduce_to, dim3(1), dim3(1), 0, s, st, w->lanes.get(), w->r.get(), w->r.get(), n);
        MYCEL_LAUNCH_CHECK();
        MYCEL_HIP(hipMemcpyAsync(w->h_st, st, sizeof(OrdState), hipMemcpyDeviceToHost, s));
        MYCEL_HIP(hipStreamSynchronize(s));
        rr = w->h_st->scratch;
    } else {  // r = P b, ||r||^2 = ||P b||^2 (:185-192)
        hipLaunchKernelGGL(k_ord_init_r, gv, b256, 0, s, n, flags, d_b, static_cast<const double *>(nullptr),
                           w->r.get());
        MYCEL_LAUNCH_CHECK();
        rr = bnorm_sq;
    }
    const double rel0 = std::sqrt(rr) / bnorm;
    if (rel0 <= tol) {  // :197-209
        MYCEL_HIP(hipStreamSynchronize(s));
        res->final_residual = rel0;
        res->converged = 1;
        res->message = MYCEL_PCG_INITIAL_CONVERGED;
        res->elapsed_ms = elapsed();
        return MYCEL_OK;
    }
    // z = M^-1 r, rho0 = r.z, p = z (:211-227).
    hipLaunchKernelGGL(k_ord_z, gv, b256, 0, s, n, static_cast<const OrdState *>(nullptr), w->r.get(),
                       op->d_inv_diag.get(), w->z.get());
    hipLaunchKernelGGL(k_ord_lanes, dim3(lane_blocks), b256, 0, s, w->r.get(), w->z.get(), n, w->lanes.get(),
                       nullptr);
    hipLaunchKernelGGL(k_ord_start, dim3(1), dim3(1), 0, s, st, w->lanes.get(), w->r.get(), w->z.get(), n, bnorm, tol,
                       rel0, max_iter);
    MYCEL_HIP(hipMemcpyAsync(w->p.get(), w->z.get(), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    MYCEL_LAUNCH_CHECK();
    if (max_iter <= 0) {  // the loop body never runs (:235,359-368)
        MYCEL_HIP(hipStreamSynchronize(s));
        res->final_residual = rel0;
        res->message = MYCEL_PCG_MAX_ITERATIONS;
        res->elapsed_ms = elapsed();
        return MYCEL_OK;
    }
    const long long max_batches = static_cast<long long>(max_iter) / kBatch + 2;
    for (long long batch = 0;; batch++) {
        if (batch > max_batches) throw Error(MYCEL_E_DEVICE, "ordered PCG batch loop did not terminate");
        if (cancel && *cancel) {
            MYCEL_HIP(hipStreamSynchronize(s));
            throw Error(MYCEL_E_CANCELED, "The operation was canceled.");
        }
        for (int it = 0; it < kBatch; it++) {
            launch_apply_k_ordered(op, w->p.get(), w->ap.get(), kApplyAProjected, s, halt);
            hipLaunchKernelGGL(k_ord_lanes, dim3(lane_blocks), b256, 0, s, w->p.get(), w->ap.get(), n,
                               w->lanes.get(), halt);
            hipLaunchKernelGGL(k_ord_alpha, dim3(1), dim3(1), 0, s, st, w->lanes.get(), w->p.get(), w->ap.get(), n);
            hipLaunchKernelGGL(k_ord_update_xr, gv, b256, 0, s, n, st, d_x, w->r.get(), w->p.get(), w->ap.get());
            hipLaunchKernelGGL(k_ord_lanes, dim3(lane_blocks), b256, 0, s, w->r.get(), w->r.get(), n,
                               w->lanes.get(), halt);
            hipLaunchKernelGGL(k_ord_check, dim3(1), dim3(1), 0, s, st, w->lanes.get(), w->r.get(), n);
            hipLaunchKernelGGL(k_ord_z, gv, b256, 0, s, n, st, w->r.get(), op->d_inv_diag.get(), w->z.get());
            hipLaunchKernelGGL(k_ord_lanes, dim3(lane_blocks), b256, 0, s, w->r.get(), w->z.get(), n,
                               w->lanes.get(), halt);
            hipLaunchKernelGGL(k_ord_beta, dim3(1), dim3(1), 0, s, st, w->lanes.get(), w->r.get(), w->z.get(), n);
            hipLaunchKernelGGL(k_ord_update_p, gv, b256, 0, s, n, st, w->z.get(), w->p.get());
        }
        MYCEL_LAUNCH_CHECK();
        MYCEL_HIP(hipMemcpyAsync(w->h_st, st, sizeof(OrdState), hipMemcpyDeviceToHost, s));
        MYCEL_HIP(hipStreamSynchronize(s));
        const OrdState &h = *w->h_st;
        if (cb) cb(h.iter, h.rel, user);
        if (h.halt) break;
    }
    const OrdState &h = *w->h_st;
    res->iterations = h.iter;
    res->final_residual = h.rel;
    res->message = h.message;
    res->message_value = h.value;
    res->converged = h.message == MYCEL_PCG_CONVERGED;
    res->elapsed_ms = elapsed();
    return MYCEL_OK;
}

// OcOptimizer.Update with sequential volume sums (OcOptimizer.cs:12-132): the bracket and bisection run on
// the host exactly as the reference writes them; each volume is one element-wise pass and one sequential sum.
double oc_update_ordered(mycel_op *op, const double *d_rho, const double *d_dc, const double *d_s,
                         const uint8_t *d_passive, double *d_rho_new, int n, double vf, double move, double rho_min,
                         double lambda_guess, int max_bisect, double vol_tol, hipStream_t s, int *evals) {
    if (n <= 0) {
        if (evals) *evals = 0;
        return lambda_guess;
    }
    OrdWork *w = ord_work(op);
    DevBuf<double> terms(n);
    const dim3 g(cdiv(n, 256)), b256(256);
    int ev = 0;
    auto volume = [&](double lambda) {
        hipLaunchKernelGGL(k_ord_oc, g, b256, 0, s, n, d_rho, d_dc, d_s, d_passive, lambda, move, rho_min, 0,
                           terms.get());
        MYCEL_LAUNCH_CHECK();
        ev++;
        return seq_sum(op, terms.get(), nullptr, nullptr, n, s);
    };
    (void)w;
    const double total_s = seq_sum(op, d_s, nullptr, nullptr, n, s);
    const double target = vf * total_s;
    const bool good = lambda_guess > 0.0 && !std::isinf(lambda_guess) && !std::isnan(lambda_guess);
    double l1 = good ? lambda_guess * 0.5 : 1e-12;
    double l2 = good ? lambda_guess * 2.0 : 1e12;
    if (l1 <= 1e-50) l1 = 1e-12;
    if (l2 <= l1) l2 = l1 * 10.0;
    double v1 = volume(l1);
    double v2 = volume(l2);
    int steps = 0;
    while (v1 < target && l1 > 1e-50 && steps++ < 50) {
        l2 = l1;
        v2 = v1;
        l1 *= 0.1;
        v1 = volume(l1);
    }
    steps = 0;
    while (v2 > target && l2 < 1e50 && steps++ < 50) {
        l1 = l2;
        v1 = v2;
        l2 *= 10.0;
        v2 = volume(l2);
    }
    double lmid = std::sqrt(l1 * l2);
    for (int step = 0; step < max_bisect; step++) {
        lmid = std::sqrt(l1 * l2);
        const double vmid = volume(lmid);
        const double rel = std::fabs(vmid - target) / std::fmax(target, 1e-12);
        if (rel <= vol_tol) break;
        if (vmid > target) l1 = lmid;
        else l2 = lmid;
        if (std::fabs(l2 - l1) <= 1e-15 * lmid) break;
    }
    hipLaunchKernelGGL(k_ord_oc, g, b256, 0, s, n, d_rho, d_dc, d_s, d_passive, lmid, move, rho_min, 1, d_rho_new);
    MYCEL_LAUNCH_CHECK();
    MYCEL_HIP(hipStreamSynchronize(s));
    if (evals) *evals = ev;
    return lmid;
}

// Sequential volume and grey sums of the SIMP snapshot (TopologyOptimizer.cs:323-338); out[1] = sum rho_bar,
// out[2] = sum g (1 - g) with g = grey_src.
void simp_sums_ordered(mycel_op *op, int na, const double *d_rho_bar, const double *d_grey_src, double *out3,
                       hipStream_t s) {
    out3[1] = seq_sum(op, d_rho_bar, nullptr, nullptr, na, s);
    DevBuf<double> t(std::max(1, na));
    hipLaunchKernelGGL(k_ord_grey_terms, dim3(cdiv(std::max(1, na), 256)), dim3(256), 0, s, na, d_grey_src, t.get());
    MYCEL_LAUNCH_CHECK();
    out3[2] = seq_sum(op, t.get(), nullptr, nullptr, na, s);
}

}  // namespace mycel
