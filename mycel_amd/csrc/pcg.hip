// Device-resident Jacobi-preconditioned CG on the projected operator A = PKP + (I - P)
// (Fem/PcgSolver.cs:80-553, Fem/JacobiPreconditioner.cs:15-35), in two precisions.
//
// One iteration is three kernels with no host involvement:
//   1. ap = A p fused with per-block partials of p.ap                       (node-centric K*u)
//   2. alpha = rz / p.ap (breakdown checks); x += alpha p; r -= alpha ap; z = r / diag;
//      per-block partials of r.r and r.z
//   3. relres, convergence, stagnation, beta = rz' / rz (breakdown check); p = z + beta p
// Every block of kernels 2 and 3 re-sums the previous kernel's partials in the same fixed order, so
// all blocks agree bit-for-bit on alpha, beta and the stopping decision without a grid barrier.
// The scalar state lives on the device in two slots indexed by iteration parity: an iteration reads
// slot s and writes slot s^1, so a batch of an even number of iterations is a fixed sequence of launches
// that is captured once into a hipGraph and replayed. Once the solve halts every kernel returns at
// entry, so the host polls the state once per batch, not once per iteration.
//
// FP64 (MYCEL_PRECISION_FP64): the reference's arithmetic; it stops on the recursive residual as
// PcgSolver does.
// Mixed (MYCEL_PRECISION_MIXED_FP32): the iteration runs on FP32 vectors and the FP32 operator, with
// FP64 dot products and scalars. The solution is accumulated in FP64 by reliable updates (residual
// replacement, van der Vorst & Ye 2000): whenever the recursive residual has fallen by a factor delta
// since the last update, or claims convergence, the iteration pauses; the FP32 increment is folded
// into the FP64 solution, the true residual r = P(b - A x) is formed with the FP64 operator, and the
// iteration resumes from it with the same search direction. Convergence is declared on the true FP64
// residual only.
// Multigrid (MYCEL_PRECISION_MIXED_GMG): the FP64 iteration above with the Jacobi step z = D^-1 r replaced
// by one FP32 geometric-multigrid V-cycle (gmg.hip). Because the FP32 V-cycle is not exactly a fixed
// linear map, beta is the flexible (Polak-Ribiere) form z'.(r' - r) / (r.z) = -alpha z'.Ap / (r.z), clamped
// at 0 (a restart) if rounding makes it negative; the operator, residual, dot products, convergence test
// and solution update are the FP64 ones of the reference path. One iteration: K*u (+ p.Ap partials),
// alpha and the x, r updates (+ r.r partials), the V-cycle (+ z.r and z.Ap partials), beta and p.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "device_common.hpp"
#include "ku_element.hpp"
#include "mycel_internal.hpp"

namespace mycel {

struct PcgState {
    double alpha[2];     // step length of the iteration that reads this slot (flexible beta)
    double rz[2];        // rho0 at the start of the iteration that reads this slot
    double prev_rel[2];  // relres of the previous iteration (stagnation test, breakdown report)
    int stag[2];
    int iter[2];         // iteration number t that reads this slot
    double tol, bnorm, delta, maxrel, beta;
    double bnorm_sq, r_norm_sq, rz0;
    int max_iter, mixed;
    int halt, done, pause, pause_slot;
    int message, iterations, nonzero_x, outer;
    double final_rel, last_rel;
    int last_iter, pad;
    double best_true;    // mixed: smallest true (FP64) relres seen at a reliable update or at the start
    int no_gain, pad2;   // mixed: consecutive reliable updates that did not improve on best_true
    double value;        // the quantity a breakdown message quotes (pAp, alpha or beta)
};

struct PcgWork {
    int n = 0;
    int nb_mv = 0, nb_vec = 0;
    int nb_rr = 0;                           // blocks (= r.r partials) of the flexible x/r update: kXrNodes nodes each
    DevBuf<double> r, z, p, ap;              // FP64 iteration vectors (r / ap double as r64 / A x64 when mixed)
    DevBuf<double> p2;                       // FP64 iterations with the fused step 4: p of the odd ones
    DevBuf<float> r32, z32, p32, ap32, x32;  // FP32 iteration vectors of the mixed path
    DevBuf<float> p32b;                      // its p of the odd iterations (fused step 4)
    DevBuf<double> part_pap, part2, part1;
    DevBuf<PcgState> state;
    PcgState *h_state = nullptr;             // pinned
    hipGraphExec_t graph64 = nullptr, graph32 = nullptr;
    int bf16 = 0;                            // the FP32 iteration rounds its search directions to bfloat16
    const double *graph64_x = nullptr;       // solution vector the FP64 graph was captured with
    DevBuf<double> part_zr, part_zap;        // V-cycle partials of z.r and z.Ap (multigrid)
    int nb_z = 0;
    hipGraphExec_t graph_gmg = nullptr;
    const double *graph_gmg_x = nullptr;
    const Gmg *graph_gmg_h = nullptr;        // hierarchy the multigrid graph was captured with
    uint64_t graph_gmg_gen = 0;
    // Several right-hand sides in lockstep (multigrid only): R slices of every vector and partial array,
    // one state per RHS.
    int b_cap = 0;
    DevBuf<double> br, bz, bp, bap, bpart_pap, bpart_rr, bpart1, bpart_zr, bpart_zap;
    DevBuf<PcgState> bstate;
    PcgState *h_bstate = nullptr;            // pinned, b_cap entries
    hipGraphExec_t graph_b = nullptr;
    const double *graph_b_x = nullptr;
    const Gmg *graph_b_h = nullptr;
    uint64_t graph_b_gen = 0;
    int graph_b_R = 0;
    // Iterations of the last converged solve in each FP64 mode (the Jacobi one, the multigrid one): the next
    // solve in that mode launches batches covering 3/4 of them before its first state read.
    int last_iters_64 = 0, last_iters_gmg = 0;
};

namespace {

constexpr int kVecPerBlock = 1024;  // DOFs per block in the vector kernels (4 per thread)
constexpr int kGraphIters = 32;     // iterations per captured batch (even: the slot parity repeats)
constexpr int kGmgGraphIters = 4;   // multigrid iterations are ~30 launches each and few in number
constexpr double kReliableDelta = 0.1;
// A reliable update is triggered once the FP32 recursive residual has dropped by kReliableDelta. When the FP32
// recurrences drift from the true residual (extreme stiffness contrast), they keep "converging" while the true
// residual does not: kReliableStall consecutive updates whose true relres is not below kReliableGain times the
// best one so far end the FP32 phase as stagnation.
constexpr double kReliableGain = 0.5;
constexpr int kReliableStall = 3;

__device__ __forceinline__ double fmad(double a, double b, double c) { return __fma_rn(a, b, c); }
__device__ __forceinline__ float fmad(float a, float b, float c) { return __fmaf_rn(a, b, c); }

__device__ __forceinline__ void halt_with(PcgState *st, int message, int iterations, double rel, double value = 0.0) {
    st->value = value;
    st->message = message;
    st->iterations = iterations;
    st->final_rel = rel;
    st->done = 1;
    st->halt = 1;
}

template <typename T>
__global__ __launch_bounds__(256) void k_pcg_update_xr(int n, int slot, PcgState *__restrict__ st,
                                                       const double *__restrict__ part_pap, int nb_mv,
                                                       T *__restrict__ x, T *__restrict__ r, const T *__restrict__ p,
                                                       const T *__restrict__ ap, T *__restrict__ z,
                                                       const T *__restrict__ inv, double *__restrict__ part2) {
    __shared__ double red[8];
    // Issue this block's vector loads first: they do not depend on alpha, so their latency overlaps the
    // halt check and the partial-sum reduction below.
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
    T xv[4], pv[4], rv[4], av[4], iv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = min(base + 256 * q, n - 1);
        xv[q] = x[i];
        pv[q] = p[i];
        rv[q] = r[i];
        av[q] = ap[i];
        iv[q] = inv[i];
    }
    // The halt flag and rho0 are read ahead of the partial sum, so their loads overlap its loads.
    const int halt = st->halt;
    const double rz0 = st->rz[slot];
    const double pap = wave_sum_partials(part_pap, nb_mv);
    if (halt) return;
    const bool curv = !(pap > 1e-30) || !isfinite(pap);  // PcgSolver.cs:245
    const double alpha = rz0 / pap;
    const bool step = !curv && (!isfinite(alpha) || alpha <= 0.0);  // :260
    if (curv || step) {
        if (blockIdx.x == 0 && threadIdx.x == 0)
            halt_with(st, curv ? MYCEL_PCG_CURVATURE_BREAKDOWN : MYCEL_PCG_STEP_BREAKDOWN, st->iter[slot],
                      st->prev_rel[slot], curv ? pap : alpha);
        return;
    }
    const T a = static_cast<T>(alpha);
    double rr = 0.0, rz = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) {
            x[i] = fmad(a, pv[q], xv[q]);
            const T ri = fmad(-a, av[q], rv[q]);
            r[i] = ri;
            const T zi = ri * iv[q];
            z[i] = zi;
            rr = __fma_rn(static_cast<double>(ri), static_cast<double>(ri), rr);
            rz = __fma_rn(static_cast<double>(ri), static_cast<double>(zi), rz);
        }
    }
    block_sum256x2(rr, rz, red);
    if (threadIdx.x == 0) {
        part2[2 * blockIdx.x] = rr;
        part2[2 * blockIdx.x + 1] = rz;
    }
}

// The state a step 4 reads, loaded at kernel entry in one go: its loads are then in flight together with the
// vector and partial loads instead of one dependent round trip each after the partial sums. Within a launch
// only the lead thread writes the state, and only fields of the other slot, the halt fields and maxrel (whose
// new value max(maxrel, rel) gives every reader the same max(maxrel, rel)).
struct Step4In {
    double bnorm, tol, delta, maxrel, prev_rel, rz, alpha;
    int iter, stag, mixed, max_iter, halt;
};
__device__ __forceinline__ Step4In step4_in(const PcgState *__restrict__ st, int slot) {
    Step4In in;
    in.bnorm = st->bnorm;
    in.tol = st->tol;
    in.delta = st->delta;
    in.maxrel = st->maxrel;
    in.prev_rel = st->prev_rel[slot];
    in.rz = st->rz[slot];
    in.alpha = st->alpha[slot];
    in.iter = st->iter[slot];
    in.stag = st->stag[slot];
    in.mixed = st->mixed;
    in.max_iter = st->max_iter;
    in.halt = st->halt;
    return in;
}

// Jacobi PCG, step 4 decisions from the per-wave sums of the interleaved (r.r, r.z) partials: relres,
// convergence (mixed: the reliable-update pause) and stagnation as PcgSolver.cs:279-313, then beta = rz' / rz
// (:334). Returns false when the solve halts or pauses here (the lead thread records why); otherwise the lead
// thread advances the other slot.
__device__ __forceinline__ bool pcg_step4(PcgState *__restrict__ st, const Step4In &in, int slot,
                                          const double *__restrict__ part2, int nb_vec, bool lead, double &beta) {
    double rr, rzn;
    wave_sum_partials2(part2, nb_vec, rr, rzn);
    if (in.halt) return false;
    const double rel = sqrt(rr) / in.bnorm;
    const int t = in.iter;
    if (in.mixed) {
        // Reliable update: the true residual decides convergence; refresh it once the recursive residual
        // has fallen by delta below its running maximum since the last refresh.
        const double mx = fmax(in.maxrel, rel);
        if (rel <= in.tol || rel < in.delta * mx) {
            if (lead) {
                st->pause = 1;
                st->pause_slot = slot;
                st->iterations = t;
                st->last_rel = rel;
                st->halt = 1;
            }
            return false;
        }
    } else if (rel <= in.tol) {  // PcgSolver.cs:279
        if (lead) halt_with(st, MYCEL_PCG_CONVERGED, t, rel);
        return false;
    }
    const int stag = (fabs(rel - in.prev_rel) < 1e-15 * rel) ? in.stag + 1 : 0;  // :293-313
    if (stag > 50) {
        if (lead) halt_with(st, MYCEL_PCG_STAGNATION, t, rel);
        return false;
    }
    beta = rzn / in.rz;
    if (!isfinite(beta) || beta < 0.0) {  // :334
        if (lead) halt_with(st, MYCEL_PCG_DIRECTION_BREAKDOWN, t, rel, beta);
        return false;
    }
    if (lead) {
        const int o = slot ^ 1;
        st->rz[o] = rzn;
        st->prev_rel[o] = rel;
        st->stag[o] = stag;
        st->iter[o] = t + 1;
        st->last_rel = rel;
        st->last_iter = t;
        if (in.mixed) st->maxrel = fmax(in.maxrel, rel);
        if (t >= in.max_iter) halt_with(st, MYCEL_PCG_MAX_ITERATIONS, t, rel);  // :359-368
    }
    return true;
}

// v rounded to the nearest bfloat16 (ties to even), as a float: the BF16 mode's search directions hold only
// such values, so every K p of that mode contracts BF16 inputs with FP32 accumulation. Finite v only.
__device__ __forceinline__ float round_bf16(float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xFFFF0000u);
}
template <typename T>
__device__ __forceinline__ T maybe_bf16(T v, int bf16) {
    if constexpr (std::is_same<T, float>::value) return bf16 ? round_bf16(v) : v;
    return v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_pcg_update_p(int n, int slot, PcgState *__restrict__ st,
                                                      const double *__restrict__ part2, int nb_vec,
                                                      T *__restrict__ p, const T *__restrict__ z, int bf16) {
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
    T pv[4], zv[4];  // prefetched before the dependent reductions
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = min(base + 256 * q, n - 1);
        pv[q] = p[i];
        zv[q] = z[i];
    }
    const Step4In in = step4_in(st, slot);
    double beta;
    if (!pcg_step4(st, in, slot, part2, nb_vec, blockIdx.x == 0 && threadIdx.x == 0, beta)) return;
    const T b = static_cast<T>(beta);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) p[i] = maybe_bf16(fmad(b, pv[q], zv[q]), bf16);
    }
}

// Flexible PCG, step 2, fused with the V-cycle's first step: alpha = rz / p.Ap (the breakdown tests of
// PcgSolver.cs:245-270; alpha is kept in the state for the flexible beta); x += alpha p; r -= alpha Ap;
// per-block partials of r.r; then, on the updated residual, b0 = float(r) (0 on constrained DOFs) and
// x0 = d0 = B^-1 b0 / theta, the finest Chebyshev smoother's first step (B^-1: the nodal block inverses).
// Block bx of one right-hand side: kXrNodes nodes, four lanes per node (lane c < 3 owns the node's component c,
// lane 3 idle) and four nodes per lane group (nodes g, g + 64, g + 128, g + 192 of the block), so the vector loads
// and stores stay nearly contiguous while the smoother's first step, which needs a node's three components, meets
// them by lane shuffles.
constexpr int kXrNodes = 256;
// Its r.r partials (one per block, nb_rr = ceil(n / 3 / kXrNodes)) go to the Jacobi partial buffer (2 per
// kVecPerBlock DOFs) in single and packed solves, so they must not outnumber that buffer's slots.
static_assert(2 * 3 * kXrNodes >= kVecPerBlock, "flexible x/r update: partials exceed the Jacobi partial buffer");
__device__ __forceinline__ void fpcg_update_xr_body(int n, int slot, PcgState *__restrict__ st,
                                                    const double *__restrict__ part_pap, int nb_mv,
                                                    double *__restrict__ x, double *__restrict__ r,
                                                    const double *__restrict__ p, const double *__restrict__ ap,
                                                    double *__restrict__ part_rr, const GmgFineView &fv, int bx) {
    __shared__ double red[4];
    constexpr int Q = kXrNodes / 64;
    const int nn = n / 3, q = threadIdx.x & 3, c = q < 3 ? q : 2;
    const int g = threadIdx.x >> 2, base = bx * kXrNodes;
    // Row c of the node's block inverse (xx, yy, zz, xy, xz, yz): (xx, xy, xz), (xy, yy, yz) or (xz, yz, zz).
    const int k0 = c == 0 ? 0 : (c == 1 ? 3 : 4), k1 = c == 0 ? 3 : (c == 1 ? 1 : 5), k2 = c == 0 ? 4 : (c == 1 ? 5 : 2);
    double xv[Q], pv[Q], rv[Q], av[Q];
    float iv[Q], b0r[Q], b1r[Q], b2r[Q];
#pragma unroll
    for (int u = 0; u < Q; u++) {
        const int node = min(base + g + 64 * u, nn - 1), i = 3 * node + c;
        xv[u] = x[i];
        pv[u] = p[i];
        rv[u] = r[i];
        av[u] = ap[i];
        iv[u] = fv.inv[i];
        b0r[u] = fv.binv[6 * node + k0];
        b1r[u] = fv.binv[6 * node + k1];
        b2r[u] = fv.binv[6 * node + k2];
    }
    const int halt = st->halt;  // read ahead of the partial sum, as in k_pcg_update_xr
    const double rz0 = st->rz[slot];
    const double pap = wave_sum_partials(part_pap, nb_mv);
    if (halt) return;
    const bool curv = !(pap > 1e-30) || !isfinite(pap);
    const double alpha = rz0 / pap;
    const bool step = !curv && (!isfinite(alpha) || alpha <= 0.0);
    if (curv || step) {
        if (bx == 0 && threadIdx.x == 0)
            halt_with(st, curv ? MYCEL_PCG_CURVATURE_BREAKDOWN : MYCEL_PCG_STEP_BREAKDOWN, st->iter[slot],
                      st->prev_rel[slot], curv ? pap : alpha);
        return;
    }
    if (bx == 0 && threadIdx.x == 0) st->alpha[slot] = alpha;
    const int lead = (threadIdx.x & 63) & ~3;
    double rr = 0.0;
#pragma unroll
    for (int u = 0; u < Q; u++) {
        const int node = base + g + 64 * u, i = 3 * node + c;
        const bool own = q < 3 && node < nn;
        const double ri = fmad(-alpha, av[u], rv[u]);
        const float bq = iv[u] > 0.0f ? static_cast<float>(ri) : 0.0f;
        // b0 of the node's three components (every lane takes part in the shuffles), then z_c = row c . b0
        const float s0 = __shfl(bq, lead, 64), s1 = __shfl(bq, lead + 1, 64), s2 = __shfl(bq, lead + 2, 64);
        const float z = (b0r[u] * s0 + b1r[u] * s1) + b2r[u] * s2;
        if (own) {
            x[i] = fmad(alpha, pv[u], xv[u]);
            r[i] = ri;
            rr = __fma_rn(ri, ri, rr);
            const float x0 = z * fv.theta_inv;
            fv.b[i] = bq;
            fv.x[i] = x0;
            fv.d[i] = x0;
        }
    }
    rr = block_sum256(rr, red);
    if (threadIdx.x == 0) part_rr[bx] = rr;
}

__global__ __launch_bounds__(256) void k_fpcg_update_xr_gmg(int n, int slot, PcgState *__restrict__ st,
                                                            const double *__restrict__ part_pap, int nb_mv,
                                                            double *__restrict__ x, double *__restrict__ r,
                                                            const double *__restrict__ p,
                                                            const double *__restrict__ ap,
                                                            double *__restrict__ part_rr, GmgFineView fv) {
    // grid dimension y: the right-hand side (vectors at y * n, partials at y * blocks, one state each)
    const long long o = static_cast<long long>(blockIdx.y) * n;
    fv.b += o;
    fv.x += o;
    fv.d += o;
    fpcg_update_xr_body(n, slot, st + blockIdx.y, part_pap + blockIdx.y * nb_mv, nb_mv, x + o, r + o, p + o, ap + o,
                        part_rr + blockIdx.y * gridDim.x, fv, blockIdx.x);
}

// Flexible PCG, step 4 decisions from the per-wave sums of the r.r, z.r and z.Ap partials: relres, convergence
// and stagnation as PcgSolver.cs:279-313, then beta = -alpha z.Ap / rz (clamped at 0). Returns false when the
// solve halts here (the lead thread records why); otherwise the lead thread advances the other slot.
__device__ __forceinline__ bool fpcg_step4(PcgState *__restrict__ st, const Step4In &in, int slot,
                                           const double *__restrict__ part_rr, int nb_vec,
                                           const double *__restrict__ part_zr, const double *__restrict__ part_zap,
                                           int nb_z, bool lead, double &beta) {
    const double *const parts[3] = {part_rr, part_zr, part_zap};
    const int counts[3] = {nb_vec, nb_z, nb_z};
    double sums[3];
    wave_sum_partials_n<3>(parts, counts, sums);
    if (in.halt) return false;
    const double rr = sums[0], zr = sums[1], zap = sums[2];
    const double rel = sqrt(rr) / in.bnorm;
    const int t = in.iter;
    if (rel <= in.tol) {
        if (lead) halt_with(st, MYCEL_PCG_CONVERGED, t, rel);
        return false;
    }
    const int stag = (fabs(rel - in.prev_rel) < 1e-15 * rel) ? in.stag + 1 : 0;
    if (stag > 50) {
        if (lead) halt_with(st, MYCEL_PCG_STAGNATION, t, rel);
        return false;
    }
    beta = -in.alpha * zap / in.rz;
    if (!isfinite(beta)) {
        if (lead) halt_with(st, MYCEL_PCG_DIRECTION_BREAKDOWN, t, rel, beta);
        return false;
    }
    beta = fmax(beta, 0.0);
    if (lead) {
        const int o = slot ^ 1;
        st->rz[o] = zr;
        st->prev_rel[o] = rel;
        st->stag[o] = stag;
        st->iter[o] = t + 1;
        st->last_rel = rel;
        st->last_iter = t;
        if (t >= in.max_iter) halt_with(st, MYCEL_PCG_MAX_ITERATIONS, t, rel);
    }
    return true;
}

// Flexible PCG, step 4: p = z + beta p.
__global__ __launch_bounds__(256) void k_fpcg_update_p(int n, int slot, PcgState *__restrict__ st,
                                                       const double *__restrict__ part_rr, int nb_vec,
                                                       const double *__restrict__ part_zr,
                                                       const double *__restrict__ part_zap, int nb_z,
                                                       double *__restrict__ p, const double *__restrict__ z) {
    {  // grid dimension y: the right-hand side
        const long long o = static_cast<long long>(blockIdx.y) * n;
        st += blockIdx.y;
        part_rr += blockIdx.y * nb_vec;
        part_zr += blockIdx.y * nb_z;
        part_zap += blockIdx.y * nb_z;
        p += o;
        z += o;
    }
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
    double pv[4], zv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = min(base + 256 * q, n - 1);
        pv[q] = p[i];
        zv[q] = z[i];
    }
    const Step4In in = step4_in(st, slot);
    double beta;
    if (!fpcg_step4(st, in, slot, part_rr, nb_vec, part_zr, part_zap, nb_z, blockIdx.x == 0 && threadIdx.x == 0,
                    beta))
        return;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) p[i] = fmad(beta, pv[q], zv[q]);
    }
}

// Step 4 (FLEX: the flexible form, FP64; else the Jacobi one, FP64 or the FP32 iteration of the reliable-update
// solver) fused with the next iteration's A p: p' = z + beta p is formed where the element-centric tile march
// reads x, and each tile writes A p', p' and the p'.A p' partials of the nodes it owns (p' goes to a second
// buffer: other tiles still read p at their halo nodes). One launch and one pass over p fewer than the p update
// followed by the K*u launch, with bit-identical values.
template <typename T, bool FLEX>
struct PcgNextP {
    const T *p, *z;
    PcgState *st;
    const double *part_rr, *part_zr, *part_zap;  // Jacobi: part_rr holds the interleaved (r.r, r.z) partials
    int nb_vec, nb_z, slot;
    double beta;
    T bt;        // beta in the vector type, as the separate p update rounds it
    Step4In in;  // loaded at kernel entry (step4_in)
    int bf16;    // FP32 only: p' rounded to bfloat16 (the BF16 mode)
    struct Raw {
        T p, z;
    };
    __device__ Raw load(size_t at, uint32_t o, int d) const {
        constexpr int sT = static_cast<int>(sizeof(T));
        return {ld_off(p + at, o, sT * d), ld_off(z + at, o, sT * d)};
    }
    __device__ T value(Raw r) const { return maybe_bf16(fmad(bt, r.p, r.z), bf16); }
    // Called once per wave after its first loads are issued, so the partials' round trip overlaps them; false
    // (the solve halted before or halts here) ends the wave's march.
    template <typename Fin>
    __device__ bool ready(Fin &fin) {
        const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
        fin.go = FLEX ? fpcg_step4(st, in, slot, part_rr, nb_vec, part_zr, part_zap, nb_z, lead, beta)
                      : pcg_step4(st, in, slot, part_rr, nb_vec, lead, beta);
        bt = static_cast<T>(beta);
        return fin.go;
    }
};

// Epilogue: the projected operator's (fixed DOFs: out = p'), stores of A p' and p' unless the solve halted.
template <typename T>
struct PcgNextFin {
    const uint8_t *flags;
    T *ap, *p_next;
    double dot = 0.0;
    bool go = false;
    struct Pre {
        uint8_t fl[3];
    };
    __device__ Pre pre(int n) const {
        Pre q;
#pragma unroll
        for (int d = 0; d < 3; d++) q.fl[d] = flags[3 * n + d];
        return q;
    }
    __device__ void operator()(int n, T (&out)[3], const T (&xc)[3], const Pre &q) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (q.fl[d] & kFlagFixed) out[d] = xc[d];
            if (go) {
                ap[3 * n + d] = out[d];
                p_next[3 * n + d] = xc[d];
            }
            dot += static_cast<double>(xc[d]) * static_cast<double>(out[d]);
        }
    }
};

template <typename T, bool FLEX, int KZ>
__global__ __launch_bounds__(256) void k_pcg_step4_ap(Dims g, const KeMono<T> km, const T *__restrict__ ep,
                                                      const uint8_t *__restrict__ flags, PcgNextP<T, FLEX> src,
                                                      T *__restrict__ ap, T *__restrict__ p_next,
                                                      double *__restrict__ part_pap, int n_tiles) {
    __shared__ double red[4];
    src.in = step4_in(src.st, src.slot);
    const int tile = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
    PcgNextFin<T> fin{flags, ap, p_next};
    // A halted solve ends each wave's march at its ready() (every wave still joins the partial sum's barriers).
    // Waves may see the halt flag differently when this launch's lead thread sets it: the solve is then over and
    // nothing reads this launch's outputs.
    if (tile < n_tiles) ec_tile_march<T, KZ>(g, km, ep, src, tile, fin);
    const double sum = block_sum256(fin.dot, red);
    if (threadIdx.x == 0 && fin.go) part_pap[blockIdx.x] = sum;  // wave 0 of every block has a tile
}

// ---- packed multigrid iterations: the outer kernels of several solves (different grids) in one launch ----
// Grid dimension y is the member, its arguments its entry of a device table, blocks beyond the member's own grid
// return at entry; each member's blocks do exactly what its own launch's blocks do. slot: the iteration parity
// (the same for every member: each joins a pack at the start of a batch, in slot 0).
struct PkXr {
    int n, nb_mv, nb;
    PcgState *st;
    const double *part_pap;
    double *x, *r;
    const double *p[2];  // p of the even / odd iterations
    const double *ap;
    double *part_rr;
    GmgFineView fv;
};
struct PkStep4 {
    Dims g;
    KeMono<double> km;
    const double *ep;
    const uint8_t *flags;
    double *p[2];        // p of the even / odd iterations; step 4 writes the other one
    const double *z;
    PcgState *st;
    const double *part_rr, *part_zr, *part_zap;
    int nb_vec, nb_z;
    double *ap, *part_pap;
    int n_tiles, nb;
};

// Table entries with every pointer marked global (gmem): read from a device table they would otherwise be generic,
// and every LDS access or lane exchange behind a flat access would wait for it.
__device__ __forceinline__ PkXr globalized(PkXr a) {
    a.st = gmem(a.st), a.part_pap = gmem(a.part_pap), a.x = gmem(a.x), a.r = gmem(a.r);
    a.p[0] = gmem(a.p[0]), a.p[1] = gmem(a.p[1]), a.ap = gmem(a.ap), a.part_rr = gmem(a.part_rr);
    a.fv.b = gmem(a.fv.b), a.fv.x = gmem(a.fv.x), a.fv.d = gmem(a.fv.d), a.fv.inv = gmem(a.fv.inv);
    a.fv.binv = gmem(a.fv.binv);
    return a;
}
__device__ __forceinline__ PkStep4 globalized(PkStep4 a) {
    a.ep = gmem(a.ep), a.flags = gmem(a.flags), a.p[0] = gmem(a.p[0]), a.p[1] = gmem(a.p[1]), a.z = gmem(a.z);
    a.st = gmem(a.st), a.part_rr = gmem(a.part_rr), a.part_zr = gmem(a.part_zr), a.part_zap = gmem(a.part_zap);
    a.ap = gmem(a.ap), a.part_pap = gmem(a.part_pap);
    return a;
}

__global__ __launch_bounds__(256) void k_pk_fpcg_update_xr(const PkXr *__restrict__ tab, int slot) {
    const PkXr a = globalized(tab[blockIdx.y]);
    if (static_cast<int>(blockIdx.x) >= a.nb) return;  // block-uniform, ahead of the block sum's barriers
    fpcg_update_xr_body(a.n, slot, a.st, a.part_pap, a.nb_mv, a.x, a.r, slot ? a.p[1] : a.p[0], a.ap, a.part_rr,
                        a.fv, blockIdx.x);
}

template <int KZ>
__global__ __launch_bounds__(256, 2) void k_pk_step4_ap(const PkStep4 *__restrict__ tab, int slot) {
    __shared__ double red[4];
    const PkStep4 a = globalized(tab[blockIdx.y]);
    if (static_cast<int>(blockIdx.x) >= a.nb) return;
    double *p_cur = slot ? a.p[1] : a.p[0], *p_next = slot ? a.p[0] : a.p[1];
    PcgNextP<double, true> src{p_cur, a.z, a.st, a.part_rr, a.part_zr, a.part_zap, a.nb_vec, a.nb_z, slot, 0.0,
                               0.0, Step4In{}, 0};
    src.in = step4_in(a.st, slot);
    const int tile = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, a.nb) * 4 + (threadIdx.x >> 6));
    PcgNextFin<double> fin{a.flags, a.ap, p_next};
    if (tile < a.n_tiles) ec_tile_march<double, KZ>(a.g, a.km, a.ep, src, tile, fin);
    const double sum = block_sum256(fin.dot, red);
    if (threadIdx.x == 0 && fin.go) a.part_pap[blockIdx.x] = sum;
}

// The members' states after a batch, gathered for one copy to the host (one block per slot).
__global__ __launch_bounds__(64) void k_pk_gather_states(PcgState *const *__restrict__ src,
                                                         PcgState *__restrict__ dst) {
    constexpr int kWords = static_cast<int>(sizeof(PcgState) / sizeof(uint32_t));
    const PcgState *from_state = gmem(src[blockIdx.x]);
    if (!from_state) return;  // an empty slot
    const uint32_t *from = reinterpret_cast<const uint32_t *>(from_state);
    uint32_t *to = reinterpret_cast<uint32_t *>(dst + blockIdx.x);
    for (int i = threadIdx.x; i < kWords; i += 64) to[i] = from[i];
}

// p = z (the first search direction of the preconditioned iteration).
__global__ __launch_bounds__(256) void k_copy(int n, const double *__restrict__ src, double *__restrict__ dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const long long o = static_cast<long long>(blockIdx.y) * n;
    if (i < n) dst[o + i] = src[o + i];
}

// x[constrained] = 0; flag a nonzero projected x; partials of ||Pb||^2 (PcgSolver.cs:113-137).
__global__ __launch_bounds__(256) void k_pcg_init(int n, const uint8_t *__restrict__ flags,
                                                  const double *__restrict__ b, double *__restrict__ x,
                                                  PcgState *__restrict__ st, double *__restrict__ part1) {
    __shared__ double red[4];
    {  // grid dimension y: the right-hand side
        const long long o = static_cast<long long>(blockIdx.y) * n;
        b += o;
        x += o;
        st += blockIdx.y;
        part1 += blockIdx.y * gridDim.x;
    }
    double v = 0.0;
    int nz = 0;
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) {
            if (flags[i] & kFlagConstrained) {
                x[i] = 0.0;
            } else {
                nz |= (x[i] != 0.0);
                v = __fma_rn(b[i], b[i], v);
            }
        }
    }
    if (nz) atomicOr(&st->nonzero_x, 1);
    v = block_sum256(v, red);
    if (threadIdx.x == 0) part1[blockIdx.x] = v;
}

// r = P(b - ap) (warm start) or r = Pb (cold start); partials of r.r.
__global__ __launch_bounds__(256) void k_pcg_init_r(int n, const uint8_t *__restrict__ flags,
                                                    const double *__restrict__ b, const double *__restrict__ ap,
                                                    double *__restrict__ r, double *__restrict__ part1) {
    __shared__ double red[4];
    {  // grid dimension y: the right-hand side
        const long long o = static_cast<long long>(blockIdx.y) * n;
        b += o;
        if (ap) ap += o;
        r += o;
        part1 += blockIdx.y * gridDim.x;
    }
    double v = 0.0;
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) {
            const double ri = (flags[i] & kFlagConstrained) ? 0.0 : (ap ? b[i] - ap[i] : b[i]);
            r[i] = ri;
            v = __fma_rn(ri, ri, v);
        }
    }
    v = block_sum256(v, red);
    if (threadIdx.x == 0) part1[blockIdx.x] = v;
}

// z = r / diag; p = z; partials of r.z. With T = float the FP64 residual is rounded into r32 first.
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_init_z(int n, const double *__restrict__ r64, T *__restrict__ r,
                                                    const T *__restrict__ inv, T *__restrict__ z, T *__restrict__ p,
                                                    double *__restrict__ part1, int bf16) {
    __shared__ double red[4];
    double v = 0.0;
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) {
            const T ri = static_cast<T>(r64[i]);
            if (r != reinterpret_cast<const T *>(r64)) r[i] = ri;
            const T zi = ri * inv[i];
            z[i] = zi;
            if (p) p[i] = maybe_bf16(zi, bf16);
            v = __fma_rn(static_cast<double>(ri), static_cast<double>(zi), v);
        }
    }
    v = block_sum256(v, red);
    if (threadIdx.x == 0) part1[blockIdx.x] = v;
}

// Reliable update, part 1: x64 += x32; x32 = 0.
__global__ __launch_bounds__(256) void k_mixed_flush(int n, double *__restrict__ x, float *__restrict__ x32) {
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) {
            x[i] += static_cast<double>(x32[i]);
            x32[i] = 0.0f;
        }
    }
}

// Reliable update, part 3 (single block): true relres, convergence, beta with the refreshed r, and
// the state for the next iteration written to slot 0 (the slot a fresh batch starts with).
__global__ __launch_bounds__(256) void k_mixed_resume(PcgState *__restrict__ st, const double *__restrict__ part_rr,
                                                      const double *__restrict__ part_rz, int nb) {
    __shared__ double red[4];
    const double rr = sum_partials256(part_rr, nb, red);
    const double rzn = sum_partials256(part_rz, nb, red);
    if (threadIdx.x != 0) return;
    const double rel = sqrt(rr) / st->bnorm;
    const int t = st->iterations;
    st->outer += 1;
    st->pause = 0;
    if (rel <= st->tol) {
        halt_with(st, MYCEL_PCG_CONVERGED, t, rel);
        return;
    }
    if (t >= st->max_iter) {
        halt_with(st, MYCEL_PCG_MAX_ITERATIONS, t, rel);
        return;
    }
    st->no_gain = rel < kReliableGain * st->best_true ? 0 : st->no_gain + 1;
    st->best_true = fmin(st->best_true, rel);
    if (st->no_gain >= kReliableStall) {
        halt_with(st, MYCEL_PCG_STAGNATION, t, rel);
        return;
    }
    const double beta = rzn / st->rz[st->pause_slot];
    if (!isfinite(beta) || beta < 0.0) {
        halt_with(st, MYCEL_PCG_DIRECTION_BREAKDOWN, t, rel, beta);
        return;
    }
    st->beta = beta;
    st->rz[0] = rzn;
    st->prev_rel[0] = rel;
    st->stag[0] = 0;
    st->iter[0] = t + 1;
    st->maxrel = rel;
    st->last_rel = rel;
    st->last_iter = t;
    st->halt = 0;
}

// Reliable update, part 4: p = z + beta p with the beta of k_mixed_resume.
// p_out = z + beta p_in (p_out may be p_in).
__global__ __launch_bounds__(256) void k_mixed_update_p(int n, const PcgState *__restrict__ st,
                                                        const float *p_in, float *p_out,
                                                        const float *__restrict__ z, int bf16) {
    if (st->done) return;
    const float b = static_cast<float>(st->beta);
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) p_out[i] = maybe_bf16(__fmaf_rn(b, p_in[i], z[i]), bf16);
    }
}

// Sums the partials into one state field (single block).
// Grid dimension y: the right-hand side (partials at y * nb, result at out + y * out_stride doubles).
__global__ __launch_bounds__(256) void k_pcg_sum_to(const double *__restrict__ part, int nb, double *__restrict__ out,
                                                    int out_stride) {
    __shared__ double red[4];
    const double s = sum_partials256(part + blockIdx.y * nb, nb, red);
    if (threadIdx.x == 0) out[static_cast<long long>(blockIdx.y) * out_stride] = s;
}

// A right-hand side that needs no iteration (zero RHS, converged start, no iterations allowed).
__global__ void k_pcg_halt_at_start(PcgState *__restrict__ st, int message, double rel) {
    st->message = message;
    st->iterations = 0;
    st->final_rel = rel;
    st->done = 1;
    st->halt = 1;
}

__global__ void k_pcg_start(PcgState *__restrict__ st, double tol, double bnorm, double rel, int max_iter, int mixed,
                            double delta) {
    st->tol = tol;
    st->bnorm = bnorm;
    st->delta = delta;
    st->maxrel = rel;
    st->max_iter = max_iter;
    st->mixed = mixed;
    st->rz[0] = st->rz0;
    st->prev_rel[0] = rel;
    st->stag[0] = 0;
    st->iter[0] = 1;
    st->done = 0;
    st->halt = 0;
    st->pause = 0;
    st->outer = 0;
    st->last_rel = rel;
    st->last_iter = 0;
    st->best_true = rel;
    st->no_gain = 0;
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

void ensure_work(mycel_op *op) {
    if (op->pcg) return;
    PcgWork *w = new PcgWork();
    op->pcg = w;
    w->n = op->d.n_dofs;
    w->nb_mv = apply_k_blocks(op);
    w->nb_vec = cdiv(w->n, kVecPerBlock);
    w->nb_rr = cdiv(w->n / 3, kXrNodes);
    w->r.resize(w->n);
    w->z.resize(w->n);
    w->p.resize(w->n);
    w->p2.resize(w->n);
    w->ap.resize(w->n);
    w->part_pap.resize(w->nb_mv);
    w->part2.resize(2 * w->nb_vec);
    w->part1.resize(2 * w->nb_vec);
    w->state.resize(1);
    MYCEL_HIP(hipHostMalloc(reinterpret_cast<void **>(&w->h_state), sizeof(PcgState), hipHostMallocDefault));
}

void ensure_mixed(PcgWork *w) {
    if (w->x32.size()) return;
    w->r32.resize(w->n);
    w->z32.resize(w->n);
    w->p32.resize(w->n);
    w->p32b.resize(w->n);
    w->ap32.resize(w->n);
    w->x32.resize(w->n);
}

// Step 4 fused with the next A p (k_pcg_step4_ap) in every solver mode: on by default; MYCEL_PCG_FUSE_P=0
// separates them (the p update, then the K*u launch). Values are identical either way.
bool fuse_step4() {
    static const bool on = [] {
        const char *e = std::getenv("MYCEL_PCG_FUSE_P");
        return !(e && e[0] == '0');
    }();
    return on;
}

// p of iteration t (fused step 4): p (p32) for t even, p2 (p32b) for t odd; step 4 of iteration t writes the
// other. A batch starts at t = 0, and so does the FP32 iteration after a reliable update.
static_assert(kGraphIters % 2 == 0 && kGmgGraphIters % 2 == 0, "fused step 4: p buffers alternate per iteration");

template <typename T, bool FLEX>
void launch_step4_ap(mycel_op *op, const KeMono<T> &km, const T *ep, const T *p, const T *z, T *p_next, T *ap,
                     PcgWork *w, int slot, hipStream_t s) {
    const int kz = ku_element_layers(op);
    const int nt = ec_tiles(op->d, kz);
    // update_xr reads the partials of this launch as those of the automatic K*u form
    if (cdiv(nt, 4) != w->nb_mv) throw Error(MYCEL_E_INVALID_OPERATION, "fused step 4: partial count mismatch");
    // part2: the interleaved (r.r, r.z) partials of the Jacobi update (nb_vec), or the flexible update's r.r (nb_rr)
    const PcgNextP<T, FLEX> src{p, z, w->state.get(), w->part2.get(), w->part_zr.get(), w->part_zap.get(),
                                FLEX ? w->nb_rr : w->nb_vec, w->nb_z, slot, 0.0, T(0), Step4In{}, w->bf16};
    const dim3 grid(cdiv(nt, 4)), b256(256);
    const uint8_t *fl = op->d_flags.get();
    double *part = w->part_pap.get();
    if (kz == 2)
        hipLaunchKernelGGL((k_pcg_step4_ap<T, FLEX, 2>), grid, b256, 0, s, op->d, km, ep, fl, src, ap, p_next, part,
                           nt);
    else if (kz == 3)
        hipLaunchKernelGGL((k_pcg_step4_ap<T, FLEX, 3>), grid, b256, 0, s, op->d, km, ep, fl, src, ap, p_next, part,
                           nt);
    else if (kz == 4)
        hipLaunchKernelGGL((k_pcg_step4_ap<T, FLEX, 4>), grid, b256, 0, s, op->d, km, ep, fl, src, ap, p_next, part,
                           nt);
    else if (kz == 8)
        hipLaunchKernelGGL((k_pcg_step4_ap<T, FLEX, 8>), grid, b256, 0, s, op->d, km, ep, fl, src, ap, p_next, part,
                           nt);
    else
        throw Error(MYCEL_E_INVALID_OPERATION, "fused step 4: unsupported tile depth");
}

template <bool FLEX>
void launch_step4_ap64(mycel_op *op, PcgWork *w, int slot, hipStream_t s) {
    launch_step4_ap<double, FLEX>(op, op->km64, op->d_ep.get(), slot ? w->p2.get() : w->p.get(), w->z.get(),
                                  slot ? w->p.get() : w->p2.get(), w->ap.get(), w, slot, s);
}

// One FP64 Jacobi iteration. Fused: its A p was formed by the step 4 before it (or, for the first, before the
// loop).
void enqueue_iteration64(mycel_op *op, PcgWork *w, double *d_x, int slot, hipStream_t s) {
    PcgState *st = w->state.get();
    if (fuse_step4()) {
        hipLaunchKernelGGL(k_pcg_update_xr<double>, dim3(w->nb_vec), dim3(256), 0, s, w->n, slot, st,
                           w->part_pap.get(), w->nb_mv, d_x, w->r.get(), slot ? w->p2.get() : w->p.get(), w->ap.get(),
                           w->z.get(), op->d_inv_diag.get(), w->part2.get());
        launch_step4_ap64<false>(op, w, slot, s);
        return;
    }
    launch_apply_k(op, w->p.get(), w->ap.get(), kApplyAProjected, s, w->part_pap.get(), &st->halt);
    hipLaunchKernelGGL(k_pcg_update_xr<double>, dim3(w->nb_vec), dim3(256), 0, s, w->n, slot, st, w->part_pap.get(),
                       w->nb_mv, d_x, w->r.get(), w->p.get(), w->ap.get(), w->z.get(), op->d_inv_diag.get(),
                       w->part2.get());
    hipLaunchKernelGGL(k_pcg_update_p<double>, dim3(w->nb_vec), dim3(256), 0, s, w->n, slot, st, w->part2.get(),
                       w->nb_vec, w->p.get(), w->z.get(), 0);
}

// One FP32 iteration of the reliable-update solver. Fused: its A p was formed by the step 4 before it (or, for
// the first after the start or a reliable update, by a plain K*u launch).
void enqueue_iteration32(mycel_op *op, PcgWork *w, int slot, hipStream_t s) {
    PcgState *st = w->state.get();
    if (fuse_step4()) {
        float *p = slot ? w->p32b.get() : w->p32.get(), *p_next = slot ? w->p32.get() : w->p32b.get();
        hipLaunchKernelGGL(k_pcg_update_xr<float>, dim3(w->nb_vec), dim3(256), 0, s, w->n, slot, st,
                           w->part_pap.get(), w->nb_mv, w->x32.get(), w->r32.get(), p, w->ap32.get(), w->z32.get(),
                           op->d_inv_diag32.get(), w->part2.get());
        launch_step4_ap<float, false>(op, op->km32, op->d_ep32.get(), p, w->z32.get(), p_next, w->ap32.get(), w, slot,
                                      s);
        return;
    }
    launch_apply_k_f32(op, w->p32.get(), w->ap32.get(), kApplyAProjected, s, w->part_pap.get(), &st->halt);
    hipLaunchKernelGGL(k_pcg_update_xr<float>, dim3(w->nb_vec), dim3(256), 0, s, w->n, slot, st, w->part_pap.get(),
                       w->nb_mv, w->x32.get(), w->r32.get(), w->p32.get(), w->ap32.get(), w->z32.get(),
                       op->d_inv_diag32.get(), w->part2.get());
    hipLaunchKernelGGL(k_pcg_update_p<float>, dim3(w->nb_vec), dim3(256), 0, s, w->n, slot, st, w->part2.get(),
                       w->nb_vec, w->p32.get(), w->z32.get(), w->bf16);
}

void ensure_gmg(mycel_op *op, PcgWork *w) {
    if (w->part_zr.size()) return;
    w->nb_z = gmg_partials(op);
    w->part_zr.resize(w->nb_z);
    w->part_zap.resize(w->nb_z);
}


// One multigrid-preconditioned iteration. Fused: its A p was formed by the step 4 before it (or, for the first,
// before the loop).
void enqueue_iteration_gmg(mycel_op *op, PcgWork *w, double *d_x, int slot, hipStream_t s) {
    PcgState *st = w->state.get();
    if (fuse_step4()) {
        hipLaunchKernelGGL(k_fpcg_update_xr_gmg, dim3(w->nb_rr), dim3(256), 0, s, w->n, slot, st, w->part_pap.get(),
                           w->nb_mv, d_x, w->r.get(), slot ? w->p2.get() : w->p.get(), w->ap.get(), w->part2.get(),
                           gmg_fine_view(op));
        gmg_enqueue_vcycle(op, w->r.get(), w->ap.get(), w->z.get(), w->part_zr.get(), w->part_zap.get(), s, true, 1,
                           &st->halt);
        launch_step4_ap64<true>(op, w, slot, s);
        return;
    }
    launch_apply_k(op, w->p.get(), w->ap.get(), kApplyAProjected, s, w->part_pap.get(), &st->halt);
    hipLaunchKernelGGL(k_fpcg_update_xr_gmg, dim3(w->nb_rr), dim3(256), 0, s, w->n, slot, st, w->part_pap.get(),
                       w->nb_mv, d_x, w->r.get(), w->p.get(), w->ap.get(), w->part2.get(), gmg_fine_view(op));
    gmg_enqueue_vcycle(op, w->r.get(), w->ap.get(), w->z.get(), w->part_zr.get(), w->part_zap.get(), s, true, 1,
                       &st->halt);
    hipLaunchKernelGGL(k_fpcg_update_p, dim3(w->nb_vec), dim3(256), 0, s, w->n, slot, st, w->part2.get(), w->nb_rr,
                       w->part_zr.get(), w->part_zap.get(), w->nb_z, w->p.get(), w->z.get());
}

void ensure_batch(mycel_op *op, PcgWork *w, int R) {
    if (R <= w->b_cap) return;
    const size_t n = static_cast<size_t>(w->n) * R;
    const int nb_node = apply_k_blocks(op);
    w->br.resize(n);
    w->bz.resize(n);
    w->bp.resize(n);
    w->bap.resize(n);
    w->bpart_pap.resize(static_cast<size_t>(nb_node) * R);
    w->bpart_rr.resize(static_cast<size_t>(w->nb_rr) * R);  // the flexible x/r update's r.r: nb_rr per RHS
    w->bpart1.resize(static_cast<size_t>(w->nb_vec) * R);
    w->bpart_zr.resize(static_cast<size_t>(gmg_partials(op)) * R);
    w->bpart_zap.resize(static_cast<size_t>(gmg_partials(op)) * R);
    w->bstate.resize(R);
    if (w->h_bstate) (void)hipHostFree(w->h_bstate);
    w->h_bstate = nullptr;
    MYCEL_HIP(hipHostMalloc(reinterpret_cast<void **>(&w->h_bstate), sizeof(PcgState) * R, hipHostMallocDefault));
    w->b_cap = R;
    if (w->graph_b) (void)hipGraphExecDestroy(w->graph_b);
    w->graph_b = nullptr;
}

// One multigrid-preconditioned iteration for R right-hand sides (the kernels of enqueue_iteration_gmg on
// grid rows y = 0..R-1).
void enqueue_iteration_gmg_batch(mycel_op *op, PcgWork *w, double *d_X, int R, int slot, hipStream_t s) {
    PcgState *st = w->bstate.get();
    const int nb_node = apply_k_blocks(op);  // partials of the batched K*u launch
    const int words = static_cast<int>(sizeof(PcgState) / sizeof(int));
    const Batch bt{R, static_cast<long long>(w->n), nb_node, words};
    if (w->bpart_rr.size() < static_cast<size_t>(w->nb_rr) * R)
        throw Error(MYCEL_E_INVALID_OPERATION, "batched flexible update: partial buffer too small");
    launch_apply_k(op, w->bp.get(), w->bap.get(), kApplyAProjected, s, w->bpart_pap.get(), &st->halt, bt);
    hipLaunchKernelGGL(k_fpcg_update_xr_gmg, dim3(w->nb_rr, R), dim3(256), 0, s, w->n, slot, st,
                       w->bpart_pap.get(), nb_node, d_X, w->br.get(), w->bp.get(), w->bap.get(), w->bpart_rr.get(),
                       gmg_fine_view(op));
    gmg_enqueue_vcycle(op, w->br.get(), w->bap.get(), w->bz.get(), w->bpart_zr.get(), w->bpart_zap.get(), s, true,
                       R, &st->halt, words);
    hipLaunchKernelGGL(k_fpcg_update_p, dim3(w->nb_vec, R), dim3(256), 0, s, w->n, slot, st, w->bpart_rr.get(),
                       w->nb_rr, w->bpart_zr.get(), w->bpart_zap.get(), w->nb_z, w->bp.get(), w->bz.get());
}

void read_state(PcgWork *w, hipStream_t s) {
    MYCEL_HIP(hipMemcpyAsync(w->h_state, w->state.get(), sizeof(PcgState), hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipStreamSynchronize(s));
}

template <typename F>
hipGraphExec_t capture(hipStream_t s, F &&enqueue) {
    hipGraph_t g;
    hipGraphExec_t ex;
    MYCEL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    enqueue();
    MYCEL_HIP(hipStreamEndCapture(s, &g));
    MYCEL_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    MYCEL_HIP(hipGraphDestroy(g));
    return ex;
}

}  // namespace

void pcg_free(PcgWork *w) {
    if (!w) return;
    if (w->graph64) (void)hipGraphExecDestroy(w->graph64);
    if (w->graph32) (void)hipGraphExecDestroy(w->graph32);
    if (w->graph_gmg) (void)hipGraphExecDestroy(w->graph_gmg);
    if (w->graph_b) (void)hipGraphExecDestroy(w->graph_b);
    if (w->h_bstate) (void)hipHostFree(w->h_bstate);
    if (w->h_state) (void)hipHostFree(w->h_state);
    delete w;
}

namespace {

using SolveClock = std::chrono::steady_clock;

double ms_since(SolveClock::time_point t0) {
    return std::chrono::duration<double, std::milli>(SolveClock::now() - t0).count();
}

// A solve's set-up, up to its first batch (PcgSolver.cs:80-233): ||Pb||, the zero right-hand side, the warm start,
// the initial residual and preconditioned residual, the initial convergence test, the state of slot 0 and (fused
// step 4) the first A p. Returns false when the solve ended here (res filled).
bool solve_setup(mycel_op *op, const double *d_b, double *d_x, double tol, int max_iter, int precision,
                 mycel_pcg_result *res, SolveClock::time_point t0) {
    const bool mixed = precision == MYCEL_PRECISION_MIXED_FP32 || precision == MYCEL_PRECISION_MIXED_BF16;
    const bool gmg = precision == MYCEL_PRECISION_MIXED_GMG;
    if (op->active_fixed < 6)  // PcgSolver.cs:95-100
        throw Error(MYCEL_E_INVALID_OPERATION,
                    "Model has fewer than 6 unique fixed DOFs on active nodes (found " +
                        std::to_string(op->active_fixed) +
                        "). At least 6 fixed DOFs are required to prevent rigid body modes in 3D.");
    ensure_work(op);
    PcgWork *w = op->pcg;
    if (mixed) ensure_mixed(w);
    const int bf16 = precision == MYCEL_PRECISION_MIXED_BF16;
    if (w->graph32 && w->bf16 != bf16) {  // the captured FP32 batch bakes in the rounding mode
        MYCEL_HIP(hipGraphExecDestroy(w->graph32));
        w->graph32 = nullptr;
    }
    w->bf16 = bf16;
    if (gmg) {
        gmg_prepare(op, op->stream);
        ensure_gmg(op, w);
    }
    hipStream_t s = op->stream;
    PcgState *st = w->state.get();
    const int n = w->n;
    const dim3 gv(w->nb_vec), b256(256);

    MYCEL_HIP(hipMemsetAsync(st, 0, sizeof(PcgState), s));
    hipLaunchKernelGGL(k_pcg_init, gv, b256, 0, s, n, op->d_flags.get(), d_b, d_x, st, w->part1.get());
    hipLaunchKernelGGL(k_pcg_sum_to, dim3(1), b256, 0, s, w->part1.get(), w->nb_vec, &st->bnorm_sq, 0);
    MYCEL_LAUNCH_CHECK();
    read_state(w, s);
    const double bnorm = std::sqrt(w->h_state->bnorm_sq);
    if (bnorm <= 1e-30) {  // PcgSolver.cs:139-151
        MYCEL_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, s));
        MYCEL_HIP(hipStreamSynchronize(s));
        res->converged = 1;
        res->message = MYCEL_PCG_RHS_ZERO;
        res->elapsed_ms = ms_since(t0);
        return false;
    }
    const bool warm = w->h_state->nonzero_x != 0;
    if (warm) launch_apply_k(op, d_x, w->ap.get(), kApplyAProjected, s);
    hipLaunchKernelGGL(k_pcg_init_r, gv, b256, 0, s, n, op->d_flags.get(), d_b, warm ? w->ap.get() : nullptr,
                       w->r.get(), w->part1.get());
    hipLaunchKernelGGL(k_pcg_sum_to, dim3(1), b256, 0, s, w->part1.get(), w->nb_vec, &st->r_norm_sq, 0);
    if (gmg) {
        gmg_enqueue_vcycle(op, w->r.get(), w->r.get(), w->z.get(), w->part_zr.get(), w->part_zap.get(), s);
        hipLaunchKernelGGL(k_copy, dim3(cdiv(n, 256)), b256, 0, s, n, w->z.get(), w->p.get());
        hipLaunchKernelGGL(k_pcg_sum_to, dim3(1), b256, 0, s, w->part_zr.get(), w->nb_z, &st->rz0, 0);
    } else if (mixed) {
        hipLaunchKernelGGL(k_pcg_init_z<float>, gv, b256, 0, s, n, w->r.get(), w->r32.get(), op->d_inv_diag32.get(),
                           w->z32.get(), w->p32.get(), w->part1.get(), w->bf16);
        MYCEL_HIP(hipMemsetAsync(w->x32.get(), 0, sizeof(float) * n, s));
    } else {
        hipLaunchKernelGGL(k_pcg_init_z<double>, gv, b256, 0, s, n, w->r.get(), w->r.get(), op->d_inv_diag.get(),
                           w->z.get(), w->p.get(), w->part1.get(), 0);
    }
    // rz0 = r.z: the Jacobi forms left its partials in part1; the multigrid form summed them above.
    if (!gmg) hipLaunchKernelGGL(k_pcg_sum_to, dim3(1), b256, 0, s, w->part1.get(), w->nb_vec, &st->rz0, 0);
    MYCEL_LAUNCH_CHECK();
    read_state(w, s);
    // A cold start uses ||Pb||^2 itself as ||r||^2 (PcgSolver.cs:185-192).
    const double rr0 = warm ? w->h_state->r_norm_sq : w->h_state->bnorm_sq;
    const double rel0 = std::sqrt(rr0) / bnorm;
    if (rel0 <= tol) {  // :197-209
        res->final_residual = rel0;
        res->converged = 1;
        res->message = MYCEL_PCG_INITIAL_CONVERGED;
        res->elapsed_ms = ms_since(t0);
        return false;
    }
    if (max_iter <= 0) {  // the reference's loop body never runs (PcgSolver.cs:235,359-368)
        res->final_residual = rel0;
        res->message = MYCEL_PCG_MAX_ITERATIONS;
        res->elapsed_ms = ms_since(t0);
        return false;
    }
    hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(1), 0, s, st, tol, bnorm, rel0, max_iter, mixed ? 1 : 0,
                       kReliableDelta);
    // The fused iterations start from A p of their first direction.
    if (fuse_step4()) {
        if (mixed)
            launch_apply_k_f32(op, w->p32.get(), w->ap32.get(), kApplyAProjected, s, w->part_pap.get(), &st->halt);
        else
            launch_apply_k(op, w->p.get(), w->ap.get(), kApplyAProjected, s, w->part_pap.get(), &st->halt);
    }
    MYCEL_LAUNCH_CHECK();
    return true;
}

// The iteration loop of one solve on its own: batches of iterations captured once as a hipGraph and replayed,
// the state read between batches; leaves the final state in w->h_state.
void solve_loop(mycel_op *op, const double *d_b, double *d_x, int max_iter, int precision,
                const volatile int32_t *cancel, mycel_pcg_progress_cb cb, void *user) {
    const bool mixed = precision == MYCEL_PRECISION_MIXED_FP32 || precision == MYCEL_PRECISION_MIXED_BF16;
    const bool gmg = precision == MYCEL_PRECISION_MIXED_GMG;
    PcgWork *w = op->pcg;
    hipStream_t s = op->stream;
    PcgState *st = w->state.get();
    const int n = w->n;
    const dim3 gv(w->nb_vec), b256(256);
    hipGraphExec_t graph;
    if (gmg) {
        if (w->graph_gmg && (w->graph_gmg_x != d_x || w->graph_gmg_h != op->gmg ||
                             w->graph_gmg_gen != gmg_generation(op))) {
            MYCEL_HIP(hipGraphExecDestroy(w->graph_gmg));
            w->graph_gmg = nullptr;
        }
        if (!w->graph_gmg) {
            w->graph_gmg = capture(s, [&] {
                for (int t = 0; t < kGmgGraphIters; t++) enqueue_iteration_gmg(op, w, d_x, t & 1, s);
            });
            w->graph_gmg_x = d_x;
            w->graph_gmg_h = op->gmg;
            w->graph_gmg_gen = gmg_generation(op);
        }
        graph = w->graph_gmg;
    } else if (mixed) {
        if (!w->graph32)
            w->graph32 = capture(s, [&] {
                for (int t = 0; t < kGraphIters; t++) enqueue_iteration32(op, w, t & 1, s);
            });
        graph = w->graph32;
    } else {
        // The captured FP64 batch bakes in the solution pointer; recapture if the caller's x moved.
        if (w->graph64 && w->graph64_x != d_x) {
            MYCEL_HIP(hipGraphExecDestroy(w->graph64));
            w->graph64 = nullptr;
        }
        if (!w->graph64) {
            w->graph64 = capture(s, [&] {
                for (int t = 0; t < kGraphIters; t++) enqueue_iteration64(op, w, d_x, t & 1, s);
            });
            w->graph64_x = d_x;
        }
        graph = w->graph64;
    }

    // Every batch advances the iteration count by at least one or halts the solve, so max_iter bounds
    // the number of batches (a guard against a logic error, not a limit).
    const long long max_batches = static_cast<long long>(max_iter) + 64;
    // Batches launched back to back before the first state read (FP64 modes), none on the first solve: each
    // state read is a host round trip that leaves the GPU idle. A batch that runs past the halt only recomputes
    // values nothing reads (every kernel that writes x or r tests the halt flag); in the Jacobi mode such a
    // batch's kernels exit at entry, in the multigrid mode its V-cycles still run, so that mode covers only half
    // of the previous converged solve's iterations and the Jacobi mode three quarters.
    int &last_iters = gmg ? w->last_iters_gmg : w->last_iters_64;
    const long long prior = std::min(last_iters, max_iter);
    // A caller that can cancel or watches progress gets a state read after every batch.
    const long long blind = (mixed || cancel || cb) ? 0 : gmg ? prior / 2 / kGmgGraphIters : prior * 3 / 4 / kGraphIters;
    for (long long batch = 0;; batch++) {
        if (batch > max_batches) throw Error(MYCEL_E_DEVICE, "PCG batch loop did not terminate");
        if (cancel && *cancel) {
            MYCEL_HIP(hipStreamSynchronize(s));
            throw Error(MYCEL_E_CANCELED, "The operation was canceled.");
        }
        MYCEL_HIP(hipGraphLaunch(graph, s));
        if (batch < blind) continue;
        read_state(w, s);
        const PcgState &h = *w->h_state;
        if (cb) cb(h.done ? h.iterations : h.last_iter, h.done ? h.final_rel : h.last_rel, user);
        if (h.done) break;
        if (h.pause) {  // reliable update with the FP64 operator
            hipLaunchKernelGGL(k_mixed_flush, gv, b256, 0, s, n, d_x, w->x32.get());
            launch_apply_k(op, d_x, w->ap.get(), kApplyAProjected, s);
            hipLaunchKernelGGL(k_pcg_init_r, gv, b256, 0, s, n, op->d_flags.get(), d_b, w->ap.get(), w->r.get(),
                               w->part1.get());
            hipLaunchKernelGGL(k_pcg_init_z<float>, gv, b256, 0, s, n, w->r.get(), w->r32.get(),
                               op->d_inv_diag32.get(), w->z32.get(), static_cast<float *>(nullptr),
                               w->part1.get() + w->nb_vec, 0);
            hipLaunchKernelGGL(k_mixed_resume, dim3(1), b256, 0, s, st, w->part1.get(), w->part1.get() + w->nb_vec,
                               w->nb_vec);
            const float *p_in = fuse_step4() && w->h_state->pause_slot ? w->p32b.get() : w->p32.get();
            hipLaunchKernelGGL(k_mixed_update_p, gv, b256, 0, s, n, st, p_in, w->p32.get(), w->z32.get(), w->bf16);
            if (fuse_step4())
                launch_apply_k_f32(op, w->p32.get(), w->ap32.get(), kApplyAProjected, s, w->part_pap.get(),
                                   &st->halt);
            MYCEL_LAUNCH_CHECK();
            read_state(w, s);
            if (w->h_state->done) break;
        }
    }
}

// The result from the final state in w->h_state, and the FP64 finish after a breakdown of the FP32
// preconditioner or iteration.
void solve_finish(mycel_op *op, const double *d_b, double *d_x, double tol, int max_iter, int precision,
                  const volatile int32_t *cancel, mycel_pcg_progress_cb cb, void *user, mycel_pcg_result *res,
                  SolveClock::time_point t0) {
    const bool mixed = precision == MYCEL_PRECISION_MIXED_FP32 || precision == MYCEL_PRECISION_MIXED_BF16;
    const bool gmg = precision == MYCEL_PRECISION_MIXED_GMG;
    PcgWork *w = op->pcg;
    hipStream_t s = op->stream;
    const int n = w->n;
    const dim3 gv(w->nb_vec), b256(256);
    const PcgState h = *w->h_state;
    int &last_iters = gmg ? w->last_iters_gmg : w->last_iters_64;
    if (!mixed) last_iters = h.message == MYCEL_PCG_CONVERGED ? h.iterations : 0;
    res->iterations = h.iterations;
    res->final_residual = h.final_rel;
    res->message = h.message;
    res->message_value = h.value;
    res->converged = h.message == MYCEL_PCG_CONVERGED;
    res->outer_corrections = h.outer;
    if (gmg) {
        const bool broke = h.message == MYCEL_PCG_CURVATURE_BREAKDOWN || h.message == MYCEL_PCG_STEP_BREAKDOWN ||
                           h.message == MYCEL_PCG_DIRECTION_BREAKDOWN || h.message == MYCEL_PCG_STAGNATION;
        if (broke && h.iterations < max_iter) {
            // The FP32 preconditioner lost positivity on this residual: x is a valid FP64 iterate, so finish
            // with the FP64 Jacobi iteration warm-started from it.
            mycel_pcg_result r64{};
            pcg_solve_device(op, d_b, d_x, tol, max_iter - h.iterations, MYCEL_PRECISION_FP64, cancel, cb, user,
                             &r64);
            res->iterations = h.iterations + r64.iterations;
            res->final_residual = r64.final_residual;
            res->message = r64.message;
            res->message_value = r64.message_value;
            res->converged = r64.converged;
            res->fallback_iterations = r64.iterations;
        }
    } else if (mixed) {
        const bool broke = h.message == MYCEL_PCG_CURVATURE_BREAKDOWN || h.message == MYCEL_PCG_STEP_BREAKDOWN ||
                           h.message == MYCEL_PCG_DIRECTION_BREAKDOWN || h.message == MYCEL_PCG_STAGNATION;
        if (broke && h.iterations < max_iter) {
            // The FP32 recurrences lost conjugacy (extreme stiffness contrast): drop the FP32 increment since
            // the last verified reliable update and finish in FP64 from that solution.
            mycel_pcg_result r64{};
            pcg_solve_device(op, d_b, d_x, tol, max_iter - h.iterations, MYCEL_PRECISION_FP64, cancel, cb, user,
                             &r64);
            res->iterations = h.iterations + r64.iterations;
            res->final_residual = r64.final_residual;
            res->message = r64.message;
            res->message_value = r64.message_value;
            res->converged = r64.converged;
            res->fallback_iterations = r64.iterations;
        } else {  // fold the last FP32 increment (zero after a converging update)
            hipLaunchKernelGGL(k_mixed_flush, gv, b256, 0, s, n, d_x, w->x32.get());
            MYCEL_LAUNCH_CHECK();
            MYCEL_HIP(hipStreamSynchronize(s));
        }
    }
    res->elapsed_ms = ms_since(t0);
}

// Whether a multigrid solve on op goes to its operator's pack broker.
bool solve_packed(const mycel_op *op, int precision, const volatile int32_t *cancel, mycel_pcg_progress_cb cb) {
    return op->broker && precision == MYCEL_PRECISION_MIXED_GMG && !cancel && !cb && fuse_step4() &&
           gmg_packable(op);
}

void broker_solve(SolveBroker *b, mycel_op *op, double *d_x);

}  // namespace

int32_t pcg_solve_device(mycel_op *op, const double *d_b, double *d_x, double tol, int max_iter, int precision,
                         const volatile int32_t *cancel, mycel_pcg_progress_cb cb, void *user,
                         mycel_pcg_result *res) {
    const auto t0 = SolveClock::now();
    *res = mycel_pcg_result{};
    if (precision == MYCEL_PRECISION_FP64_ORDERED)
        return pcg_solve_ordered(op, d_b, d_x, tol, max_iter, cancel, cb, user, res);
    if (precision != MYCEL_PRECISION_FP64 && precision != MYCEL_PRECISION_MIXED_FP32 &&
        precision != MYCEL_PRECISION_MIXED_BF16 && precision != MYCEL_PRECISION_MIXED_GMG)
        throw Error(MYCEL_E_ARGUMENT, "unknown precision mode " + std::to_string(precision));
    if (precision == MYCEL_PRECISION_MIXED_GMG && !gmg_supported(op)) {
        // A grid too thin for the hierarchy (an axis reaches one element above the dense coarsest size): the
        // reference's FP64 Jacobi iteration, reported as fallback iterations.
        const int32_t st = pcg_solve_device(op, d_b, d_x, tol, max_iter, MYCEL_PRECISION_FP64, cancel, cb, user, res);
        res->fallback_iterations = res->iterations;
        return st;
    }
    if (!solve_setup(op, d_b, d_x, tol, max_iter, precision, res, t0)) return MYCEL_OK;
    if (solve_packed(op, precision, cancel, cb))
        broker_solve(op->broker, op, d_x);  // the same iterations, in launches shared with other resident solves
    else
        solve_loop(op, d_b, d_x, max_iter, precision, cancel, cb, user);
    solve_finish(op, d_b, d_x, tol, max_iter, precision, cancel, cb, user, res, t0);
    return MYCEL_OK;
}

// ---- packed multigrid solves -------------------------------------------------------------------------
// The iterations of several resident multigrid solves (different grids: jobs of the server, or the systems of one
// mycel_pcg_solve_packed call) run together: per iteration one launch of each outer kernel and one per V-cycle
// phase for all members (k_pk_*, gmg_pack_enqueue), 4 iterations per batch, the members' states read once per
// batch. Members join and leave between batches (continuous batching); a member's arithmetic is its solve's alone.

constexpr int kPackMax = 16;  // members of one pack

struct PackTables {
    GmgPack *gp = nullptr;
    DevBuf<uint8_t> dev;      // outer-kernel tables and the members' state pointers
    uint8_t *host = nullptr;  // pinned staging of the same
    size_t host_cap = 0, bytes = 0;
    DevBuf<PcgState> d_states;
    PcgState *h_states = nullptr;  // pinned, one per slot
    int xr_grid = 0;
    size_t off_xr = 0, off_src = 0;
    // Step 4 launches, one per tile depth any member has brought (each a table of one entry per slot).
    struct Step4Group {
        int kz, grid;
        size_t off;
    };
    std::vector<Step4Group> s4;
    uint64_t shape = 0, gmg_shape = 0;  // the outer launches' shape, and the V-cycle pack's at the last build
    int launches_per_batch = 0;
    int slots = 0;  // member slots: grid dimension y of every launch

    explicit PackTables(int n_slots) : slots(n_slots) {
        if (slots < 1 || slots > kPackMax) throw Error(MYCEL_E_OUT_OF_RANGE, "pack: 1 to 16 slots");
        gp = gmg_pack_create(slots);
        d_states.resize(slots);
        MYCEL_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_states), sizeof(PcgState) * slots, hipHostMallocDefault));
    }
    ~PackTables() {
        gmg_pack_free(gp);
        if (host) (void)hipHostFree(host);
        if (h_states) (void)hipHostFree(h_states);
    }
    PackTables(const PackTables &) = delete;
    PackTables &operator=(const PackTables &) = delete;

    // Slot q (of `slots`) holds operator ops[q] (set up by solve_setup, its state in slot 0) with solution xs[q],
    // or nothing (ops[q] null). Returns true when the launch shape changed (a captured batch is stale).
    bool build(mycel_op *const *ops, double *const *xs, hipStream_t s) {
        MYCEL_HIP(hipStreamSynchronize(s));
        std::vector<GmgPackMember> gm(slots, GmgPackMember{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                                           nullptr});
        std::vector<PkXr> xr(slots, PkXr{});
        std::vector<std::pair<int, PkStep4>> st4(slots, {0, PkStep4{}});
        std::vector<PcgState *> src(slots, nullptr);
        bool grown = false;
        for (int q = 0; q < slots; q++) {
            mycel_op *op = ops[q];
            if (!op) continue;
            PcgWork *w = op->pcg;
            PcgState *st = w->state.get();
            gm[q] = GmgPackMember{op, w->r.get(), w->ap.get(), w->z.get(), w->part_zr.get(), w->part_zap.get(),
                                  &st->halt};
            xr[q] = PkXr{w->n, w->nb_mv, w->nb_rr, st, w->part_pap.get(), xs[q], w->r.get(), {w->p.get(), w->p2.get()},
                         w->ap.get(), w->part2.get(), gmg_fine_view(op)};
            if (w->nb_rr > xr_grid) {
                xr_grid = w->nb_rr;
                grown = true;
            }
            const int kz = ku_element_layers(op), nt = ec_tiles(op->d, kz);
            if (cdiv(nt, 4) != w->nb_mv) throw Error(MYCEL_E_INVALID_OPERATION, "pack: partial count mismatch");
            st4[q] = {kz, PkStep4{op->d, op->km64, op->d_ep.get(), op->d_flags.get(), {w->p.get(), w->p2.get()},
                                  w->z.get(), st, w->part2.get(), w->part_zr.get(), w->part_zap.get(), w->nb_rr,
                                  w->nb_z, w->ap.get(), w->part_pap.get(), nt, cdiv(nt, 4)}};
            auto g = std::find_if(s4.begin(), s4.end(), [&](const Step4Group &x) { return x.kz == kz; });
            if (g == s4.end()) {
                s4.push_back(Step4Group{kz, cdiv(nt, 4), 0});
                grown = true;
            } else if (cdiv(nt, 4) > g->grid) {
                g->grid = cdiv(nt, 4);
                grown = true;
            }
            src[q] = st;
        }
        gmg_pack_build(gp, gm.data(), slots, s);
        if (grown) {
            std::sort(s4.begin(), s4.end(), [](const Step4Group &a, const Step4Group &b) { return a.kz < b.kz; });
            size_t off = 0;
            auto place = [&](size_t n) {
                const size_t at = off;
                off = (off + n + 63) / 64 * 64;
                return at;
            };
            off_xr = place(sizeof(PkXr) * slots);
            off_src = place(sizeof(PcgState *) * slots);
            for (Step4Group &g : s4) g.off = place(sizeof(PkStep4) * slots);
            bytes = off;
            shape++;
        }
        if (bytes > host_cap) {
            if (host) (void)hipHostFree(host);
            host = nullptr;
            host_cap = 0;
            MYCEL_HIP(hipHostMalloc(reinterpret_cast<void **>(&host), bytes, hipHostMallocDefault));
            host_cap = bytes;
        }
        if (bytes > dev.size()) {
            dev.resize(bytes);
            shape++;
        }
        std::memset(host, 0, bytes);
        std::memcpy(host + off_xr, xr.data(), sizeof(PkXr) * slots);
        std::memcpy(host + off_src, src.data(), sizeof(PcgState *) * slots);
        for (int q = 0; q < slots; q++) {
            if (!ops[q]) continue;
            const auto g = std::find_if(s4.begin(), s4.end(), [&](const Step4Group &x) { return x.kz == st4[q].first; });
            std::memcpy(host + g->off + sizeof(PkStep4) * q, &st4[q].second, sizeof(PkStep4));
        }
        MYCEL_HIP(hipMemcpyAsync(dev.get(), host, bytes, hipMemcpyHostToDevice, s));
        launches_per_batch = kGmgGraphIters * (1 + gmg_pack_launches(gp) + static_cast<int>(s4.size()));
        const bool changed = grown || gmg_pack_shape(gp) != gmg_shape;
        gmg_shape = gmg_pack_shape(gp);
        return changed;
    }

    // One batch: kGmgGraphIters iterations of every member.
    void enqueue_batch(hipStream_t s) const {
        const dim3 b256(256);
        const PkXr *txr = reinterpret_cast<const PkXr *>(dev.get() + off_xr);
        for (int t = 0; t < kGmgGraphIters; t++) {
            const int slot = t & 1;
            hipLaunchKernelGGL(k_pk_fpcg_update_xr, dim3(xr_grid, slots), b256, 0, s, txr, slot);
            gmg_pack_enqueue(gp, s);
            for (const Step4Group &g : s4) {
                const PkStep4 *ts = reinterpret_cast<const PkStep4 *>(dev.get() + g.off);
                const dim3 grid(g.grid, slots);
                if (g.kz == 2) hipLaunchKernelGGL(k_pk_step4_ap<2>, grid, b256, 0, s, ts, slot);
                else if (g.kz == 3) hipLaunchKernelGGL(k_pk_step4_ap<3>, grid, b256, 0, s, ts, slot);
                else if (g.kz == 4) hipLaunchKernelGGL(k_pk_step4_ap<4>, grid, b256, 0, s, ts, slot);
                else if (g.kz == 8) hipLaunchKernelGGL(k_pk_step4_ap<8>, grid, b256, 0, s, ts, slot);
                else throw Error(MYCEL_E_INVALID_OPERATION, "pack: unsupported tile depth");
            }
        }
        MYCEL_LAUNCH_CHECK();
    }

    // The states of every slot after the batches enqueued so far, into h_states (synchronises s; an empty slot's
    // entry is left as it was).
    void read_states(hipStream_t s) const {
        hipLaunchKernelGGL(k_pk_gather_states, dim3(slots), dim3(64), 0, s,
                           reinterpret_cast<PcgState *const *>(dev.get() + off_src), d_states.get());
        MYCEL_LAUNCH_CHECK();
        MYCEL_HIP(hipMemcpyAsync(h_states, d_states.get(), sizeof(PcgState) * slots, hipMemcpyDeviceToHost, s));
        MYCEL_HIP(hipStreamSynchronize(s));
    }
};

struct PackReq {
    mycel_op *op;
    double *d_x;
    bool done = false;
    int32_t status = MYCEL_OK;
    std::string error;
    PcgState final_state{};
};

struct SolveBroker {
    int device = 0;
    int slots = 1;  // members of its pack (the server's workers)
    hipStream_t s = nullptr;
    std::mutex mu;
    std::condition_variable cv_in, cv_done;
    std::deque<PackReq *> incoming;
    bool stop = false;
    std::thread th;
    mycel_pack_stats stats{};  // under mu
};

namespace {

void broker_loop(SolveBroker *b) {
    std::vector<PackReq *> slot(b->slots, nullptr);  // a member keeps its slot until it leaves
    std::unique_ptr<PackTables> T;
    hipGraphExec_t graph = nullptr;
    bool dirty = true;
    auto drop_graph = [&] {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    };
    auto count = [&] { return static_cast<int>(std::count_if(slot.begin(), slot.end(), [](PackReq *r) { return r; })); };
    for (;;) {
        int members = 0;
        {
            std::unique_lock<std::mutex> lk(b->mu);
            b->cv_in.wait(lk, [&] { return b->stop || !b->incoming.empty() || count() > 0; });
            if (count() == 0 && b->incoming.empty()) break;  // stopping, nothing in flight
            for (int q = 0; q < b->slots && !b->incoming.empty(); q++)
                if (!slot[q]) {
                    slot[q] = b->incoming.front();
                    b->incoming.pop_front();
                    dirty = true;
                }
            members = count();
        }
        std::vector<PackReq *> finished;
        int32_t status = MYCEL_OK;
        std::string error;
        bool captured = false;
        int launches = 0;
        double t_build = 0.0, t_enqueue = 0.0, t_wait = 0.0;
        try {
            auto t0 = SolveClock::now();
            if (!T) {
                MYCEL_HIP(hipSetDevice(b->device));
                T.reset(new PackTables(b->slots));
            }
            if (dirty) {
                std::vector<mycel_op *> ops(b->slots, nullptr);
                std::vector<double *> xs(b->slots, nullptr);
                for (int q = 0; q < b->slots; q++)
                    if (slot[q]) {
                        ops[q] = slot[q]->op;
                        xs[q] = slot[q]->d_x;
                    }
                if (T->build(ops.data(), xs.data(), b->s)) drop_graph();
                dirty = false;
            }
            // The batch is a graph captured for the pack's shape; a new membership of the same shape only rewrites
            // the tables it reads.
            if (!graph) {
                graph = capture(b->s, [&] { T->enqueue_batch(b->s); });
                captured = true;
            }
            auto t1 = SolveClock::now();
            MYCEL_HIP(hipGraphLaunch(graph, b->s));
            launches = T->launches_per_batch;
            auto t2 = SolveClock::now();
            T->read_states(b->s);
            auto t3 = SolveClock::now();
            t_build = std::chrono::duration<double, std::milli>(t1 - t0).count();
            t_enqueue = std::chrono::duration<double, std::milli>(t2 - t1).count();
            t_wait = std::chrono::duration<double, std::milli>(t3 - t2).count();
            for (int q = 0; q < b->slots; q++)
                if (slot[q] && T->h_states[q].done) {
                    slot[q]->final_state = T->h_states[q];
                    finished.push_back(slot[q]);
                }
        } catch (const Error &e) {
            status = e.status;
            error = e.what();
            for (PackReq *r : slot)
                if (r) finished.push_back(r);
        } catch (const std::exception &e) {
            status = MYCEL_E_DEVICE;
            error = e.what();
            for (PackReq *r : slot)
                if (r) finished.push_back(r);
        }
        {
            std::lock_guard<std::mutex> lk(b->mu);
            if (status == MYCEL_OK) {
                b->stats.batches++;
                b->stats.member_batches += members;
                b->stats.launches += launches;
                if (!captured) b->stats.graph_batches++;
                b->stats.build_ms += t_build;
                b->stats.enqueue_ms += t_enqueue;
                b->stats.wait_ms += t_wait;
            }
            for (PackReq *r : finished) {
                r->status = status;
                r->error = error;
                r->done = true;
                *std::find(slot.begin(), slot.end(), r) = nullptr;
                b->stats.solves++;
                dirty = true;
            }
            if (!finished.empty()) b->stats.rebuilds++;
        }
        if (status != MYCEL_OK) {
            drop_graph();
            T.reset();  // rebuilt from scratch for the next members
        }
        if (!finished.empty()) b->cv_done.notify_all();
    }
    drop_graph();
    T.reset();
}

// Hands a solve (set up by solve_setup) to the broker and waits for it; leaves the final state in w->h_state.
void broker_solve(SolveBroker *b, mycel_op *op, double *d_x) {
    MYCEL_HIP(hipStreamSynchronize(op->stream));  // the set-up is complete before the broker's stream reads it
    PackReq r{op, d_x};
    {
        std::lock_guard<std::mutex> lk(b->mu);
        b->incoming.push_back(&r);
    }
    b->cv_in.notify_all();
    {
        std::unique_lock<std::mutex> lk(b->mu);
        b->cv_done.wait(lk, [&] { return r.done; });
    }
    if (r.status != MYCEL_OK) throw Error(r.status, r.error);
    *op->pcg->h_state = r.final_state;
}

}  // namespace

SolveBroker *broker_create(int device, int slots) {
    if (slots < 1 || slots > kPackMax) throw Error(MYCEL_E_OUT_OF_RANGE, "broker: 1 to 16 slots");
    std::unique_ptr<SolveBroker> b(new SolveBroker());
    b->device = device;
    b->slots = slots;
    MYCEL_HIP(hipSetDevice(device));
    MYCEL_HIP(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
    b->th = std::thread(broker_loop, b.get());
    return b.release();
}

void broker_destroy(SolveBroker *b) {
    if (!b) return;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        b->stop = true;
    }
    b->cv_in.notify_all();
    if (b->th.joinable()) b->th.join();
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
}

mycel_pack_stats broker_stats(SolveBroker *b) {
    std::lock_guard<std::mutex> lk(b->mu);
    return b->stats;
}

// Several independent systems on different operators, solved together in packed launches (the caller's thread
// drives the batches). Each system's result is bit-identical to its multigrid solve alone.
void pcg_solve_packed(mycel_op *const *ops, const double *const *d_b, double *const *d_x, int count, double tol,
                      int max_iter, mycel_pcg_result *res) {
    if (count < 1 || count > kPackMax) throw Error(MYCEL_E_OUT_OF_RANGE, "count must be in [1, 16]");
    const auto t0 = SolveClock::now();
    const int P = MYCEL_PRECISION_MIXED_GMG;
    std::vector<int> live;  // members that iterate (set-up did not end them)
    for (int i = 0; i < count; i++) {
        for (int j = 0; j < i; j++)
            if (ops[j] == ops[i]) throw Error(MYCEL_E_ARGUMENT, "each system needs its own operator handle");
        res[i] = mycel_pcg_result{};
        if (!gmg_supported(ops[i])) {  // as pcg_solve_device: the FP64 Jacobi iteration
            pcg_solve_device(ops[i], d_b[i], d_x[i], tol, max_iter, P, nullptr, nullptr, nullptr, &res[i]);
            continue;
        }
        if (!solve_setup(ops[i], d_b[i], d_x[i], tol, max_iter, P, &res[i], t0)) continue;
        if (!fuse_step4() || !gmg_packable(ops[i])) {
            solve_loop(ops[i], d_b[i], d_x[i], max_iter, P, nullptr, nullptr, nullptr);
            solve_finish(ops[i], d_b[i], d_x[i], tol, max_iter, P, nullptr, nullptr, nullptr, &res[i], t0);
            continue;
        }
        MYCEL_HIP(hipStreamSynchronize(ops[i]->stream));
        live.push_back(i);
    }
    if (live.empty()) return;
    hipStream_t s = ops[live[0]]->stream;
    const int n_slots = static_cast<int>(live.size());
    PackTables T(n_slots);
    std::vector<mycel_op *> slot_op(n_slots, nullptr);
    std::vector<double *> slot_x(n_slots, nullptr);
    std::vector<int> slot_sys(n_slots, -1);
    for (int k = 0; k < static_cast<int>(live.size()); k++) {  // system live[k] in slot k
        slot_op[k] = ops[live[k]];
        slot_x[k] = d_x[live[k]];
        slot_sys[k] = live[k];
    }
    hipGraphExec_t graph = nullptr;
    bool dirty = true;
    int left = static_cast<int>(live.size());
    const long long max_batches = static_cast<long long>(max_iter) + 64;
    try {
        for (long long batch = 0; left > 0; batch++) {
            if (batch > max_batches) throw Error(MYCEL_E_DEVICE, "packed PCG batch loop did not terminate");
            if (dirty) {
                if (T.build(slot_op.data(), slot_x.data(), s) && graph) {
                    MYCEL_HIP(hipGraphExecDestroy(graph));
                    graph = nullptr;
                }
                dirty = false;
            }
            if (!graph) graph = capture(s, [&] { T.enqueue_batch(s); });
            MYCEL_HIP(hipGraphLaunch(graph, s));
            T.read_states(s);
            for (int q = 0; q < n_slots; q++) {
                if (!slot_op[q] || !T.h_states[q].done) continue;
                const int i = slot_sys[q];
                *ops[i]->pcg->h_state = T.h_states[q];
                solve_finish(ops[i], d_b[i], d_x[i], tol, max_iter, P, nullptr, nullptr, nullptr, &res[i], t0);
                slot_op[q] = nullptr;
                slot_x[q] = nullptr;
                left--;
                dirty = true;
            }
        }
    } catch (...) {
        if (graph) (void)hipGraphExecDestroy(graph);
        throw;
    }
    if (graph) (void)hipGraphExecDestroy(graph);
}

int32_t pcg_solve_batch(mycel_op *op, const double *d_B, double *d_X, int R, double tol, int max_iter, int precision,
                        mycel_pcg_result *res) {
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    if (R < 1) throw Error(MYCEL_E_ARGUMENT, "at least one right-hand side is required");
    if (precision != MYCEL_PRECISION_MIXED_GMG)
        throw Error(MYCEL_E_ARGUMENT, "several right-hand sides in one solve need MYCEL_PRECISION_MIXED_GMG");
    if (!gmg_supported(op)) {  // no hierarchy on this grid: one FP64 Jacobi solve per right-hand side
        for (int k = 0; k < R; k++) {
            pcg_solve_device(op, d_B + static_cast<size_t>(k) * op->d.n_dofs, d_X + static_cast<size_t>(k) * op->d.n_dofs,
                             tol, max_iter, MYCEL_PRECISION_MIXED_GMG, nullptr, nullptr, nullptr, &res[k]);
        }
        return MYCEL_OK;
    }
    if (op->active_fixed < 6)  // PcgSolver.cs:95-100
        throw Error(MYCEL_E_INVALID_OPERATION,
                    "Model has fewer than 6 unique fixed DOFs on active nodes (found " +
                        std::to_string(op->active_fixed) +
                        "). At least 6 fixed DOFs are required to prevent rigid body modes in 3D.");
    for (int k = 0; k < R; k++) res[k] = mycel_pcg_result{};
    ensure_work(op);
    PcgWork *w = op->pcg;
    hipStream_t s = op->stream;
    gmg_prepare(op, s);
    ensure_gmg(op, w);
    gmg_reserve(op, R);
    ensure_batch(op, w, R);
    const int n = w->n;
    const long long nn = n;
    PcgState *st = w->bstate.get();
    const dim3 b256(256);
    const int dstride = static_cast<int>(sizeof(PcgState) / sizeof(double));
    const Batch bt_plain{R, nn, 0, 0};

    // The single-RHS set-up of pcg_solve_device, for every RHS at once (PcgSolver.cs:113-209): x[c] = 0,
    // ||Pb||^2, r = P(b - A x) (exactly Pb for a zero start), z = M r, p = z, rz.
    MYCEL_HIP(hipMemsetAsync(st, 0, sizeof(PcgState) * R, s));
    hipLaunchKernelGGL(k_pcg_init, dim3(w->nb_vec, R), b256, 0, s, n, op->d_flags.get(), d_B, d_X, st,
                       w->bpart1.get());
    hipLaunchKernelGGL(k_pcg_sum_to, dim3(1, R), b256, 0, s, w->bpart1.get(), w->nb_vec, &st->bnorm_sq, dstride);
    launch_apply_k(op, d_X, w->bap.get(), kApplyAProjected, s, nullptr, nullptr, bt_plain);
    hipLaunchKernelGGL(k_pcg_init_r, dim3(w->nb_vec, R), b256, 0, s, n, op->d_flags.get(), d_B, w->bap.get(),
                       w->br.get(), w->bpart1.get());
    hipLaunchKernelGGL(k_pcg_sum_to, dim3(1, R), b256, 0, s, w->bpart1.get(), w->nb_vec, &st->r_norm_sq, dstride);
    gmg_enqueue_vcycle(op, w->br.get(), w->br.get(), w->bz.get(), w->bpart_zr.get(), w->bpart_zap.get(), s, false, R);
    hipLaunchKernelGGL(k_copy, dim3(cdiv(n, 256), R), b256, 0, s, n, w->bz.get(), w->bp.get());
    hipLaunchKernelGGL(k_pcg_sum_to, dim3(1, R), b256, 0, s, w->bpart_zr.get(), w->nb_z, &st->rz0, dstride);
    MYCEL_LAUNCH_CHECK();
    MYCEL_HIP(hipMemcpyAsync(w->h_bstate, st, sizeof(PcgState) * R, hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipStreamSynchronize(s));
    int active = 0;
    for (int k = 0; k < R; k++) {
        const PcgState &h = w->h_bstate[k];
        const double bnorm = std::sqrt(h.bnorm_sq);
        if (bnorm <= 1e-30) {  // :139-151
            MYCEL_HIP(hipMemsetAsync(d_X + k * nn, 0, sizeof(double) * n, s));
            hipLaunchKernelGGL(k_pcg_halt_at_start, dim3(1), dim3(1), 0, s, st + k, MYCEL_PCG_RHS_ZERO, 0.0);
            continue;
        }
        const double rr0 = h.nonzero_x ? h.r_norm_sq : h.bnorm_sq;
        const double rel0 = std::sqrt(rr0) / bnorm;
        if (rel0 <= tol) {
            hipLaunchKernelGGL(k_pcg_halt_at_start, dim3(1), dim3(1), 0, s, st + k, MYCEL_PCG_INITIAL_CONVERGED, rel0);
        } else if (max_iter <= 0) {
            hipLaunchKernelGGL(k_pcg_halt_at_start, dim3(1), dim3(1), 0, s, st + k, MYCEL_PCG_MAX_ITERATIONS, rel0);
        } else {
            hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(1), 0, s, st + k, tol, bnorm, rel0, max_iter, 0,
                               kReliableDelta);
            active++;
        }
    }
    MYCEL_LAUNCH_CHECK();
    if (active) {
        if (w->graph_b && (w->graph_b_x != d_X || w->graph_b_h != op->gmg || w->graph_b_gen != gmg_generation(op) ||
                           w->graph_b_R != R)) {
            MYCEL_HIP(hipGraphExecDestroy(w->graph_b));
            w->graph_b = nullptr;
        }
        if (!w->graph_b) {
            w->graph_b = capture(s, [&] {
                for (int t = 0; t < kGmgGraphIters; t++) enqueue_iteration_gmg_batch(op, w, d_X, R, t & 1, s);
            });
            w->graph_b_x = d_X;
            w->graph_b_h = op->gmg;
            w->graph_b_gen = gmg_generation(op);
            w->graph_b_R = R;
        }
        const long long max_batches = static_cast<long long>(max_iter) + 64;
        for (long long batch = 0;; batch++) {
            if (batch > max_batches) throw Error(MYCEL_E_DEVICE, "PCG batch loop did not terminate");
            MYCEL_HIP(hipGraphLaunch(w->graph_b, s));
            MYCEL_HIP(hipMemcpyAsync(w->h_bstate, st, sizeof(PcgState) * R, hipMemcpyDeviceToHost, s));
            MYCEL_HIP(hipStreamSynchronize(s));
            bool all = true;
            for (int k = 0; k < R; k++) all = all && w->h_bstate[k].done;
            if (all) break;
        }
    } else {
        MYCEL_HIP(hipMemcpyAsync(w->h_bstate, st, sizeof(PcgState) * R, hipMemcpyDeviceToHost, s));
        MYCEL_HIP(hipStreamSynchronize(s));
    }
    for (int k = 0; k < R; k++) {
        const PcgState h = w->h_bstate[k];
        mycel_pcg_result &r = res[k];
        r.iterations = h.iterations;
        r.final_residual = h.final_rel;
        r.message = h.message;
        r.message_value = h.value;
        r.converged = h.message == MYCEL_PCG_CONVERGED || h.message == MYCEL_PCG_RHS_ZERO ||
                      h.message == MYCEL_PCG_INITIAL_CONVERGED;
        const bool broke = h.message == MYCEL_PCG_CURVATURE_BREAKDOWN || h.message == MYCEL_PCG_STEP_BREAKDOWN ||
                           h.message == MYCEL_PCG_DIRECTION_BREAKDOWN || h.message == MYCEL_PCG_STAGNATION;
        if (broke && h.iterations < max_iter) {  // as the single-RHS multigrid solve: finish in FP64 Jacobi
            mycel_pcg_result r64{};
            pcg_solve_device(op, d_B + k * nn, d_X + k * nn, tol, max_iter - h.iterations, MYCEL_PRECISION_FP64,
                             nullptr, nullptr, nullptr, &r64);
            r.iterations = h.iterations + r64.iterations;
            r.final_residual = r64.final_residual;
            r.message = r64.message;
            r.message_value = r64.message_value;
            r.converged = r64.converged;
            r.fallback_iterations = r64.iterations;
        }
    }
    const double ms = elapsed();
    for (int k = 0; k < R; k++) res[k].elapsed_ms = ms;
    return MYCEL_OK;
}

}  // namespace mycel
