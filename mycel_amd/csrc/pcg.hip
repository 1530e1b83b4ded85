// Device-resident Jacobi-preconditioned CG on the projected operator A = PKP + (I - P)
// (Fem/PcgSolver.cs:80-553, Fem/JacobiPreconditioner.cs:15-35).
//
// One iteration is three kernels with no host involvement:
//   1. ap = A p fused with per-block partials of p.ap                       (node-centric K*u)
//   2. alpha = rz / p.ap (breakdown checks); x += alpha p; r -= alpha ap; z = r / diag;
//      per-block partials of r.r and r.z
//   3. relres, convergence, stagnation, beta = rz' / rz (breakdown check); p = z + beta p
// Every block of kernels 2 and 3 re-sums the previous kernel's partials in the same fixed order, so
// all blocks agree bit-for-bit on alpha, beta and the stopping decision without a grid barrier.
// The scalar state lives on the device in two slots indexed by iteration parity: iteration t reads
// slot (t-1)&1 and writes slot t&1, so a batch of an even number of iterations is a fixed sequence
// of launches that is captured once into a hipGraph and replayed. After termination every kernel
// returns at entry, so the host polls the state once per batch, not once per iteration.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>

#include "device_common.hpp"
#include "mycel_internal.hpp"

namespace mycel {

struct PcgState {
    double rz[2];        // rho0 at the start of the iteration in this slot
    double prev_rel[2];  // relres of the previous iteration (stagnation test, breakdown report)
    int stag[2];
    int iter[2];         // iteration number t that reads this slot
    double tol, bnorm;
    double bnorm_sq, r_norm_sq, rz0;
    int done, message, iterations, nonzero_x;
    double final_rel;
    double last_rel;
    int last_iter, pad;
};

struct PcgWork {
    int n = 0;
    int nb_mv = 0, nb_vec = 0;
    DevBuf<double> r, z, p, ap;
    DevBuf<double> part_pap, part2, part1;
    DevBuf<PcgState> state;
    PcgState *h_state = nullptr;  // pinned
    hipGraphExec_t graph = nullptr;
    int graph_iters = 0;
    const double *graph_x = nullptr;  // solution vector the graph was captured with
};

namespace {

constexpr int kVecPerBlock = 1024;  // DOFs per block in the vector kernels (4 per thread)

__global__ __launch_bounds__(256) void k_pcg_update_xr(int n, int slot, PcgState *__restrict__ st,
                                                       const double *__restrict__ part_pap, int nb_mv,
                                                       double *__restrict__ x, double *__restrict__ r,
                                                       const double *__restrict__ p, const double *__restrict__ ap,
                                                       double *__restrict__ z, const double *__restrict__ inv,
                                                       double *__restrict__ part2) {
    __shared__ double red[4];
    if (st->done) return;
    const double pap = sum_partials256(part_pap, nb_mv, red);
    const double rz0 = st->rz[slot];
    const bool curv = !(pap > 1e-30) || !isfinite(pap);  // PcgSolver.cs:245
    const double alpha = rz0 / pap;
    const bool step = !curv && (!isfinite(alpha) || alpha <= 0.0);  // :260
    if (curv || step) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->message = curv ? MYCEL_PCG_CURVATURE_BREAKDOWN : MYCEL_PCG_STEP_BREAKDOWN;
            st->iterations = st->iter[slot];
            st->final_rel = st->prev_rel[slot];
            st->done = 1;
        }
        return;
    }
    double rr = 0.0, rz = 0.0;
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) {
            x[i] = __fma_rn(alpha, p[i], x[i]);
            const double ri = __fma_rn(-alpha, ap[i], r[i]);
            r[i] = ri;
            const double zi = ri * inv[i];
            z[i] = zi;
            rr = __fma_rn(ri, ri, rr);
            rz = __fma_rn(ri, zi, rz);
        }
    }
    rr = block_sum256(rr, red);
    rz = block_sum256(rz, red);
    if (threadIdx.x == 0) {
        part2[2 * blockIdx.x] = rr;
        part2[2 * blockIdx.x + 1] = rz;
    }
}

__device__ __forceinline__ double sum_strided2(const double *__restrict__ part, int n, int which, double *lds4) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) v += part[2 * i + which];
    return block_sum256(v, lds4);
}

__global__ __launch_bounds__(256) void k_pcg_update_p(int n, int slot, PcgState *__restrict__ st,
                                                      const double *__restrict__ part2, int nb_vec,
                                                      double *__restrict__ p, const double *__restrict__ z) {
    __shared__ double red[4];
    if (st->done) return;
    const double rr = sum_strided2(part2, nb_vec, 0, red);
    const double rzn = sum_strided2(part2, nb_vec, 1, red);
    const double rel = sqrt(rr) / st->bnorm;
    const int t = st->iter[slot];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (rel <= st->tol) {  // PcgSolver.cs:279
        if (lead) {
            st->message = MYCEL_PCG_CONVERGED;
            st->iterations = t;
            st->final_rel = rel;
            st->done = 1;
        }
        return;
    }
    const int stag = (fabs(rel - st->prev_rel[slot]) < 1e-15 * rel) ? st->stag[slot] + 1 : 0;  // :293-313
    if (stag > 50) {
        if (lead) {
            st->message = MYCEL_PCG_STAGNATION;
            st->iterations = t;
            st->final_rel = rel;
            st->done = 1;
        }
        return;
    }
    const double beta = rzn / st->rz[slot];
    if (!isfinite(beta) || beta < 0.0) {  // :334
        if (lead) {
            st->message = MYCEL_PCG_DIRECTION_BREAKDOWN;
            st->iterations = t;
            st->final_rel = rel;
            st->done = 1;
        }
        return;
    }
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) p[i] = __fma_rn(beta, p[i], z[i]);
    }
    if (lead) {
        const int o = slot ^ 1;
        st->rz[o] = rzn;
        st->prev_rel[o] = rel;
        st->stag[o] = stag;
        st->iter[o] = t + 1;
        st->last_rel = rel;
        st->last_iter = t;
    }
}

// x[constrained] = 0; flag a nonzero projected x; partials of ||Pb||^2 (PcgSolver.cs:113-137).
__global__ __launch_bounds__(256) void k_pcg_init(int n, const uint8_t *__restrict__ flags,
                                                  const double *__restrict__ b, double *__restrict__ x,
                                                  PcgState *__restrict__ st, double *__restrict__ part1) {
    __shared__ double red[4];
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

// z = r / diag; p = z; partials of r.z.
__global__ __launch_bounds__(256) void k_pcg_init_z(int n, const double *__restrict__ r,
                                                    const double *__restrict__ inv, double *__restrict__ z,
                                                    double *__restrict__ p, double *__restrict__ part1) {
    __shared__ double red[4];
    double v = 0.0;
    const int base = blockIdx.x * kVecPerBlock + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) {
            const double zi = r[i] * inv[i];
            z[i] = zi;
            p[i] = zi;
            v = __fma_rn(r[i], zi, v);
        }
    }
    v = block_sum256(v, red);
    if (threadIdx.x == 0) part1[blockIdx.x] = v;
}

// Sums the partials into one state field (single block).
__global__ __launch_bounds__(256) void k_pcg_sum_to(const double *__restrict__ part, int nb, double *__restrict__ out) {
    __shared__ double red[4];
    const double s = sum_partials256(part, nb, red);
    if (threadIdx.x == 0) *out = s;
}

__global__ void k_pcg_start(PcgState *__restrict__ st, double tol, double bnorm, double rel) {
    st->tol = tol;
    st->bnorm = bnorm;
    st->rz[0] = st->rz0;
    st->prev_rel[0] = rel;
    st->stag[0] = 0;
    st->iter[0] = 1;
    st->done = 0;
    st->last_rel = rel;
    st->last_iter = 0;
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

void ensure_work(mycel_op *op) {
    if (op->pcg) return;
    PcgWork *w = new PcgWork();
    op->pcg = w;
    w->n = op->d.n_dofs;
    w->nb_mv = apply_k_blocks(op);
    w->nb_vec = cdiv(w->n, kVecPerBlock);
    w->r.resize(w->n);
    w->z.resize(w->n);
    w->p.resize(w->n);
    w->ap.resize(w->n);
    w->part_pap.resize(w->nb_mv);
    w->part2.resize(2 * w->nb_vec);
    w->part1.resize(w->nb_vec);
    w->state.resize(1);
    MYCEL_HIP(hipHostMalloc(reinterpret_cast<void **>(&w->h_state), sizeof(PcgState), hipHostMallocDefault));
}

void enqueue_iteration(mycel_op *op, PcgWork *w, double *d_x, int slot, hipStream_t s) {
    PcgState *st = w->state.get();
    launch_apply_k(op, w->p.get(), w->ap.get(), kApplyAProjected, s, w->part_pap.get(), &st->done);
    hipLaunchKernelGGL(k_pcg_update_xr, dim3(w->nb_vec), dim3(256), 0, s, w->n, slot, st, w->part_pap.get(),
                       w->nb_mv, d_x, w->r.get(), w->p.get(), w->ap.get(), w->z.get(), op->d_inv_diag.get(),
                       w->part2.get());
    hipLaunchKernelGGL(k_pcg_update_p, dim3(w->nb_vec), dim3(256), 0, s, w->n, slot, st, w->part2.get(), w->nb_vec,
                       w->p.get(), w->z.get());
}

void read_state(PcgWork *w, hipStream_t s) {
    MYCEL_HIP(hipMemcpyAsync(w->h_state, w->state.get(), sizeof(PcgState), hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipStreamSynchronize(s));
}

constexpr int kGraphIters = 32;  // iterations per captured batch (even: slot parity repeats)

}  // namespace

void pcg_free(PcgWork *w) {
    if (!w) return;
    if (w->graph) (void)hipGraphExecDestroy(w->graph);
    if (w->h_state) (void)hipHostFree(w->h_state);
    delete w;
}

int32_t pcg_solve_fp64(mycel_op *op, const double *d_b, double *d_x, double tol, int max_iter,
                       const volatile int32_t *cancel, mycel_pcg_progress_cb cb, void *user, mycel_pcg_result *res) {
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
    ensure_work(op);
    PcgWork *w = op->pcg;
    hipStream_t s = op->stream;
    PcgState *st = w->state.get();
    const int n = w->n;

    MYCEL_HIP(hipMemsetAsync(st, 0, sizeof(PcgState), s));
    hipLaunchKernelGGL(k_pcg_init, dim3(w->nb_vec), dim3(256), 0, s, n, op->d_flags.get(), d_b, d_x, st,
                       w->part1.get());
    hipLaunchKernelGGL(k_pcg_sum_to, dim3(1), dim3(256), 0, s, w->part1.get(), w->nb_vec, &st->bnorm_sq);
    MYCEL_LAUNCH_CHECK();
    read_state(w, s);
    const double bnorm = std::sqrt(w->h_state->bnorm_sq);
    if (bnorm <= 1e-30) {  // PcgSolver.cs:139-151
        MYCEL_HIP(hipMemsetAsync(d_x, 0, sizeof(double) * n, s));
        MYCEL_HIP(hipStreamSynchronize(s));
        res->iterations = 0;
        res->final_residual = 0.0;
        res->converged = 1;
        res->message = MYCEL_PCG_RHS_ZERO;
        res->elapsed_ms = elapsed();
        return MYCEL_OK;
    }
    const bool warm = w->h_state->nonzero_x != 0;
    if (warm) launch_apply_k(op, d_x, w->ap.get(), kApplyAProjected, s);
    hipLaunchKernelGGL(k_pcg_init_r, dim3(w->nb_vec), dim3(256), 0, s, n, op->d_flags.get(), d_b,
                       warm ? w->ap.get() : nullptr, w->r.get(), w->part1.get());
    hipLaunchKernelGGL(k_pcg_sum_to, dim3(1), dim3(256), 0, s, w->part1.get(), w->nb_vec, &st->r_norm_sq);
    hipLaunchKernelGGL(k_pcg_init_z, dim3(w->nb_vec), dim3(256), 0, s, n, w->r.get(), op->d_inv_diag.get(),
                       w->z.get(), w->p.get(), w->part1.get());
    hipLaunchKernelGGL(k_pcg_sum_to, dim3(1), dim3(256), 0, s, w->part1.get(), w->nb_vec, &st->rz0);
    MYCEL_LAUNCH_CHECK();
    read_state(w, s);
    // A cold start uses ||Pb||^2 itself as ||r||^2 (PcgSolver.cs:185-192).
    const double rr0 = warm ? w->h_state->r_norm_sq : w->h_state->bnorm_sq;
    const double rel0 = std::sqrt(rr0) / bnorm;
    if (rel0 <= tol) {  // :197-209
        res->iterations = 0;
        res->final_residual = rel0;
        res->converged = 1;
        res->message = MYCEL_PCG_INITIAL_CONVERGED;
        res->elapsed_ms = elapsed();
        return MYCEL_OK;
    }
    hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(1), 0, s, st, tol, bnorm, rel0);
    MYCEL_LAUNCH_CHECK();

    // The captured batch bakes in the solution pointer; recapture if the caller's x moved.
    if (w->graph && w->graph_x != d_x) {
        MYCEL_HIP(hipGraphExecDestroy(w->graph));
        w->graph = nullptr;
    }
    if (!w->graph) {
        hipGraph_t g;
        MYCEL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int t = 0; t < kGraphIters; t++) enqueue_iteration(op, w, d_x, t & 1, s);
        MYCEL_HIP(hipStreamEndCapture(s, &g));
        MYCEL_HIP(hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0));
        MYCEL_HIP(hipGraphDestroy(g));
        w->graph_iters = kGraphIters;
        w->graph_x = d_x;
    }

    int launched = 0;
    bool done = false;
    while (launched < max_iter) {
        if (cancel && *cancel) {
            MYCEL_HIP(hipStreamSynchronize(s));
            res->message = MYCEL_PCG_CANCELED;
            res->elapsed_ms = elapsed();
            throw Error(MYCEL_E_CANCELED, "The operation was canceled.");
        }
        const int remaining = max_iter - launched;
        if (remaining >= w->graph_iters) {
            MYCEL_HIP(hipGraphLaunch(w->graph, s));
            launched += w->graph_iters;
        } else {
            for (int t = 0; t < remaining; t++) enqueue_iteration(op, w, d_x, (launched + t) & 1, s);
            MYCEL_LAUNCH_CHECK();
            launched = max_iter;
        }
        read_state(w, s);
        if (cb) cb(w->h_state->done ? w->h_state->iterations : w->h_state->last_iter,
                   w->h_state->done ? w->h_state->final_rel : w->h_state->last_rel, user);
        if (w->h_state->done) {
            done = true;
            break;
        }
    }
    const PcgState &h = *w->h_state;
    if (done) {
        res->iterations = h.iterations;
        res->final_residual = h.final_rel;
        res->message = h.message;
        res->converged = h.message == MYCEL_PCG_CONVERGED;
    } else {  // PcgSolver.cs:359-368
        res->iterations = max_iter;
        res->final_residual = h.last_rel;
        res->message = MYCEL_PCG_MAX_ITERATIONS;
        res->converged = 0;
    }
    res->elapsed_ms = elapsed();
    return MYCEL_OK;
}

int32_t pcg_solve_device(mycel_op *op, const double *d_b, double *d_x, double tol, int max_iter, int precision,
                         const volatile int32_t *cancel, mycel_pcg_progress_cb cb, void *user,
                         mycel_pcg_result *res) {
    if (precision != MYCEL_PRECISION_FP64)
        throw Error(MYCEL_E_ARGUMENT, "precision mode not available in this build");
    return pcg_solve_fp64(op, d_b, d_x, tol, max_iter, cancel, cb, user, res);
}

}  // namespace mycel
