// SIMP optimisation kernels on gfx950: linear cone density filter (forward and adjoint),
// Optimality Criteria bisection, and the per-iteration helpers of the SIMP loop.
//
// Filter (Optimization/DensityFilter.cs:34-303). The reference stores W and W^T as CSR. On a
// structured grid the weight of a tap depends only on its offset, H = rmin - |d|, and on the row
// normaliser 1/Hs_e, so the GPU keeps a table of offsets (ordered dz, dy, dx ascending = the
// reference's jz, jy, jx loop order) plus one 1/Hs per active element. For each element the taps
// are visited in that order with the same products (H * 1/Hs) and the same sequential sum, so the
// forward filter, the adjoint and S are bit-identical to the reference's CSR evaluation. The
// adjoint row of element i lists the rows r that contain i in ascending r, which is again the
// ascending box order of the neighbours of i.
//
// OC (Optimization/OcOptimizer.cs:12-175). The bracket expansion and log-bisection form a state
// machine that runs on the device: every launch sums the partial volumes of the previous
// evaluation, advances the machine identically in every block, and evaluates the volume at the next
// lambda. The host only polls for completion once per batch of launches.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "mycel_internal.hpp"
#include "mycel_opt.hpp"

namespace mycel {

namespace {
inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ---- filter ------------------------------------------------------------------------------------
struct FilterGeom {
    int nx, ny, nz, n_active, n_taps;
};

__global__ __launch_bounds__(256) void k_filter_inv_hs(FilterGeom g, const int4 *__restrict__ taps,
                                                       const double *__restrict__ tap_h,
                                                       const int32_t *__restrict__ active,
                                                       const int32_t *__restrict__ to_active,
                                                       double *__restrict__ inv_hs) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n_active) return;
    const int e = active[i];
    const int ex = e % g.nx, t = e / g.nx, ey = t % g.ny, ez = t / g.ny;
    double hs = 0.0;
    for (int k = 0; k < g.n_taps; k++) {
        const int4 d = taps[k];
        const int jx = ex + d.x, jy = ey + d.y, jz = ez + d.z;
        if (jx < 0 || jy < 0 || jz < 0 || jx >= g.nx || jy >= g.ny || jz >= g.nz) continue;
        if (to_active[jx + g.nx * (jy + g.ny * jz)] < 0) continue;
        hs += tap_h[k];
    }
    inv_hs[i] = hs > 0.0 ? 1.0 / hs : 1.0;  // DensityFilter.cs:121
}

// ADJOINT = false: out_i = sum_k (H_k / Hs_i) in_nbr ; ADJOINT = true: out_i = sum_k (H_k / Hs_nbr) in_nbr.
template <bool ADJOINT>
__global__ __launch_bounds__(256) void k_filter_apply(FilterGeom g, const int4 *__restrict__ taps,
                                                      const double *__restrict__ tap_h,
                                                      const int32_t *__restrict__ active,
                                                      const int32_t *__restrict__ to_active,
                                                      const double *__restrict__ inv_hs,
                                                      const double *__restrict__ in, double *__restrict__ out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= g.n_active) return;
    const int e = active[i];
    const int ex = e % g.nx, t = e / g.nx, ey = t % g.ny, ez = t / g.ny;
    const double own = ADJOINT ? 0.0 : inv_hs[i];
    double sum = 0.0;
    for (int k = 0; k < g.n_taps; k++) {
        const int4 d = taps[k];
        const int jx = ex + d.x, jy = ey + d.y, jz = ez + d.z;
        if (jx < 0 || jy < 0 || jz < 0 || jx >= g.nx || jy >= g.ny || jz >= g.nz) continue;
        const int j = to_active[jx + g.nx * (jy + g.ny * jz)];
        if (j < 0) continue;
        const double w = tap_h[k] * (ADJOINT ? inv_hs[j] : own);
        sum += w * (in ? in[j] : 1.0);
    }
    out[i] = sum;
}

// ---- OC ----------------------------------------------------------------------------------------
enum : int { kOcPendNone = 0, kOcPendInit = 1, kOcPendLower = 2, kOcPendUpper = 3, kOcPendMid = 4 };
enum : int { kOcPhaseLower = 1, kOcPhaseUpper = 2, kOcPhaseBisect = 3, kOcPhaseDone = 4 };

__device__ __forceinline__ double oc_trial(double r, double dc, double si, double lambda, double move,
                                           double rho_min, double *vol_term) {
#pragma clang fp contract(off)
    const double lo = fmax(rho_min, r - move), hi = fmin(1.0, r + move);
    const double neg = -dc;
    double v;
    if (neg <= 0.0 || si <= 0.0) {
        v = lo;
    } else {
        const double b = neg / (lambda * si);
        const double tr = r * sqrt(b);
        v = tr < lo ? lo : (tr > hi ? hi : tr);  // Math.Clamp
    }
    *vol_term = si * v;
    return v;
}

struct OcArgs {
    const double *rho, *dc, *s;
    const uint8_t *passive;
    double *rho_new;
    int n;
    double move, rho_min, vol_tol;
    int max_bisect;
};

__device__ void oc_eval_partials(const OcArgs &a, double l0, double l1, bool two, double *part, double *red) {
    double v0 = 0.0, v1 = 0.0;
    const int base = blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < a.n) {
            if (a.passive && a.passive[i]) {
                v0 += a.s[i];
                v1 += a.s[i];
            } else {
                double t;
                oc_trial(a.rho[i], a.dc[i], a.s[i], l0, a.move, a.rho_min, &t);
                v0 += t;
                if (two) {
                    oc_trial(a.rho[i], a.dc[i], a.s[i], l1, a.move, a.rho_min, &t);
                    v1 += t;
                }
            }
        }
    }
    v0 = block_sum256(v0, red);
    if (two) v1 = block_sum256(v1, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = v0;
        part[2 * blockIdx.x + 1] = v1;
    }
}

__global__ __launch_bounds__(256) void k_oc_sum_s(const double *__restrict__ s, int n, double *__restrict__ part) {
    __shared__ double red[4];
    double v = 0.0;
    const int base = blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < n) v += s[i];
    }
    v = block_sum256(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Launch k reads state slot k&1 and partials buffer k&1; writes slot (k+1)&1.
__global__ __launch_bounds__(256) void k_oc_step(OcArgs a, int slot, OcState *__restrict__ states,
                                                 double *__restrict__ parts, int nb, OcResult *__restrict__ res) {
    __shared__ double red[4];
    OcState S = states[slot];
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (S.phase == kOcPhaseDone) {  // finished in an earlier launch: carry the state forward
        if (lead) states[slot ^ 1] = S;
        return;
    }
    const double *pin = parts + (size_t)slot * 2 * nb;
    double *pout = parts + (size_t)(slot ^ 1) * 2 * nb;
    double sum0 = 0.0, sum1 = 0.0;
    if (S.pending != kOcPendNone) {
        double v = 0.0, w = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) {
            v += pin[2 * b];
            w += pin[2 * b + 1];
        }
        sum0 = block_sum256(v, red);
        sum1 = block_sum256(w, red);
    }
    if (S.pending == kOcPendInit) {
        S.v1 = sum0;
        S.v2 = sum1;
    } else if (S.pending == kOcPendLower) {
        S.v1 = sum0;
    } else if (S.pending == kOcPendUpper) {
        S.v2 = sum0;
    }
    const int pending_was = S.pending;
    S.pending = kOcPendNone;
    double eval_l = 0.0;
    // OcOptimizer.cs:51-100, advanced until the next volume evaluation is needed.
    if (S.phase == kOcPhaseLower) {
        if (S.v1 < S.target && S.l1 > 1e-50 && S.steps++ < 50) {
            S.l2 = S.l1;
            S.v2 = S.v1;
            S.l1 *= 0.1;
            S.pending = kOcPendLower;
            eval_l = S.l1;
        } else {
            S.phase = kOcPhaseUpper;
            S.steps = 0;
        }
    }
    if (S.pending == kOcPendNone && S.phase == kOcPhaseUpper) {
        if (S.v2 > S.target && S.l2 < 1e50 && S.steps++ < 50) {
            S.l1 = S.l2;
            S.v1 = S.v2;
            S.l2 *= 10.0;
            S.pending = kOcPendUpper;
            eval_l = S.l2;
        } else {
            S.phase = kOcPhaseBisect;
            S.bis = 0;
            S.lmid = sqrt(S.l1 * S.l2);
        }
    }
    if (S.pending == kOcPendNone && S.phase == kOcPhaseBisect) {
        bool stop = false;
        if (pending_was == kOcPendMid) {
            const double vmid = sum0;
            const double rel = fabs(vmid - S.target) / fmax(S.target, 1e-12);
            if (rel <= a.vol_tol) {
                stop = true;
            } else {
                if (vmid > S.target) S.l1 = S.lmid;
                else S.l2 = S.lmid;
                if (fabs(S.l2 - S.l1) <= 1e-15 * S.lmid) stop = true;
            }
            S.bis++;
        }
        if (!stop && S.bis < a.max_bisect) {
            S.lmid = sqrt(S.l1 * S.l2);
            S.pending = kOcPendMid;
            eval_l = S.lmid;
        } else {
            S.phase = kOcPhaseDone;
        }
    }
    if (S.phase == kOcPhaseDone) {
        // Final densities with lMid (OcOptimizer.cs:102-129); rho_new may alias rho.
        const int base = blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int i = base + 256 * q;
            if (i < a.n) {
                if (a.passive && a.passive[i]) {
                    a.rho_new[i] = 1.0;
                } else {
                    double t;
                    a.rho_new[i] = oc_trial(a.rho[i], a.dc[i], a.s[i], S.lmid, a.move, a.rho_min, &t);
                }
            }
        }
        if (lead) {
            states[slot ^ 1] = S;
            res->lambda = S.lmid;
            res->evals = S.evals;
            res->done = 1;
        }
        return;
    }
    oc_eval_partials(a, eval_l, 0.0, false, pout, red);
    if (lead) {
        S.evals++;
        states[slot ^ 1] = S;
    }
}

// First launch: target volume, initial bracket (OcOptimizer.cs:29-49), volumes at l1 and l2.
__global__ __launch_bounds__(256) void k_oc_begin(OcArgs a, const double *__restrict__ spart, int nbs, double vf,
                                                  double lambda_guess, OcState *__restrict__ states,
                                                  double *__restrict__ parts, OcResult *__restrict__ res) {
    __shared__ double red[4];
    const double total_s = sum_partials256(spart, nbs, red);
    OcState S{};
    S.target = vf * total_s;
    const bool good = lambda_guess > 0.0 && !isinf(lambda_guess) && !isnan(lambda_guess);
    S.l1 = good ? lambda_guess * 0.5 : 1e-12;
    S.l2 = good ? lambda_guess * 2.0 : 1e12;
    if (S.l1 <= 1e-50) S.l1 = 1e-12;
    if (S.l2 <= S.l1) S.l2 = S.l1 * 10.0;
    S.phase = kOcPhaseLower;
    S.steps = 0;
    S.pending = kOcPendInit;
    S.evals = 2;
    oc_eval_partials(a, S.l1, S.l2, true, parts, red);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        states[0] = S;
        res->done = 0;
    }
}

// ---- SIMP helpers --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter_density(int na, const int32_t *__restrict__ active,
                                                         const double *__restrict__ rho_bar,
                                                         double *__restrict__ density_box) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < na) density_box[active[i]] = rho_bar[i];
}

__global__ __launch_bounds__(256) void k_force_passive(int na, const uint8_t *__restrict__ passive,
                                                       double *__restrict__ rho) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < na && passive[i]) rho[i] = 1.0;
}

__global__ __launch_bounds__(256) void k_project(int n, const uint8_t *__restrict__ flags, const double *__restrict__ f,
                                                 double *__restrict__ b) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = (flags[i] & kFlagConstrained) ? 0.0 : f[i];
}

// Per-block partials: max|rho - rho_old|, sum rho_bar, sum rho_bar (1 - rho_bar) (TopologyOptimizer.cs:323-338).
// grey_src: the field the grey level is measured on (rho_bar, or rho_tilde under Heaviside projection,
// TopologyOptimizer.cs:333).
__global__ __launch_bounds__(256) void k_simp_metrics(int na, const double *__restrict__ rho,
                                                      const double *__restrict__ rho_old,
                                                      const double *__restrict__ rho_bar,
                                                      const double *__restrict__ grey_src, double *__restrict__ part) {
    __shared__ double red[4];
    __shared__ double redm[4];
    double mx = 0.0, vol = 0.0, grey = 0.0;
    const int base = blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = base + 256 * q;
        if (i < na) {
            mx = fmax(mx, fabs(rho[i] - rho_old[i]));
            vol += rho_bar[i];
            grey += grey_src[i] * (1.0 - grey_src[i]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) redm[threadIdx.x >> 6] = mx;
    vol = block_sum256(vol, red);
    grey = block_sum256(grey, red);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = fmax(fmax(redm[0], redm[1]), fmax(redm[2], redm[3]));
        part[3 * blockIdx.x + 1] = vol;
        part[3 * blockIdx.x + 2] = grey;
    }
}

// Per-block partials: F.u and max over active nodes of |u|^2 (TopologyOptimizer.cs:256-276).
__global__ __launch_bounds__(256) void k_compliance_maxdisp(int n_nodes, const uint8_t *__restrict__ active_nodes,
                                                            const double *__restrict__ f, const double *__restrict__ u,
                                                            double *__restrict__ part) {
    __shared__ double red[4];
    __shared__ double redm[4];
    double c = 0.0, mx = 0.0;
    const int base = blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int nd = base + 256 * q;
        if (nd < n_nodes) {
            const double ux = u[3 * nd], uy = u[3 * nd + 1], uz = u[3 * nd + 2];
            c += f[3 * nd] * ux + f[3 * nd + 1] * uy + f[3 * nd + 2] * uz;
            if (active_nodes[nd]) mx = fmax(mx, ux * ux + uy * uy + uz * uz);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) redm[threadIdx.x >> 6] = mx;
    c = block_sum256(c, red);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = c;
        part[2 * blockIdx.x + 1] = fmax(fmax(redm[0], redm[1]), fmax(redm[2], redm[3]));
    }
}

// Final fixed-order combine of K-wide partials (sums, or max for the columns flagged in max_mask).
__global__ __launch_bounds__(256) void k_combine(const double *__restrict__ part, int nb, int k, int max_mask,
                                                 double *__restrict__ out) {
    __shared__ double red[4];
    for (int c = 0; c < k; c++) {
        const bool is_max = (max_mask >> c) & 1;
        double v = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) v = is_max ? fmax(v, part[k * b + c]) : v + part[k * b + c];
        if (is_max) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
            __syncthreads();
            v = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
            __syncthreads();
        } else {
            v = block_sum256(v, red);
        }
        if (threadIdx.x == 0) out[c] = v;
    }
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// Filter (host side)
// ------------------------------------------------------------------------------------------------
DeviceFilter::DeviceFilter(int nx_, int ny_, int nz_, const uint8_t *domain, const int32_t *active_, int n_active,
                           double rmin_, hipStream_t s)
    : nx(nx_), ny(ny_), nz(nz_), n(n_active), rmin(rmin_), stream(s) {
    const int total = nx * ny * nz;
    std::vector<int32_t> to_active(total, -1);
    for (int i = 0; i < n; i++) {
        const int e = active_[i];
        if (e < 0 || e >= total) throw Error(MYCEL_E_OUT_OF_RANGE, "active element index out of range");
        if (domain[e]) to_active[e] = i;  // the reference requires domain != 0 and an active index
    }
    // Tap table in the reference's jz, jy, jx loop order (DensityFilter.cs:85-118).
    const int rc = static_cast<int>(std::ceil(rmin));
    const double rsq = rmin * rmin;
    std::vector<int4> taps;
    std::vector<double> th;
    for (int dz = -rc; dz <= rc; dz++)
        for (int dy = -rc; dy <= rc; dy++)
            for (int dx = -rc; dx <= rc; dx++) {
                const int dsq = dz * dz + dy * dy + dx * dx;
                if (!(dsq < rsq)) continue;
                const double dist = std::sqrt(static_cast<double>(dsq));
                if (!(dist < rmin)) continue;
                taps.push_back(make_int4(dx, dy, dz, 0));
                th.push_back(rmin - dist);
            }
    n_taps = static_cast<int>(taps.size());
    d_taps.resize(n_taps > 0 ? n_taps : 1);
    d_tap_h.resize(n_taps > 0 ? n_taps : 1);
    if (n_taps) {
        d_taps.upload(taps.data(), n_taps, s);
        d_tap_h.upload(th.data(), n_taps, s);
    }
    d_to_active.resize(total);
    d_to_active.upload(to_active.data(), total, s);
    d_active.resize(n > 0 ? n : 1);
    if (n) d_active.upload(active_, n, s);
    d_inv_hs.resize(n > 0 ? n : 1);
    d_s.resize(n > 0 ? n : 1);
    d_in.resize(n > 0 ? n : 1);
    d_out.resize(n > 0 ? n : 1);
    if (n) {
        const FilterGeom g{nx, ny, nz, n, n_taps};
        hipLaunchKernelGGL(k_filter_inv_hs, dim3(cdiv(n, 256)), dim3(256), 0, s, g, d_taps.get(), d_tap_h.get(),
                           d_active.get(), d_to_active.get(), d_inv_hs.get());
        MYCEL_LAUNCH_CHECK();
        apply(nullptr, d_s.get(), true, s);  // S_i = sum_e W_ei (DensityFilter.cs:188-198)
    }
    // Count the non-zeros of W once on the host from the same rules (diagnostic only).
    nnz = 0;
    for (int i = 0; i < n; i++) {
        const int e = active_[i], ex = e % nx, t = e / nx, ey = t % ny, ez = t / ny;
        for (const int4 &d : taps) {
            const int jx = ex + d.x, jy = ey + d.y, jz = ez + d.z;
            if (jx < 0 || jy < 0 || jz < 0 || jx >= nx || jy >= ny || jz >= nz) continue;
            if (to_active[jx + nx * (jy + ny * jz)] >= 0) nnz++;
        }
    }
    MYCEL_HIP(hipStreamSynchronize(s));
}

void DeviceFilter::apply(const double *d_in_, double *d_out_, bool adjoint, hipStream_t s) {
    if (!n) return;
    const FilterGeom g{nx, ny, nz, n, n_taps};
    if (adjoint)
        hipLaunchKernelGGL(k_filter_apply<true>, dim3(cdiv(n, 256)), dim3(256), 0, s, g, d_taps.get(), d_tap_h.get(),
                           d_active.get(), d_to_active.get(), d_inv_hs.get(), d_in_, d_out_);
    else
        hipLaunchKernelGGL(k_filter_apply<false>, dim3(cdiv(n, 256)), dim3(256), 0, s, g, d_taps.get(),
                           d_tap_h.get(), d_active.get(), d_to_active.get(), d_inv_hs.get(), d_in_, d_out_);
    MYCEL_LAUNCH_CHECK();
}

// ------------------------------------------------------------------------------------------------
// OC (host side)
// ------------------------------------------------------------------------------------------------
OcWork::OcWork() {
    states.resize(2);
    res.resize(1);
    MYCEL_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_res), sizeof(OcResult), hipHostMallocDefault));
}

OcWork::~OcWork() {
    if (h_res) (void)hipHostFree(h_res);
}

namespace {
// Device memory of one OC update: the two-slot state, the volume partials (two buffers of 2 * nb), the
// partial sums of s (nb) and the result the host polls.
struct OcScratch {
    OcState *states;
    double *parts, *spart;
    OcResult *res;
};

// The OC state machine on the stream: sum of s, the initial bracket, then batches of 16 steps with one
// poll of the result after each batch. h_res receives the polled result (pinned or pageable host memory).
double oc_run(const OcScratch &w, OcResult *h_res, const double *d_rho, const double *d_dc, const double *d_s,
              const uint8_t *d_passive, double *d_rho_new, int n, double vf, double move, double rho_min,
              double lambda_guess, int max_bisect, double vol_tol, hipStream_t s, int *evals) {
    const int nb = cdiv(n, 1024);
    OcArgs a{d_rho, d_dc, d_s, d_passive, d_rho_new, n, move, rho_min, vol_tol, max_bisect};
    hipLaunchKernelGGL(k_oc_sum_s, dim3(nb), dim3(256), 0, s, d_s, n, w.spart);
    hipLaunchKernelGGL(k_oc_begin, dim3(nb), dim3(256), 0, s, a, w.spart, nb, vf, lambda_guess, w.states, w.parts,
                       w.res);
    MYCEL_LAUNCH_CHECK();
    // Worst case: 2 + 50 + 50 + max_bisect evaluations plus the final write.
    const int max_launches = 2 + 50 + 50 + max_bisect + 2;
    int launched = 0, slot = 0;
    const int batch = 16;
    while (true) {
        for (int k = 0; k < batch; k++, launched++, slot ^= 1)
            hipLaunchKernelGGL(k_oc_step, dim3(nb), dim3(256), 0, s, a, slot, w.states, w.parts, nb, w.res);
        MYCEL_LAUNCH_CHECK();
        MYCEL_HIP(hipMemcpyAsync(h_res, w.res, sizeof(OcResult), hipMemcpyDeviceToHost, s));
        MYCEL_HIP(hipStreamSynchronize(s));
        if (h_res->done) break;
        if (launched > max_launches) throw Error(MYCEL_E_DEVICE, "OC bisection did not terminate");
    }
    if (evals) *evals = h_res->evals;
    return h_res->lambda;
}

// ---- the OC update that allocates nothing ------------------------------------------------------
// Its scratch lives in this file-scope device array: every device the library runs on holds one copy until
// the process ends, and every call on a device shares it (calls that overlap on one device race). In host
// code the array's name is not a device address: each device's copy is reached only through the address
// hipGetSymbolAddress returns for it, kept in g_oc_scratch_ptr.
constexpr int kOcStateBytes = 256;  // 2 OcState + OcResult, padded to the array's alignment
constexpr size_t kOcScratchBytes = kOcStateBytes + sizeof(double) * 5 * static_cast<size_t>(kOcNoAllocMaxBlocks);
static_assert(2 * sizeof(OcState) + sizeof(OcResult) <= kOcStateBytes, "OC state does not fit its slot");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kOcScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_oc_scratch_mutex;
unsigned char *g_oc_scratch_ptr[kMaxDevices] = {};

// The current device's copy of the scratch array, looked up once per device (with that device current).
unsigned char *oc_scratch_ptr() {
    int dev = 0;
    MYCEL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(MYCEL_E_DEVICE, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_oc_scratch_mutex);
    if (!g_oc_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Error(MYCEL_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_oc_scratch_ptr[dev] = static_cast<unsigned char *>(ptr);
    }
    return g_oc_scratch_ptr[dev];
}
}  // namespace

double oc_update_device(OcWork &w, const double *d_rho, const double *d_dc, const double *d_s,
                        const uint8_t *d_passive, double *d_rho_new, int n, double vf, double move, double rho_min,
                        double lambda_guess, int max_bisect, double vol_tol, hipStream_t s, int *evals) {
    if (n <= 0) {
        if (evals) *evals = 0;
        return lambda_guess;
    }
    const int nb = cdiv(n, 1024);
    if (w.parts.size() < static_cast<size_t>(4 * nb)) w.parts.resize(4 * nb);
    if (w.spart.size() < static_cast<size_t>(nb)) w.spart.resize(nb);
    const OcScratch sc{w.states.get(), w.parts.get(), w.spart.get(), w.res.get()};
    return oc_run(sc, w.h_res, d_rho, d_dc, d_s, d_passive, d_rho_new, n, vf, move, rho_min, lambda_guess, max_bisect,
                  vol_tol, s, evals);
}

double oc_update_noalloc(const double *d_rho, const double *d_dc, const double *d_s, const uint8_t *d_passive,
                         double *d_rho_new, int n, double vf, double move, double rho_min, double lambda_guess,
                         int max_bisect, double vol_tol, hipStream_t s) {
    if (n > kOcNoAllocMaxElements)
        throw Error(MYCEL_E_OUT_OF_RANGE, "n = " + std::to_string(n) + " exceeds the no-allocation OC update's limit of " +
                                              std::to_string(kOcNoAllocMaxElements) + " elements");
    if (n <= 0) return lambda_guess;
    unsigned char *base = oc_scratch_ptr();
    const int nb = cdiv(n, 1024);
    OcScratch sc;
    sc.states = reinterpret_cast<OcState *>(base);
    sc.res = reinterpret_cast<OcResult *>(base + 2 * sizeof(OcState));
    sc.spart = reinterpret_cast<double *>(base + kOcStateBytes);
    sc.parts = sc.spart + nb;
    // Every word the kernels read is written earlier in the same call (k_oc_sum_s writes spart, k_oc_begin the
    // state, the result's done flag and the first partials), so nothing left by an earlier call is read.
    OcResult h{};
    return oc_run(sc, &h, d_rho, d_dc, d_s, d_passive, d_rho_new, n, vf, move, rho_min, lambda_guess, max_bisect,
                  vol_tol, s, nullptr);
}

// ------------------------------------------------------------------------------------------------
// SIMP helpers (host launchers)
// ------------------------------------------------------------------------------------------------
void launch_scatter_density(int na, const int32_t *d_active, const double *d_rho_bar, double *d_density,
                            hipStream_t s) {
    if (!na) return;
    hipLaunchKernelGGL(k_scatter_density, dim3(cdiv(na, 256)), dim3(256), 0, s, na, d_active, d_rho_bar, d_density);
    MYCEL_LAUNCH_CHECK();
}

void launch_force_passive(int na, const uint8_t *d_passive, double *d_rho, hipStream_t s) {
    if (!na) return;
    hipLaunchKernelGGL(k_force_passive, dim3(cdiv(na, 256)), dim3(256), 0, s, na, d_passive, d_rho);
    MYCEL_LAUNCH_CHECK();
}

// HeavisideProjection.Project (Optimization/HeavisideProjection.cs:12-24):
// rho_bar = clamp((tanh(beta eta) + tanh(beta (rt - eta))) * (1 / denom), 0, 1).
__global__ __launch_bounds__(256) void k_heaviside_project(int n, const double *__restrict__ rt,
                                                           double *__restrict__ rb, double beta, double eta) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double tbe = tanh(beta * eta);
    const double inv_denom = 1.0 / (tbe + tanh(beta * (1.0 - eta)));
    const double v = (tbe + tanh(beta * (rt[i] - eta))) * inv_denom;
    rb[i] = fmin(fmax(v, 0.0), 1.0);
}

// dc *= max(beta (1 - t^2) / denom, 1e-12), t = tanh(beta (rt - eta)): the chain rule through the
// projection (HeavisideProjection.cs:26-37, TopologyOptimizer.cs:281-289).
__global__ __launch_bounds__(256) void k_heaviside_chain(int n, const double *__restrict__ rt, double *__restrict__ dc,
                                                         double beta, double eta) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double inv_denom = 1.0 / (tanh(beta * eta) + tanh(beta * (1.0 - eta)));
    const double t = tanh(beta * (rt[i] - eta));
    dc[i] = dc[i] * fmax(beta * (1.0 - t * t) * inv_denom, 1e-12);
}

void launch_heaviside_project(int n, const double *d_rt, double *d_rb, double beta, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_heaviside_project, dim3(cdiv(n, 256)), dim3(256), 0, s, n, d_rt, d_rb, beta, 0.5);
    MYCEL_LAUNCH_CHECK();
}

void launch_heaviside_chain(int n, const double *d_rt, double *d_dc, double beta, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_heaviside_chain, dim3(cdiv(n, 256)), dim3(256), 0, s, n, d_rt, d_dc, beta, 0.5);
    MYCEL_LAUNCH_CHECK();
}

// dst += src (the per-load-case sensitivities summed in case order).
__global__ __launch_bounds__(256) void k_add_into(int n, const double *__restrict__ src, double *__restrict__ dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] += src[i];
}

void launch_add_into(int n, const double *d_src, double *d_dst, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_add_into, dim3(cdiv(n, 256)), dim3(256), 0, s, n, d_src, d_dst);
    MYCEL_LAUNCH_CHECK();
}

void launch_project(int n, const uint8_t *d_flags, const double *d_f, double *d_b, hipStream_t s) {
    hipLaunchKernelGGL(k_project, dim3(cdiv(n, 256)), dim3(256), 0, s, n, d_flags, d_f, d_b);
    MYCEL_LAUNCH_CHECK();
}

void simp_metrics(int na, const double *d_rho, const double *d_rho_old, const double *d_rho_bar,
                  const double *d_grey_src, DevBuf<double> &part, double *h_out3, hipStream_t s) {
    const int nb = cdiv(na, 1024);
    if (part.size() < static_cast<size_t>(3 * nb + 3)) part.resize(3 * nb + 3);
    hipLaunchKernelGGL(k_simp_metrics, dim3(nb), dim3(256), 0, s, na, d_rho, d_rho_old, d_rho_bar,
                       d_grey_src ? d_grey_src : d_rho_bar, part.get());
    hipLaunchKernelGGL(k_combine, dim3(1), dim3(256), 0, s, part.get(), nb, 3, 1, part.get() + 3 * nb);
    MYCEL_LAUNCH_CHECK();
    MYCEL_HIP(hipMemcpyAsync(h_out3, part.get() + 3 * nb, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipStreamSynchronize(s));
}

void compliance_maxdisp(int n_nodes, const uint8_t *d_active_nodes, const double *d_f, const double *d_u,
                        DevBuf<double> &part, double *h_out2, hipStream_t s) {
    const int nb = cdiv(n_nodes, 1024);
    if (part.size() < static_cast<size_t>(2 * nb + 2)) part.resize(2 * nb + 2);
    hipLaunchKernelGGL(k_compliance_maxdisp, dim3(nb), dim3(256), 0, s, n_nodes, d_active_nodes, d_f, d_u,
                       part.get());
    hipLaunchKernelGGL(k_combine, dim3(1), dim3(256), 0, s, part.get(), nb, 2, 2, part.get() + 2 * nb);
    MYCEL_LAUNCH_CHECK();
    MYCEL_HIP(hipMemcpyAsync(h_out2, part.get() + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipStreamSynchronize(s));
}

}  // namespace mycel
