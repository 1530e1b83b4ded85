// Geometric multigrid preconditioner for the projected operator A = PKP + (I - P), applied in FP32.
//
// The reference ships only the Jacobi preconditioner behind IPreconditioner (Fem/IPreconditioner.cs:9-20,
// Fem/JacobiPreconditioner.cs:15-35). This is a second implementation of that interface: one V-cycle
// z = M(r) per outer PCG iteration, with the outer iteration, its operator, residual and dot products kept
// in FP64 (pcg.hip, MYCEL_PRECISION_MIXED_GMG).
//
// Hierarchy. The voxel grid coarsens by 2 in every axis (ceil(n/2) elements) while the level has more than
// kCoarseMaxDofs DOFs and every axis still has >= 2 elements. A coarse element is a cube of twice the pitch,
// so its stiffness is the same Ke scaled by h_l / h (Hex8 elasticity scales linearly with h), and its
// modulus is the mean of its 8 children on the zero-padded fine element grid (rediscretisation). A coarse
// DOF is constrained when no element with E > 0 touches its node or when the collocated fine DOF
// (2I, 2J, 2K) is constrained. Constrained DOFs carry inv = 0, and every vector of the V-cycle is zero on
// them, so the smoother, residual, restriction and prolongation never branch on flags.
//
// Smoother: Chebyshev iteration of degree 2 on D^-1 A over [lmax / 30, lmax], where lmax is the
// Gershgorin bound max_r sum_c |Ke_rc| / Ke_rr of the element matrix: a row of D^-1 A is an E-weighted
// average of element rows, so the bound holds on every level for every density field and costs nothing.
// Each step is one fused kernel: the node-centric gather y = K x (the same slab gather as the K*u
// kernels), r = b - y, d' = cd d + cr D^-1 r, x' = x + d'.
//
// Transfer: trilinear prolongation x_f += P x_c (node-centric, <= 8 coarse nodes per fine node) and
// restriction b_c = P^T r_f (27 fine nodes per coarse node with weights 1, 1/2, 1/4, 1/8); the restriction
// also starts the coarse pre-smoother (x = d = D^-1 b / theta), which saves a launch per level.
//
// Coarsest level (<= kCoarseMaxDofs DOFs): one workgroup assembles the dense FP64 matrix into LDS (one row
// per thread, fixed element order), inverts it by Gauss-Jordan elimination (SPD, no pivoting) and writes
// the inverse; each V-cycle then applies it as a dense product. Everything is deterministic: no atomics.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>

#include "apply_k.hpp"
#include "device_common.hpp"
#include "mycel_internal.hpp"

namespace mycel {

namespace {

constexpr int kCoarseMaxDofs = 140;  // 140^2 doubles = 156,800 B of LDS (gfx950: 160 KiB per workgroup)
constexpr int kMaxLevels = 16;
constexpr double kSmoothRange = 30.0;  // Chebyshev interval [lmax / 30, lmax]

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ void node_ijk(const Dims &g, int n, int &i, int &j, int &k) {
    i = n % g.nx1;
    const int t = n / g.nx1;
    j = t % g.ny1;
    k = t / g.ny1;
}

// One Chebyshev step on level g: r = b - K x; d' = cd d + cr inv r; x' = x + d'. FINAL (finest level, last
// step) writes x' as the FP64 preconditioned residual z and per-block partials of z.r64 and z.ap64 (the
// flexible-CG inner products); otherwise x' goes to x_out and d' back to d.
template <bool FINAL>
__global__ __launch_bounds__(256, 2) void k_gmg_cheb(Dims g, KeSym<float> ks, const float *__restrict__ ep,
                                                     const float *__restrict__ x, const float *__restrict__ b,
                                                     float *__restrict__ d, const float *__restrict__ inv, float cd,
                                                     float cr, float *__restrict__ x_out, double *__restrict__ z64,
                                                     const double *__restrict__ r64, const double *__restrict__ ap64,
                                                     double *__restrict__ part_zr, double *__restrict__ part_zap) {
    __shared__ double red[8];
    const int n = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    double zr = 0.0, zap = 0.0;
    if (n < g.n_nodes) {
        int i, j, k;
        node_ijk(g, n, i, j, k);
        float kx[3], xc[3];
        ku_gather<float, kApplyK>(g, ks, ep, x, nullptr, i, j, k, kx, xc);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int dof = 3 * n + c;
            const float r = b[dof] - kx[c];
            const float dn = cd * d[dof] + cr * (inv[dof] * r);
            const float xn = xc[c] + dn;
            if (FINAL) {
                const double z = static_cast<double>(xn);
                z64[dof] = z;
                zr = __fma_rn(z, r64[dof], zr);
                zap = __fma_rn(z, ap64[dof], zap);
            } else {
                x_out[dof] = xn;
                d[dof] = dn;
            }
        }
    }
    if (FINAL) {
        block_sum256x2(zr, zap, red);
        if (threadIdx.x == 0) {
            part_zr[blockIdx.x] = zr;
            part_zap[blockIdx.x] = zap;
        }
    }
}

// r = b - K x on unconstrained DOFs, 0 on constrained ones.
__global__ __launch_bounds__(256, 2) void k_gmg_residual(Dims g, KeSym<float> ks, const float *__restrict__ ep,
                                                         const float *__restrict__ x, const float *__restrict__ b,
                                                         const float *__restrict__ inv, float *__restrict__ r) {
    const int n = xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (n >= g.n_nodes) return;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    float kx[3], xc[3];
    ku_gather<float, kApplyK>(g, ks, ep, x, nullptr, i, j, k, kx, xc);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dof = 3 * n + c;
        r[dof] = inv[dof] > 0.0f ? b[dof] - kx[c] : 0.0f;
    }
}

// b0 = float(r64) on unconstrained DOFs; the finest pre-smoother's first step x = d = inv b0 / theta.
__global__ __launch_bounds__(256) void k_gmg_init_fine(int n, const double *__restrict__ r64,
                                                       const float *__restrict__ inv, float theta_inv,
                                                       float *__restrict__ b, float *__restrict__ x,
                                                       float *__restrict__ d) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float iv = inv[i];
    const float bi = iv > 0.0f ? static_cast<float>(r64[i]) : 0.0f;
    const float xi = (iv * bi) * theta_inv;
    b[i] = bi;
    x[i] = xi;
    d[i] = xi;
}

// b_c = P^T r_f, then the coarse pre-smoother's first step x_c = d_c = inv_c b_c / theta.
__global__ __launch_bounds__(256) void k_gmg_restrict(Dims gf, const float *__restrict__ rf, Dims gc,
                                                      const float *__restrict__ invc, float theta_inv,
                                                      float *__restrict__ bc, float *__restrict__ xcv,
                                                      float *__restrict__ dc) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= gc.n_nodes) return;
    int I, J, K;
    node_ijk(gc, n, I, J, K);
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (int dk = -1; dk <= 1; dk++) {
        const int k = 2 * K + dk;
        if (k < 0 || k > gf.nz) continue;
        const float wz = dk == 0 ? 1.0f : 0.5f;
        for (int dj = -1; dj <= 1; dj++) {
            const int j = 2 * J + dj;
            if (j < 0 || j > gf.ny) continue;
            const float wyz = wz * (dj == 0 ? 1.0f : 0.5f);
            for (int di = -1; di <= 1; di++) {
                const int i = 2 * I + di;
                if (i < 0 || i > gf.nx) continue;
                const float w = wyz * (di == 0 ? 1.0f : 0.5f);
                const int m = 3 * (i + gf.nx1 * (j + gf.ny1 * k));
                s0 += w * rf[m];
                s1 += w * rf[m + 1];
                s2 += w * rf[m + 2];
            }
        }
    }
    const float s[3] = {s0, s1, s2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dof = 3 * n + c;
        const float iv = invc[dof];
        const float bi = iv > 0.0f ? s[c] : 0.0f;
        const float xi = (iv * bi) * theta_inv;
        bc[dof] = bi;
        xcv[dof] = xi;
        dc[dof] = xi;
    }
}

// x_f += P x_c on unconstrained fine DOFs (trilinear: weight 1 at an even fine index, 1/2 from each
// neighbour at an odd one).
__global__ __launch_bounds__(256) void k_gmg_prolong(Dims gf, float *__restrict__ xf, const float *__restrict__ invf,
                                                     Dims gc, const float *__restrict__ xcv) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= gf.n_nodes) return;
    int i, j, k;
    node_ijk(gf, n, i, j, k);
    const int i0 = i >> 1, i1 = (i + 1) >> 1, j0 = j >> 1, j1 = (j + 1) >> 1, k0 = k >> 1, k1 = (k + 1) >> 1;
    const float wi = (i & 1) ? 0.5f : 1.0f, wj = (j & 1) ? 0.5f : 1.0f, wk = (k & 1) ? 0.5f : 1.0f;
    const int is[2] = {i0, i1}, js[2] = {j0, j1}, ks_[2] = {k0, k1};
    const int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (int c = 0; c < nk; c++)
        for (int b = 0; b < nj; b++)
            for (int a = 0; a < ni; a++) {
                const int m = 3 * (is[a] + gc.nx1 * (js[b] + gc.ny1 * ks_[c]));
                const float w = wi * wj * wk;
                s0 += w * xcv[m];
                s1 += w * xcv[m + 1];
                s2 += w * xcv[m + 2];
            }
    const float s[3] = {s0, s1, s2};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dof = 3 * n + c;
        if (invf[dof] > 0.0f) xf[dof] += s[c];
    }
}

// ---- set-up ----------------------------------------------------------------------------------------

// Finest-level inverse diagonal of the smoother: 1 / DiagA on unconstrained DOFs, 0 on constrained ones.
__global__ __launch_bounds__(256) void k_gmg_inv_fine(int n, const uint8_t *__restrict__ flags,
                                                      const double *__restrict__ diag, float *__restrict__ inv) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double dv = diag[i];
    inv[i] = ((flags[i] & kFlagConstrained) || !(dv > 1e-30)) ? 0.0f : static_cast<float>(1.0 / dv);
}

// Coarse padded modulus grid: the mean of the 8 children of each coarse element; 0 on the halo.
__global__ __launch_bounds__(256) void k_gmg_coarsen_e(Dims gf, const float *__restrict__ epf, Dims gc,
                                                       float *__restrict__ epc) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= gc.n_pad) return;
    const int px = q % gc.px, t = q / gc.px, py = t % gc.py, pz = t / gc.py;
    float e = 0.0f;
    if (px >= 1 && px <= gc.nx && py >= 1 && py <= gc.ny && pz >= 1 && pz <= gc.nz) {
        const int ex = 2 * (px - 1), ey = 2 * (py - 1), ez = 2 * (pz - 1);
        float s = 0.0f;
        for (int c = 0; c < 2; c++)
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++)
                    s += epf[(ex + a + 1) + gf.px * ((ey + b + 1) + gf.py * (ez + c + 1))];
        e = 0.125f * s;
    }
    epc[q] = e;
}

// Coarse constraints and inverse diagonal (all diagonal entries of the cube's Ke are equal: kdiag).
__global__ __launch_bounds__(256) void k_gmg_inv_coarse(Dims gc, const float *__restrict__ epc, float kdiag, Dims gf,
                                                        const float *__restrict__ invf, float *__restrict__ invc) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= gc.n_nodes) return;
    int I, J, K;
    node_ijk(gc, n, I, J, K);
    const int pb = (I + 1) + gc.px * ((J + 1) + gc.py * (K + 1));
    float se = 0.0f;
#pragma unroll
    for (int a = 0; a < 8; a++) se += epc[pb - corner_x(a) - gc.px * (corner_y(a) + gc.py * corner_z(a))];
    const bool inside = 2 * I <= gf.nx && 2 * J <= gf.ny && 2 * K <= gf.nz;
    const int fm = inside ? 3 * (2 * I + gf.nx1 * (2 * J + gf.ny1 * 2 * K)) : 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const bool cons = !(se > 0.0f) || (inside && !(invf[fm + c] > 0.0f));
        invc[3 * n + c] = cons ? 0.0f : 1.0f / (kdiag * se);
    }
}

// Coarsest level: assemble the dense matrix (identity rows and zero columns on constrained DOFs) into LDS,
// invert it in place by Gauss-Jordan elimination and store the inverse. One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_gmg_coarse_factor(Dims g, const float *__restrict__ ep,
                                                            const double *__restrict__ ke, double scale,
                                                            const float *__restrict__ inv, double *__restrict__ ainv) {
    extern __shared__ __align__(16) double A[];
    __shared__ double col[kCoarseMaxDofs];
    const int n = g.n_dofs;
    for (int q = threadIdx.x; q < n * n; q += blockDim.x) A[q] = 0.0;
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        const int node = r / 3, c = r % 3;
        if (!(inv[r] > 0.0f)) {
            A[r * n + r] = 1.0;
            continue;
        }
        int I, J, K;
        node_ijk(g, node, I, J, K);
        for (int a = 0; a < 8; a++) {  // the element that holds this node as its corner a
            const int ox = I - corner_x(a), oy = J - corner_y(a), oz = K - corner_z(a);
            const double E = static_cast<double>(ep[(ox + 1) + g.px * ((oy + 1) + g.py * (oz + 1))]);
            if (E == 0.0) continue;
            for (int b = 0; b < 8; b++) {
                const int m = (ox + corner_x(b)) + g.nx1 * ((oy + corner_y(b)) + g.ny1 * (oz + corner_z(b)));
                for (int d = 0; d < 3; d++) {
                    const int col_dof = 3 * m + d;
                    if (!(inv[col_dof] > 0.0f)) continue;
                    A[r * n + col_dof] += E * scale * ke[(3 * a + c) * 24 + 3 * b + d];
                }
            }
        }
    }
    __syncthreads();
    for (int k = 0; k < n; k++) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) col[i] = A[i * n + k];
        __syncthreads();
        const double piv = 1.0 / col[k];
        for (int j = threadIdx.x; j < n; j += blockDim.x) A[k * n + j] = (j == k) ? piv : A[k * n + j] * piv;
        __syncthreads();
        for (int q = threadIdx.x; q < n * n; q += blockDim.x) {
            const int i = q / n, j = q % n;
            if (i == k) continue;
            A[q] = (j == k) ? -col[i] * piv : A[q] - col[i] * A[k * n + j];
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < n * n; q += blockDim.x) ainv[q] = A[q];
}

// x = A^-1 b on the coarsest level: one wavefront per row, FP64 accumulation.
__global__ __launch_bounds__(256) void k_gmg_coarse_solve(int n, const double *__restrict__ ainv,
                                                          const float *__restrict__ b, float *__restrict__ x) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s = __fma_rn(ainv[row * n + j], static_cast<double>(b[j]), s);
    s = wave_sum(s);
    if (lane == 0) x[row] = static_cast<float>(s);
}

}  // namespace

struct GmgLevel {
    Dims d{};
    KeSym<float> ks{};
    double scale = 1.0;  // h_l / h
    float kdiag = 0.0f;  // diagonal entry of the level's Ke
    const float *ep = nullptr;
    DevBuf<float> ep_own, inv, b, x, x2, dv, r;
    int blocks = 0;
};

struct Gmg {
    GmgLevel lv[kMaxLevels];
    int levels = 0;
    DevBuf<double> ainv;
    uint64_t epoch = ~0ull;
    float theta_inv = 0, cd1 = 0, cr1 = 0;
};

void gmg_free(Gmg *g) { delete g; }

int gmg_partials(const mycel_op *op) { return cdiv(op->d.n_nodes, 256); }

// Builds the level structure; op->gmg is set only once the whole hierarchy exists.
static void gmg_build(mycel_op *op) {
    std::unique_ptr<Gmg> owner(new Gmg());
    Gmg *g = owner.get();
    // Chebyshev constants from the Gershgorin bound of D^-1 A.
    double lmax = 0.0;
    for (int r = 0; r < 24; r++) {
        double s = 0.0;
        for (int c = 0; c < 24; c++) s += std::fabs(op->ke[r * 24 + c]);
        lmax = std::fmax(lmax, s / op->ke[r * 24 + r]);
    }
    for (int r = 1; r < 24; r++)
        if (std::fabs(op->ke[r * 25] - op->ke[0]) > 1e-12 * std::fabs(op->ke[0]))
            throw Error(MYCEL_E_INVALID_OPERATION, "multigrid: element matrix diagonal is not uniform");
    const double lmin = lmax / kSmoothRange, theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta, rho0 = 1.0 / sigma, rho1 = 1.0 / (2.0 * sigma - rho0);
    g->theta_inv = static_cast<float>(1.0 / theta);
    g->cd1 = static_cast<float>(rho1 * rho0);
    g->cr1 = static_cast<float>(2.0 * rho1 / delta);

    Dims d = op->d;
    double scale = 1.0;
    for (int l = 0;; l++) {
        if (l >= kMaxLevels) throw Error(MYCEL_E_INVALID_OPERATION, "multigrid: too many levels");
        GmgLevel &L = g->lv[l];
        L.d = d;
        L.scale = scale;
        for (int q = 0; q < kKeClasses; q++) L.ks.v[q] = static_cast<float>(op->ks32.v[q] * scale);
        L.kdiag = static_cast<float>(op->ke[0] * scale);
        if (l == 0) {
            L.ep = op->d_ep32.get();
        } else {
            L.ep_own.resize(d.n_pad);
            L.ep = L.ep_own.get();
        }
        L.inv.resize(d.n_dofs);
        L.b.resize(d.n_dofs);
        L.x.resize(d.n_dofs);
        L.x2.resize(d.n_dofs);
        L.dv.resize(d.n_dofs);
        L.r.resize(d.n_dofs);
        L.blocks = cdiv(d.n_nodes, 256);
        g->levels = l + 1;
        if (d.n_dofs <= kCoarseMaxDofs) break;
        if (d.nx < 2 || d.ny < 2 || d.nz < 2)
            throw Error(MYCEL_E_ARGUMENT,
                        "multigrid preconditioner: the grid cannot be coarsened to <= " +
                            std::to_string(kCoarseMaxDofs) +
                            " DOFs with >= 2 elements per axis; use the FP64 or mixed FP32 solver");
        d = make_dims((d.nx + 1) / 2, (d.ny + 1) / 2, (d.nz + 1) / 2);
        scale *= 2.0;
    }
    const int nc = g->lv[g->levels - 1].d.n_dofs;
    g->ainv.resize(static_cast<size_t>(nc) * nc);
    op->gmg = owner.release();
}

void gmg_prepare(mycel_op *op, hipStream_t s) {
    if (!op->gmg) gmg_build(op);
    Gmg *g = op->gmg;
    if (g->epoch == op->epoch) return;
    const dim3 b256(256);
    GmgLevel &F = g->lv[0];
    hipLaunchKernelGGL(k_gmg_inv_fine, dim3(cdiv(F.d.n_dofs, 256)), b256, 0, s, F.d.n_dofs, op->d_flags.get(),
                       op->d_diag.get(), F.inv.get());
    for (int l = 1; l < g->levels; l++) {
        GmgLevel &P = g->lv[l - 1], &L = g->lv[l];
        hipLaunchKernelGGL(k_gmg_coarsen_e, dim3(cdiv(L.d.n_pad, 256)), b256, 0, s, P.d, P.ep, L.d, L.ep_own.get());
        hipLaunchKernelGGL(k_gmg_inv_coarse, dim3(cdiv(L.d.n_nodes, 256)), b256, 0, s, L.d, L.ep, L.kdiag, P.d,
                           P.inv.get(), L.inv.get());
    }
    GmgLevel &C = g->lv[g->levels - 1];
    const size_t lds = sizeof(double) * C.d.n_dofs * C.d.n_dofs;
    MYCEL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_gmg_coarse_factor),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(k_gmg_coarse_factor, dim3(1), dim3(1024), lds, s, C.d, C.ep, op->d_ke.get(), C.scale,
                       C.inv.get(), g->ainv.get());
    MYCEL_LAUNCH_CHECK();
    g->epoch = op->epoch;
}

// z = M(r): one V-cycle. Also writes per-block partials of z.r and z.ap (gmg_partials(op) of each).
void gmg_enqueue_vcycle(mycel_op *op, const double *r64, const double *ap64, double *z64, double *part_zr,
                        double *part_zap, hipStream_t s) {
    Gmg *g = op->gmg;
    const dim3 b256(256);
    const int top = g->levels - 1;
    float *cur[kMaxLevels], *nxt[kMaxLevels];
    for (int l = 0; l <= top; l++) {
        cur[l] = g->lv[l].x.get();
        nxt[l] = g->lv[l].x2.get();
    }
    auto cheb = [&](int l, float cd, float cr, bool final_step) {
        GmgLevel &L = g->lv[l];
        if (final_step) {
            hipLaunchKernelGGL(k_gmg_cheb<true>, dim3(L.blocks), b256, 0, s, L.d, L.ks, L.ep, cur[l], L.b.get(),
                               L.dv.get(), L.inv.get(), cd, cr, nxt[l], z64, r64, ap64, part_zr, part_zap);
        } else {
            hipLaunchKernelGGL(k_gmg_cheb<false>, dim3(L.blocks), b256, 0, s, L.d, L.ks, L.ep, cur[l], L.b.get(),
                               L.dv.get(), L.inv.get(), cd, cr, nxt[l], static_cast<double *>(nullptr),
                               static_cast<const double *>(nullptr), static_cast<const double *>(nullptr),
                               static_cast<double *>(nullptr), static_cast<double *>(nullptr));
            std::swap(cur[l], nxt[l]);
        }
    };
    GmgLevel &F = g->lv[0];
    hipLaunchKernelGGL(k_gmg_init_fine, dim3(cdiv(F.d.n_dofs, 256)), b256, 0, s, F.d.n_dofs, r64, F.inv.get(),
                       g->theta_inv, F.b.get(), cur[0], F.dv.get());
    for (int l = 0; l < top; l++) {
        GmgLevel &L = g->lv[l], &C = g->lv[l + 1];
        cheb(l, g->cd1, g->cr1, false);
        hipLaunchKernelGGL(k_gmg_residual, dim3(L.blocks), b256, 0, s, L.d, L.ks, L.ep, cur[l], L.b.get(),
                           L.inv.get(), L.r.get());
        hipLaunchKernelGGL(k_gmg_restrict, dim3(C.blocks), b256, 0, s, L.d, L.r.get(), C.d, C.inv.get(),
                           g->theta_inv, C.b.get(), cur[l + 1], C.dv.get());
    }
    GmgLevel &T = g->lv[top];
    hipLaunchKernelGGL(k_gmg_coarse_solve, dim3(cdiv(T.d.n_dofs, 4)), b256, 0, s, T.d.n_dofs, g->ainv.get(),
                       T.b.get(), cur[top]);
    for (int l = top - 1; l >= 0; l--) {
        GmgLevel &L = g->lv[l], &C = g->lv[l + 1];
        hipLaunchKernelGGL(k_gmg_prolong, dim3(L.blocks), b256, 0, s, L.d, cur[l], L.inv.get(), C.d, cur[l + 1]);
        cheb(l, 0.0f, g->theta_inv, false);
        cheb(l, g->cd1, g->cr1, l == 0);
    }
    MYCEL_LAUNCH_CHECK();
}

}  // namespace mycel
