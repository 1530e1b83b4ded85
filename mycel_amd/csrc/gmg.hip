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
// (2I, 2J, 2K), clamped to the fine grid, is constrained. Constrained DOFs carry inv = 0, and every vector of
// the V-cycle is zero on them, so the smoother, residual, restriction and prolongation never branch on flags.
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
// Coarsest level (<= kCoarseMaxDofs DOFs): one workgroup assembles the dense FP64 Galerkin matrix into LDS
// (fixed element order), inverts it by Gauss-Jordan elimination (SPD, no pivoting) and writes the inverse
// column-major; each V-cycle then applies it as a dense product spread over all waves of the tail
// workgroup. Everything is deterministic: no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <tuple>
#include <vector>

#include "apply_k.hpp"
#include "ku_element.hpp"
#include "device_common.hpp"
#include "mycel_internal.hpp"

namespace mycel {

namespace {

constexpr int kCoarseMaxDofs = 140;  // 140^2 doubles = 156,800 B of LDS (gfx950: 160 KiB per workgroup)
constexpr int kMaxLevels = 16;
constexpr double kSmoothRange = 30.0;  // Chebyshev interval [lmax / 30, lmax] (MYCEL_GMG_SMOOTH_RANGE: 2 .. 1000)
constexpr int kFineDegree = 3;          // Chebyshev degree on the finest level (2 on the coarser ones) ...
constexpr int kMaxDegree = 8;           // ... unless MYCEL_GMG_FINE_DEGREE asks for another in [2, kMaxDegree]

int fine_degree() {
    static const int v = [] {
        const char *e = std::getenv("MYCEL_GMG_FINE_DEGREE");
        return e && *e ? std::min(std::max(2, std::atoi(e)), kMaxDegree) : kFineDegree;
    }();
    return v;
}

double smooth_range() {
    static const double v = [] {
        const char *e = std::getenv("MYCEL_GMG_SMOOTH_RANGE");
        return e && *e ? std::min(std::max(2.0, std::atof(e)), 1000.0) : kSmoothRange;
    }();
    return v;
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ void node_ijk(const Dims &g, int n, int &i, int &j, int &k) {
    i = n % g.nx1;
    const int t = n / g.nx1;
    j = t % g.ny1;
    k = t / g.ny1;
}

// (q, r) = n divided by d, for 0 <= n < 2^16, from rd = 1 / d rounded to FP32: (n + 1/2) / d lies at least
// 1/(2 d) from an integer, and the product's error, below (n + 1/2) / d * 2^-23, is smaller (n + 1/2 < 2^22).
// Three operations instead of an integer division's ~30.
__device__ __forceinline__ void small_divmod(int n, int d, float rd, int &q, int &r) {
    q = static_cast<int>((static_cast<float>(n) + 0.5f) * rd);
    r = n - q * d;
}
// Node decompositions: node_ijk, or small_divmod with the grid's reciprocal node counts (grids < 2^16 nodes).
struct NodeIjk {
    __device__ void operator()(const Dims &g, int n, int &i, int &j, int &k) const { node_ijk(g, n, i, j, k); }
};
struct SmallIjk {
    float rx, ry;  // 1 / nx1, 1 / ny1
    __device__ void operator()(const Dims &g, int n, int &i, int &j, int &k) const {
        int t;
        small_divmod(n, g.nx1, rx, t, i);
        small_divmod(t, g.ny1, ry, k, j);
    }
};

// Per-node bodies of the V-cycle phases, shared by the one-launch-per-phase kernels of the large levels and
// by the single-workgroup tail kernel of the small ones.

// The smoother's preconditioner is the nodal block Jacobi B^-1: per node the inverse of the 3x3 diagonal block of
// the level's operator, stored as (xx, yy, zz, xy, xz, yz), with zero rows and columns on constrained components.
constexpr int kBinv = 6;  // floats per node

__device__ __forceinline__ void bload(const float *__restrict__ binv, int n, float (&bi)[kBinv]) {
#pragma unroll
    for (int q = 0; q < kBinv; q++) bi[q] = binv[kBinv * n + q];
}

// z = B^-1 r at one node.
__device__ __forceinline__ void bapply(const float (&bi)[kBinv], const float (&r)[3], float (&z)[3]) {
    z[0] = (bi[0] * r[0] + bi[3] * r[1]) + bi[4] * r[2];
    z[1] = (bi[3] * r[0] + bi[1] * r[1]) + bi[5] * r[2];
    z[2] = (bi[4] * r[0] + bi[5] * r[1]) + bi[2] * r[2];
}

// The Chebyshev direction update d' = cd d + cr z (two products and a sum, never contracted): the launched
// coarse levels' one-step and two-step kernels form it alike, so both give the same bits.
__device__ __forceinline__ float cheb_dn(float cd, float d, float cr, float z) { return cd * d + cr * z; }

// The first Chebyshev step of a level at node n from its right-hand side sv (0 on constrained components, whose
// inv is 0): b = sv, x = d = B^-1 b / theta.
__device__ __forceinline__ void first_step(const float *__restrict__ inv, const float *__restrict__ binv,
                                           float theta_inv, const float (&sv)[3], float *__restrict__ b,
                                           float *__restrict__ x, float *__restrict__ d, int n) {
    float bi[kBinv], bv[3], z[3];
    bload(binv, n, bi);
#pragma unroll
    for (int c = 0; c < 3; c++) bv[c] = inv[3 * n + c] > 0.0f ? sv[c] : 0.0f;
    bapply(bi, bv, z);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dof = 3 * n + c;
        const float xi = z[c] * theta_inv;
        b[dof] = bv[c];
        x[dof] = xi;
        d[dof] = xi;
    }
}

// One Chebyshev step at node n: r = b - K x; d' = cd d + cr B^-1 r; x' = x + d' (x' to x_out, d' back to d).
__device__ __forceinline__ void cheb_node(const Dims &g, const KeSym<float> &ks, const float *__restrict__ ep,
                                          const float *__restrict__ x, const float *__restrict__ b,
                                          float *__restrict__ d, const float *__restrict__ binv, float cd, float cr,
                                          float *__restrict__ x_out, int n, int k0) {
    int i, j, k;
    node_ijk(g, n, i, j, k);
    float kx[3], xc[3];
    ku_gather<float, kApplyK>(g, ks, ep, x, nullptr, i, j, k, k0, kx, xc);
    float bi[kBinv], r[3], z[3];
    bload(binv, n, bi);
#pragma unroll
    for (int c = 0; c < 3; c++) r[c] = b[3 * n + c] - kx[c];
    bapply(bi, r, z);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dof = 3 * n + c;
        const float dn = cd * d[dof] + cr * z[c];
        x_out[dof] = xc[c] + dn;
        d[dof] = dn;
    }
}

// r = b - K x on unconstrained DOFs, 0 on constrained ones.
__device__ __forceinline__ void residual_node(const Dims &g, const KeSym<float> &ks, const float *__restrict__ ep,
                                              const float *__restrict__ x, const float *__restrict__ b,
                                              const float *__restrict__ inv, float *__restrict__ r, int n, int k0) {
    int i, j, k;
    node_ijk(g, n, i, j, k);
    float kx[3], xc[3];
    ku_gather<float, kApplyK>(g, ks, ep, x, nullptr, i, j, k, k0, kx, xc);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int dof = 3 * n + c;
        r[dof] = inv[dof] > 0.0f ? b[dof] - kx[c] : 0.0f;
    }
}

// Restriction taps along one axis: fine index 2C + d for d = -1, 0, 1 with weights 1/2, 1, 1/2; a tap
// outside [0, nf] gets weight 0 and a clamped (valid) index, so all 27 loads of a coarse node are
// independent and issued together.
__device__ __forceinline__ void restrict_taps(int C, int nf, int (&idx)[3], float (&w)[3]) {
#pragma unroll
    for (int t = 0; t < 3; t++) {
        const int f = 2 * C + t - 1;
        const bool in = f >= 0 && f <= nf;
        idx[t] = in ? f : (f < 0 ? 0 : nf);
        w[t] = in ? (t == 1 ? 1.0f : 0.5f) : 0.0f;
    }
}

// b_c = P^T r_f at coarse node n, then the coarse pre-smoother's first step x_c = d_c = B_c^-1 b_c / theta.
// The taps are summed z-major, then y, then x; taps outside the fine grid add nothing.
template <typename IJK = NodeIjk>
__device__ __forceinline__ void restrict_node(const Dims &gf, const float *__restrict__ rf, const Dims &gc,
                                              const float *__restrict__ invc, const float *__restrict__ binvc,
                                              float theta_inv, float *__restrict__ bc, float *__restrict__ xcv,
                                              float *__restrict__ dc, int n, IJK ijk = {}) {
    int I, J, K;
    ijk(gc, n, I, J, K);
    int ix[3], iy[3], iz[3];
    float wx[3], wy[3], wz[3];
    restrict_taps(I, gf.nx, ix, wx);
    restrict_taps(J, gf.ny, iy, wy);
    restrict_taps(K, gf.nz, iz, wz);
    float v[27][3];  // all taps loaded first, then summed: one memory round trip
#pragma unroll
    for (int t = 0; t < 27; t++) {
        const int m = 3 * (ix[t % 3] + gf.nx1 * (iy[(t / 3) % 3] + gf.ny1 * iz[t / 9]));
        v[t][0] = rf[m];
        v[t][1] = rf[m + 1];
        v[t][2] = rf[m + 2];
    }
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int t = 0; t < 27; t++) {
        const float w = wz[t / 9] * wy[(t / 3) % 3] * wx[t % 3];
        s0 += w != 0.0f ? w * v[t][0] : 0.0f;
        s1 += w != 0.0f ? w * v[t][1] : 0.0f;
        s2 += w != 0.0f ? w * v[t][2] : 0.0f;
    }
    const float sv[3] = {s0, s1, s2};
    first_step(invc, binvc, theta_inv, sv, bc, xcv, dc, n);
}

// B^-1 at node n of a level from its zero-padded moduli ep: the node's 3x3 diagonal block is the sum over its <= 8
// elements of E_e times that element's corner block, kd on the diagonal and ko s_i s_j off it (s = +1 / -1 for a
// corner on the low / high side of the element along that axis). Constrained components (inv == 0) get zero rows and
// columns; the free ones the inverse of their sub-block, formed in FP64 and stored in FP32.
__device__ __forceinline__ void binv_node(const Dims &g, const float *__restrict__ ep, float kd, float ko,
                                          const float *__restrict__ inv, int n, float (&bi)[kBinv]) {
    using R = double;
    int i, j, k;
    node_ijk(g, n, i, j, k);
    const int pb = (i + 1) + g.px * ((j + 1) + g.py * (k + 1));
    float sd = 0.0f, sxy = 0.0f, sxz = 0.0f, syz = 0.0f;
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const float e = ep[pb - corner_x(a) - g.px * (corner_y(a) + g.py * corner_z(a))];
        const float sx = corner_x(a) ? -1.0f : 1.0f, sy = corner_y(a) ? -1.0f : 1.0f, sz = corner_z(a) ? -1.0f : 1.0f;
        sd += e;
        sxy += e * (sx * sy);
        sxz += e * (sx * sz);
        syz += e * (sy * sz);
    }
    bool f[3];
#pragma unroll
    for (int c = 0; c < 3; c++) f[c] = inv[3 * n + c] > 0.0f;
    // Constrained rows and columns replaced by the identity, inverted, then zeroed again.
    const R dg = static_cast<R>(kd) * static_cast<R>(sd);
    R m[3][3] = {{dg, static_cast<R>(ko) * static_cast<R>(sxy), static_cast<R>(ko) * static_cast<R>(sxz)},
                 {static_cast<R>(ko) * static_cast<R>(sxy), dg, static_cast<R>(ko) * static_cast<R>(syz)},
                 {static_cast<R>(ko) * static_cast<R>(sxz), static_cast<R>(ko) * static_cast<R>(syz), dg}};
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            if (!f[r] || !f[c]) m[r][c] = r == c ? R(1) : R(0);
    const R c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1], c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2],
            c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const R det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    const R id = det > R(0) ? R(1) / det : R(0);
    const R q[6] = {c00 * id, (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id, (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id,
                    c01 * id, c02 * id, (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id};
    // (xx, yy, zz, xy, xz, yz): entry (r, c) is kept when both components are free.
    const bool keep[6] = {f[0], f[1], f[2], f[0] && f[1], f[0] && f[2], f[1] && f[2]};
#pragma unroll
    for (int t = 0; t < kBinv; t++) bi[t] = keep[t] ? static_cast<float>(q[t]) : 0.0f;
}

// x_f += P x_c at fine node n, unconstrained DOFs only (trilinear: weight 1 at an even fine index, 1/2 from
// each neighbour at an odd one). All 8 candidate coarse nodes are read (indices always valid); those that
// do not contribute have weight 0 and add nothing.
// (P x_c) at fine node (i, j, k), into sv.
__device__ __forceinline__ void prolong_sums(const Dims &gc, const float *__restrict__ xcv, int i, int j, int k,
                                             float (&sv)[3]) {
    const int is[2] = {i >> 1, (i + 1) >> 1}, js[2] = {j >> 1, (j + 1) >> 1}, ks_[2] = {k >> 1, (k + 1) >> 1};
    const float wx[2] = {(i & 1) ? 0.5f : 1.0f, (i & 1) ? 0.5f : 0.0f};
    const float wy[2] = {(j & 1) ? 0.5f : 1.0f, (j & 1) ? 0.5f : 0.0f};
    const float wz[2] = {(k & 1) ? 0.5f : 1.0f, (k & 1) ? 0.5f : 0.0f};
    float v[8][3];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const int m = 3 * (is[t & 1] + gc.nx1 * (js[(t >> 1) & 1] + gc.ny1 * ks_[t >> 2]));
        v[t][0] = xcv[m];
        v[t][1] = xcv[m + 1];
        v[t][2] = xcv[m + 2];
    }
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const float w = wx[t & 1] * wy[(t >> 1) & 1] * wz[t >> 2];
        s0 += w != 0.0f ? w * v[t][0] : 0.0f;
        s1 += w != 0.0f ? w * v[t][1] : 0.0f;
        s2 += w != 0.0f ? w * v[t][2] : 0.0f;
    }
    sv[0] = s0;
    sv[1] = s1;
    sv[2] = s2;
}

template <typename IJK = NodeIjk>
__device__ __forceinline__ void prolong_node(const Dims &gf, float *__restrict__ xf, const float *__restrict__ invf,
                                             const Dims &gc, const float *__restrict__ xcv, int n, IJK ijk = {}) {
    int i, j, k;
    ijk(gf, n, i, j, k);
    float iv[3], xo[3], sv[3];  // every load first, then the sums: one memory round trip
#pragma unroll
    for (int c = 0; c < 3; c++) {
        iv[c] = invf[3 * n + c];
        xo[c] = xf[3 * n + c];
    }
    prolong_sums(gc, xcv, i, j, k, sv);
#pragma unroll
    for (int c = 0; c < 3; c++)
        if (iv[c] > 0.0f) xf[3 * n + c] = xo[c] + sv[c];
}

// One Chebyshev step (one node per thread).
// Grid dimension y of every V-cycle kernel is the right-hand side: the level's vectors of RHS k start at
// k * n_dofs (the moduli, D^-1 and B^-1 are shared).
__global__ __launch_bounds__(256, 2) void k_gmg_cheb(Dims g, KeSym<float> ks, const float *__restrict__ ep,
                                                     const float *__restrict__ x, const float *__restrict__ b,
                                                     float *__restrict__ d, const float *__restrict__ binv, float cd,
                                                     float cr, float *__restrict__ x_out) {
    const int n0 = xcd_block(blockIdx.x, gridDim.x) * 256, n = n0 + threadIdx.x;
    const long long o = static_cast<long long>(blockIdx.y) * g.n_dofs;
    if (n < g.n_nodes) cheb_node(g, ks, ep, x + o, b + o, d + o, binv, cd, cr, x_out + o, n, first_layer(g, n0));
}

// Island deflation (enabled per operator): every floating solid island (a face-connected cluster of elements
// with E >= ratio E0 that touches no fixed DOF) adds its six rigid-body modes Z to the preconditioner as an additive
// coarse correction, z = V(r) + Z (Z^T A Z)^-1 Z^T r. Such an island is held only by the void around it, so its
// rigid motions are near-null modes of A (eigenvalues of the order of the stiffness contrast) that neither the
// smoother nor the geometric coarse levels represent. The correction coefficients of island c are
// (t, w): t the translation, w the rotation about the island's node centroid.
struct Defl {
    const int *ncl;      // per node: island or -1
    const double *coef;  // 6 per island, kDeflStride per right-hand side
    const double *cen;   // island centroids
    double h;
    int nx1, ny1;
    int m;  // islands in use (0: off)
};
constexpr int kMaxIslands = 64;
constexpr int kDeflStride = 6 * kMaxIslands;

// z (FP64, the node's three components) += t + w x (p - centroid) of the node's island.
__device__ __forceinline__ void defl_add(const Defl &df, const double *coef, int n, int cl, double (&z)[3]) {
    const int i = n % df.nx1, t = n / df.nx1, j = t % df.ny1, k = t / df.ny1;
    const double dx = i * df.h - df.cen[3 * cl], dy = j * df.h - df.cen[3 * cl + 1], dz = k * df.h - df.cen[3 * cl + 2];
    const double *c = coef + 6 * cl;
    z[0] += c[0] + (c[4] * dz - c[5] * dy);
    z[1] += c[1] + (c[5] * dx - c[3] * dz);
    z[2] += c[2] + (c[3] * dy - c[4] * dx);
}

// y_c = Z_c^T r for island c (grid x) and right-hand side (grid y): the sums over the island's nodes (ascending)
// of r and of (p - centroid) x r, each in a fixed block tree.
__global__ __launch_bounds__(256) void k_defl_zt(int nx1, int ny1, double h, const int *__restrict__ off,
                                                 const int *__restrict__ nodes, const double *__restrict__ cen,
                                                 const double *__restrict__ r64, long long rstride,
                                                 double *__restrict__ y) {
    __shared__ double red[4];
    const int c = blockIdx.x;
    r64 += blockIdx.y * rstride;
    y += static_cast<long long>(blockIdx.y) * kDeflStride;
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = off[c] + threadIdx.x; q < off[c + 1]; q += 256) {
        const int n = nodes[q];
        const int i = n % nx1, t = n / nx1, j = t % ny1, k = t / ny1;
        const double dx = i * h - cen[3 * c], dy = j * h - cen[3 * c + 1], dz = k * h - cen[3 * c + 2];
        const double rx = r64[3 * n], ry = r64[3 * n + 1], rz = r64[3 * n + 2];
        s[0] += rx;
        s[1] += ry;
        s[2] += rz;
        s[3] += dy * rz - dz * ry;  // rotation about x: (0, -dz, dy)
        s[4] += dz * rx - dx * rz;  // about y: (dz, 0, -dx)
        s[5] += dx * ry - dy * rx;  // about z: (-dy, dx, 0)
    }
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const double v = block_sum256(s[q], red);
        if (threadIdx.x == 0) y[6 * c + q] = v;
    }
}

// coef = (Z^T A Z)^-1 y per right-hand side (grid y), rows in a fixed order.
__global__ __launch_bounds__(256) void k_defl_solve(int m6, const double *__restrict__ einv,
                                                    const double *__restrict__ y, double *__restrict__ coef) {
    y += static_cast<long long>(blockIdx.y) * kDeflStride;
    coef += static_cast<long long>(blockIdx.y) * kDeflStride;
    for (int q = threadIdx.x; q < m6; q += 256) {
        double acc = 0.0;
        for (int k = 0; k < m6; k++) acc = __fma_rn(einv[q * m6 + k], y[k], acc);
        coef[q] = acc;
    }
}

// The last post-smoothing step on the finest level: writes x' as the FP64 preconditioned residual z and
// per-block partials of z.r64 and z.ap64 (the flexible-CG inner products).
__global__ __launch_bounds__(256, 2) void k_gmg_cheb_final(Dims g, KeSym<float> ks, const float *__restrict__ ep,
                                                           const float *__restrict__ x, const float *__restrict__ b,
                                                           const float *__restrict__ d, const float *__restrict__ binv,
                                                           float cd, float cr, double *__restrict__ z64,
                                                           const double *__restrict__ r64,
                                                           const double *__restrict__ ap64,
                                                           double *__restrict__ part_zr, double *__restrict__ part_zap,
                                                           Defl df) {
    __shared__ double red[8];
    const int n0 = xcd_block(blockIdx.x, gridDim.x) * 256, n = n0 + threadIdx.x;
    const long long o = static_cast<long long>(blockIdx.y) * g.n_dofs;
    x += o;
    b += o;
    d += o;
    z64 += o;
    r64 += o;
    ap64 += o;
    part_zr += blockIdx.y * gridDim.x;
    part_zap += blockIdx.y * gridDim.x;
    double zr = 0.0, zap = 0.0;
    if (n < g.n_nodes) {
        int i, j, k;
        node_ijk(g, n, i, j, k);
        float kx[3], xc[3];
        ku_gather<float, kApplyK>(g, ks, ep, x, nullptr, i, j, k, first_layer(g, n0), kx, xc);
        double z[3];
        float bi[kBinv], rv[3], zv[3];
        bload(binv, n, bi);
#pragma unroll
        for (int c = 0; c < 3; c++) rv[c] = b[3 * n + c] - kx[c];
        bapply(bi, rv, zv);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int dof = 3 * n + c;
            const float dn = cd * d[dof] + cr * zv[c];
            z[c] = static_cast<double>(xc[c] + dn);
        }
        const int cl = df.m ? df.ncl[n] : -1;
        if (cl >= 0) defl_add(df, df.coef + static_cast<long long>(blockIdx.y) * kDeflStride, n, cl, z);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int dof = 3 * n + c;
            z64[dof] = z[c];
            zr = __fma_rn(z[c], r64[dof], zr);
            zap = __fma_rn(z[c], ap64[dof], zap);
        }
    }
    block_sum256x2(zr, zap, red);
    if (threadIdx.x == 0) {
        part_zr[blockIdx.x] = zr;
        part_zap[blockIdx.x] = zap;
    }
}

__global__ __launch_bounds__(256, 2) void k_gmg_residual(Dims g, KeSym<float> ks, const float *__restrict__ ep,
                                                         const float *__restrict__ x, const float *__restrict__ b,
                                                         const float *__restrict__ inv, float *__restrict__ r) {
    const int n0 = xcd_block(blockIdx.x, gridDim.x) * 256, n = n0 + threadIdx.x;
    const long long o = static_cast<long long>(blockIdx.y) * g.n_dofs;
    if (n < g.n_nodes) residual_node(g, ks, ep, x + o, b + o, inv, r + o, n, first_layer(g, n0));
}

// Element-centric forms of the finest level's three gather kernels (ec_tile_run: each element's product once,
// 8x8 element-column tiles, KZ node layers per tile, four tiles per 256-thread block); the same epilogues, with
// their loads issued at the start of the step that completes the node.
__device__ __forceinline__ int ec_wave_tile() {
    return __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
}

struct ChebFin {  // x' = x + d', d' = cd d + cr B^-1 (b - K x)
    const float *b, *binv;
    float *d, *x_out;
    float cd, cr;
    struct Pre {
        float b[3], d[3], bi[kBinv];
    };
    __device__ Pre pre(int n) const {
        Pre p;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            p.b[c] = b[3 * n + c];
            p.d[c] = d[3 * n + c];
        }
        bload(binv, n, p.bi);
        return p;
    }
    __device__ void operator()(int n, float (&kx)[3], const float (&xc)[3], const Pre &p) const {
        float r[3], z[3];
#pragma unroll
        for (int c = 0; c < 3; c++) r[c] = p.b[c] - kx[c];
        bapply(p.bi, r, z);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float dn = cd * p.d[c] + cr * z[c];
            x_out[3 * n + c] = xc[c] + dn;
            d[3 * n + c] = dn;
        }
    }
};

struct ResidualFin {  // r = b - K x on unconstrained DOFs, 0 on constrained ones
    const float *b, *inv;
    float *r;
    struct Pre {
        float b[3], inv[3];
    };
    __device__ Pre pre(int n) const {
        Pre p;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            p.b[c] = b[3 * n + c];
            p.inv[c] = inv[3 * n + c];
        }
        return p;
    }
    __device__ void operator()(int n, float (&kx)[3], const float (&)[3], const Pre &p) const {
#pragma unroll
        for (int c = 0; c < 3; c++) r[3 * n + c] = p.inv[c] > 0.0f ? p.b[c] - kx[c] : 0.0f;
    }
};

struct FinalFin {  // the last smoothing step written as the FP64 z, with the z.r64 and z.ap64 contributions
    const float *b, *d, *binv;
    double *z64;
    const double *r64, *ap64;
    float cd, cr;
    Defl df;
    const double *coef;  // this right-hand side's island coefficients
    double zr = 0.0, zap = 0.0;
    struct Pre {
        float b[3], d[3], bi[kBinv];
        double r[3], ap[3];
        int cl;
    };
    __device__ Pre pre(int n) const {
        Pre p;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            p.b[c] = b[3 * n + c];
            p.d[c] = d[3 * n + c];
            p.r[c] = r64[3 * n + c];
            p.ap[c] = ap64[3 * n + c];
        }
        bload(binv, n, p.bi);
        p.cl = df.m ? df.ncl[n] : -1;
        return p;
    }
    __device__ void operator()(int n, float (&kx)[3], const float (&xc)[3], const Pre &p) {
        double z[3];
        float rv[3], zv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) rv[c] = p.b[c] - kx[c];
        bapply(p.bi, rv, zv);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float dn = cd * p.d[c] + cr * zv[c];
            z[c] = static_cast<double>(xc[c] + dn);
        }
        if (p.cl >= 0) defl_add(df, coef, n, p.cl, z);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            z64[3 * n + c] = z[c];
            zr = __fma_rn(z[c], p.r[c], zr);
            zap = __fma_rn(z[c], p.ap[c], zap);
        }
    }
};

// The finest level's iterate as the tile march's source, with the outer iteration's halt flag of this launch's
// right-hand side (HALT instantiations, launched for several right-hand sides): a converged case's V-cycles end
// each wave's march before its arithmetic while the other cases run; nothing reads their outputs, and the flag
// cannot change while a V-cycle runs. A single solve launches the instantiations without the test: there it
// cost more than the iterations past convergence it skips.
struct FineX {
    const float *x;
    int halted;
    using Raw = float;
    __device__ Raw load(size_t at, uint32_t o, int d) const { return ld_off(x + at, o, d * 4); }
    __device__ float value(Raw r) const { return r; }
    template <typename Fin>
    __device__ bool ready(Fin &) const {
        return !halted;
    }
};

template <bool HALT, int KZ, typename Fin>
__device__ __forceinline__ void fine_march(const Dims &g, const KeMono<float> &km, const float *ep, const float *x,
                                           const int *halt, int halt_stride, int tile, Fin &fin) {
    if constexpr (HALT) {
        FineX src{x, halt[blockIdx.y * halt_stride]};
        ec_tile_march<float, KZ>(g, km, ep, src, tile, fin);
    } else {
        ec_tile_run<float, false, KZ>(g, km, ep, x, nullptr, tile, fin);
    }
}

template <int KZ, bool HALT>
__global__ __launch_bounds__(256) void k_gmg_cheb_ec(Dims g, KeMono<float> km, const float *__restrict__ ep,
                                                     const float *__restrict__ x, const float *__restrict__ b,
                                                     float *__restrict__ d, const float *__restrict__ binv, float cd,
                                                     float cr, float *__restrict__ x_out, int n_tiles,
                                                     const int *__restrict__ halt, int halt_stride) {
    const long long o = static_cast<long long>(blockIdx.y) * g.n_dofs;
    const int tile = ec_wave_tile();
    if (tile >= n_tiles) return;  // wave-uniform; no barrier follows
    ChebFin fin{b + o, binv, d + o, x_out + o, cd, cr};
    fine_march<HALT, KZ>(g, km, ep, x + o, halt, halt_stride, tile, fin);
}

template <int KZ, bool HALT>
__global__ __launch_bounds__(256) void k_gmg_residual_ec(Dims g, KeMono<float> km, const float *__restrict__ ep,
                                                         const float *__restrict__ x, const float *__restrict__ b,
                                                         const float *__restrict__ inv, float *__restrict__ r,
                                                         int n_tiles, const int *__restrict__ halt,
                                                         int halt_stride) {
    const long long o = static_cast<long long>(blockIdx.y) * g.n_dofs;
    const int tile = ec_wave_tile();
    if (tile >= n_tiles) return;
    ResidualFin fin{b + o, inv, r + o};
    fine_march<HALT, KZ>(g, km, ep, x + o, halt, halt_stride, tile, fin);
}

template <int KZ, bool HALT>
__global__ __launch_bounds__(256) void k_gmg_cheb_final_ec(Dims g, KeMono<float> km, const float *__restrict__ ep,
                                                           const float *__restrict__ x, const float *__restrict__ b,
                                                           const float *__restrict__ d, const float *__restrict__ binv,
                                                           float cd, float cr, double *__restrict__ z64,
                                                           const double *__restrict__ r64,
                                                           const double *__restrict__ ap64,
                                                           double *__restrict__ part_zr, double *__restrict__ part_zap,
                                                           int n_tiles, const int *__restrict__ halt,
                                                           int halt_stride, Defl df) {
    __shared__ double red[8];
    const long long o = static_cast<long long>(blockIdx.y) * g.n_dofs;
    FinalFin fin{b + o, d + o, binv, z64 + o, r64 + o, ap64 + o, cd, cr, df,
                 df.coef + static_cast<long long>(blockIdx.y) * kDeflStride};
    const int tile = ec_wave_tile();
    if (tile < n_tiles) fine_march<HALT, KZ>(g, km, ep, x + o, halt, halt_stride, tile, fin);
    double zr = fin.zr, zap = fin.zap;
    block_sum256x2(zr, zap, red);
    if (threadIdx.x == 0) {
        part_zr[blockIdx.y * gridDim.x + blockIdx.x] = zr;
        part_zap[blockIdx.y * gridDim.x + blockIdx.x] = zap;
    }
}

// Slab-split forms of the three gather kernels (two lanes per node, ku_gather_half): a 256-thread block covers
// 128 nodes, wave w of it nodes t0 + 32 w .. + 31, lanes 0-31 the elements below their node and lanes 32-63
// the elements above; the two halves meet with one cross-lane add. Against one lane per node this halves
// the serial instruction stream of a wave and needs one memory round trip instead of two, which is what
// bounds these launches on the coarse levels (few waves per SIMD).
constexpr int kSplitNodes = 128;  // nodes per 256-thread block

// Combined K x (kx) and own x values (xc) for the node of this lane pair; true on the lane that finishes it.
__device__ __forceinline__ bool split_gather(const Dims &g, const KeSym<float> &ks, const float *__restrict__ ep,
                                             const float *__restrict__ x, int t0, float (&kx)[3], float (&xc)[3],
                                             int &n) {
    const int lane = threadIdx.x & 63, half = lane >> 5;
    n = t0 + (threadIdx.x >> 6) * 32 + (lane & 31);
#pragma unroll
    for (int c = 0; c < 3; c++) kx[c] = xc[c] = 0.0f;
    if (n < g.n_nodes) {
        int i, j, k;
        node_ijk(g, n, i, j, k);
        ku_gather_half<float>(g, ks, ep, x, i, j, k, first_layer(g, t0), half, kx, xc);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) kx[c] += __shfl_down(kx[c], 32, 64);  // every lane takes part
    return half == 0 && n < g.n_nodes;
}

// Block bx of nb (the XCD remap runs over nb) of one right-hand side's vectors.
__device__ __forceinline__ void cheb_s_body(const Dims &g, const KeSym<float> &ks, const float *__restrict__ ep,
                                            const float *__restrict__ x, const float *__restrict__ b,
                                            float *__restrict__ d, const float *__restrict__ binv, float cd, float cr,
                                            float *__restrict__ x_out, int bx, int nb) {
    float kx[3], xc[3];
    int n;
    if (split_gather(g, ks, ep, x, xcd_block(bx, nb) * kSplitNodes, kx, xc, n)) {
        float bi[kBinv], r[3], z[3];
        bload(binv, n, bi);
#pragma unroll
        for (int c = 0; c < 3; c++) r[c] = b[3 * n + c] - kx[c];
        bapply(bi, r, z);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int dof = 3 * n + c;
            const float dn = cheb_dn(cd, d[dof], cr, z[c]);
            x_out[dof] = xc[c] + dn;
            d[dof] = dn;
        }
    }
}

__global__ __launch_bounds__(256) void k_gmg_cheb_s(Dims g, KeSym<float> ks, const float *__restrict__ ep,
                                                    const float *__restrict__ x, const float *__restrict__ b,
                                                    float *__restrict__ d, const float *__restrict__ binv, float cd,
                                                    float cr, float *__restrict__ x_out) {
    const long long o = static_cast<long long>(blockIdx.y) * g.n_dofs;
    cheb_s_body(g, ks, ep, x + o, b + o, d + o, binv, cd, cr, x_out + o, blockIdx.x, gridDim.x);
}

// ---- Two Chebyshev steps of a launched coarse level in one launch (the post-smoothing pair) ----
// A block owns a T^3 tile of nodes (T = 4: 512 threads; T = 2: 256). It stages the iterate on the tile and two node layers around
// it in LDS (the prolongation from the level below, when that level is launched too, applied on the staged
// copy), forms the first step on the tile and one layer around it (the neighbours the second step reads; the
// halo's nodes are formed again by the neighbouring tiles) and the second step on the tile. Every global load
// (the staged iterate, and b, B^-1, d and the moduli of both steps' nodes) is issued before the first barrier.
// Each node's prolongation sums are k_gmg_prolong's, its half sums ku_gather_half's (from LDS) and its updates
// cheb_s_body's: the same bits as k_gmg_prolong and two k_gmg_cheb_s launches. The level's d is not written:
// after the post-smoothing it is dead until the next cycle's restriction sets it again.
// Tile edge T: first-step frame A = T + 2 and staged frame X = T + 4 node edges; NT threads.
template <int T>
struct Cheb2Tile {
    static constexpr int A = T + 2, X = T + 4;
    static constexpr int NT = ((X * X * X > 2 * A * A * A ? X * X * X : 2 * A * A * A) + 63) / 64 * 64;
    static_assert(X * X * X <= NT && A * A * A <= NT / 2 && NT <= 1024, "one staged node per thread, one "
                  "first-step node per lane pair");
};

// The moduli of the four elements of half `half` at node (i, j, k), as ku_gather_half reads them.
__device__ __forceinline__ void half_moduli(const Dims &g, const float *__restrict__ ep, int i, int j, int k,
                                            int half, float (&E)[4]) {
    const int e0 = (i + 1) + g.px * ((j + 1) + g.py * (k + half));
#pragma unroll
    for (int q = 0; q < 4; q++) E[q] = ep[e0 - (corner_x(q) ? 1 : 0) - (corner_y(q) ? g.px : 0)];
}

// ku_gather_half at global node (i, j, k) (its neighbours clamped to the grid as there) with x from an LDS frame
// of edge S in which the node sits at (li, lj, lk).
template <int S>
__device__ __forceinline__ void gather_half_lds(const Dims &g, const KeSym<float> &ks, const float *xl,
                                                const float (&E)[4], int i, int j, int k, int li, int lj, int lk,
                                                int half, float (&kx)[3], float (&xc)[3]) {
    const int xs[3] = {i > 0 ? li - 1 : li, li, i < g.nx ? li + 1 : li};
    const int ys[3] = {j > 0 ? lj - 1 : lj, lj, j < g.ny ? lj + 1 : lj};
    const int kl = half ? (k < g.nz ? lk + 1 : lk) : (k > 0 ? lk - 1 : lk);  // the (mirrored) lower layer
    float lo[9][3], up[9][3];
#pragma unroll
    for (int dy = 0; dy < 3; dy++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
            const int ml = 3 * (xs[dx] + S * (ys[dy] + S * kl)), mu = 3 * (xs[dx] + S * (ys[dy] + S * lk));
#pragma unroll
            for (int c = 0; c < 3; c++) {
                lo[dx + 3 * dy][c] = xl[ml + c];
                up[dx + 3 * dy][c] = xl[mu + c];
            }
        }
#pragma unroll
    for (int c = 0; c < 3; c++) xc[c] = up[4][c];
    const float sgn = half ? -1.0f : 1.0f;
#pragma unroll
    for (int q = 0; q < 9; q++) {
        lo[q][2] *= sgn;
        up[q][2] *= sgn;
    }
    float acc0 = 0, acc1 = 0, acc2 = 0;
    corner_rows_slab<4>(ks, lo, up, E[0], acc0, acc1, acc2);
    corner_rows_slab<5>(ks, lo, up, E[1], acc0, acc1, acc2);
    corner_rows_slab<6>(ks, lo, up, E[2], acc0, acc1, acc2);
    corner_rows_slab<7>(ks, lo, up, E[3], acc0, acc1, acc2);
    kx[0] = acc0;
    kx[1] = acc1;
    kx[2] = acc2 * sgn;
}

__device__ __forceinline__ bool node_inside(const Dims &g, int i, int j, int k) {
    return i >= 0 && i <= g.nx && j >= 0 && j <= g.ny && k >= 0 && k <= g.nz;
}

// Steps (cd0, cr0) and (cd1, cr1) from x into x_out; grid (tiles, right-hand sides), tx, ty: tiles along x, y.
// PRO: the staged iterate is x + P xcv on the unconstrained DOFs (inv > 0), the coarse correction of level gc
// prolonged as k_gmg_prolong forms it (one launch fewer; x itself is left as it was).
template <int T, bool PRO>
__global__ __launch_bounds__(Cheb2Tile<T>::NT) void k_gmg_cheb2_s(Dims g, KeSym<float> ks, const float *__restrict__ ep,
                                                            const float *__restrict__ x, const float *__restrict__ b,
                                                            const float *__restrict__ d,
                                                            const float *__restrict__ binv, float cd0, float cr0,
                                                            float cd1, float cr1, float *__restrict__ x_out, int tx,
                                                            int ty, const float *__restrict__ inv, Dims gc,
                                                            const float *__restrict__ xcv) {
    constexpr int kC2T = T, kC2A = Cheb2Tile<T>::A, kC2X = Cheb2Tile<T>::X;
    __shared__ float xs[3 * kC2X * kC2X * kC2X];
    __shared__ float x1[3 * kC2A * kC2A * kC2A], d1[3 * kC2A * kC2A * kC2A];
    const long long o = static_cast<long long>(blockIdx.y) * g.n_dofs;
    x += o;
    b += o;
    d += o;
    x_out += o;
    if (PRO) xcv += static_cast<long long>(blockIdx.y) * gc.n_dofs;
    const int tile = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = kC2T * (tile % tx), j0 = kC2T * ((tile / tx) % ty), k0 = kC2T * (tile / (tx * ty));
    const int t = threadIdx.x, lane = t & 63, half = lane >> 5, slot = (t >> 6) * 32 + (lane & 31);

    // Staged frame: node (i0 - 2 + sx, j0 - 2 + sy, k0 - 2 + sz) at sx + kC2X (sy + kC2X sz).
    const int sx = t % kC2X, sy = (t / kC2X) % kC2X, sz = t / (kC2X * kC2X);
    const bool vs = t < kC2X * kC2X * kC2X && node_inside(g, i0 - 2 + sx, j0 - 2 + sy, k0 - 2 + sz);
    float xv[3] = {0.0f, 0.0f, 0.0f};
    if (vs) {
        const int m = 3 * ((i0 - 2 + sx) + g.nx1 * ((j0 - 2 + sy) + g.ny1 * (k0 - 2 + sz)));
#pragma unroll
        for (int c = 0; c < 3; c++) xv[c] = x[m + c];
        if (PRO) {
            float iv[3], sv[3];
#pragma unroll
            for (int c = 0; c < 3; c++) iv[c] = inv[m + c];
            prolong_sums(gc, xcv, i0 - 2 + sx, j0 - 2 + sy, k0 - 2 + sz, sv);
#pragma unroll
            for (int c = 0; c < 3; c++)
                if (iv[c] > 0.0f) xv[c] = xv[c] + sv[c];
        }
    }
    // First step: lane pair `slot` of the frame around the tile, node (i0 - 1 + ax, ...).
    const int ax = slot % kC2A, ay = (slot / kC2A) % kC2A, az = slot / (kC2A * kC2A);
    const int ai = i0 - 1 + ax, aj = j0 - 1 + ay, ak = k0 - 1 + az;
    const bool va = slot < kC2A * kC2A * kC2A && node_inside(g, ai, aj, ak);
    const int an = ai + g.nx1 * (aj + g.ny1 * ak);
    float ea[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ba[3] = {0.0f, 0.0f, 0.0f}, da[3] = {0.0f, 0.0f, 0.0f}, bia[kBinv] = {};
    if (va) {
        half_moduli(g, ep, ai, aj, ak, half, ea);
        if (half == 0) {
            bload(binv, an, bia);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ba[c] = b[3 * an + c];
                da[c] = d[3 * an + c];
            }
        }
    }
    // Second step: lane pair `slot` of the tile, node (i0 + bx_, ...).
    const int bx_ = slot % kC2T, by_ = (slot / kC2T) % kC2T, bz_ = slot / (kC2T * kC2T);
    const int bi_ = i0 + bx_, bj_ = j0 + by_, bk_ = k0 + bz_;
    const bool vb = slot < kC2T * kC2T * kC2T && node_inside(g, bi_, bj_, bk_);
    const int bn = bi_ + g.nx1 * (bj_ + g.ny1 * bk_);
    float eb[4] = {0.0f, 0.0f, 0.0f, 0.0f}, bb[3] = {0.0f, 0.0f, 0.0f}, bib[kBinv] = {};
    if (vb) {
        half_moduli(g, ep, bi_, bj_, bk_, half, eb);
        if (half == 0) {
            bload(binv, bn, bib);
#pragma unroll
            for (int c = 0; c < 3; c++) bb[c] = b[3 * bn + c];
        }
    }
    if (t < kC2X * kC2X * kC2X) {
#pragma unroll
        for (int c = 0; c < 3; c++) xs[3 * t + c] = xv[c];
    }
    __syncthreads();

    float kx[3] = {0.0f, 0.0f, 0.0f}, xc[3] = {0.0f, 0.0f, 0.0f};
    if (va) gather_half_lds<kC2X>(g, ks, xs, ea, ai, aj, ak, ax + 1, ay + 1, az + 1, half, kx, xc);
#pragma unroll
    for (int c = 0; c < 3; c++) kx[c] += __shfl_down(kx[c], 32, 64);  // every lane takes part
    if (va && half == 0) {
        float r[3], z[3];
#pragma unroll
        for (int c = 0; c < 3; c++) r[c] = ba[c] - kx[c];
        bapply(bia, r, z);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float dn = cheb_dn(cd0, da[c], cr0, z[c]);
            x1[3 * slot + c] = xc[c] + dn;
            d1[3 * slot + c] = dn;
        }
    }
    __syncthreads();

#pragma unroll
    for (int c = 0; c < 3; c++) kx[c] = xc[c] = 0.0f;
    const int as = (bx_ + 1) + kC2A * ((by_ + 1) + kC2A * (bz_ + 1));  // the tile node's first-step slot
    if (vb) gather_half_lds<kC2A>(g, ks, x1, eb, bi_, bj_, bk_, bx_ + 1, by_ + 1, bz_ + 1, half, kx, xc);
#pragma unroll
    for (int c = 0; c < 3; c++) kx[c] += __shfl_down(kx[c], 32, 64);
    if (vb && half == 0) {
        float r[3], z[3];
#pragma unroll
        for (int c = 0; c < 3; c++) r[c] = bb[c] - kx[c];
        bapply(bib, r, z);
#pragma unroll
        for (int c = 0; c < 3; c++) x_out[3 * bn + c] = xc[c] + cheb_dn(cd1, d1[3 * as + c], cr1, z[c]);
    }
}

// Whether the post-smoothing pair of a launched coarse level (and the prolongation before it) runs as one
// k_gmg_cheb2_s (MYCEL_GMG_CHEB2=0: k_gmg_prolong and two k_gmg_cheb_s launches, the same bits).
bool gmg_cheb2() {
    static const bool v = [] {
        const char *e = std::getenv("MYCEL_GMG_CHEB2");
        return !(e && *e == '0');
    }();
    return v;
}

// Levels with fewer nodes than this take 2x2x2 tiles in k_gmg_cheb2_s (more blocks), the others 4x4x4 (fewer
// recomputed halo nodes) (MYCEL_GMG_CHEB2_SMALL=n). Holed box, paired runs: 2x2x2 tiles on its 2,541-node
// level 165.7-165.8 vs 165.3-165.5 solves/s; on the 18,081-node level as well 159.7.
int gmg_cheb2_small() {
    static const int v = [] {
        const char *e = std::getenv("MYCEL_GMG_CHEB2_SMALL");
        return e && *e ? std::atoi(e) : 8192;
    }();
    return v;
}

__global__ __launch_bounds__(256) void k_gmg_residual_s(Dims g, KeSym<float> ks, const float *__restrict__ ep,
                                                        const float *__restrict__ x, const float *__restrict__ b,
                                                        const float *__restrict__ inv, float *__restrict__ r) {
    const long long o = static_cast<long long>(blockIdx.y) * g.n_dofs;
    x += o;
    b += o;
    r += o;
    float kx[3], xc[3];
    int n;
    if (split_gather(g, ks, ep, x, xcd_block(blockIdx.x, gridDim.x) * kSplitNodes, kx, xc, n)) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int dof = 3 * n + c;
            r[dof] = inv[dof] > 0.0f ? b[dof] - kx[c] : 0.0f;
        }
    }
}

// Whether the residual on every multi-block level and the Chebyshev steps below the finest level use the
// slab-split gathers (MYCEL_GMG_SPLIT=0 turns them off). Measured per kernel on the holed box: residual 7.8 ->
// 7.4 us on the finest level and ~4.4 -> ~4.1 us below it, coarse Chebyshev steps ~4.5 -> ~4.4 us; on the
// finest level the Chebyshev steps, which run ~2 waves per SIMD, were no faster split.
bool gmg_split() {
    static const bool v = [] {
        const char *e = std::getenv("MYCEL_GMG_SPLIT");
        return !(e && *e == '0');
    }();
    return v;
}

// b0 = float(r64) on unconstrained DOFs; the finest pre-smoother's first step x = d = inv b0 / theta.
// One thread per node (the block smoother's first step needs the node's three components). n: nodes.
__global__ __launch_bounds__(256) void k_gmg_init_fine(int n, const double *__restrict__ r64,
                                                       const float *__restrict__ inv, const float *__restrict__ binv,
                                                       float theta_inv, float *__restrict__ b, float *__restrict__ x,
                                                       float *__restrict__ d) {
    const int node = blockIdx.x * 256 + threadIdx.x;
    if (node >= n) return;
    const long long o = static_cast<long long>(blockIdx.y) * 3 * n;
    const float sv[3] = {static_cast<float>(r64[o + 3 * node]), static_cast<float>(r64[o + 3 * node + 1]),
                         static_cast<float>(r64[o + 3 * node + 2])};
    first_step(inv, binv, theta_inv, sv, b + o, x + o, d + o, node);
}


// Restriction from the finest level: four lanes per coarse node (64 nodes per block), lane c < 3 summing component
// c over the 27 taps (all 27 loads first, then the sum: one memory round trip), lane 3 idle; the node's three sums
// meet by lane shuffles and lane c writes component c of the coarse pre-smoother's first step b, x = d = B^-1 b /
// theta. The same sums in the same order as the tail's restrict_node.
__device__ __forceinline__ void restrict_dof_body(const Dims &gf, const float *__restrict__ rf, const Dims &gc,
                                                  const float *__restrict__ invc, const float *__restrict__ binvc,
                                                  float theta_inv, float *__restrict__ bc, float *__restrict__ xcv,
                                                  float *__restrict__ dc, int bx) {
    const int q = threadIdx.x & 3, c = q < 3 ? q : 2;
    const int n = (bx * 256 + threadIdx.x) >> 2;
    const bool live = n < gc.n_nodes;
    float sum = 0.0f;
    if (live) {
        int I, J, K;
        node_ijk(gc, n, I, J, K);
        int ix[3], iy[3], iz[3];
        float wx[3], wy[3], wz[3];
        restrict_taps(I, gf.nx, ix, wx);
        restrict_taps(J, gf.ny, iy, wy);
        restrict_taps(K, gf.nz, iz, wz);
        float v[27];
#pragma unroll
        for (int t = 0; t < 27; t++) v[t] = rf[3 * (ix[t % 3] + gf.nx1 * (iy[(t / 3) % 3] + gf.ny1 * iz[t / 9])) + c];
#pragma unroll
        for (int t = 0; t < 27; t++) {
            const float w = wz[t / 9] * wy[(t / 3) % 3] * wx[t % 3];
            sum += w != 0.0f ? w * v[t] : 0.0f;
        }
    }
    const int lead = (threadIdx.x & 63) & ~3;
    float bv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) bv[k] = __shfl(sum, lead + k, 64);  // every lane takes part
    if (live && q < 3) {
        float bi[kBinv], z[3];
        bload(binvc, n, bi);
#pragma unroll
        for (int k = 0; k < 3; k++) bv[k] = invc[3 * n + k] > 0.0f ? bv[k] : 0.0f;
        bapply(bi, bv, z);
        const float xi = z[q] * theta_inv;
        bc[3 * n + q] = bv[q];
        xcv[3 * n + q] = xi;
        dc[3 * n + q] = xi;
    }
}

__global__ __launch_bounds__(256) void k_gmg_restrict_dof(Dims gf, const float *__restrict__ rf, Dims gc,
                                                          const float *__restrict__ invc,
                                                          const float *__restrict__ binvc, float theta_inv,
                                                          float *__restrict__ bc, float *__restrict__ xcv,
                                                          float *__restrict__ dc) {
    const long long of = static_cast<long long>(blockIdx.y) * gf.n_dofs;
    const long long oc = static_cast<long long>(blockIdx.y) * gc.n_dofs;
    restrict_dof_body(gf, rf + of, gc, invc, binvc, theta_inv, bc + oc, xcv + oc, dc + oc, blockIdx.x);
}

// The residual of a coarse level fused with the restriction to the next one (the kernels of those levels are
// bound by their fixed cost, not by arithmetic). Half-wave h of a block owns coarse node 8 blockIdx.x + h; its
// lane t < 27 forms the residual at restriction tap t (the fine node's own gather, so taps shared by neighbouring
// coarse nodes are formed again), weighted; the half-wave adds its 32 lanes in a fixed butterfly, and the
// restriction's coarse pre-smoothing first step follows (x = d = theta^-1 D^-1 b).
__device__ __forceinline__ void residual_restrict_body(const Dims &gf, const KeSym<float> &ks,
                                                       const float *__restrict__ ep, const float *__restrict__ xf,
                                                       const float *__restrict__ bf, const float *__restrict__ invf,
                                                       const Dims &gc, const float *__restrict__ invc,
                                                       const float *__restrict__ binvc,
                                                       float theta_inv, float *__restrict__ bc,
                                                       float *__restrict__ xcv, float *__restrict__ dc, int bx) {
    const int t = threadIdx.x & 31;
    const int n = (bx * 256 + threadIdx.x) >> 5;
    float v[3] = {0.0f, 0.0f, 0.0f};
    if (n < gc.n_nodes && t < 27) {
        int I, J, K;
        node_ijk(gc, n, I, J, K);
        const int dx = t % 3, dy = (t / 3) % 3, dz = t / 9;
        const int fi = 2 * I - 1 + dx, fj = 2 * J - 1 + dy, fk = 2 * K - 1 + dz;
        if (fi >= 0 && fi <= gf.nx && fj >= 0 && fj <= gf.ny && fk >= 0 && fk <= gf.nz) {
            const float w = (dz == 1 ? 1.0f : 0.5f) * (dy == 1 ? 1.0f : 0.5f) * (dx == 1 ? 1.0f : 0.5f);
            float kx[3], xn[3];
            ku_gather<float, kApplyK>(gf, ks, ep, xf, nullptr, fi, fj, fk, 0, kx, xn);
            const int m = 3 * (fi + gf.nx1 * (fj + gf.ny1 * fk));
#pragma unroll
            for (int c = 0; c < 3; c++) v[c] = w * (invf[m + c] > 0.0f ? bf[m + c] - kx[c] : 0.0f);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) v[c] += __shfl_xor(v[c], off, 32);
    if (t == 0 && n < gc.n_nodes) first_step(invc, binvc, theta_inv, v, bc, xcv, dc, n);
}

__global__ __launch_bounds__(256) void k_gmg_residual_restrict(Dims gf, KeSym<float> ks, const float *__restrict__ ep,
                                                               const float *__restrict__ xf,
                                                               const float *__restrict__ bf,
                                                               const float *__restrict__ invf, Dims gc,
                                                               const float *__restrict__ invc,
                                                               const float *__restrict__ binvc, float theta_inv,
                                                               float *__restrict__ bc, float *__restrict__ xcv,
                                                               float *__restrict__ dc) {
    const long long of = static_cast<long long>(blockIdx.y) * gf.n_dofs, oc = static_cast<long long>(blockIdx.y) * gc.n_dofs;
    residual_restrict_body(gf, ks, ep, xf + of, bf + of, invf, gc, invc, binvc, theta_inv, bc + oc, xcv + oc, dc + oc,
                           blockIdx.x);
}

// Whether the coarse levels use k_gmg_residual_restrict (MYCEL_GMG_FUSE_RR=0: separate residual and restriction).
bool fuse_rr() {
    static const bool v = [] {
        const char *e = std::getenv("MYCEL_GMG_FUSE_RR");
        return !(e && *e == '0');
    }();
    return v;
}

__global__ __launch_bounds__(256) void k_gmg_prolong(Dims gf, float *__restrict__ xf, const float *__restrict__ invf,
                                                     Dims gc, const float *__restrict__ xcv) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n < gf.n_nodes)
        prolong_node(gf, xf + static_cast<long long>(blockIdx.y) * gf.n_dofs, invf, gc,
                     xcv + static_cast<long long>(blockIdx.y) * gc.n_dofs, n);
}

// The small levels of the V-cycle in one workgroup, resident in LDS. On entry every thread loads its share
// of lv[0]'s b, x, d (written by the restriction from the level above) and of the tail image (every tail level's
// moduli, inverse diagonal and block inverses, and the coarsest inverse, in one buffer), all loads in flight
// together (one memory round trip), and stores them into LDS; every
// phase after that (smoothing, residual, restriction, coarsest solve, prolongation) reads and writes LDS
// only, separated by workgroup barriers. At the end the tail prolongs lv[0]'s correction into the level above
// (or, when lv[0] is the finest level, writes lv[0]'s iterate back). A chain of global accesses would cost a
// memory round trip (~1 us) per phase instead. lv[count-1] is the coarsest level, solved with its dense FP32
// inverse (column-major; staged into LDS with the rest when it fits).
constexpr int kMaxTail = 4;  // levels in the tail: from <= 512 nodes to the dense coarsest one it is at most 3
constexpr int kTailThreads = 1024;   // 16 waves (up to 128 VGPRs each): two lanes per node on the tail levels
constexpr int kTailPer = 2;          // a tail level's arrays hold <= kTailPer * kTailThreads floats each
constexpr int kTailNodes = 512;      // levels up to this many nodes run in the tail (larger ones are FMA-bound
                                     // on one CU); MYCEL_GMG_TAIL_NODES lowers it (at most kTailThreads / 2)
constexpr int kTailLdsMax = 144 << 10;  // dynamic LDS of the tail (gfx950: 160 KiB per workgroup)

int tail_nodes() {
    static const int v = [] {
        const char *e = std::getenv("MYCEL_GMG_TAIL_NODES");
        return e && *e ? std::min(std::max(0, std::atoi(e)), kTailThreads / 2) : kTailNodes;
    }();
    return v;
}
// Tail levels with more than kTailThreads / 8 nodes form their products element by element (tail_ec);
// MYCEL_GMG_TAIL_EC=0 keeps the two-lanes-per-node gathers on every tail level.
bool tail_ec() {
    static const bool on = [] {
        const char *e = std::getenv("MYCEL_GMG_TAIL_EC");
        return !(e && *e == '0');
    }();
    return on;
}
struct TailLevel {
    Dims d;
    KeSym<float> ks;
    KeMono<float> km;  // the same Ke in the monomial basis (ec levels)
    int o_e, o_inv, o_b, o_x, o_x2, o_d, o_r;  // float offsets of the level's arrays in LDS
    int o_binv;        // and of its nodal block inverses
    int ec;            // 1: the level's products are formed element by element (tail_ec), into the scratch at o_scr
    int o_scr;
    float rnx1, rny1, rnx, rny;  // 1 / node and element counts along x and y (small_divmod)
};
constexpr int kImgPer = kTailLdsMax / 4 / kTailThreads;  // tail image floats per thread, at most
// lv[0]'s b, x or d, staged into LDS on entry: n floats from src + the RHS offset to LDS offset dst.
struct TailSeg {
    const float *src;
    int n, dst;
};
// The tail's levels [t0, t0 + count) and its LDS image: first the constant part, which is also the layout of the
// hierarchy's tail image buffer (per level: padded moduli (not the coarsest level's: the tail never smooths it),
// D^-1, B^-1; then the coarsest inverse if it fits), then per level the vectors b, x, x2, d, r.
struct TailLayout {
    int t0 = 0, count = 0;
    int o_e[kMaxTail] = {}, o_inv[kMaxTail] = {}, o_binv[kMaxTail] = {}, o_vec[kMaxTail] = {};
    int ec[kMaxTail] = {};  // levels whose products are formed element by element, in the scratch at o_scr
    int o_scr = 0;
    int ainv = -1;  // offset of the coarsest inverse, or -1: read from global memory
    int img_n = 0, lds_floats = 0;
};
struct TailArgs {
    // The staging table comes first and is contiguous, so its scalar loads merge and are all issued before
    // the first vector load (fields scattered over the levels cost one scalar-load latency each).
    TailSeg seg[3];          // lv[0]'s b, x, d (RHS 0; RHS k at k * n_dofs)
    const float *img;        // the constant part of the LDS image (img_n floats to LDS offset 0)
    int img_n;
    int ainv_lds;            // float offset of the coarsest inverse in LDS, or -1 to read it from global memory
    const float *ainv;
    float *out0;             // where lv[0]'s corrected iterate goes (global, RHS layout as above), unless:
    float *pro_x;            // non-null: the level above lv[0] (its iterate, RHS layout as above), into which the
    const float *pro_inv;    //   tail prolongs lv[0]'s correction itself (k_gmg_prolong's arithmetic) from LDS
    Dims pro_d;
    float pro_rnx1, pro_rny1;
    float theta_inv, cd1, cr1;
    int count;
    // The outer iteration's halt flag of right-hand side k at halt[k * halt_stride], or null. A halted solve's
    // tail returns after staging: the V-cycles a batch runs past the halt compute values nothing reads, and the
    // flag cannot change while a V-cycle runs (only the outer-iteration kernels around it write it).
    const int *halt;
    int halt_stride;
    TailLevel lv[kMaxTail];
};

static_assert(sizeof(TailArgs) % 4 == 0 && sizeof(TailArgs) / 4 <= kTailThreads, "tail arguments: one word per thread");

// The tail of right-hand side rhs (its vectors at rhs * n_dofs of each level) from the arguments at targ: the
// kernel's own arguments (one workgroup per right-hand side) or a pack member's table entry.
__device__ __forceinline__ void tail_run(const TailArgs *__restrict__ targ, int rhs) {
    extern __shared__ __align__(16) float sm[];
    constexpr int kParts = kTailThreads / kCoarseMaxDofs, kSpan = (kCoarseMaxDofs + kParts - 1) / kParts;
    __shared__ double bsh[kCoarseMaxDofs];
    __shared__ double psum[kParts][kCoarseMaxDofs];
    // The arguments (~1.6 KB) are copied into LDS with one vector load per thread: read field by field they
    // are scalar loads from the kernel-argument segment, each wait a memory latency, in every phase. The few
    // the staging needs are read as scalar loads first (the constant address space lets the compiler do so for
    // the kernel arguments and for a pack's table alike), so that the staging loads are in flight before the
    // copy, and the copy's round trip overlaps theirs.
    __shared__ TailArgs t;
    const auto *ct = (const __attribute__((address_space(4))) TailArgs *)targ;
    if (ct->count == 0) return;  // an empty slot of a pack (the same value in every thread)
    const long long o0 = static_cast<long long>(rhs) * ct->lv[0].d.n_dofs;
    {
        // The image (<= kImgPer * kTailThreads floats) and lv[0]'s vectors (<= kTailPer * kTailThreads each; both
        // checked on the host) are loaded before the first LDS store, so every load is in flight together. A chunk
        // of kTailThreads floats is loaded only if the array reaches it (a uniform branch), at a clamped index.
        float v[3][kTailPer], w[kImgPer];
        const float *img = gmem(ct->img);
        const int img_n = ct->img_n;
#pragma unroll
        for (int u = 0; u < kImgPer; u++) {
            const int i = threadIdx.x + u * kTailThreads;
            if (u * kTailThreads < img_n) w[u] = img[i < img_n - 1 ? i : img_n - 1];
        }
        int nv[3];
#pragma unroll
        for (int sg = 0; sg < 3; sg++) {
            const float *src = gmem(ct->seg[sg].src) + o0;
            const int n = nv[sg] = ct->seg[sg].n;
#pragma unroll
            for (int u = 0; u < kTailPer; u++) {
                const int i = threadIdx.x + u * kTailThreads;
                if (u * kTailThreads < n) v[sg][u] = src[i < n - 1 ? i : n - 1];
            }
        }
        const int *halt = ct->halt;
        const int halted = halt ? gmem(halt)[rhs * ct->halt_stride] : 0;
        if (threadIdx.x < sizeof(TailArgs) / 4)
            reinterpret_cast<uint32_t *>(&t)[threadIdx.x] = reinterpret_cast<const uint32_t *>(targ)[threadIdx.x];
#pragma unroll
        for (int u = 0; u < kImgPer; u++) {
            const int i = threadIdx.x + u * kTailThreads;
            if (u * kTailThreads < img_n && i < img_n) sm[i] = w[u];
        }
#pragma unroll
        for (int sg = 0; sg < 3; sg++) {
            float *dst = sm + ct->seg[sg].dst;
#pragma unroll
            for (int u = 0; u < kTailPer; u++) {
                const int i = threadIdx.x + u * kTailThreads;
                if (u * kTailThreads < nv[sg] && i < nv[sg]) dst[i] = v[sg][u];
            }
        }
        if (halted) return;  // the same value in every thread of the workgroup
    }
    __syncthreads();
    // Bit l of `flip` set: level l's current iterate is in x2 (its scratch in x).
    unsigned flip = 0;
    auto S = [&](int off) { return sm + off; };
    auto cur = [&](int l) { return sm + (((flip >> l) & 1u) ? t.lv[l].o_x2 : t.lv[l].o_x); };
    auto nxt = [&](int l) { return sm + (((flip >> l) & 1u) ? t.lv[l].o_x : t.lv[l].o_x2); };
    const int last = t.count - 1;
    // One Chebyshev step and the residual of a tail level. On a level of <= kTailThreads / 8 nodes (and on every
    // level with MYCEL_GMG_TAIL_EC=0) each node gets two lanes (lanes w*64 + l and w*64 + l + 32, the elements below
    // and above the node: ku_gather_half), which halves each lane's dependent FMA chain: with a few waves the
    // phase is bound by that chain. Half 0 finishes the node. Larger levels (L.ec) form each element's product
    // once instead (ec_kx below): with 12+ waves on the CU their phases are bound by VALU issue.
    auto split_kx = [&](const TailLevel &L, int l, int &n, float (&kx)[3], float (&xc)[3]) {
        const int lane = threadIdx.x & 63, half = lane >> 5;
        n = (threadIdx.x >> 6) * 32 + (lane & 31);
#pragma unroll
        for (int c = 0; c < 3; c++) kx[c] = xc[c] = 0.0f;
        if (n < L.d.n_nodes) {
            int i, j, k;
            SmallIjk{L.rnx1, L.rny1}(L.d, n, i, j, k);
            ku_gather_half<float>(L.d, L.ks, S(L.o_e), cur(l), i, j, k, 0, half, kx, xc);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) kx[c] += __shfl_down(kx[c], 32, 64);  // every lane takes part
        return half == 0 && n < L.d.n_nodes;
    };
    // An ec level: lane e forms element e's product E Ke x_e once, in the monomial basis (the tile march's
    // transforms for one element: its lower and upper faces, the sparse product, each face's corner values),
    // and stores its 8 corners' values in the scratch (component-major: scr[(3 a + c) ne + e]); after a barrier
    // lane n sums node n's corner values of its <= 8 elements in a fixed order (z, y, x side). About 200
    // operations per element instead of 4 x 72 multiply-adds per half-node lane: these levels' phases are bound
    // by VALU issue, not by their dependent chains.
    auto ec_kx = [&](const TailLevel &L, int l, int &n, float (&kx)[3], float (&xc)[3]) {
        const Dims &g = L.d;
        const float *x = cur(l), *ep = S(L.o_e);
        float *scr = S(L.o_scr);
        const int ne = g.n_elem, e = threadIdx.x;
        if (e < ne) {
            int ex, et, ey, ez;
            small_divmod(e, g.nx, L.rnx, et, ex);
            small_divmod(et, g.ny, L.rny, ez, ey);
            const float E = ep[(ex + 1) + g.px * ((ey + 1) + g.py * (ez + 1))];
            float v[8][3];
#pragma unroll
            for (int a = 0; a < 8; a++) {
                const int m = 3 * ((ex + corner_x(a)) + g.nx1 * ((ey + corner_y(a)) + g.ny1 * (ez + corner_z(a))));
#pragma unroll
                for (int c = 0; c < 3; c++) v[a][c] = x[m + c];
            }
            FaceSums<float> lo[3], hi[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                lo[c] = face_sums(v[0][c], v[1][c], v[2][c], v[3][c]);
                hi[c] = face_sums(v[4][c], v[5][c], v[6][c], v[7][c]);
            }
            float cf[3][7], gm[3][7];
            mono_coefficients(lo, hi, cf);
            mono_product(L.km, E, cf, gm);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float gx = gm[c][0], gy = gm[c][1], gz = gm[c][2], gxy = gm[c][3], gxz = gm[c][4],
                            gyz = gm[c][5], gxyz = gm[c][6];
                const FaceSums<float> flo{-gz, gx - gxz, gy - gyz, gxy - gxyz}, fhi{gz, gx + gxz, gy + gyz, gxy + gxyz};
                float y[8];
                face_corners(flo, y[0], y[1], y[2], y[3]);
                face_corners(fhi, y[4], y[5], y[6], y[7]);
#pragma unroll
                for (int a = 0; a < 8; a++) scr[(3 * a + c) * ne + e] = y[a];
            }
        }
        __syncthreads();
        n = threadIdx.x;
#pragma unroll
        for (int c = 0; c < 3; c++) kx[c] = xc[c] = 0.0f;
        if (n >= g.n_nodes) return false;
        int i, j, k;
        SmallIjk{L.rnx1, L.rny1}(g, n, i, j, k);
#pragma unroll
        for (int c = 0; c < 3; c++) xc[c] = x[3 * n + c];
#pragma unroll
        for (int sz = 0; sz < 2; sz++)
#pragma unroll
            for (int sy = 0; sy < 2; sy++)
#pragma unroll
                for (int sx = 0; sx < 2; sx++) {
                    const int ex = i - 1 + sx, ey = j - 1 + sy, ez = k - 1 + sz;
                    if (ex < 0 || ex >= g.nx || ey < 0 || ey >= g.ny || ez < 0 || ez >= g.nz) continue;
                    // the node is corner (1 - sx, 1 - sy, 1 - sz) of element (ex, ey, ez)
                    const int cx = 1 - sx, cy = 1 - sy, a = 4 * (1 - sz) + (cy ? (cx ? 2 : 3) : cx);
                    const int el = ex + g.nx * (ey + g.ny * ez);
#pragma unroll
                    for (int c = 0; c < 3; c++) kx[c] += scr[(3 * a + c) * ne + el];
                }
        return true;
    };
    auto level_kx = [&](const TailLevel &L, int l, int &n, float (&kx)[3], float (&xc)[3]) {
        return L.ec ? ec_kx(L, l, n, kx, xc) : split_kx(L, l, n, kx, xc);  // L.ec: the same in every thread
    };
    auto cheb_phase = [&](const TailLevel &L, int l, float cd, float cr) {
        {
            int n;
            float kx[3], xc[3];
            if (level_kx(L, l, n, kx, xc)) {
                const float *b = S(L.o_b);
                float *d = S(L.o_d), *xo = nxt(l);
                float bi[kBinv], r[3], z[3];
                bload(S(L.o_binv), n, bi);
#pragma unroll
                for (int c = 0; c < 3; c++) r[c] = b[3 * n + c] - kx[c];
                bapply(bi, r, z);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const int dof = 3 * n + c;
                    const float dn = cd * d[dof] + cr * z[c];
                    xo[dof] = xc[c] + dn;
                    d[dof] = dn;
                }
            }
        }
        __syncthreads();
        flip ^= 1u << l;
    };
    auto residual_phase = [&](const TailLevel &L, int l) {
        {
            int n;
            float kx[3], xc[3];
            if (level_kx(L, l, n, kx, xc)) {
                const float *b = S(L.o_b), *inv = S(L.o_inv);
                float *r = S(L.o_r);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const int dof = 3 * n + c;
                    r[dof] = inv[dof] > 0.0f ? b[dof] - kx[c] : 0.0f;
                }
            }
        }
        __syncthreads();
    };
    for (int l = 0; l < last; l++) {
        const TailLevel &L = t.lv[l], &C = t.lv[l + 1];
        cheb_phase(L, l, t.cd1, t.cr1);
        residual_phase(L, l);
        for (int n = threadIdx.x; n < C.d.n_nodes; n += kTailThreads)
            restrict_node(L.d, S(L.o_r), C.d, S(C.o_inv), S(C.o_binv), t.theta_inv, S(C.o_b), cur(l + 1), S(C.o_d),
                          n, SmallIjk{C.rnx1, C.rny1});
        __syncthreads();
    }
    {
        // Coarsest level: x = A^-1 b with all waves. Thread (part, row) sums columns [part * span, +span) of its
        // row (the inverse is column-major, so a wave's loads are contiguous and all of a thread's loads are
        // in flight together); the kParts partial sums of a row are added in a fixed order.
        const TailLevel &T = t.lv[last];
        const int n = T.d.n_dofs, span = (n + kParts - 1) / kParts;
        float *xo = cur(last);
        if (threadIdx.x < n) bsh[threadIdx.x] = static_cast<double>(sm[T.o_b + threadIdx.x]);
        __syncthreads();
        const int part = threadIdx.x / kCoarseMaxDofs, row = threadIdx.x - part * kCoarseMaxDofs;
        if (part < kParts && row < n) {
            double s = 0.0;
            if (t.ainv_lds >= 0) {  // staged in LDS on entry
                const float *ai = sm + t.ainv_lds;
#pragma unroll
                for (int q = 0; q < kSpan; q++) {
                    const int j = part * span + q;
                    if (q < span && j < n) s = __fma_rn(static_cast<double>(ai[j * n + row]), bsh[j], s);
                }
            } else {
#pragma unroll
                for (int q = 0; q < kSpan; q++) {
                    const int j = part * span + q;
                    if (q < span && j < n)
                        s = __fma_rn(static_cast<double>(gmem(t.ainv)[static_cast<size_t>(j) * n + row]), bsh[j], s);
                }
            }
            psum[part][row] = s;
        }
        __syncthreads();
        if (threadIdx.x < n) {
            double s = psum[0][threadIdx.x];
            for (int p = 1; p < kParts; p++) s += psum[p][threadIdx.x];
            xo[threadIdx.x] = static_cast<float>(s);
        }
        __syncthreads();
    }
    for (int l = last - 1; l >= 0; l--) {
        const TailLevel &L = t.lv[l], &C = t.lv[l + 1];
        for (int n = threadIdx.x; n < L.d.n_nodes; n += kTailThreads)
            prolong_node(L.d, cur(l), S(L.o_inv), C.d, cur(l + 1), n, SmallIjk{L.rnx1, L.rny1});
        __syncthreads();
        cheb_phase(L, l, 0.0f, t.theta_inv);
        cheb_phase(L, l, t.cd1, t.cr1);
    }
    const float *res = cur(0);
    if (t.pro_x) {
        float *xf = t.pro_x + static_cast<long long>(rhs) * t.pro_d.n_dofs;
        for (int n = threadIdx.x; n < t.pro_d.n_nodes; n += kTailThreads)
            prolong_node(t.pro_d, xf, t.pro_inv, t.lv[0].d, res, n, SmallIjk{t.pro_rnx1, t.pro_rny1});
    } else {
        for (int q = threadIdx.x; q < t.lv[0].d.n_dofs; q += kTailThreads) gmem(t.out0)[o0 + q] = res[q];
    }
}

// One workgroup per right-hand side (blockIdx.x).
__global__ __launch_bounds__(kTailThreads) void k_gmg_tail(TailArgs targ) { tail_run(&targ, blockIdx.x); }

// ---- packed V-cycles -------------------------------------------------------------------------------
// The V-cycles of several operators (each with its own grid, hierarchy and vectors) in one launch per phase:
// grid dimension y is the member, its arguments are its entry of a device table, and blocks beyond the member's
// own grid return at entry. A member's blocks do exactly what the blocks of its own launch do (the same block
// indices, the XCD remap over its own block count, the same partial slots), so its results are bit-identical to
// its V-cycle alone. The finest-level kernels and the tail test the member's halt flag, so a member that
// converged within a batch skips the rest of it.

struct PkEc {  // finest level, element-centric: Chebyshev step, residual, last step (z and its partials)
    Dims g;
    KeMono<float> km;
    const float *ep, *x, *b, *inv, *binv;  // inv: the residual's constraint mask; binv: the smoother's B^-1
    float *d, *out;  // Chebyshev: d and x'; residual: out = r
    float cd, cr;
    double *z64;
    const double *r64, *ap64;
    double *part_zr, *part_zap;
    const int *halt;
    int n_tiles, nb;
};
struct PkS {  // slab-split Chebyshev step of a coarse level
    Dims g;
    KeSym<float> ks;
    const float *ep, *x, *b, *binv;
    float *d, *x_out;
    float cd, cr;
    int nb;
};
struct PkRR {  // coarse residual fused with the restriction
    Dims gf, gc;
    KeSym<float> ks;
    const float *ep, *xf, *bf, *invf, *invc, *binvc;
    float *bc, *xcv, *dc;
    float theta_inv;
    int nb;
};
struct PkRD {  // restriction from the finest level, one thread per coarse DOF
    Dims gf, gc;
    const float *rf, *invc, *binvc;
    float *bc, *xcv, *dc;
    float theta_inv;
    int nb;
};
struct PkPro {  // prolongation
    Dims gf, gc;
    float *xf;
    const float *invf, *xcv;
    int nb;
};

// A table entry with every pointer marked global (gmem): read from a device table they would otherwise be generic,
// and the element-centric march's lane exchanges would wait for every flat access in flight.
__device__ __forceinline__ PkEc globalized(PkEc a) {
    a.ep = gmem(a.ep), a.x = gmem(a.x), a.b = gmem(a.b), a.inv = gmem(a.inv), a.binv = gmem(a.binv);
    a.d = gmem(a.d), a.out = gmem(a.out), a.z64 = gmem(a.z64), a.r64 = gmem(a.r64), a.ap64 = gmem(a.ap64);
    a.part_zr = gmem(a.part_zr), a.part_zap = gmem(a.part_zap), a.halt = gmem(a.halt);
    return a;
}
__device__ __forceinline__ PkS globalized(PkS a) {
    a.ep = gmem(a.ep), a.x = gmem(a.x), a.b = gmem(a.b), a.binv = gmem(a.binv), a.d = gmem(a.d);
    a.x_out = gmem(a.x_out);
    return a;
}
__device__ __forceinline__ PkRR globalized(PkRR a) {
    a.ep = gmem(a.ep), a.xf = gmem(a.xf), a.bf = gmem(a.bf), a.invf = gmem(a.invf), a.invc = gmem(a.invc);
    a.binvc = gmem(a.binvc), a.bc = gmem(a.bc), a.xcv = gmem(a.xcv), a.dc = gmem(a.dc);
    return a;
}
__device__ __forceinline__ PkRD globalized(PkRD a) {
    a.rf = gmem(a.rf), a.invc = gmem(a.invc), a.binvc = gmem(a.binvc), a.bc = gmem(a.bc), a.xcv = gmem(a.xcv);
    a.dc = gmem(a.dc);
    return a;
}
__device__ __forceinline__ PkPro globalized(PkPro a) {
    a.xf = gmem(a.xf), a.invf = gmem(a.invf), a.xcv = gmem(a.xcv);
    return a;
}

__device__ __forceinline__ int pk_wave_tile(int nb) {
    return __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, nb) * 4 + (threadIdx.x >> 6));
}

template <int KZ>
__global__ __launch_bounds__(256) void k_pk_cheb_ec(const PkEc *__restrict__ tab) {
    const PkEc a = globalized(tab[blockIdx.y]);
    if (static_cast<int>(blockIdx.x) >= a.nb) return;
    const int tile = pk_wave_tile(a.nb);
    if (tile >= a.n_tiles) return;  // wave-uniform; no barrier follows
    ChebFin fin{a.b, a.binv, a.d, a.out, a.cd, a.cr};
    FineX src{a.x, *a.halt};
    ec_tile_march<float, KZ>(a.g, a.km, a.ep, src, tile, fin);
}

template <int KZ>
__global__ __launch_bounds__(256) void k_pk_residual_ec(const PkEc *__restrict__ tab) {
    const PkEc a = globalized(tab[blockIdx.y]);
    if (static_cast<int>(blockIdx.x) >= a.nb) return;
    const int tile = pk_wave_tile(a.nb);
    if (tile >= a.n_tiles) return;
    ResidualFin fin{a.b, a.inv, a.out};
    FineX src{a.x, *a.halt};
    ec_tile_march<float, KZ>(a.g, a.km, a.ep, src, tile, fin);
}

template <int KZ>
__global__ __launch_bounds__(256) void k_pk_cheb_final_ec(const PkEc *__restrict__ tab) {
    __shared__ double red[8];
    const PkEc a = globalized(tab[blockIdx.y]);
    if (static_cast<int>(blockIdx.x) >= a.nb) return;  // block-uniform, ahead of the block sum's barriers
    const Defl none{nullptr, nullptr, nullptr, 0.0, 0, 0, 0};
    FinalFin fin{a.b, a.d, a.binv, a.z64, a.r64, a.ap64, a.cd, a.cr, none, nullptr};
    const int tile = pk_wave_tile(a.nb);
    if (tile < a.n_tiles) {
        FineX src{a.x, *a.halt};
        ec_tile_march<float, KZ>(a.g, a.km, a.ep, src, tile, fin);
    }
    double zr = fin.zr, zap = fin.zap;
    block_sum256x2(zr, zap, red);
    if (threadIdx.x == 0) {
        a.part_zr[blockIdx.x] = zr;
        a.part_zap[blockIdx.x] = zap;
    }
}

__global__ __launch_bounds__(256) void k_pk_cheb_s(const PkS *__restrict__ tab) {
    const PkS a = globalized(tab[blockIdx.y]);
    if (static_cast<int>(blockIdx.x) >= a.nb) return;
    cheb_s_body(a.g, a.ks, a.ep, a.x, a.b, a.d, a.binv, a.cd, a.cr, a.x_out, blockIdx.x, a.nb);
}

__global__ __launch_bounds__(256) void k_pk_residual_restrict(const PkRR *__restrict__ tab) {
    const PkRR a = globalized(tab[blockIdx.y]);
    if (static_cast<int>(blockIdx.x) >= a.nb) return;
    residual_restrict_body(a.gf, a.ks, a.ep, a.xf, a.bf, a.invf, a.gc, a.invc, a.binvc, a.theta_inv, a.bc, a.xcv,
                           a.dc, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_pk_restrict_dof(const PkRD *__restrict__ tab) {
    const PkRD a = globalized(tab[blockIdx.y]);
    if (static_cast<int>(blockIdx.x) >= a.nb) return;
    restrict_dof_body(a.gf, a.rf, a.gc, a.invc, a.binvc, a.theta_inv, a.bc, a.xcv, a.dc, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_pk_prolong(const PkPro *__restrict__ tab) {
    const PkPro a = globalized(tab[blockIdx.y]);
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (static_cast<int>(blockIdx.x) < a.nb && n < a.gf.n_nodes) prolong_node(a.gf, a.xf, a.invf, a.gc, a.xcv, n);
}

// One workgroup per slot (blockIdx.x), its arguments from the table (an empty slot's entry has count 0).
__global__ __launch_bounds__(kTailThreads) void k_pk_tail(const TailArgs *__restrict__ tab) {
    tail_run(tab + blockIdx.x, 0);
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
    // The fine node the coarse node takes its constraint from: the collocated one (2I, 2J, 2K), clamped to the fine
    // grid. A fine grid with an odd element count along an axis has its last node layer at an odd index, between
    // the last two coarse layers, the outer of which (2I = n + 1) lies outside it; that outer layer takes the
    // constraints of the fine grid's last layer, so a support on that face (or on a face that reaches an odd index
    // further down the hierarchy) still holds every level. Without the clamp such a level had no support at all:
    // singular, and the FP32 V-cycle broke down within a few iterations.
    const int fm = 3 * (min(2 * I, gf.nx) + gf.nx1 * (min(2 * J, gf.ny) + gf.ny1 * min(2 * K, gf.nz)));
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const bool cons = !(se > 0.0f) || !(invf[fm + c] > 0.0f);
        invc[3 * n + c] = cons ? 0.0f : 1.0f / (kdiag * se);
    }
}

// The nodal block inverses of a level (binv_node), once its moduli and constraints are set.
__global__ __launch_bounds__(256) void k_gmg_binv(Dims g, const float *__restrict__ ep, float kd, float ko,
                                                  const float *__restrict__ inv, float *__restrict__ binv) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= g.n_nodes) return;
    float bi[kBinv];
    binv_node(g, ep, kd, ko, inv, n, bi);
#pragma unroll
    for (int t = 0; t < kBinv; t++) binv[kBinv * n + t] = bi[t];
}

// Galerkin blocks of the second-coarsest level: for each of its elements e, B_e = E_e P_e^T Ke P_e (24 x 24),
// where P_e maps the 8 coarsest nodes around e (local index q = qx + 2 qy + 4 qz, origin (ex/2, ey/2, ez/2))
// to e's 8 corners with the trilinear weights, and rows of constrained DOFs of e are zero. One block per
// element.
__global__ __launch_bounds__(256) void k_gmg_galerkin_blocks(Dims gs, const float *__restrict__ eps,
                                                             const float *__restrict__ invs,
                                                             const double *__restrict__ ke, double scale,
                                                             double *__restrict__ blocks) {
    __shared__ double w[8][8];  // w[a][q]: weight of coarse node q at corner a
    __shared__ double T[576];
    __shared__ double mask[24];
    const int e = blockIdx.x;
    const int ex = e % gs.nx, t = e / gs.nx, ey = t % gs.ny, ez = t / gs.ny;
    const double E = static_cast<double>(eps[(ex + 1) + gs.px * ((ey + 1) + gs.py * (ez + 1))]);
    double *out = blocks + static_cast<size_t>(e) * 576;
    if (threadIdx.x < 64) {
        const int a = threadIdx.x >> 3, q = threadIdx.x & 7;
        const int d[3] = {ex + corner_x(a) - 2 * ((ex >> 1) + (q & 1)), ey + corner_y(a) - 2 * ((ey >> 1) + ((q >> 1) & 1)),
                          ez + corner_z(a) - 2 * ((ez >> 1) + (q >> 2))};
        double v = 1.0;
        for (int k = 0; k < 3; k++) v *= d[k] == 0 ? 1.0 : (d[k] == 1 || d[k] == -1) ? 0.5 : 0.0;
        w[a][q] = v;
    } else if (threadIdx.x < 88) {
        const int sdof = threadIdx.x - 64, a = sdof / 3, c = sdof % 3;
        const int node = (ex + corner_x(a)) + gs.nx1 * ((ey + corner_y(a)) + gs.ny1 * (ez + corner_z(a)));
        mask[sdof] = invs[3 * node + c] > 0.0f ? 1.0 : 0.0;
    }
    __syncthreads();
    if (E == 0.0) {
        for (int k = threadIdx.x; k < 576; k += 256) out[k] = 0.0;
        return;
    }
    for (int k = threadIdx.x; k < 576; k += 256) {  // T = Ke P_e
        const int r = k / 24, col = k % 24, q = col / 3, c = col % 3;
        double v = 0.0;
        for (int b = 0; b < 8; b++) v += ke[r * 24 + 3 * b + c] * (w[b][q] * mask[3 * b + c]);
        T[k] = v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 576; k += 256) {  // B = E P_e^T T
        const int row = k / 24, col = k % 24, q = row / 3, c = row % 3;
        double v = 0.0;
        for (int a = 0; a < 8; a++) v += (w[a][q] * mask[3 * a + c]) * T[(3 * a + c) * 24 + col];
        out[k] = E * scale * v;
    }
}

// Coarsest level: assemble the dense matrix into LDS (one row per thread, fixed order; identity rows and zero
// columns on constrained DOFs), invert it by Gauss-Jordan elimination (SPD, no pivoting) and store the
// inverse column-major in FP32 (ainv[j * n + i] = (A^-1)_ij; the V-cycle is FP32: the tail stages half the
// bytes and still accumulates the product in FP64; outer iteration counts were unchanged on every test case). With `blocks` the matrix is the Galerkin product P^T A_s P summed from the second-coarsest
// level's element blocks; without (a single-level hierarchy) it is assembled from Ke and the level's moduli.
// The elimination keeps the matrix in registers: thread (ti, tj) = (t / 8, t % 8) owns rows ti + 128 s and
// columns tj + 8 c; each step broadcasts the pivot column and the scaled pivot row through LDS.
__global__ __launch_bounds__(1024) void k_gmg_coarse_factor(Dims g, const float *__restrict__ ep,
                                                            const double *__restrict__ ke, double scale,
                                                            const float *__restrict__ inv, Dims gs,
                                                            const double *__restrict__ blocks,
                                                            float *__restrict__ ainv) {
    extern __shared__ __align__(16) double A[];
    const int n = g.n_dofs;
    for (int q = threadIdx.x; q < n * n; q += blockDim.x) A[q] = 0.0;
    __syncthreads();
    if (blocks) {
        // Galerkin: one thread per (row node I, neighbour node m) pair sums the 3x3 block A[I][m] over the
        // second-coarsest elements that hold both nodes, in a fixed element order (deterministic, no atomics).
        const int nn = g.n_nodes;
        for (int w = threadIdx.x; w < nn * 27; w += blockDim.x) {
            const int node = w / 27, nb = w - 27 * (w / 27);
            int I, J, K;
            node_ijk(g, node, I, J, K);
            const int mi = I + nb % 3 - 1, mj = J + (nb / 3) % 3 - 1, mk = K + nb / 9 - 1;
            if (mi < 0 || mi > g.nx || mj < 0 || mj > g.ny || mk < 0 || mk > g.nz) continue;
            const int m = mi + g.nx1 * (mj + g.ny1 * mk);
            double acc[3][3] = {};
            for (int ez = max(2 * max(K, mk) - 2, 0); ez <= min(2 * min(K, mk) + 1, gs.nz - 1); ez++)
                for (int ey = max(2 * max(J, mj) - 2, 0); ey <= min(2 * min(J, mj) + 1, gs.ny - 1); ey++)
                    for (int ex = max(2 * max(I, mi) - 2, 0); ex <= min(2 * min(I, mi) + 1, gs.nx - 1); ex++) {
                        const int bx = ex >> 1, by = ey >> 1, bz = ez >> 1;
                        const int q1 = (I - bx) + 2 * (J - by) + 4 * (K - bz);
                        const int q2 = (mi - bx) + 2 * (mj - by) + 4 * (mk - bz);
                        const double *blk = blocks + static_cast<size_t>(ex + gs.nx * (ey + gs.ny * ez)) * 576;
                        for (int c = 0; c < 3; c++)
                            for (int d = 0; d < 3; d++) acc[c][d] += blk[(3 * q1 + c) * 24 + 3 * q2 + d];
                    }
            for (int c = 0; c < 3; c++) {
                const int r = 3 * node + c;
                if (!(inv[r] > 0.0f)) continue;
                for (int d = 0; d < 3; d++)
                    if (inv[3 * m + d] > 0.0f) A[r * n + 3 * m + d] = acc[c][d];
            }
        }
        __syncthreads();
        for (int r = threadIdx.x; r < n; r += blockDim.x)
            if (!(inv[r] > 0.0f) || !(A[r * n + r] > 0.0)) A[r * n + r] = 1.0;  // constrained or vanished row
    } else {
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            const int node = r / 3, c = r % 3;
            double *row = A + r * n;
            if (!(inv[r] > 0.0f)) {
                row[r] = 1.0;
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
                    for (int d = 0; d < 3; d++)
                        if (inv[3 * m + d] > 0.0f) row[3 * m + d] += E * scale * ke[(3 * a + c) * 24 + 3 * b + d];
                }
            }
        }
    }
    __syncthreads();
    // Gauss-Jordan inversion in registers: thread (ti, tj) holds rows ti + 64 s and columns tj + 16 c of the
    // matrix padded with the identity to 64 RS x 16 CS (the padding stays the identity: no pivot reaches it; row
    // slices that hold only padding for this n are skipped; padded columns of a pivot row are 0, so their
    // updates subtract 0 and need no test). Pivot k = 16 C + tk runs with C
    // unrolled, so column k is register a[s][C] of lane tk of each 16-lane group (a lane shuffle fetches it) and
    // row k is in slice s = C / 4. Per pivot, one barrier: the lanes holding row k write it to an LDS buffer
    // (two, alternating with k). Same products and differences as the textbook in-place form: row k <- row k / p
    // (its pivot entry 1/p); row i <- row i - a_ik (row k / p) (its column-k entry -a_ik / p).
    constexpr int RS = (kCoarseMaxDofs + 63) / 64, CS = (kCoarseMaxDofs + 15) / 16;
    static_assert(2 * 16 * CS <= 4 * kCoarseMaxDofs, "row buffers within the launch's LDS");
    const int ti = threadIdx.x >> 4, tj = threadIdx.x & 15, lane = threadIdx.x & 63;
    const int rs_n = (n + 63) >> 6;  // row slices holding rows < n
    double a[RS][CS];
#pragma unroll
    for (int s = 0; s < RS; s++)
#pragma unroll
        for (int c = 0; c < CS; c++) {
            const int i = ti + 64 * s, j = tj + 16 * c;
            a[s][c] = (i < n && j < n) ? A[i * n + j] : (i == j ? 1.0 : 0.0);
        }
    __syncthreads();  // A's LDS now holds the row buffers
#pragma unroll
    for (int C = 0; C < CS; C++) {
        const int SK = C >> 2;  // slice of rows 16 C .. 16 C + 15
        for (int tk = 0; tk < 16 && 16 * C + tk < n; tk++) {
            const int k = 16 * C + tk;
            double *rb = A + (k & 1) * (16 * CS);
            const bool own = ti == (k & 63);  // this lane holds part of row k, in a[SK]
            double col[RS];
#pragma unroll
            for (int s = 0; s < RS; s++) col[s] = __shfl(a[s][C], (lane & ~15) | tk, 64);
            if (own) {
#pragma unroll
                for (int c = 0; c < CS; c++) rb[tj + 16 * c] = a[SK][c];
            }
            __syncthreads();
            const double piv = 1.0 / rb[k];
            const bool kc = tj == tk;  // this lane holds column k (register C)
            double r[CS];
#pragma unroll
            for (int c = 0; c < CS; c++) r[c] = rb[tj + 16 * c] * piv;  // row k / p
            if (kc) r[C] = piv;
#pragma unroll
            for (int s = 0; s < RS; s++) {
                if (s >= rs_n) continue;
#pragma unroll
                for (int c = 0; c < CS; c++) a[s][c] = a[s][c] - col[s] * r[c];
                if (kc) a[s][C] = 0.0 - col[s] * piv;  // column k: -a_ik / p
            }
            if (own) {
#pragma unroll
                for (int c = 0; c < CS; c++) a[SK][c] = r[c];  // row k / p, its pivot entry 1/p
            }
        }
    }
#pragma unroll
    for (int s = 0; s < RS; s++)
#pragma unroll
        for (int c = 0; c < CS; c++) {
            const int i = ti + 64 * s, j = tj + 16 * c;
            if (i < n && j < n) ainv[j * n + i] = static_cast<float>(a[s][c]);  // column-major
        }
}

}  // namespace

struct GmgLevel {
    Dims d{};
    KeSym<float> ks{};
    KeMono<float> km{};  // the same Ke in the monomial basis (element-centric kernels)
    double scale = 1.0;  // h_l / h
    float kdiag = 0.0f;  // diagonal entry of the level's Ke
    float koff = 0.0f;   // off-diagonal entry (x, y) of its corner-0 block
    const float *ep = nullptr;  // padded moduli: the operator's (finest level) or ep_w
    float *ep_w = nullptr;      // a coarse level's padded moduli (written by the coarsening kernel)
    float *inv = nullptr;       // inverse diagonal (the constraint mask)
    float *binv = nullptr;      // the smoother's nodal block inverses, kBinv per node
    // ep_w, inv and binv live in the tail image (Gmg::tail_img) for the tail's levels, else in these buffers
    DevBuf<float> ep_own, inv_own, binv_own, b, x, x2, dv, r;
    int blocks = 0;    // 256 nodes per block (one lane per node)
    int blocks_s = 0;  // kSplitNodes per block (the slab-split gathers)
};

// Island deflation state (host set-up on every density update when the operator enables it).
struct Deflation {
    int m = 0;                  // islands in use
    DevBuf<int> ncl;            // per node: island or -1
    DevBuf<int> off, nodes;     // each island's nodes, ascending (CSR)
    DevBuf<double> cen;         // 3 per island
    DevBuf<double> einv;        // (6m)^2, row-major inverse of Z^T A Z
    DevBuf<double> y, coef;     // kDeflStride per right-hand side
    int rhs_cap = 0;
};

struct Gmg {
    Deflation defl;
    GmgLevel lv[kMaxLevels];
    int levels = 0;
    float *ainv = nullptr;  // coarsest inverse, column-major: in the tail image when the tail stages it, else ainv_own
    DevBuf<float> ainv_own;
    TailLayout tl;           // the tail's levels and its LDS image
    DevBuf<float> tail_img;  // the constant part of that image (tl.img_n floats), in the same order
    DevBuf<double> blocks;  // Galerkin element blocks of the second-coarsest level
    uint64_t epoch = ~0ull;
    float theta_inv = 0;
    float cd[kMaxDegree], cr[kMaxDegree];  // Chebyshev step k >= 1: d' = cd[k] d + cr[k] D^-1 r
    int fine_degree = kFineDegree;
    int smooth_kz = 0;  // tile depth of the finest level's Chebyshev and residual launches (autotuned)
    int rhs_cap = 1;    // right-hand sides the level vectors hold
    uint64_t gen = 0;   // bumped when the level vectors move (captured graphs must be rebuilt)
};

void gmg_reserve(mycel_op *op, int n_rhs) {
    Gmg *g = op->gmg;
    if (n_rhs > g->defl.rhs_cap) {
        g->defl.y.resize(static_cast<size_t>(kDeflStride) * n_rhs);
        g->defl.coef.resize(static_cast<size_t>(kDeflStride) * n_rhs);
        g->defl.rhs_cap = n_rhs;
        g->gen++;
    }
    if (n_rhs <= g->rhs_cap) return;
    for (int l = 0; l < g->levels; l++) {
        GmgLevel &L = g->lv[l];
        const size_t n = static_cast<size_t>(L.d.n_dofs) * n_rhs;
        L.b.resize(n);
        L.x.resize(n);
        L.x2.resize(n);
        L.dv.resize(n);
        L.r.resize(n);
    }
    g->rhs_cap = n_rhs;
    g->gen++;
}

uint64_t gmg_generation(const mycel_op *op) { return op->gmg ? op->gmg->gen : 0; }

void gmg_free(Gmg *g) { delete g; }

// Node layers per tile of the finest level's element-centric kernels (the operator's automatic K*u rule), or 0
// for the node gathers (MYCEL_GMG_FINE_EC=0).
static int fine_ec_kz(const mycel_op *op) {
    static const bool on = [] {
        const char *e = std::getenv("MYCEL_GMG_FINE_EC");
        return !(e && *e == '0');
    }();
    return on ? ku_element_layers(op) : 0;
}

int gmg_partials(const mycel_op *op) {
    const int kz = fine_ec_kz(op);
    return kz ? cdiv(ec_tiles(op->d, kz), 4) : cdiv(op->d.n_nodes, 256);
}

// Whether the grid coarsens by 2 per axis (every axis keeping >= 2 elements) down to <= kCoarseMaxDofs DOFs, the
// hierarchy gmg_build makes.
bool gmg_supported(const mycel_op *op) {
    Dims d = op->d;
    for (int l = 0; l < kMaxLevels; l++) {
        if (d.n_dofs <= kCoarseMaxDofs) return true;
        if (d.nx < 2 || d.ny < 2 || d.nz < 2) return false;
        d = make_dims((d.nx + 1) / 2, (d.ny + 1) / 2, (d.nz + 1) / 2);
    }
    return false;
}

// The tail's levels: from the coarsest up, as many as have <= tail_nodes() nodes (never the finest: it has the
// FP64 entry and exit kernels) and fit the LDS image, at most kMaxTail; the layout of that image.
static TailLayout tail_layout(const Gmg *g) {
    const int top = g->levels - 1;
    auto img_floats = [&](int l) { return (l < top ? g->lv[l].d.n_pad : 0) + 3 * g->lv[l].d.n_dofs; };
    auto level_floats = [&](int l) { return img_floats(l) + 5 * g->lv[l].d.n_dofs; };
    auto fits = [&](int l) { return g->lv[l].d.n_dofs <= kTailPer * kTailThreads; };
    TailLayout tl;
    int t0 = top;
    long long total = level_floats(top);
    while (t0 > 1 && top - t0 + 1 < kMaxTail && g->lv[t0 - 1].d.n_nodes <= tail_nodes() && fits(t0 - 1) &&
           4 * (total + level_floats(t0 - 1)) <= kTailLdsMax) {
        t0--;
        total += level_floats(t0);
    }
    if (top == 0) t0 = 0;
    tl.t0 = t0;
    tl.count = top - t0 + 1;
    if (!fits(top) || 4 * total > kTailLdsMax)
        throw Error(MYCEL_E_INVALID_OPERATION, "multigrid: coarsest level does not fit the tail kernel");
    int off = 0;
    for (int l = t0; l <= top; l++) {
        const int k = l - t0, nd = g->lv[l].d.n_dofs;
        tl.o_e[k] = off;
        if (l < top) off += g->lv[l].d.n_pad;
        tl.o_inv[k] = off;
        off += nd;
        tl.o_binv[k] = off;  // kBinv per node = 2 per DOF
        off += 2 * nd;
    }
    const int nc = g->lv[top].d.n_dofs;
    if (4 * (total + static_cast<long long>(nc) * nc) <= kTailLdsMax) {
        tl.ainv = off;
        off += nc * nc;
    }
    tl.img_n = off;
    for (int l = t0; l <= top; l++) {
        tl.o_vec[l - t0] = off;
        off += 5 * g->lv[l].d.n_dofs;
    }
    // Element-by-element products on the smoothed levels with many nodes (one lane per element, 24 corner values
    // each in a scratch shared by those levels), where their LDS fits.
    tl.o_scr = off;
    int scr = 0;
    for (int l = t0; l < top && tail_ec(); l++) {
        const Dims &d = g->lv[l].d;
        if (d.n_nodes > kTailThreads / 8 && d.n_elem <= kTailThreads &&
            4LL * (off + std::max(scr, 24 * d.n_elem)) <= kTailLdsMax) {
            tl.ec[l - t0] = 1;
            scr = std::max(scr, 24 * d.n_elem);
        }
    }
    tl.lds_floats = off + scr;
    return tl;
}

// Builds the level structure; op->gmg is set only once the whole hierarchy exists.
static void gmg_build(mycel_op *op) {
    std::unique_ptr<Gmg> owner(new Gmg());
    Gmg *g = owner.get();
    // The smoother's preconditioner is the nodal block Jacobi B^-1 (the 3x3 diagonal blocks of A). Every corner
    // block of the cube's Ke is kd on the diagonal and ko s_i s_j off it (s: the corner's side per axis), which the
    // per-node inverses rely on; checked here.
    const double kd = op->ke[0], ko = op->ke[1];
    for (int a = 0; a < 8; a++) {
        const double sgn[3] = {corner_x(a) ? -1.0 : 1.0, corner_y(a) ? -1.0 : 1.0, corner_z(a) ? -1.0 : 1.0};
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) {
                const double want = r == c ? kd : ko * sgn[r] * sgn[c];
                if (std::fabs(op->ke[(3 * a + r) * 24 + 3 * a + c] - want) > 1e-12 * std::fabs(kd))
                    throw Error(MYCEL_E_INVALID_OPERATION, "multigrid: element matrix blocks are not the cube's");
            }
    }
    // Chebyshev interval [lmax / 30, lmax]: lmax = lmax(B_e^-1 K_e) of one element (B_e its block diagonal) bounds
    // lmax(B^-1 A) on every level for every density field: x'Ax / x'Bx = sum_e E_e x_e'K_e x_e / sum_e E_e x_e'B_e x_e.
    // Power iteration on the 24 x 24 pencil; 0.1% added for the iteration's own error.
    double lmax = 0.0;
    {
        double binv_e[8][3][3];
        for (int a = 0; a < 8; a++) {
            double m[3][3];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) m[r][c] = op->ke[(3 * a + r) * 24 + 3 * a + c];
            const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1], c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2],
                         c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
            const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
            const double adj[3][3] = {{c00, m[0][2] * m[2][1] - m[0][1] * m[2][2], m[0][1] * m[1][2] - m[0][2] * m[1][1]},
                                      {c01, m[0][0] * m[2][2] - m[0][2] * m[2][0], m[0][2] * m[1][0] - m[0][0] * m[1][2]},
                                      {c02, m[0][1] * m[2][0] - m[0][0] * m[2][1], m[0][0] * m[1][1] - m[0][1] * m[1][0]}};
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) binv_e[a][r][c] = adj[r][c] / det;
        }
        double v[24], w[24];
        for (int q = 0; q < 24; q++) v[q] = 1.0 + 0.1 * q;
        for (int it = 0; it < 5000; it++) {
            double kv[24];
            for (int r = 0; r < 24; r++) {
                double t = 0.0;
                for (int c = 0; c < 24; c++) t += op->ke[r * 24 + c] * v[c];
                kv[r] = t;
            }
            double nrm = 0.0;
            for (int a = 0; a < 8; a++)
                for (int r = 0; r < 3; r++) {
                    double t = 0.0;
                    for (int c = 0; c < 3; c++) t += binv_e[a][r][c] * kv[3 * a + c];
                    w[3 * a + r] = t;
                    nrm += t * t;
                }
            nrm = std::sqrt(nrm);
            for (int q = 0; q < 24; q++) v[q] = w[q] / nrm;
        }
        double vkv = 0.0, vbv = 0.0;
        for (int r = 0; r < 24; r++) {
            double t = 0.0;
            for (int c = 0; c < 24; c++) t += op->ke[r * 24 + c] * v[c];
            vkv += v[r] * t;
        }
        for (int a = 0; a < 8; a++)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) vbv += v[3 * a + r] * op->ke[(3 * a + r) * 24 + 3 * a + c] * v[3 * a + c];
        lmax = 1.001 * vkv / vbv;
    }
    const double lmin = lmax / smooth_range(), theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    g->theta_inv = static_cast<float>(1.0 / theta);
    double rho = 1.0 / sigma;
    g->cd[0] = 0.0f;
    g->cr[0] = g->theta_inv;
    g->fine_degree = fine_degree();
    for (int k = 1; k < kMaxDegree; k++) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        g->cd[k] = static_cast<float>(rn * rho);
        g->cr[k] = static_cast<float>(2.0 * rn / delta);
        rho = rn;
    }

    Dims d = op->d;
    double scale = 1.0;
    for (int l = 0;; l++) {
        if (l >= kMaxLevels) throw Error(MYCEL_E_INVALID_OPERATION, "multigrid: too many levels");
        GmgLevel &L = g->lv[l];
        L.d = d;
        L.scale = scale;
        for (int q = 0; q < kKeClasses; q++) L.ks.v[q] = static_cast<float>(op->ks32.v[q] * scale);
        {
            const KeMono<double> &m = op->km64;
            L.km = {static_cast<float>(m.a * scale), static_cast<float>(m.b * scale), static_cast<float>(m.c * scale),
                    static_cast<float>(m.d * scale), static_cast<float>(m.e * scale), static_cast<float>(m.f * scale),
                    static_cast<float>(m.h * scale), static_cast<float>(m.q * scale)};
        }
        L.kdiag = static_cast<float>(op->ke[0] * scale);
        L.koff = static_cast<float>(op->ke[1] * scale);
        L.b.resize(d.n_dofs);
        L.x.resize(d.n_dofs);
        L.x2.resize(d.n_dofs);
        L.dv.resize(d.n_dofs);
        L.r.resize(d.n_dofs);
        L.blocks = cdiv(d.n_nodes, 256);
        L.blocks_s = cdiv(d.n_nodes, kSplitNodes);
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
    // The tail's levels keep their constant arrays (moduli, D^-1, B^-1, the coarsest inverse) in one buffer laid out
    // as the constant part of the tail's LDS image: the tail stages it with one contiguous copy. (Gathered from a
    // dozen separate arrays, the staging took ~11k cycles, three times a contiguous copy of the same bytes: the
    // per-array address, clamp and mask arithmetic of 16 waves made it VALU-bound.)
    g->tl = tail_layout(g);
    g->tail_img.resize(g->tl.img_n);
    const int top = g->levels - 1;
    for (int l = 0; l < g->levels; l++) {
        GmgLevel &L = g->lv[l];
        const int k = l - g->tl.t0;
        float *img = g->tail_img.get();
        if (l == 0) {
            L.ep = op->d_ep32.get();
        } else {
            if (k >= 0 && l < top) {
                L.ep_w = img + g->tl.o_e[k];
            } else {
                L.ep_own.resize(L.d.n_pad);
                L.ep_w = L.ep_own.get();
            }
            L.ep = L.ep_w;
        }
        if (k >= 0) {
            L.inv = img + g->tl.o_inv[k];
            L.binv = img + g->tl.o_binv[k];
        } else {
            L.inv_own.resize(L.d.n_dofs);
            L.binv_own.resize(static_cast<size_t>(kBinv) * L.d.n_nodes);
            L.inv = L.inv_own.get();
            L.binv = L.binv_own.get();
        }
    }
    const int nc = g->lv[top].d.n_dofs;
    if (g->tl.ainv >= 0) {
        g->ainv = g->tail_img.get() + g->tl.ainv;
    } else {
        g->ainv_own.resize(static_cast<size_t>(nc) * nc);
        g->ainv = g->ainv_own.get();
    }
    if (g->levels >= 2) g->blocks.resize(static_cast<size_t>(g->lv[g->levels - 2].d.n_elem) * 576);
    op->gmg = owner.release();
}

// Finds the floating solid islands of the current moduli and forms the coarse correction's operator
// (Z^T A Z, inverted by Cholesky) on the host from the element moduli, in a fixed order. Islands beyond
// kMaxIslands (the smallest ones) are left to the iteration.
static void gmg_setup_deflation(mycel_op *op, Gmg *g, hipStream_t s) {
    Deflation &D = g->defl;
    const Dims &d = op->d;
    const int m_old = D.m;
    D.m = 0;
    std::vector<double> ep(d.n_pad);
    MYCEL_HIP(hipMemcpyAsync(ep.data(), op->d_ep.get(), sizeof(double) * d.n_pad, hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipStreamSynchronize(s));
    auto E = [&](int ex, int ey, int ez) { return ep[(ex + 1) + d.px * ((ey + 1) + d.py * (ez + 1))]; };
    const double thr = op->island_ratio * op->e0;
    // Face-connected clusters of solid elements, labelled in ascending element order. Clusters that meet only at
    // an edge or a corner node stay apart: such a joint is a hinge with almost no stiffness, so each side's rigid
    // modes are near-null modes of their own (the bimodal cantilever of seed 11: 266 iterations against 700 with
    // 26-connected clusters). A node shared by two clusters belongs to the one labelled last.
    std::vector<int> lab(d.n_elem, -1), size, stack;
    int ncomp = 0;
    for (int e = 0; e < d.n_elem; e++) {
        const int ex = e % d.nx, ey = (e / d.nx) % d.ny, ez = e / (d.nx * d.ny);
        if (lab[e] >= 0 || !(E(ex, ey, ez) >= thr) || E(ex, ey, ez) == 0.0) continue;
        int cnt = 0;
        lab[e] = ncomp;
        stack.assign(1, e);
        while (!stack.empty()) {
            const int q = stack.back();
            stack.pop_back();
            cnt++;
            const int qx = q % d.nx, qy = (q / d.nx) % d.ny, qz = q / (d.nx * d.ny);
            for (int oz = -1; oz <= 1; oz++)
                for (int oy = -1; oy <= 1; oy++)
                    for (int ox = -1; ox <= 1; ox++) {
                        const int nx_ = qx + ox, ny_ = qy + oy, nz_ = qz + oz;
                        if (std::abs(ox) + std::abs(oy) + std::abs(oz) != 1) continue;
                        if (nx_ < 0 || ny_ < 0 || nz_ < 0 || nx_ >= d.nx || ny_ >= d.ny || nz_ >= d.nz) continue;
                        const int r = nx_ + d.nx * (ny_ + d.ny * nz_);
                        if (lab[r] >= 0 || !(E(nx_, ny_, nz_) >= thr) || E(nx_, ny_, nz_) == 0.0) continue;
                        lab[r] = ncomp;
                        stack.push_back(r);
                    }
        }
        size.push_back(cnt);
        ncomp++;
    }
    std::vector<int> ncl(d.n_nodes, -1);
    for (int e = 0; e < d.n_elem; e++) {
        if (lab[e] < 0) continue;
        const int ex = e % d.nx, ey = (e / d.nx) % d.ny, ez = e / (d.nx * d.ny);
        const int n0 = ex + d.nx1 * (ey + d.ny1 * ez);
        for (int a = 0; a < 8; a++) ncl[n0 + corner_x(a) + d.nx1 * (corner_y(a) + d.ny1 * corner_z(a))] = lab[e];
    }
    std::vector<char> anchored(ncomp, 0);
    for (int n = 0; n < d.n_nodes; n++)
        if (ncl[n] >= 0)
            for (int c = 0; c < 3; c++)
                if (op->dof_flags[3 * n + c] & kFlagConstrained) anchored[ncl[n]] = 1;
    std::vector<int> floating;
    for (int c = 0; c < ncomp; c++)
        if (!anchored[c]) floating.push_back(c);
    std::stable_sort(floating.begin(), floating.end(), [&](int a, int b) { return size[a] > size[b]; });
    if (floating.size() > static_cast<size_t>(kMaxIslands)) floating.resize(kMaxIslands);
    std::vector<int> newid(ncomp, -1);
    std::sort(floating.begin(), floating.end());  // islands numbered in order of their first element
    const int m = static_cast<int>(floating.size());
    for (int k = 0; k < m; k++) newid[floating[k]] = k;
    for (int n = 0; n < d.n_nodes; n++) ncl[n] = ncl[n] >= 0 ? newid[ncl[n]] : -1;
    if (m > 0) {
        const double h = op->h;
        std::vector<int> off(m + 1, 0), nodes;
        std::vector<double> cen(3 * m, 0.0);
        for (int n = 0; n < d.n_nodes; n++)
            if (ncl[n] >= 0) off[ncl[n] + 1]++;
        for (int k = 0; k < m; k++) off[k + 1] += off[k];
        nodes.assign(off[m], 0);
        std::vector<int> fill(off.begin(), off.end() - 1);
        for (int n = 0; n < d.n_nodes; n++) {
            const int c = ncl[n];
            if (c < 0) continue;
            nodes[fill[c]++] = n;
            cen[3 * c] += (n % d.nx1) * h;
            cen[3 * c + 1] += ((n / d.nx1) % d.ny1) * h;
            cen[3 * c + 2] += (n / (d.nx1 * d.ny1)) * h;
        }
        for (int c = 0; c < m; c++)
            for (int q = 0; q < 3; q++) cen[3 * c + q] /= (off[c + 1] - off[c]);
        // Z^T K Z element by element: sum over the elements touching an island of R_a^T (E_e Ke_ab) R_b.
        const int m6 = 6 * m;
        std::vector<double> A(static_cast<size_t>(m6) * m6, 0.0);
        auto rbm = [&](int n, int c, double (&R)[3][6]) {
            const double dx = (n % d.nx1) * h - cen[3 * c], dy = ((n / d.nx1) % d.ny1) * h - cen[3 * c + 1],
                         dz = (n / (d.nx1 * d.ny1)) * h - cen[3 * c + 2];
            const double R_[3][6] = {{1, 0, 0, 0, dz, -dy}, {0, 1, 0, -dz, 0, dx}, {0, 0, 1, dy, -dx, 0}};
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 6; j++) R[i][j] = R_[i][j];
        };
        for (int e = 0; e < d.n_elem; e++) {
            const int ex = e % d.nx, ey = (e / d.nx) % d.ny, ez = e / (d.nx * d.ny);
            const double ee = E(ex, ey, ez);
            if (ee == 0.0) continue;
            const int n0 = ex + d.nx1 * (ey + d.ny1 * ez);
            int nd[8], cl[8];
            bool any = false;
            for (int a = 0; a < 8; a++) {
                nd[a] = n0 + corner_x(a) + d.nx1 * (corner_y(a) + d.ny1 * corner_z(a));
                cl[a] = ncl[nd[a]];
                any |= cl[a] >= 0;
            }
            if (!any) continue;
            for (int a = 0; a < 8; a++) {
                if (cl[a] < 0) continue;
                double Ra[3][6];
                rbm(nd[a], cl[a], Ra);
                for (int b = 0; b < 8; b++) {
                    if (cl[b] < 0) continue;
                    double Rb[3][6], KR[3][6];
                    rbm(nd[b], cl[b], Rb);
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 6; j++) {
                            double acc = 0.0;
                            for (int k = 0; k < 3; k++) acc += op->ke[(3 * a + i) * 24 + 3 * b + k] * Rb[k][j];
                            KR[i][j] = ee * acc;
                        }
                    for (int i = 0; i < 6; i++)
                        for (int j = 0; j < 6; j++) {
                            double acc = 0.0;
                            for (int k = 0; k < 3; k++) acc += Ra[k][i] * KR[k][j];
                            A[static_cast<size_t>(6 * cl[a] + i) * m6 + 6 * cl[b] + j] += acc;
                        }
                }
            }
        }
        // Cholesky A = L L^T, then A^-1 column by column.
        std::vector<double> L(A);
        bool ok = true;
        for (int j = 0; j < m6 && ok; j++) {
            double dsum = L[static_cast<size_t>(j) * m6 + j];
            for (int k = 0; k < j; k++) dsum -= L[static_cast<size_t>(j) * m6 + k] * L[static_cast<size_t>(j) * m6 + k];
            if (!(dsum > 0.0)) {
                ok = false;
                break;
            }
            const double ljj = std::sqrt(dsum);
            L[static_cast<size_t>(j) * m6 + j] = ljj;
            for (int i = j + 1; i < m6; i++) {
                double v = L[static_cast<size_t>(i) * m6 + j];
                for (int k = 0; k < j; k++) v -= L[static_cast<size_t>(i) * m6 + k] * L[static_cast<size_t>(j) * m6 + k];
                L[static_cast<size_t>(i) * m6 + j] = v / ljj;
            }
        }
        if (ok) {
            std::vector<double> inv(static_cast<size_t>(m6) * m6, 0.0), col(m6);
            for (int c = 0; c < m6; c++) {
                for (int i = 0; i < m6; i++) {  // L y = e_c
                    double v = i == c ? 1.0 : 0.0;
                    for (int k = 0; k < i; k++) v -= L[static_cast<size_t>(i) * m6 + k] * col[k];
                    col[i] = v / L[static_cast<size_t>(i) * m6 + i];
                }
                for (int i = m6 - 1; i >= 0; i--) {  // L^T x = y
                    double v = col[i];
                    for (int k = i + 1; k < m6; k++) v -= L[static_cast<size_t>(k) * m6 + i] * col[k];
                    col[i] = v / L[static_cast<size_t>(i) * m6 + i];
                }
                for (int i = 0; i < m6; i++) inv[static_cast<size_t>(i) * m6 + c] = col[i];
            }
            if (D.ncl.size() == 0) {  // fixed capacities: the captured graphs keep valid pointers
                D.ncl.resize(d.n_nodes);
                D.nodes.resize(d.n_nodes);
                D.off.resize(kMaxIslands + 1);
                D.cen.resize(3 * kMaxIslands);
                D.einv.resize(static_cast<size_t>(kDeflStride) * kDeflStride);
                g->gen++;
            }
            D.ncl.upload(ncl.data(), d.n_nodes, s);
            D.off.upload(off.data(), m + 1, s);
            D.nodes.upload(nodes.data(), nodes.size(), s);
            D.cen.upload(cen.data(), 3 * m, s);
            D.einv.upload(inv.data(), inv.size(), s);
            MYCEL_HIP(hipStreamSynchronize(s));
            D.m = m;
        }
    }
    if (D.m != m_old) g->gen++;  // the island count is a kernel argument of the captured iteration
}

// The finest level's Chebyshev and residual launches write no partials, so every tile depth gives the same bits:
// the autotuner picks the fastest for this grid (the last smoothing step, which writes the z.r partials, keeps the
// size rule's depth). Timed on the level's own vectors, zeroed first (the V-cycle writes them before reading).
static void gmg_tune_smoothing(mycel_op *op, Gmg *g, hipStream_t s) {
    const int rule = fine_ec_kz(op);
    if (!rule) return;
    GmgLevel &F = g->lv[0];
    const size_t bytes = sizeof(float) * static_cast<size_t>(F.d.n_dofs);
    MYCEL_HIP(hipMemsetAsync(F.b.get(), 0, bytes, s));
    MYCEL_HIP(hipMemsetAsync(F.x.get(), 0, bytes, s));
    MYCEL_HIP(hipMemsetAsync(F.x2.get(), 0, bytes, s));
    MYCEL_HIP(hipMemsetAsync(F.dv.get(), 0, bytes, s));
    int cand[4];
    const int nc = tile_depth_candidates(rule, cand);
    const dim3 b256(256);
    g->smooth_kz = autotune(kTuneSmooth, op->device, F.d, cand, nc, s, [&](int kz) {
        const int nt = ec_tiles(F.d, kz);
        const dim3 grid(cdiv(nt, 4));
        const float cd = g->cd[1], cr = g->cr[1];
        if (kz == 2)
            hipLaunchKernelGGL((k_gmg_cheb_ec<2, false>), grid, b256, 0, s, F.d, F.km, F.ep, F.x.get(), F.b.get(),
                               F.dv.get(), F.binv, cd, cr, F.x2.get(), nt, nullptr, 0);
        else if (kz == 3)
            hipLaunchKernelGGL((k_gmg_cheb_ec<3, false>), grid, b256, 0, s, F.d, F.km, F.ep, F.x.get(), F.b.get(),
                               F.dv.get(), F.binv, cd, cr, F.x2.get(), nt, nullptr, 0);
        else if (kz == 4)
            hipLaunchKernelGGL((k_gmg_cheb_ec<4, false>), grid, b256, 0, s, F.d, F.km, F.ep, F.x.get(), F.b.get(),
                               F.dv.get(), F.binv, cd, cr, F.x2.get(), nt, nullptr, 0);
        else
            hipLaunchKernelGGL((k_gmg_cheb_ec<8, false>), grid, b256, 0, s, F.d, F.km, F.ep, F.x.get(), F.b.get(),
                               F.dv.get(), F.binv, cd, cr, F.x2.get(), nt, nullptr, 0);
        MYCEL_LAUNCH_CHECK();
    });
}

void gmg_prepare(mycel_op *op, hipStream_t s) {
    if (!op->gmg) gmg_build(op);
    Gmg *g = op->gmg;
    gmg_reserve(op, 1);
    if (g->epoch == op->epoch) return;
    const dim3 b256(256);
    GmgLevel &F = g->lv[0];
    hipLaunchKernelGGL(k_gmg_inv_fine, dim3(cdiv(F.d.n_dofs, 256)), b256, 0, s, F.d.n_dofs, op->d_flags.get(),
                       op->d_diag.get(), F.inv);
    for (int l = 1; l < g->levels; l++) {
        GmgLevel &P = g->lv[l - 1], &L = g->lv[l];
        hipLaunchKernelGGL(k_gmg_coarsen_e, dim3(cdiv(L.d.n_pad, 256)), b256, 0, s, P.d, P.ep, L.d, L.ep_w);
        hipLaunchKernelGGL(k_gmg_inv_coarse, dim3(cdiv(L.d.n_nodes, 256)), b256, 0, s, L.d, L.ep, L.kdiag, P.d,
                           P.inv, L.inv);
    }
    for (int l = 0; l < g->levels; l++) {
        GmgLevel &L = g->lv[l];
        hipLaunchKernelGGL(k_gmg_binv, dim3(cdiv(L.d.n_nodes, 256)), b256, 0, s, L.d, L.ep, L.kdiag, L.koff,
                           L.inv, L.binv);
    }
    GmgLevel &C = g->lv[g->levels - 1];
    const int nc = C.d.n_dofs;
    const size_t lds = sizeof(double) * std::max(nc * nc, 4 * kCoarseMaxDofs);
    MYCEL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_gmg_coarse_factor),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    MYCEL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_gmg_tail),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsMax));
    if (g->levels >= 2) {  // Galerkin coarsest operator from the second-coarsest level's elements
        GmgLevel &S = g->lv[g->levels - 2];
        hipLaunchKernelGGL(k_gmg_galerkin_blocks, dim3(S.d.n_elem), b256, 0, s, S.d, S.ep, S.inv,
                           op->d_ke.get(), S.scale, g->blocks.get());
        hipLaunchKernelGGL(k_gmg_coarse_factor, dim3(1), dim3(1024), lds, s, C.d, C.ep, op->d_ke.get(), C.scale,
                           C.inv, S.d, static_cast<const double *>(g->blocks.get()), g->ainv);
    } else {
        hipLaunchKernelGGL(k_gmg_coarse_factor, dim3(1), dim3(1024), lds, s, C.d, C.ep, op->d_ke.get(), C.scale,
                           C.inv, C.d, static_cast<const double *>(nullptr), g->ainv);
    }
    MYCEL_LAUNCH_CHECK();
    if (op->island_ratio > 0.0) {
        gmg_setup_deflation(op, g, s);
    } else if (g->defl.m != 0) {
        g->defl.m = 0;
        g->gen++;
    }
    g->epoch = op->epoch;
    if (!g->smooth_kz) gmg_tune_smoothing(op, g, s);
}

GmgFineView gmg_fine_view(mycel_op *op) {
    GmgLevel &F = op->gmg->lv[0];
    return GmgFineView{F.b.get(), F.x.get(), F.dv.get(), F.inv, F.binv, op->gmg->theta_inv};
}

// The tail kernel's arguments: levels [t0, top] run in the single-workgroup tail (the finest level never does: it
// has the FP64 entry and exit kernels), as many as fit its node limit, its per-array limit and its LDS. cur / nxt:
// each level's current iterate and scratch before the tail; tail_lds: the dynamic LDS bytes of its launch.
static TailArgs make_tail_args(const Gmg *g, float *const *cur, float *const *nxt, const int *halt, int halt_stride,
                               int &t0, size_t &tail_lds) {
    const TailLayout &tl = g->tl;
    t0 = tl.t0;
    TailArgs ta{};
    ta.count = tl.count;
    for (int k = 0; k < tl.count; k++) {
        const GmgLevel &L = g->lv[t0 + k];
        TailLevel &T = ta.lv[k];
        T.d = L.d;
        T.ks = L.ks;
        T.km = L.km;
        T.ec = tl.ec[k];
        T.o_scr = tl.o_scr;
        T.rnx1 = 1.0f / static_cast<float>(L.d.nx1);
        T.rny1 = 1.0f / static_cast<float>(L.d.ny1);
        T.rnx = 1.0f / static_cast<float>(L.d.nx);
        T.rny = 1.0f / static_cast<float>(L.d.ny);
        const int nd = L.d.n_dofs, ov = tl.o_vec[k];
        T.o_e = tl.o_e[k];
        T.o_inv = tl.o_inv[k];
        T.o_binv = tl.o_binv[k];
        T.o_b = ov;
        T.o_x = ov + nd;
        T.o_x2 = ov + 2 * nd;
        T.o_d = ov + 3 * nd;
        T.o_r = ov + 4 * nd;
    }
    ta.ainv_lds = tl.ainv;
    tail_lds = sizeof(float) * static_cast<size_t>(tl.lds_floats);
    ta.img = g->tail_img.get();
    ta.img_n = tl.img_n;
    {
        const TailLevel &T0 = ta.lv[0];
        ta.seg[0] = TailSeg{g->lv[t0].b.get(), T0.d.n_dofs, T0.o_b};
        ta.seg[1] = TailSeg{cur[t0], T0.d.n_dofs, T0.o_x};
        ta.seg[2] = TailSeg{g->lv[t0].dv.get(), T0.d.n_dofs, T0.o_d};
    }
    ta.out0 = ta.count > 1 ? nxt[t0] : cur[t0];  // the host-side swap below names it the current iterate
    ta.ainv = g->ainv;
    ta.theta_inv = g->theta_inv;
    ta.cd1 = g->cd[1];
    ta.cr1 = g->cr[1];
    ta.halt = halt;
    ta.halt_stride = halt_stride;
    return ta;
}

// z = M(r): one V-cycle. Also writes per-block partials of z.r and z.ap (gmg_partials(op) of each).
// fine_ready: the caller has already written the finest level's b, x = d = D^-1 b / theta (gmg_fine_view).
void gmg_enqueue_vcycle(mycel_op *op, const double *r64, const double *ap64, double *z64, double *part_zr,
                        double *part_zap, hipStream_t s, bool fine_ready, int n_rhs, const int *halt,
                        int halt_stride) {
    Gmg *g = op->gmg;
    const dim3 b256(256);
    if (n_rhs < 1 || n_rhs > g->rhs_cap) throw Error(MYCEL_E_ARGUMENT, "multigrid: RHS count beyond the reserved");
    const int R = n_rhs;
    const int top = g->levels - 1;
    const bool split = gmg_split();
    float *cur[kMaxLevels], *nxt[kMaxLevels];
    for (int l = 0; l <= top; l++) {
        cur[l] = g->lv[l].x.get();
        nxt[l] = g->lv[l].x2.get();
    }
    // Island deflation: the correction's coefficients from r, ahead of the V-cycle; its last step adds Z coef.
    Deflation &D = g->defl;
    if (D.m > 0 && D.rhs_cap < R) throw Error(MYCEL_E_ARGUMENT, "multigrid: deflation storage not reserved");
    Defl df{D.ncl.get(), D.coef.get(), D.cen.get(), op->h, op->d.nx1, op->d.ny1, D.m};
    if (D.m > 0) {
        hipLaunchKernelGGL(k_defl_zt, dim3(D.m, R), b256, 0, s, op->d.nx1, op->d.ny1, op->h, D.off.get(),
                           D.nodes.get(), D.cen.get(), r64, static_cast<long long>(op->d.n_dofs), D.y.get());
        hipLaunchKernelGGL(k_defl_solve, dim3(1, R), b256, 0, s, 6 * D.m, D.einv.get(), D.y.get(), D.coef.get());
    }
    const int fkz = fine_ec_kz(op);  // finest level: element-centric kernels with fkz layers per tile (0: gathers)
    const bool fh = n_rhs > 1 && halt;  // the finest level's kernels test the halt flag (several right-hand sides)
    const int f_tiles = fkz ? ec_tiles(g->lv[0].d, fkz) : 0;
    const dim3 f_grid(cdiv(f_tiles, 4), R);
    // Element-centric kernels on the finest level only: on the coarse levels, whose few waves each march a
    // chain of dependent z-steps, they measured slower than the slab-split gathers (holed box: 126.1 solves/s
    // with the finest level alone, 123.7 with level 1 as well, 122.5 with levels 1 and 2).
    // The finest level's smoothing and residual launches take the autotuned depth (no partials: the same bits at
    // every depth); the last step, which writes the partials, takes fkz.
    auto lkz = [&](int l) { return l == 0 && fkz ? (g->smooth_kz ? g->smooth_kz : fkz) : 0; };
    auto cheb = [&](int l, float cd, float cr) {
        GmgLevel &L = g->lv[l];
        if (const int kz = lkz(l)) {
            const int nt = ec_tiles(L.d, kz);
            const dim3 grid(cdiv(nt, 4), R);
            if (kz == 2)
                hipLaunchKernelGGL((fh ? k_gmg_cheb_ec<2, true> : k_gmg_cheb_ec<2, false>), grid, b256, 0, s,
                                   L.d, L.km, L.ep, cur[l], L.b.get(),
                                   L.dv.get(), L.binv, cd, cr, nxt[l], nt, halt, halt_stride);
            else if (kz == 3)
                hipLaunchKernelGGL((fh ? k_gmg_cheb_ec<3, true> : k_gmg_cheb_ec<3, false>), grid, b256, 0, s,
                                   L.d, L.km, L.ep, cur[l], L.b.get(),
                                   L.dv.get(), L.binv, cd, cr, nxt[l], nt, halt, halt_stride);
            else if (kz == 4)
                hipLaunchKernelGGL((fh ? k_gmg_cheb_ec<4, true> : k_gmg_cheb_ec<4, false>), grid, b256, 0, s,
                                   L.d, L.km, L.ep, cur[l], L.b.get(),
                                   L.dv.get(), L.binv, cd, cr, nxt[l], nt, halt, halt_stride);
            else
                hipLaunchKernelGGL((fh ? k_gmg_cheb_ec<8, true> : k_gmg_cheb_ec<8, false>), grid, b256, 0, s,
                                   L.d, L.km, L.ep, cur[l], L.b.get(),
                                   L.dv.get(), L.binv, cd, cr, nxt[l], nt, halt, halt_stride);
        } else if (split && l > 0)
            hipLaunchKernelGGL(k_gmg_cheb_s, dim3(L.blocks_s, R), b256, 0, s, L.d, L.ks, L.ep, cur[l], L.b.get(),
                               L.dv.get(), L.binv, cd, cr, nxt[l]);
        else
            hipLaunchKernelGGL(k_gmg_cheb, dim3(L.blocks, R), b256, 0, s, L.d, L.ks, L.ep, cur[l], L.b.get(),
                               L.dv.get(), L.binv, cd, cr, nxt[l]);
        std::swap(cur[l], nxt[l]);
    };
    int t0 = 0;
    size_t tail_lds = 0;
    TailArgs ta = make_tail_args(g, cur, nxt, halt, halt_stride, t0, tail_lds);

    GmgLevel &F = g->lv[0];
    if (!fine_ready)
        hipLaunchKernelGGL(k_gmg_init_fine, dim3(cdiv(F.d.n_nodes, 256), R), b256, 0, s, F.d.n_nodes, r64,
                           F.inv, F.binv, g->theta_inv, F.b.get(), cur[0], F.dv.get());
    for (int l = 0; l < t0; l++) {
        GmgLevel &L = g->lv[l], &C = g->lv[l + 1];
        for (int k = 1; k < (l == 0 ? g->fine_degree : 2); k++) cheb(l, g->cd[k], g->cr[k]);
        if (l > 0 && fuse_rr()) {
            hipLaunchKernelGGL(k_gmg_residual_restrict, dim3(cdiv(C.d.n_nodes, 8), R), b256, 0, s, L.d, L.ks, L.ep,
                               cur[l], L.b.get(), L.inv, C.d, C.inv, C.binv, g->theta_inv, C.b.get(),
                               cur[l + 1],
                               C.dv.get());
            continue;
        }
        if (const int kz = lkz(l)) {
            const int nt = ec_tiles(L.d, kz);
            const dim3 grid(cdiv(nt, 4), R);
            if (kz == 2)
                hipLaunchKernelGGL((fh ? k_gmg_residual_ec<2, true> : k_gmg_residual_ec<2, false>), grid, b256, 0, s,
                                   L.d, L.km, L.ep, cur[l], L.b.get(),
                                   L.inv, L.r.get(), nt, halt, halt_stride);
            else if (kz == 3)
                hipLaunchKernelGGL((fh ? k_gmg_residual_ec<3, true> : k_gmg_residual_ec<3, false>), grid, b256, 0, s,
                                   L.d, L.km, L.ep, cur[l], L.b.get(),
                                   L.inv, L.r.get(), nt, halt, halt_stride);
            else if (kz == 4)
                hipLaunchKernelGGL((fh ? k_gmg_residual_ec<4, true> : k_gmg_residual_ec<4, false>), grid, b256, 0, s,
                                   L.d, L.km, L.ep, cur[l], L.b.get(),
                                   L.inv, L.r.get(), nt, halt, halt_stride);
            else
                hipLaunchKernelGGL((fh ? k_gmg_residual_ec<8, true> : k_gmg_residual_ec<8, false>), grid, b256, 0, s,
                                   L.d, L.km, L.ep, cur[l], L.b.get(),
                                   L.inv, L.r.get(), nt, halt, halt_stride);
        } else if (split)
            hipLaunchKernelGGL(k_gmg_residual_s, dim3(L.blocks_s, R), b256, 0, s, L.d, L.ks, L.ep, cur[l], L.b.get(),
                               L.inv, L.r.get());
        else
            hipLaunchKernelGGL(k_gmg_residual, dim3(L.blocks, R), b256, 0, s, L.d, L.ks, L.ep, cur[l], L.b.get(),
                               L.inv, L.r.get());
        hipLaunchKernelGGL(k_gmg_restrict_dof, dim3(cdiv(C.d.n_nodes, 64), R), b256, 0, s, L.d, L.r.get(), C.d,
                           C.inv, C.binv, g->theta_inv, C.b.get(), cur[l + 1], C.dv.get());
    }
    if (t0 > 0) {
        // The tail prolongs into level t0 - 1 itself (one launch fewer): into that level's current iterate as
        // the pre-smoothing above has left it (the smoothing steps swap the level's iterate buffers).
        ta.pro_x = cur[t0 - 1];
        ta.pro_inv = g->lv[t0 - 1].inv;
        ta.pro_d = g->lv[t0 - 1].d;
        ta.pro_rnx1 = 1.0f / static_cast<float>(ta.pro_d.nx1);
        ta.pro_rny1 = 1.0f / static_cast<float>(ta.pro_d.ny1);
    }
    hipLaunchKernelGGL(k_gmg_tail, dim3(R), dim3(kTailThreads), tail_lds, s, ta);
    if (ta.count > 1) std::swap(cur[t0], nxt[t0]);  // the tail leaves its first level's iterate in x2
    for (int l = t0 - 1; l >= 0; l--) {
        GmgLevel &L = g->lv[l], &C = g->lv[l + 1];
        // Level t0 - 1 was prolonged into by the tail; in a single solve a launched coarse level takes its
        // prolongation and its two post-smoothing steps in one launch. With several right-hand sides these
        // launches are no longer bound by their fixed cost, and the tiles' recomputed halo costs more than the
        // launches save (16 load cases of a 100x50x50 box: 3.97 vs 3.68 ms per case-solve): separate launches.
        if (l > 0 && split && R == 1 && gmg_cheb2()) {
            const bool small = L.d.n_nodes < gmg_cheb2_small();  // 2x2x2 tiles: more blocks on a small level
            const int T = small ? 2 : 4;
            const int tx = cdiv(L.d.nx1, T), ty = cdiv(L.d.ny1, T), tz = cdiv(L.d.nz1, T);
            const dim3 grid(tx * ty * tz, R), blk(small ? Cheb2Tile<2>::NT : Cheb2Tile<4>::NT);
            const bool pro = l != t0 - 1;
            auto kern = small ? (pro ? k_gmg_cheb2_s<2, true> : k_gmg_cheb2_s<2, false>)
                              : (pro ? k_gmg_cheb2_s<4, true> : k_gmg_cheb2_s<4, false>);
            hipLaunchKernelGGL(kern, grid, blk, 0, s, L.d, L.ks, L.ep, cur[l], L.b.get(), L.dv.get(), L.binv, 0.0f,
                               g->theta_inv, g->cd[1], g->cr[1], nxt[l], tx, ty, L.inv, C.d, cur[l + 1]);
            std::swap(cur[l], nxt[l]);
            continue;
        }
        if (l != t0 - 1)
            hipLaunchKernelGGL(k_gmg_prolong, dim3(L.blocks, R), b256, 0, s, L.d, cur[l], L.inv, C.d,
                               cur[l + 1]);
        cheb(l, 0.0f, g->theta_inv);
        for (int k = 1; k < (l == 0 ? g->fine_degree - 1 : 2); k++) cheb(l, g->cd[k], g->cr[k]);
    }
    // Finest level: the last post-smoothing step writes z (FP64) and the partials. A single-level hierarchy
    // (a grid of <= kCoarseMaxDofs DOFs) was solved directly above: the step is then the identity (cd = cr = 0).
    const int fd = g->fine_degree;
    const float cd = top == 0 ? 0.0f : g->cd[fd - 1], cr = top == 0 ? 0.0f : g->cr[fd - 1];
    if (fkz == 2)
        hipLaunchKernelGGL((fh ? k_gmg_cheb_final_ec<2, true> : k_gmg_cheb_final_ec<2, false>), f_grid, b256, 0, s,
                           F.d, F.km, F.ep, cur[0], F.b.get(), F.dv.get(),
                           F.binv, cd, cr, z64, r64, ap64, part_zr, part_zap, f_tiles, halt,
                           halt_stride, df);
    else if (fkz == 3)
        hipLaunchKernelGGL((fh ? k_gmg_cheb_final_ec<3, true> : k_gmg_cheb_final_ec<3, false>), f_grid, b256, 0, s,
                           F.d, F.km, F.ep, cur[0], F.b.get(), F.dv.get(),
                           F.binv, cd, cr, z64, r64, ap64, part_zr, part_zap, f_tiles, halt,
                           halt_stride, df);
    else if (fkz == 4)
        hipLaunchKernelGGL((fh ? k_gmg_cheb_final_ec<4, true> : k_gmg_cheb_final_ec<4, false>), f_grid, b256, 0, s,
                           F.d, F.km, F.ep, cur[0], F.b.get(), F.dv.get(),
                           F.binv, cd, cr, z64, r64, ap64, part_zr, part_zap, f_tiles, halt,
                           halt_stride, df);
    else if (fkz == 8)
        hipLaunchKernelGGL((fh ? k_gmg_cheb_final_ec<8, true> : k_gmg_cheb_final_ec<8, false>), f_grid, b256, 0, s,
                           F.d, F.km, F.ep, cur[0], F.b.get(), F.dv.get(),
                           F.binv, cd, cr, z64, r64, ap64, part_zr, part_zap, f_tiles, halt,
                           halt_stride, df);
    else
        hipLaunchKernelGGL(k_gmg_cheb_final, dim3(F.blocks, R), b256, 0, s, F.d, F.ks, F.ep, cur[0], F.b.get(),
                           F.dv.get(), F.binv, cd, cr, z64, r64, ap64, part_zr, part_zap, df);
    MYCEL_LAUNCH_CHECK();
}


// ---- packed V-cycles: the host side ----------------------------------------------------------------

namespace {

enum PkKind : int { kPkChebEc, kPkResidualEc, kPkFinalEc, kPkChebS, kPkResidualRestrict, kPkRestrictDof, kPkProlong,
                    kPkTail };

// One launch of one member: its place in the V-cycle (key: the members' sequences are merged by key), the packed
// kernel, its tile depth (finest level), its own grid and its table entry.
struct PkRec {
    long long key;
    int kind, kz, nb;
    size_t lds;
    std::vector<uint8_t> entry;
};

template <typename A>
PkRec pk_rec(long long key, int kind, int kz, int nb, const A &a, size_t lds = 0) {
    PkRec r{key, kind, kz, nb, lds, std::vector<uint8_t>(sizeof(A))};
    std::memcpy(r.entry.data(), &a, sizeof(A));
    return r;
}

// Keys: the down sweep by level, then the tail, then the up sweep from the coarse levels to the finest; within a
// level the steps in their order. A member with fewer launched levels has a subsequence of a deeper one's keys.
constexpr long long kKeyTail = 1LL << 20, kKeyUp = 2LL << 20;
long long key_down(int l, int k) { return static_cast<long long>(l) * 64 + k; }
long long key_up(int l, int k) { return kKeyUp + static_cast<long long>(kMaxLevels - l) * 64 + k; }

// The launches of gmg_enqueue_vcycle (fine_ready, one right-hand side, the default kernel set) for one member,
// with the same buffers in the same order.
void pk_member_records(const GmgPackMember &mb, std::vector<PkRec> &out) {
    Gmg *g = mb.op->gmg;
    const int top = g->levels - 1;
    float *cur[kMaxLevels], *nxt[kMaxLevels];
    for (int l = 0; l <= top; l++) {
        cur[l] = g->lv[l].x.get();
        nxt[l] = g->lv[l].x2.get();
    }
    const GmgLevel &F = g->lv[0];
    const int fkz = fine_ec_kz(mb.op), skz = g->smooth_kz ? g->smooth_kz : fkz;
    const int f_tiles = ec_tiles(F.d, fkz), f_nb = cdiv(f_tiles, 4);
    const int s_tiles = ec_tiles(F.d, skz), s_nb = cdiv(s_tiles, 4);
    auto fine = [&](float cd, float cr, float *outv, bool smoothing) {
        PkEc a{};
        a.g = F.d;
        a.km = F.km;
        a.ep = F.ep;
        a.x = cur[0];
        a.b = F.b.get();
        a.inv = F.inv;
        a.binv = F.binv;
        a.d = F.dv.get();
        a.out = outv;
        a.cd = cd;
        a.cr = cr;
        a.halt = mb.halt;
        a.n_tiles = smoothing ? s_tiles : f_tiles;
        a.nb = smoothing ? s_nb : f_nb;
        return a;
    };
    auto cheb = [&](int l, float cd, float cr, long long key) {
        const GmgLevel &L = g->lv[l];
        if (l == 0) {
            out.push_back(pk_rec(key, kPkChebEc, skz, s_nb, fine(cd, cr, nxt[0], true)));
        } else {
            PkS a{};
            a.g = L.d;
            a.ks = L.ks;
            a.ep = L.ep;
            a.x = cur[l];
            a.b = L.b.get();
            a.binv = L.binv;
            a.d = L.dv.get();
            a.x_out = nxt[l];
            a.cd = cd;
            a.cr = cr;
            a.nb = L.blocks_s;
            out.push_back(pk_rec(key, kPkChebS, 0, L.blocks_s, a));
        }
        std::swap(cur[l], nxt[l]);
    };
    int t0 = 0;
    size_t tail_lds = 0;
    TailArgs ta = make_tail_args(g, cur, nxt, mb.halt, 0, t0, tail_lds);
    for (int l = 0; l < t0; l++) {
        const GmgLevel &L = g->lv[l], &C = g->lv[l + 1];
        for (int k = 1; k < (l == 0 ? g->fine_degree : 2); k++) cheb(l, g->cd[k], g->cr[k], key_down(l, k));
        if (l > 0) {
            PkRR a{};
            a.gf = L.d;
            a.gc = C.d;
            a.ks = L.ks;
            a.ep = L.ep;
            a.xf = cur[l];
            a.bf = L.b.get();
            a.invf = L.inv;
            a.invc = C.inv;
            a.binvc = C.binv;
            a.bc = C.b.get();
            a.xcv = cur[l + 1];
            a.dc = C.dv.get();
            a.theta_inv = g->theta_inv;
            a.nb = cdiv(C.d.n_nodes, 8);
            out.push_back(pk_rec(key_down(l, 40), kPkResidualRestrict, 0, a.nb, a));
        } else {
            out.push_back(pk_rec(key_down(0, 40), kPkResidualEc, skz, s_nb, fine(0.0f, 0.0f, L.r.get(), true)));
            PkRD a{};
            a.gf = L.d;
            a.gc = C.d;
            a.rf = L.r.get();
            a.invc = C.inv;
            a.binvc = C.binv;
            a.bc = C.b.get();
            a.xcv = cur[1];
            a.dc = C.dv.get();
            a.theta_inv = g->theta_inv;
            a.nb = cdiv(C.d.n_nodes, 64);
            out.push_back(pk_rec(key_down(0, 41), kPkRestrictDof, 0, a.nb, a));
        }
    }
    ta.pro_x = cur[t0 - 1];  // t0 >= 1: every packable hierarchy has at least two levels
    ta.pro_inv = g->lv[t0 - 1].inv;
    ta.pro_d = g->lv[t0 - 1].d;
    ta.pro_rnx1 = 1.0f / static_cast<float>(ta.pro_d.nx1);
    ta.pro_rny1 = 1.0f / static_cast<float>(ta.pro_d.ny1);
    out.push_back(pk_rec(kKeyTail, kPkTail, 0, 1, ta, tail_lds));
    if (ta.count > 1) std::swap(cur[t0], nxt[t0]);
    for (int l = t0 - 1; l >= 0; l--) {
        const GmgLevel &L = g->lv[l], &C = g->lv[l + 1];
        if (l != t0 - 1) {
            PkPro a{};
            a.gf = L.d;
            a.gc = C.d;
            a.xf = cur[l];
            a.invf = L.inv;
            a.xcv = cur[l + 1];
            a.nb = L.blocks;
            out.push_back(pk_rec(key_up(l, 0), kPkProlong, 0, L.blocks, a));
        }
        cheb(l, 0.0f, g->theta_inv, key_up(l, 1));
        for (int k = 1; k < (l == 0 ? g->fine_degree - 1 : 2); k++) cheb(l, g->cd[k], g->cr[k], key_up(l, 1 + k));
    }
    const int fd = g->fine_degree;
    PkEc a = fine(g->cd[fd - 1], g->cr[fd - 1], nullptr, false);
    a.z64 = mb.z64;
    a.r64 = mb.r64;
    a.ap64 = mb.ap64;
    a.part_zr = mb.part_zr;
    a.part_zap = mb.part_zap;
    out.push_back(pk_rec(key_up(0, 63), kPkFinalEc, fkz, f_nb, a));
}

}  // namespace

// A pack has a fixed number of member slots (grid dimension y of every launch) and a launch list that only grows:
// a launch per (step key, kernel, tile depth) that any member has brought, with the largest grid and LDS seen.
// A membership change then rewrites only the tables (empty entries: zero, so every block returns at entry), and a
// graph captured for the current shape stays valid until the shape grows.
struct GmgPack {
    struct Launch {
        long long key;
        int kind, kz, grid_x;
        size_t lds, entry, off;  // entry: table entry size; off: the table's offset (slots entries)
    };
    std::vector<Launch> launches;  // in (key, kind, tile depth) order
    int slots = 0;
    uint64_t shape = 0;            // bumped whenever launches, grid sizes, LDS or table offsets change
    DevBuf<uint8_t> dev;           // the tables
    uint8_t *host = nullptr;       // pinned staging of the tables
    size_t host_cap = 0, bytes = 0;
    ~GmgPack() {
        if (host) (void)hipHostFree(host);
    }
};

bool gmg_packable(const mycel_op *op) {
    const Gmg *g = op->gmg;
    return g && g->levels >= 2 && g->defl.m == 0 && op->island_ratio == 0.0 && fine_ec_kz(op) != 0 && gmg_split() &&
           fuse_rr() && g->rhs_cap >= 1;
}

GmgPack *gmg_pack_create(int slots) {
    if (slots < 1) throw Error(MYCEL_E_OUT_OF_RANGE, "multigrid pack: at least one slot");
    MYCEL_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_pk_tail),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kTailLdsMax));
    GmgPack *p = new GmgPack();
    p->slots = slots;
    return p;
}

void gmg_pack_free(GmgPack *p) { delete p; }

int gmg_pack_launches(const GmgPack *p) { return p ? static_cast<int>(p->launches.size()) : 0; }

uint64_t gmg_pack_shape(const GmgPack *p) { return p->shape; }

void gmg_pack_build(GmgPack *p, const GmgPackMember *m, int count, hipStream_t s) {
    if (count != p->slots) throw Error(MYCEL_E_INVALID_OPERATION, "multigrid pack: one entry per slot");
    MYCEL_HIP(hipStreamSynchronize(s));  // nothing in flight still reads the old tables
    std::vector<std::vector<PkRec>> recs(count);
    auto find = [&](const PkRec &r) {
        return std::find_if(p->launches.begin(), p->launches.end(), [&](const GmgPack::Launch &L) {
            return L.key == r.key && L.kind == r.kind && L.kz == r.kz;
        });
    };
    bool grown = false;
    for (int q = 0; q < count; q++) {
        if (!m[q].op) continue;
        if (!gmg_packable(m[q].op)) throw Error(MYCEL_E_INVALID_OPERATION, "multigrid pack: member not packable");
        pk_member_records(m[q], recs[q]);
        for (const PkRec &r : recs[q]) {
            auto it = find(r);
            if (it == p->launches.end()) {
                p->launches.push_back(GmgPack::Launch{r.key, r.kind, r.kz, r.nb, r.lds, r.entry.size(), 0});
                grown = true;
            } else {
                if (r.nb > it->grid_x || r.lds > it->lds) grown = true;
                it->grid_x = std::max(it->grid_x, r.nb);
                it->lds = std::max(it->lds, r.lds);
            }
        }
    }
    if (grown) {
        std::sort(p->launches.begin(), p->launches.end(), [](const GmgPack::Launch &a, const GmgPack::Launch &b) {
            return std::make_tuple(a.key, a.kind, a.kz) < std::make_tuple(b.key, b.kind, b.kz);
        });
        size_t off = 0;
        for (GmgPack::Launch &L : p->launches) {
            L.off = off;
            off = (off + L.entry * static_cast<size_t>(p->slots) + 63) / 64 * 64;
        }
        p->bytes = off;
        p->shape++;
    }
    if (p->bytes > p->host_cap) {
        if (p->host) (void)hipHostFree(p->host);
        p->host = nullptr;
        p->host_cap = 0;
        MYCEL_HIP(hipHostMalloc(reinterpret_cast<void **>(&p->host), p->bytes, hipHostMallocDefault));
        p->host_cap = p->bytes;
    }
    if (p->bytes > p->dev.size()) {
        p->dev.resize(p->bytes);
        p->shape++;  // the tables moved
    }
    std::memset(p->host, 0, p->bytes);
    for (int q = 0; q < count; q++)
        for (const PkRec &r : recs[q]) {
            const auto it = find(r);
            std::memcpy(p->host + it->off + it->entry * static_cast<size_t>(q), r.entry.data(), r.entry.size());
        }
    MYCEL_HIP(hipMemcpyAsync(p->dev.get(), p->host, p->bytes, hipMemcpyHostToDevice, s));
}

void gmg_pack_enqueue(const GmgPack *p, hipStream_t s) {
    const dim3 b256(256);
    for (const GmgPack::Launch &L : p->launches) {
        const void *t = p->dev.get() + L.off;
        const dim3 grid(L.grid_x, p->slots);
        const PkEc *ec = static_cast<const PkEc *>(t);
        switch (L.kind) {
            case kPkChebEc:
                if (L.kz == 2) hipLaunchKernelGGL(k_pk_cheb_ec<2>, grid, b256, 0, s, ec);
                else if (L.kz == 3) hipLaunchKernelGGL(k_pk_cheb_ec<3>, grid, b256, 0, s, ec);
                else if (L.kz == 4) hipLaunchKernelGGL(k_pk_cheb_ec<4>, grid, b256, 0, s, ec);
                else hipLaunchKernelGGL(k_pk_cheb_ec<8>, grid, b256, 0, s, ec);
                break;
            case kPkResidualEc:
                if (L.kz == 2) hipLaunchKernelGGL(k_pk_residual_ec<2>, grid, b256, 0, s, ec);
                else if (L.kz == 3) hipLaunchKernelGGL(k_pk_residual_ec<3>, grid, b256, 0, s, ec);
                else if (L.kz == 4) hipLaunchKernelGGL(k_pk_residual_ec<4>, grid, b256, 0, s, ec);
                else hipLaunchKernelGGL(k_pk_residual_ec<8>, grid, b256, 0, s, ec);
                break;
            case kPkFinalEc:
                if (L.kz == 2) hipLaunchKernelGGL(k_pk_cheb_final_ec<2>, grid, b256, 0, s, ec);
                else if (L.kz == 3) hipLaunchKernelGGL(k_pk_cheb_final_ec<3>, grid, b256, 0, s, ec);
                else if (L.kz == 4) hipLaunchKernelGGL(k_pk_cheb_final_ec<4>, grid, b256, 0, s, ec);
                else hipLaunchKernelGGL(k_pk_cheb_final_ec<8>, grid, b256, 0, s, ec);
                break;
            case kPkChebS:
                hipLaunchKernelGGL(k_pk_cheb_s, grid, b256, 0, s, static_cast<const PkS *>(t));
                break;
            case kPkResidualRestrict:
                hipLaunchKernelGGL(k_pk_residual_restrict, grid, b256, 0, s, static_cast<const PkRR *>(t));
                break;
            case kPkRestrictDof:
                hipLaunchKernelGGL(k_pk_restrict_dof, grid, b256, 0, s, static_cast<const PkRD *>(t));
                break;
            case kPkProlong:
                hipLaunchKernelGGL(k_pk_prolong, grid, b256, 0, s, static_cast<const PkPro *>(t));
                break;
            case kPkTail:
                hipLaunchKernelGGL(k_pk_tail, dim3(p->slots), dim3(kTailThreads), L.lds, s,
                                   static_cast<const TailArgs *>(t));
                break;
            default:
                throw Error(MYCEL_E_DEVICE, "multigrid pack: unknown launch kind");
        }
    }
    MYCEL_LAUNCH_CHECK();
}

}  // namespace mycel
