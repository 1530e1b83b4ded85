// FemOperator on gfx950: SIMP stiffness update, Jacobi diagonal, node-centric matrix-free K*u,
// element post-processing (von Mises, sensitivities) and deterministic FP64 reductions.
//
// Layout: vectors over DOFs are the reference's box-sized interleaved [ux,uy,uz] per node.
// E_e lives on a zero-padded (nx+2)(ny+2)(nz+2) element grid: 0 for inactive elements and for
// the halo, so a node's 8 incident elements are always addressable and inactive ones add nothing
// (Fem/FeModel.cs:103-107: inactive elements are simply absent).
//
// K*u is a gather: one thread owns one node and forms its three output rows from the 8 incident
// elements, y_n = sum_a E_e(a) * Ke[3a:3a+3, :] * u_e(a). Each element row is computed exactly once
// (by the node that owns it), so there is no redundant arithmetic, no colouring and no atomics;
// the 8-colour lattice of the reference (Fem/FemOperator.cs:214-222) exists only to avoid write
// races on a CPU. Ke is read with wave-uniform addresses and lands in scalar registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

#include "apply_k.hpp"
#include "device_common.hpp"
#include "ku_element.hpp"
#include "mycel_internal.hpp"

namespace mycel {

// ------------------------------------------------------------------------------------------------
// Hex8 element, host side (Fem/Hex8Element.cs:48-251). Computed once per handle.
// ------------------------------------------------------------------------------------------------
namespace {
const double kXi[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
const double kEta[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
const double kZeta[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// Strain-displacement matrix at (xi, eta, zeta), Voigt order xx,yy,zz,yz,zx,xy (Hex8Element.cs:183-215).
void strain_displacement(double xi, double eta, double zeta, double h, double *b) {
    std::fill(b, b + 144, 0.0);
    const double s = 0.25 * (1.0 / h);
    for (int a = 0; a < 8; a++) {
        const double gx = s * kXi[a] * (1.0 + kEta[a] * eta) * (1.0 + kZeta[a] * zeta);
        const double gy = s * kEta[a] * (1.0 + kXi[a] * xi) * (1.0 + kZeta[a] * zeta);
        const double gz = s * kZeta[a] * (1.0 + kXi[a] * xi) * (1.0 + kEta[a] * eta);
        double *c = b + 3 * a;
        c[0 * 24 + 0] = gx;
        c[1 * 24 + 1] = gy;
        c[2 * 24 + 2] = gz;
        c[3 * 24 + 1] = gz;
        c[3 * 24 + 2] = gy;
        c[4 * 24 + 0] = gz;
        c[4 * 24 + 2] = gx;
        c[5 * 24 + 0] = gy;
        c[5 * 24 + 1] = gx;
    }
}
}  // namespace

void constitutive(double e, double nu, double *D) {
    const double f = e / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double c1 = f * (1.0 - nu), c2 = f * nu, g = e / (2.0 * (1.0 + nu));
    std::fill(D, D + 36, 0.0);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) D[r * 6 + c] = (r == c) ? c1 : c2;
    D[21] = D[28] = D[35] = g;
}

// Ke = h * sum_q B^T D B / 8 over the 2x2x2 Gauss points +-1/sqrt(3) (Hex8Element.cs:98-178).
void hex8_ke(double h, double nu, double *ke) {
    const double gp[2] = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)};
    double D[36], B[144], DB[144];
    constitutive(1.0, nu, D);
    std::fill(ke, ke + 576, 0.0);
    for (int ix = 0; ix < 2; ix++)
        for (int iy = 0; iy < 2; iy++)
            for (int iz = 0; iz < 2; iz++) {
                strain_displacement(gp[ix], gp[iy], gp[iz], 1.0, B);
                for (int r = 0; r < 6; r++)
                    for (int c = 0; c < 24; c++) {
                        double s = 0.0;
                        for (int k = 0; k < 6; k++) s += D[r * 6 + k] * B[k * 24 + c];
                        DB[r * 24 + c] = s;
                    }
                for (int i = 0; i < 24; i++)
                    for (int j = 0; j < 24; j++) {
                        double s = 0.0;
                        for (int k = 0; k < 6; k++) s += B[k * 24 + i] * DB[k * 24 + j];
                        ke[i * 24 + j] += s * (1.0 / 8.0);
                    }
            }
    // The reference scales by h on its nu = 0.3 fast path always, and otherwise only when h != 1;
    // both give quad * h.
    if (std::fabs(nu - 0.3) < 1e-15 || std::fabs(h - 1.0) > 1e-15)
        for (int i = 0; i < 576; i++) ke[i] *= h;
}

void hex8_bc(double h, double *bc) {
    double base[144];
    strain_displacement(0.0, 0.0, 0.0, 1.0, base);
    const double inv = 1.0 / h;
    for (int i = 0; i < 144; i++) bc[i] = base[i] * inv;
}

// ------------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------------
namespace {

// E_e = Emin + rho^p (E0 - Emin) on the padded grid (FemOperator.cs:114-138).
__global__ __launch_bounds__(256) void k_update_e(Dims g, const double *__restrict__ density,
                                                  const uint8_t *__restrict__ domain, double e0, double emin,
                                                  double penal, double *__restrict__ ep, float *__restrict__ ep32) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= g.n_pad) return;
    const int ip = p % g.px, t = p / g.px, jp = t % g.py, kp = t / g.py;
    double e = 0.0;
    if (ip >= 1 && ip <= g.nx && jp >= 1 && jp <= g.ny && kp >= 1 && kp <= g.nz) {
        const int el = (ip - 1) + g.nx * ((jp - 1) + g.ny * (kp - 1));
        if (domain[el]) e = emin + pow_penal(density[el], penal) * (e0 - emin);
    }
    ep[p] = e;
    ep32[p] = static_cast<float>(e);
}

// DiagA in colour order (FemOperator.cs:143-172): DOF (n, c) receives E_e * DiagKe[3a+c] from the
// element of colour 0..7 that contains node n as its corner a. Constrained DOFs get 1.
__global__ __launch_bounds__(256) void k_diag(Dims g, const double *__restrict__ ke, const double *__restrict__ ep,
                                              const uint8_t *__restrict__ flags, double *__restrict__ diag,
                                              double *__restrict__ inv_diag, float *__restrict__ inv_diag32) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= g.n_nodes) return;
    const int i = n % g.nx1, t = n / g.nx1, j = t % g.ny1, k = t / g.ny1;
    const int corner_of[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // (ax + 2ay + 4az) -> Hex8 corner index
    const double *dke = ke + 576;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < 8; c++) {
        const int ax = (i ^ c) & 1, ay = (j ^ (c >> 1)) & 1, az = (k ^ (c >> 2)) & 1;
        const double e = ep[(i - ax + 1) + g.px * ((j - ay + 1) + g.py * (k - az + 1))];
        const int a = corner_of[ax + 2 * ay + 4 * az];
        if (e != 0.0)
            for (int d = 0; d < 3; d++) acc[d] += e * dke[3 * a + d];
    }
    for (int d = 0; d < 3; d++) {
        const int dof = 3 * n + d;
        const double v = (flags[dof] & kFlagConstrained) ? 1.0 : acc[d];
        diag[dof] = v;
        const double inv = fabs(v) > 1e-30 ? 1.0 / v : 1.0;  // JacobiPreconditioner.cs:25
        inv_diag[dof] = inv;
        inv_diag32[dof] = static_cast<float>(inv);
    }
}

// Node-centric K*u in storage type T (double: the FP64 path; float: the inner operator of the mixed
// path). MODE: kApplyK (y = Kx), kApplyAProjected (y = Kx, y[fixed] = x[fixed]), kApplyAGeneral
// (y = K(Px), y[constrained] = x). DOT adds a per-block FP64 partial of x.y; `stop` (PCG state) turns
// the launch into a no-op once the solve has halted.
template <typename T, int MODE, bool DOT>
__global__ __launch_bounds__(256, 2) void k_apply_k(Dims g, const KeSym<T> ks, const T *__restrict__ ep,
                                                    const T *__restrict__ x, const uint8_t *__restrict__ flags,
                                                    T *__restrict__ y, double *__restrict__ part,
                                                    const int *__restrict__ stop, Batch bt) {
    __shared__ double red[4];
    const int rhs = blockIdx.y;  // Batch: one RHS per grid row
    x += rhs * bt.vstride;
    y += rhs * bt.vstride;
    if (DOT) part += rhs * bt.pstride;
    if (stop) stop += rhs * bt.sstride;
    if (DOT && stop && *stop) return;
    const int n0 = xcd_block(blockIdx.x, gridDim.x) * 256, n = n0 + threadIdx.x;
    const double dotv = apply_k_node<T, MODE>(g, ks, ep, x, flags, y, n, first_layer(g, n0));
    if (DOT) {
        const double s = block_sum256(dotv, red);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
    }
}

// Element-centric K*u (ec_tile): one 8 x 8 tile of element columns per wavefront, four tiles per block, KZ node
// layers per tile. Per-block partials of x.y as k_apply_k.
template <typename T, int MODE, bool DOT, int KZ>
__global__ __launch_bounds__(256) void k_apply_k_ec(Dims g, const KeMono<T> km, const T *__restrict__ ep,
                                                    const T *__restrict__ x, const uint8_t *__restrict__ flags,
                                                    T *__restrict__ y, double *__restrict__ part,
                                                    const int *__restrict__ stop, Batch bt, int n_tiles) {
    __shared__ double red[4];
    const int rhs = blockIdx.y;
    x += rhs * bt.vstride;
    y += rhs * bt.vstride;
    if (DOT) part += rhs * bt.pstride;
    if (stop) stop += rhs * bt.sstride;
    if (DOT && stop && *stop) return;
    // Wave-uniform by construction; readfirstlane lets the tile arithmetic and addressing live in SGPRs.
    const int tile = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
    double dotv = 0.0;
    if (tile < n_tiles) dotv = ec_tile<T, MODE, KZ>(g, km, ep, x, flags, y, tile);
    if (DOT) {
        const double s = block_sum256(dotv, red);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
    }
}

// z-marching K*u (apply_k_column): thread t owns column (i, j) = t mod (nx1 ny1) of z-block t / (nx1 ny1),
// nodes k0 = KZ * z-block onwards; the lanes of a wave are consecutive columns of one z-block (or two, where
// a wave straddles), so their loads stay coalesced. Per-block partials of x.y as k_apply_k.
template <typename T, int MODE, bool DOT, int KZ>
__global__ __launch_bounds__(256) void k_apply_k_zm(Dims g, const KeSym<T> ks, const T *__restrict__ ep,
                                                    const T *__restrict__ x, const uint8_t *__restrict__ flags,
                                                    T *__restrict__ y, double *__restrict__ part,
                                                    const int *__restrict__ stop, Batch bt) {
    __shared__ double red[4];
    const int rhs = blockIdx.y;
    x += rhs * bt.vstride;
    y += rhs * bt.vstride;
    if (DOT) part += rhs * bt.pstride;
    if (stop) stop += rhs * bt.sstride;
    if (DOT && stop && *stop) return;
    const int cols = g.nx1 * g.ny1;
    const int t0 = xcd_block(blockIdx.x, gridDim.x) * 256, t = t0 + threadIdx.x;
    const int zb = t / cols, c = t - zb * cols;
    const int k0 = zb * KZ;
    double dotv = 0.0;
    if (k0 <= g.nz)
        dotv = apply_k_column<T, MODE, KZ>(g, ks, ep, x, flags, y, c % g.nx1, c / g.nx1, k0, (t0 / cols) * KZ);
    if (DOT) {
        const double s = block_sum256(dotv, red);
        if (threadIdx.x == 0) part[blockIdx.x] = s;
    }
}

// Per-block partial of sum a_i b_i (optionally only over unconstrained i).
template <bool MASKED>
__global__ __launch_bounds__(256) void k_dot_partial(const double *__restrict__ a, const double *__restrict__ b,
                                                     int n, const uint8_t *__restrict__ flags,
                                                     double *__restrict__ part) {
    __shared__ double red[4];
    double v = 0.0;
    const int base = blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int idx = base + 256 * q;
        if (idx < n && (!MASKED || !(flags[idx] & kFlagConstrained))) v += a[idx] * b[idx];
    }
    const double s = block_sum256(v, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ part, int n,
                                                      double *__restrict__ out) {
    __shared__ double red[4];
    const double s = sum_partials256(part, n, red);
    if (threadIdx.x == 0) *out = s;
}

// Element von Mises at the centroid (StaticAnalysis.cs:144-216). cst: Bc(1) [144] then D(1, nu) [36].
__global__ __launch_bounds__(256) void k_von_mises(Dims g, const int32_t *__restrict__ active, int na,
                                                   const double *__restrict__ cst, const double *__restrict__ ep,
                                                   const double *__restrict__ u, double inv_h,
                                                   double *__restrict__ vm) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= na) return;
    const int e = active[q];
    const int ix = e % g.nx, t = e / g.nx, iy = t % g.ny, iz = t / g.ny;
    const int n0 = ix + g.nx1 * (iy + g.ny1 * iz);
    double ue[24];
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const int nd = n0 + corner_x(a) + g.nx1 * (corner_y(a) + g.ny1 * corner_z(a));
#pragma unroll
        for (int d = 0; d < 3; d++) ue[3 * a + d] = u[3 * nd + d];
    }
    const double ee = ep[(ix + 1) + g.px * ((iy + 1) + g.py * (iz + 1))];
    double eps[6], sig[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 24; c++) s += cst[r * 24 + c] * ue[c];
        eps[r] = s * inv_h;
    }
#pragma unroll
    for (int r = 0; r < 6; r++) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 6; c++) s += cst[144 + r * 6 + c] * eps[c];
        sig[r] = s * ee;
    }
    const double d1 = sig[0] - sig[1], d2 = sig[1] - sig[2], d3 = sig[2] - sig[0];
    vm[e] = sqrt(0.5 * (d1 * d1 + d2 * d2 + d3 * d3) + 3.0 * (sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5]));
}

// dc_e = -p rho^(p-1) (E0 - Emin) u_e^T Ke u_e, in the reference's summation order
// (SensitivityAnalysis.cs:80-95); no FMA contraction so the per-element value matches it.
__global__ __launch_bounds__(256) void k_sensitivities(Dims g, const int32_t *__restrict__ active, int na,
                                                       const double *__restrict__ ke, const double *__restrict__ u,
                                                       const double *__restrict__ rho, double penal, double de,
                                                       double *__restrict__ dc) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= na) return;
    const int e = active[q];
    const int ix = e % g.nx, t = e / g.nx, iy = t % g.ny, iz = t / g.ny;
    const int n0 = ix + g.nx1 * (iy + g.ny1 * iz);
    double ue[24];
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const int nd = n0 + corner_x(a) + g.nx1 * (corner_y(a) + g.ny1 * corner_z(a));
#pragma unroll
        for (int d = 0; d < 3; d++) ue[3 * a + d] = u[3 * nd + d];
    }
    double se = 0.0;
#pragma unroll
    for (int r = 0; r < 24; r++) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < 24; c++) s += ke[r * 24 + c] * ue[c];
        se += ue[r] * s;
    }
    const double pm1 = penal - 1.0;
    const double pw = pm1 == 2.0 ? rho[q] * rho[q] : (pm1 == 1.0 ? rho[q] : (pm1 == 0.0 ? 1.0 : pow(rho[q], pm1)));
    const double dedr = penal * pw * de;
    dc[q] = -dedr * se;
}

// StaticAnalysis.AverageElementToNode (StaticAnalysis.cs:222-272): per node, active incident
// elements summed in ascending element order.
__global__ __launch_bounds__(256) void k_avg_e2n(Dims g, const uint8_t *__restrict__ domain,
                                                 const double *__restrict__ elem, double *__restrict__ nodal) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= g.n_nodes) return;
    const int i = n % g.nx1, t = n / g.nx1, j = t % g.ny1, k = t / g.ny1;
    double s = 0.0;
    int cnt = 0;
    for (int oz = 1; oz >= 0; oz--)
        for (int oy = 1; oy >= 0; oy--)
            for (int ox = 1; ox >= 0; ox--) {
                const int ex = i - ox, ey = j - oy, ez = k - oz;
                if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
                const int e = ex + g.nx * (ey + g.ny * ez);
                if (!domain[e]) continue;
                s += elem[e];
                cnt++;
            }
    nodal[n] = cnt > 0 ? s / cnt : 0.0;
}

// StaticAnalysis.SampleNodalScalar (StaticAnalysis.cs:278-366), one thread per point: the cell's local
// coordinates in float32 as the reference forms them from its Vector3 point, origin and float pitch, the trilinear
// weights in FP64 over the cell's active nodes (added in corner order 0..7). A point outside the grid, or whose
// cell has no active node weight above 1e-12, is marked for the nearest-node fallback (fallback[q] = 1).
__global__ __launch_bounds__(256) void k_sample_trilinear(Dims g, const uint8_t *__restrict__ active_nodes,
                                                          const double *__restrict__ nodal, int n_points,
                                                          const float *__restrict__ pts, float ox, float oy, float oz,
                                                          float h, double *__restrict__ out,
                                                          int *__restrict__ fallback) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n_points) return;
    const float rx = (pts[3 * q] - ox) / h, ry = (pts[3 * q + 1] - oy) / h, rz = (pts[3 * q + 2] - oz) / h;
    const double fx = floor(static_cast<double>(rx)), fy = floor(static_cast<double>(ry)),
                 fz = floor(static_cast<double>(rz));
    int fb = 1;
    double result = 0.0;
    // (int)Math.Floor(r) of a finite value inside the grid; anything else (NaN, huge) is outside.
    if (fx >= 0.0 && fx < g.nx && fy >= 0.0 && fy < g.ny && fz >= 0.0 && fz < g.nz) {
        const int ix = static_cast<int>(fx), iy = static_cast<int>(fy), iz = static_cast<int>(fz);
        const int n0 = ix + g.nx1 * (iy + g.ny1 * iz);
        const double xi = static_cast<double>(rx) - ix, eta = static_cast<double>(ry) - iy,
                     zeta = static_cast<double>(rz) - iz;
        const double w[8] = {(1.0 - xi) * (1.0 - eta) * (1.0 - zeta), xi * (1.0 - eta) * (1.0 - zeta),
                             xi * eta * (1.0 - zeta),                 (1.0 - xi) * eta * (1.0 - zeta),
                             (1.0 - xi) * (1.0 - eta) * zeta,         xi * (1.0 - eta) * zeta,
                             xi * eta * zeta,                         (1.0 - xi) * eta * zeta};
        double ws = 0.0, vs = 0.0;
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const int nd = n0 + corner_x(a) + g.nx1 * (corner_y(a) + g.ny1 * corner_z(a));
            if (active_nodes[nd]) {
                ws += w[a];
                vs += w[a] * nodal[nd];
            }
        }
        if (ws > 1e-12) {
            result = vs / ws;
            fb = 0;
        }
    }
    out[q] = result;
    fallback[q] = fb;
}

// The fallback (StaticAnalysis.cs:345-365): the value at the active node nearest to the point, distances squared in
// float32 from the node positions origin + (i h, j h, k h); the first node in index order wins a tie, as the
// reference's strict < does. One workgroup per point; points that need no fallback return at once.
__global__ __launch_bounds__(256) void k_sample_nearest(Dims g, const uint8_t *__restrict__ active_nodes,
                                                        const double *__restrict__ nodal,
                                                        const float *__restrict__ pts, float ox, float oy, float oz,
                                                        float h, const int *__restrict__ fallback,
                                                        double *__restrict__ out) {
#pragma clang fp contract(off)
    __shared__ float sd[256];
    __shared__ int si[256];
    const int q = blockIdx.x;
    if (!fallback[q]) return;
    const float px = pts[3 * q], py = pts[3 * q + 1], pz = pts[3 * q + 2];
    float best = __builtin_huge_valf();
    int bi = -1;
    for (int n = threadIdx.x; n < g.n_nodes; n += 256) {
        if (!active_nodes[n]) continue;
        const int i = n % g.nx1, t = n / g.nx1, j = t % g.ny1, k = t / g.ny1;
        const float dx = px - (ox + static_cast<float>(i) * h), dy = py - (oy + static_cast<float>(j) * h),
                    dz = pz - (oz + static_cast<float>(k) * h);
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < best) {  // ascending n within the thread: the first minimum stays
            best = d2;
            bi = n;
        }
    }
    sd[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            const float d2 = sd[threadIdx.x + off];
            const int i2 = si[threadIdx.x + off];
            // (distance, index) lexicographic: the smallest distance, then the lowest index
            if (i2 >= 0 && (si[threadIdx.x] < 0 || d2 < sd[threadIdx.x] ||
                            (d2 == sd[threadIdx.x] && i2 < si[threadIdx.x]))) {
                sd[threadIdx.x] = d2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = si[0] >= 0 ? nodal[si[0]] : 0.0;
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

// ------------------------------------------------------------------------------------------------
// Host launchers
// ------------------------------------------------------------------------------------------------
void op_update_e(mycel_op *op, const double *d_density_box, double penal, hipStream_t s) {
    static_assert(sizeof(uint8_t) == 1, "");
    op->penal = penal;
    op->epoch++;
    hipLaunchKernelGGL(k_update_e, dim3(cdiv(op->d.n_pad, 256)), dim3(256), 0, s, op->d, d_density_box,
                       op->d_domain.get(), op->e0, op->emin, penal,
                       op->d_ep.get(), op->d_ep32.get());
    MYCEL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_diag, dim3(cdiv(op->d.n_nodes, 256)), dim3(256), 0, s, op->d, op->d_ke.get(),
                       op->d_ep.get(), op->d_flags.get(), op->d_diag.get(), op->d_inv_diag.get(),
                       op->d_inv_diag32.get());
    MYCEL_LAUNCH_CHECK();
}

// Column length of the z-marching gather: its offsets (2 KZ + 2 node layers from a block's base) must fit in
// 32 bits.
template <typename T>
static bool zm_fits(const mycel_op *op, int kz) {
    return (2.0 * kz + 2.0) * 3.0 * sizeof(T) * op->d.nx1 * op->d.ny1 < 4294967296.0;
}

static int zm_blocks(const mycel_op *op, int kz) { return cdiv(op->d.nx1 * op->d.ny1 * cdiv(op->d.nz1, kz), 256); }
static int ec_blocks(const mycel_op *op, int kz) { return cdiv(ec_tiles(op->d, kz), 4); }

struct KuForm {
    int form;  // MYCEL_MATVEC_NODE_GATHER or MYCEL_MATVEC_ELEMENT
    int kz;    // node layers per thread (gather: 0 = one node per thread) or per tile (element: 2, 4, 8)
};

// The default form, by grid size. The element-centric product does about a third of the gather's arithmetic;
// measured on MI355X (FP64 K*u, best gather form -> element form): 9.1 -> 5.4 us on the 80x40x40 holed box
// (136k nodes, 3 layers per tile; 6.0 us with 2), 13.7 -> 7.5 us at 263k nodes (3 layers; 8.4 with 4),
// 35.0 -> 18.3 us at 1.06M nodes and 115 -> 63 us at 4.3M nodes (8 layers). Longer columns amortise the extra element layer each tile forms below
// its first node layer; shorter ones keep enough waves on small grids. The K*u launches that write dot
// partials always use this rule, so an operator's partial count never changes.
static KuForm ku_auto(const mycel_op *op) {
    const int n = op->d.n_nodes;
    return {MYCEL_MATVEC_ELEMENT, n < (1 << 19) ? 3 : n < (1 << 20) ? 4 : 8};
}

// The form of a launch: the automatic one for launches with dot partials, else the operator's setting
// (mycel_op_set_matvec_form); a gather column length whose offsets would not fit falls back to one node per
// thread.
template <typename T>
static KuForm ku_form(const mycel_op *op, bool dot) {
    // The reference-order form has an FP64 kernel of its own (launch_apply_k); other launches take the default.
    if (dot || op->ku_form == MYCEL_MATVEC_AUTO || op->ku_form == MYCEL_MATVEC_REFERENCE_ORDER) return ku_auto(op);
    if (op->ku_form == MYCEL_MATVEC_NODE_GATHER && op->ku_kz && !zm_fits<T>(op, op->ku_kz))
        return {MYCEL_MATVEC_NODE_GATHER, 0};
    return {op->ku_form, op->ku_kz};
}

static int ku_blocks(const mycel_op *op, const KuForm &f) {
    if (f.form == MYCEL_MATVEC_ELEMENT) return ec_blocks(op, f.kz);
    return f.kz ? zm_blocks(op, f.kz) : cdiv(op->d.n_nodes, 256);
}

int apply_k_blocks(const mycel_op *op) { return ku_blocks(op, ku_auto(op)); }

namespace {
std::mutex g_tune_mu;
std::map<std::array<int, 5>, int> g_tuned;  // (family, device, nx, ny, nz) -> chosen variant

bool autotune_enabled() {
    static const bool on = [] {
        const char *e = std::getenv("MYCEL_AUTOTUNE");
        return !(e && *e == '0');
    }();
    return on;
}
}  // namespace

int tile_depth_candidates(int rule_kz, int *out) {
    static const int all[4] = {2, 3, 4, 8};
    int n = 0;
    out[n++] = rule_kz;
    for (int k : all)
        if (k != rule_kz) out[n++] = k;
    return n;
}

int autotune(int family, int device, const Dims &d, const int *cand, int n, hipStream_t s,
             const std::function<void(int)> &launch) {
    if (n <= 1 || !autotune_enabled()) return cand[0];
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MYCEL_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return cand[0];
    const std::array<int, 5> key{family, device, d.nx, d.ny, d.nz};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) return it->second;
    }
    // Two untimed launches, then the mean of kReps launches back to back, per candidate; the smallest mean wins
    // (ties: the earlier candidate, so the rule's choice unless another is faster).
    constexpr int kReps = 8;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MYCEL_HIP(hipEventCreate(&e0));
    MYCEL_HIP(hipEventCreate(&e1));
    int best = cand[0];
    float best_ms = 0.0f;
    try {
        for (int q = 0; q < n; q++) {
            launch(cand[q]);
            launch(cand[q]);
            MYCEL_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < kReps; r++) launch(cand[q]);
            MYCEL_HIP(hipEventRecord(e1, s));
            MYCEL_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            MYCEL_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (q == 0 || ms < best_ms) {
                best_ms = ms;
                best = cand[q];
            }
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return g_tuned.emplace(key, best).first->second;  // a concurrent tuner of the same key keeps its entry
}

int ku_element_layers(const mycel_op *op) { return ku_auto(op).kz; }

template <typename T, int MODE, bool DOT>
static void launch_ku(mycel_op *op, const KuForm &f, const KeSym<T> &ke, const KeMono<T> &km, const T *ep,
                      const T *d_x, T *d_y, hipStream_t s, double *part, const int *stop, const Batch &bt) {
    const dim3 grid(ku_blocks(op, f), bt.count), block(256);
    const uint8_t *fl = op->d_flags.get();
    if (f.form == MYCEL_MATVEC_ELEMENT) {
        const int nt = ec_tiles(op->d, f.kz);
        if (f.kz == 2)
            hipLaunchKernelGGL((k_apply_k_ec<T, MODE, DOT, 2>), grid, block, 0, s, op->d, km, ep, d_x, fl, d_y, part,
                               stop, bt, nt);
        else if (f.kz == 3)
            hipLaunchKernelGGL((k_apply_k_ec<T, MODE, DOT, 3>), grid, block, 0, s, op->d, km, ep, d_x, fl, d_y, part,
                               stop, bt, nt);
        else if (f.kz == 4)
            hipLaunchKernelGGL((k_apply_k_ec<T, MODE, DOT, 4>), grid, block, 0, s, op->d, km, ep, d_x, fl, d_y, part,
                               stop, bt, nt);
        else
            hipLaunchKernelGGL((k_apply_k_ec<T, MODE, DOT, 8>), grid, block, 0, s, op->d, km, ep, d_x, fl, d_y, part,
                               stop, bt, nt);
    } else if (f.kz == 2) {
        hipLaunchKernelGGL((k_apply_k_zm<T, MODE, DOT, 2>), grid, block, 0, s, op->d, ke, ep, d_x, fl, d_y, part, stop,
                           bt);
    } else if (f.kz == 4) {
        hipLaunchKernelGGL((k_apply_k_zm<T, MODE, DOT, 4>), grid, block, 0, s, op->d, ke, ep, d_x, fl, d_y, part, stop,
                           bt);
    } else if (f.kz == 8) {
        hipLaunchKernelGGL((k_apply_k_zm<T, MODE, DOT, 8>), grid, block, 0, s, op->d, ke, ep, d_x, fl, d_y, part, stop,
                           bt);
    } else {
        hipLaunchKernelGGL((k_apply_k<T, MODE, DOT>), grid, block, 0, s, op->d, ke, ep, d_x, fl, d_y, part, stop, bt);
    }
    MYCEL_LAUNCH_CHECK();
}

template <typename T>
static void launch_apply_k_t(mycel_op *op, const KeSym<T> &ke, const KeMono<T> &km, const T *ep, const T *d_x,
                             T *d_y, int mode, hipStream_t s, double *d_dot_partial, const int *d_stop,
                             const Batch &bt) {
    KuForm f = ku_form<T>(op, d_dot_partial != nullptr);
    if (!d_dot_partial && op->ku_form == MYCEL_MATVEC_AUTO && f.form == MYCEL_MATVEC_ELEMENT && bt.count == 1) {
        // Launches without dot partials: the tile depth is the autotuner's (every depth gives the same bits).
        int &tuned = op->ku_tuned[std::is_same<T, float>::value ? 1 : 0];
        if (!tuned) {
            int cand[4];
            const int nc = tile_depth_candidates(f.kz, cand);
            tuned = autotune(std::is_same<T, float>::value ? kTuneKu32 : kTuneKu64, op->device, op->d, cand, nc, s,
                             [&](int kz) {
                                 const KuForm c{MYCEL_MATVEC_ELEMENT, kz};
                                 if (mode == kApplyK)
                                     launch_ku<T, kApplyK, false>(op, c, ke, km, ep, d_x, d_y, s, nullptr, nullptr, bt);
                                 else if (mode == kApplyAProjected)
                                     launch_ku<T, kApplyAProjected, false>(op, c, ke, km, ep, d_x, d_y, s, nullptr,
                                                                          nullptr, bt);
                                 else
                                     launch_ku<T, kApplyAGeneral, false>(op, c, ke, km, ep, d_x, d_y, s, nullptr,
                                                                        nullptr, bt);
                             });
        }
        f.kz = tuned;
    }
    if (d_dot_partial) {
        if (mode != kApplyAProjected) throw Error(MYCEL_E_ARGUMENT, "fused dot only on the projected operator");
        launch_ku<T, kApplyAProjected, true>(op, f, ke, km, ep, d_x, d_y, s, d_dot_partial, d_stop, bt);
    } else if (mode == kApplyK) {
        launch_ku<T, kApplyK, false>(op, f, ke, km, ep, d_x, d_y, s, nullptr, nullptr, bt);
    } else if (mode == kApplyAProjected) {
        launch_ku<T, kApplyAProjected, false>(op, f, ke, km, ep, d_x, d_y, s, nullptr, nullptr, bt);
    } else {
        launch_ku<T, kApplyAGeneral, false>(op, f, ke, km, ep, d_x, d_y, s, nullptr, nullptr, bt);
    }
}

void launch_apply_k(mycel_op *op, const double *d_x, double *d_y, int mode, hipStream_t s, double *d_dot_partial,
                    const int *d_stop, const Batch &bt) {
    if (op->ku_form == MYCEL_MATVEC_REFERENCE_ORDER && !d_dot_partial && bt.count == 1) {
        launch_apply_k_ordered(op, d_x, d_y, mode, s, d_stop);
        return;
    }
    launch_apply_k_t<double>(op, op->ks64, op->km64, op->d_ep.get(), d_x, d_y, mode, s, d_dot_partial, d_stop, bt);
}

void launch_apply_k_f32(mycel_op *op, const float *d_x, float *d_y, int mode, hipStream_t s, double *d_dot_partial,
                        const int *d_stop) {
    launch_apply_k_t<float>(op, op->ks32, op->km32, op->d_ep32.get(), d_x, d_y, mode, s, d_dot_partial, d_stop, Batch{});
}

void launch_von_mises(mycel_op *op, const double *d_u, double *d_vm, hipStream_t s) {
    const int na = static_cast<int>(op->active.size());
    MYCEL_HIP(hipMemsetAsync(d_vm, 0, sizeof(double) * op->d.n_elem, s));
    if (!na) return;
    hipLaunchKernelGGL(k_von_mises, dim3(cdiv(na, 256)), dim3(256), 0, s, op->d, op->d_active.get(), na,
                       op->d_ke.get() + 600, op->d_ep.get(), d_u, 1.0 / op->h, d_vm);
    MYCEL_LAUNCH_CHECK();
}

void launch_sensitivities(mycel_op *op, const double *d_u, const double *d_rho, double *d_dc, double penal,
                          hipStream_t s) {
    const int na = static_cast<int>(op->active.size());
    if (!na) return;
    hipLaunchKernelGGL(k_sensitivities, dim3(cdiv(na, 256)), dim3(256), 0, s, op->d, op->d_active.get(), na,
                       op->d_ke.get(), d_u, d_rho, penal, op->e0 - op->emin, d_dc);
    MYCEL_LAUNCH_CHECK();
}

void launch_avg_e2n(mycel_op *op, const double *d_elem, double *d_nodal, hipStream_t s) {
    hipLaunchKernelGGL(k_avg_e2n, dim3(cdiv(op->d.n_nodes, 256)), dim3(256), 0, s, op->d, op->d_domain.get(), d_elem,
                       d_nodal);
    MYCEL_LAUNCH_CHECK();
}

void launch_sample_nodal(mycel_op *op, const double *d_nodal, int n_points, const float *d_pts, const float *origin,
                         double *d_out, int *d_fallback, hipStream_t s) {
    if (n_points <= 0) return;
    const float h = static_cast<float>(op->h);
    hipLaunchKernelGGL(k_sample_trilinear, dim3(cdiv(n_points, 256)), dim3(256), 0, s, op->d,
                       op->d_active_nodes.get(), d_nodal, n_points, d_pts, origin[0], origin[1], origin[2], h, d_out,
                       d_fallback);
    hipLaunchKernelGGL(k_sample_nearest, dim3(n_points), dim3(256), 0, s, op->d, op->d_active_nodes.get(), d_nodal,
                       d_pts, origin[0], origin[1], origin[2], h, d_fallback, d_out);
    MYCEL_LAUNCH_CHECK();
}

static double reduce_impl(mycel_op *op, const double *a, const double *b, int n, const uint8_t *flags,
                          hipStream_t s) {
    const int nb = cdiv(n, 1024);
    if (op->d_part.size() < static_cast<size_t>(nb) + 1) op->d_part.resize(nb + 1);
    if (flags)
        hipLaunchKernelGGL(k_dot_partial<true>, dim3(nb), dim3(256), 0, s, a, b, n, flags, op->d_part.get());
    else
        hipLaunchKernelGGL(k_dot_partial<false>, dim3(nb), dim3(256), 0, s, a, b, n, flags, op->d_part.get());
    MYCEL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, op->d_part.get(), nb, op->d_part.get() + nb);
    MYCEL_LAUNCH_CHECK();
    double out = 0.0;
    MYCEL_HIP(hipMemcpyAsync(&out, op->d_part.get() + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    MYCEL_HIP(hipStreamSynchronize(s));
    return out;
}

double reduce_dot(mycel_op *op, const double *a, const double *b, int n, hipStream_t s) {
    return reduce_impl(op, a, b, n, nullptr, s);
}

double reduce_dot_masked(mycel_op *op, const double *a, const double *b, int n, const uint8_t *flags,
                         hipStream_t s) {
    return reduce_impl(op, a, b, n, flags, s);
}

}  // namespace mycel
