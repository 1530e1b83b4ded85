// Node-centric Hex8 K*u for one node, shared by the stand-alone K*u kernels and the PCG.
//
// One thread owns one node and forms its three output rows from its <= 8 incident elements,
// y_n = sum_a E_e(a) Ke[3a:3a+3, :] u_e(a) (a = the node's corner in element e). Every element row is
// computed exactly once, by the node that owns it: no colouring, no atomics, deterministic.
//
// Ke never leaves scalar registers: by the cube's symmetry every entry is +-v[k] for one of 14 class
// values (ke_sym.hpp), which arrive as kernel arguments; class and sign are compile-time constants of the
// fully unrolled loops, so each of the 576 products is one FMA with an SGPR operand and a sign modifier,
// and no load of Ke ever stalls the wave. The 27-node neighbourhood is held in registers. E lives on a
// zero-padded element grid, so inactive and out-of-box elements contribute nothing and the boundary needs
// no branches; neighbour coordinates are clamped and only ever feed such elements.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "ke_sym.hpp"
#include "mycel_internal.hpp"

namespace mycel {

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __fma_rn(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __fmaf_rn(a, b, c); }

// Rows 3A..3A+2 of Ke times the element vector of the element that holds the node as corner A.
template <int A, typename T>
__device__ __forceinline__ void corner_rows(const KeSym<T> &ks, const T (&u)[27][3], T E, T &acc0, T &acc1,
                                            T &acc2) {
    T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const int s = (corner_x(b) - corner_x(A) + 1) + 3 * (corner_y(b) - corner_y(A) + 1) +
                      9 * (corner_z(b) - corner_z(A) + 1);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int c = 3 * b + d;
            const T v = u[s][d];
            const T k0 = ks.v[kKeClass[3 * A + 0][c]], k1 = ks.v[kKeClass[3 * A + 1][c]],
                    k2 = ks.v[kKeClass[3 * A + 2][c]];
            t0 = fma_t(kKeSign[3 * A + 0][c] > 0 ? k0 : -k0, v, t0);
            t1 = fma_t(kKeSign[3 * A + 1][c] > 0 ? k1 : -k1, v, t1);
            t2 = fma_t(kKeSign[3 * A + 2][c] > 0 ? k2 : -k2, v, t2);
        }
    }
    acc0 = fma_t(E, t0, acc0);
    acc1 = fma_t(E, t1, acc1);
    acc2 = fma_t(E, t2, acc2);
}

template <int A, typename T>
__device__ __forceinline__ void all_corners(const KeSym<T> &ks, const T (&u)[27][3], const T (&E)[8], T &acc0, T &acc1,
                                            T &acc2) {
    if (E[A] != T(0)) corner_rows<A>(ks, u, E[A], acc0, acc1, acc2);
    if constexpr (A < 7) all_corners<A + 1>(ks, u, E, acc0, acc1, acc2);
}

// Rows 3A..3A+2 of Ke times the element vector of the element that holds the node as corner A, with the
// element's nodes taken from two 3x3 node layers: `lo` at the element's lower z, `up` at its upper z. The
// product joins the accumulators only when E != 0, so an element outside the domain contributes nothing even
// if x holds non-finite values there (a select, not a branch: it does not split the basic block).
template <int A, typename T>
__device__ __forceinline__ void corner_rows_slab(const KeSym<T> &ks, const T (&lo)[9][3], const T (&up)[9][3], T E,
                                                 T &acc0, T &acc1, T &acc2) {
    T t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const int s = (corner_x(b) - corner_x(A) + 1) + 3 * (corner_y(b) - corner_y(A) + 1);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int c = 3 * b + d;
            const T v = corner_z(b) ? up[s][d] : lo[s][d];
            const T k0 = ks.v[kKeClass[3 * A + 0][c]], k1 = ks.v[kKeClass[3 * A + 1][c]],
                    k2 = ks.v[kKeClass[3 * A + 2][c]];
            t0 = fma_t(kKeSign[3 * A + 0][c] > 0 ? k0 : -k0, v, t0);
            t1 = fma_t(kKeSign[3 * A + 1][c] > 0 ? k1 : -k1, v, t1);
            t2 = fma_t(kKeSign[3 * A + 2][c] > 0 ? k2 : -k2, v, t2);
        }
    }
    const bool on = E != T(0);
    acc0 = on ? fma_t(E, t0, acc0) : acc0;
    acc1 = on ? fma_t(E, t1, acc1) : acc1;
    acc2 = on ? fma_t(E, t2, acc2) : acc2;
}

// Load of a T at a byte offset from a wave-uniform base: the base stays in scalar registers and the offset is
// one 32-bit VGPR, so each load costs no 64-bit address arithmetic and constant offsets fold into the
// instruction.
template <typename T>
__device__ __forceinline__ T ld_off(const T *__restrict__ base, uint32_t byte_off, int imm = 0) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + static_cast<size_t>(byte_off) + imm);
}

// One 3x3 layer of node values: row byte offsets rows[dy] (including z) and column byte offsets xs[dx], both
// relative to xb (and, for the flags, fb with the same node numbering at one byte per DOF).
template <typename T, int MODE>
__device__ __forceinline__ void load_layer(const T *__restrict__ xb, const uint8_t *__restrict__ fb,
                                           const uint32_t (&xs)[3], const uint32_t (&rows)[3], T (&L)[9][3]) {
#pragma unroll
    for (int dy = 0; dy < 3; dy++)
#pragma unroll
        for (int dx = 0; dx < 3; dx++) {
            const uint32_t o = rows[dy] + xs[dx];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                T v = ld_off(xb, o, d * static_cast<int>(sizeof(T)));
                if (MODE == kApplyAGeneral && (ld_off(fb, o / static_cast<uint32_t>(sizeof(T)), d) & kFlagConstrained))
                    v = T(0);
                L[dx + 3 * dy][d] = v;
            }
        }
}

// First node layer of a launch's node range [n0, ...): the k0 argument of ku_gather for every node of it.
__device__ __forceinline__ int first_layer(const Dims &g, int n0) { return n0 / (g.nx1 * g.ny1); }

// Raw (K x) at node n = (i, j, k) into kx[3], and the node's own x values into xc[3] (masked like the
// neighbours in kApplyAGeneral mode).
//
// The gather runs in two z-slabs so that only two node layers are live at a time: the four elements below
// the node (corners 4..7, node layers k-1 and k), then the four above it (corners 0..3, layers k and k+1),
// whose upper layer reuses the registers of the first slab's lower one. The E values and both first layers
// are requested together up front, so the memory latencies overlap instead of chaining (no branch on the
// moduli before the gather).
//
// Addressing: k0 <= k is a node layer shared by the whole wave (the first layer of the caller's node range).
// x, flags and ep are rebased at layers max(k0 - 1, 0) and k0, which keeps every offset of the gather within
// four node layers, and each neighbour is read at a 32-bit byte offset from that uniform base.
template <typename T, int MODE>
__device__ __forceinline__ void ku_gather(const Dims &g, const KeSym<T> &ks, const T *__restrict__ ep,
                                          const T *__restrict__ x, const uint8_t *__restrict__ flags, int i, int j,
                                          int k, int k0, T (&kx)[3], T (&xc)[3]) {
    const int zb = k0 > 0 ? k0 - 1 : 0;
    const size_t layer = static_cast<size_t>(g.nx1) * g.ny1;
    const T *xb = x + 3 * layer * zb;
    const uint8_t *fb = MODE == kApplyAGeneral ? flags + 3 * layer * zb : nullptr;
    const T *eb = ep + static_cast<size_t>(g.px) * g.py * k0;
    constexpr uint32_t sT = sizeof(T);
    const uint32_t ex = sT, ey = sT * g.px, ez = ey * g.py;  // E strides (bytes)
    const uint32_t e0 = sT * ((i + 1) + g.px * ((j + 1) + g.py * (k + 1 - k0)));
    T E[8];
#pragma unroll
    for (int a = 0; a < 8; a++)
        E[a] = ld_off(eb, e0 - (corner_x(a) ? ex : 0u) - (corner_y(a) ? ey : 0u) - (corner_z(a) ? ez : 0u));
    const uint32_t sx = 3 * sT, sy = sx * g.nx1, sz = sy * g.ny1;  // node strides (bytes)
    const uint32_t oc = sx * i;
    const uint32_t xs[3] = {i > 0 ? oc - sx : oc, oc, i < g.nx ? oc + sx : oc};
    const uint32_t oy[3] = {j > 0 ? sy * (j - 1) : 0u, sy * j, j < g.ny ? sy * (j + 1) : sy * j};
    const uint32_t zc = sz * (k - zb);
    const uint32_t oz[3] = {k > 0 ? zc - sz : zc, zc, k < g.nz ? zc + sz : zc};
    uint32_t rows[3];
    T lo[9][3], mid[9][3];
#pragma unroll
    for (int dy = 0; dy < 3; dy++) rows[dy] = oy[dy] + oz[0];
    load_layer<T, MODE>(xb, fb, xs, rows, lo);
#pragma unroll
    for (int dy = 0; dy < 3; dy++) rows[dy] = oy[dy] + oz[1];
    load_layer<T, MODE>(xb, fb, xs, rows, mid);
    T acc0 = 0, acc1 = 0, acc2 = 0;
    corner_rows_slab<4>(ks, lo, mid, E[4], acc0, acc1, acc2);
    corner_rows_slab<5>(ks, lo, mid, E[5], acc0, acc1, acc2);
    corner_rows_slab<6>(ks, lo, mid, E[6], acc0, acc1, acc2);
    corner_rows_slab<7>(ks, lo, mid, E[7], acc0, acc1, acc2);
#pragma unroll
    for (int dy = 0; dy < 3; dy++) rows[dy] = oy[dy] + oz[2];
    load_layer<T, MODE>(xb, fb, xs, rows, lo);  // layer k+1
    corner_rows_slab<0>(ks, mid, lo, E[0], acc0, acc1, acc2);
    corner_rows_slab<1>(ks, mid, lo, E[1], acc0, acc1, acc2);
    corner_rows_slab<2>(ks, mid, lo, E[2], acc0, acc1, acc2);
    corner_rows_slab<3>(ks, mid, lo, E[3], acc0, acc1, acc2);
    kx[0] = acc0;
    kx[1] = acc1;
    kx[2] = acc2;
#pragma unroll
    for (int d = 0; d < 3; d++) xc[d] = mid[4][d];
}

// Mode handling and store of one node's three rows (out = raw K x at the node, xc = its own x values as the
// gather read them); returns the node's x.y contribution.
template <typename T, int MODE>
__device__ __forceinline__ double finish_node(const T *__restrict__ x, const uint8_t *__restrict__ flags,
                                              T *__restrict__ y, int n, T (&out)[3], const T (&xc)[3]) {
    double dotv = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int dof = 3 * n + d;
        if (MODE == kApplyAProjected) {
            if (flags[dof] & kFlagFixed) out[d] = xc[d];
        } else if (MODE == kApplyAGeneral) {
            if (flags[dof] & kFlagConstrained) out[d] = x[dof];
        }
        y[dof] = out[d];
        dotv += static_cast<double>(MODE == kApplyAGeneral ? x[dof] : xc[d]) * static_cast<double>(out[d]);
    }
    return dotv;
}

// z-marching form for large grids: one thread owns a column of KZ nodes (i, j, k0 .. k0 + KZ - 1) and keeps
// the three node layers of the current node in registers, so each further node costs one new 3x3 layer and the
// four moduli above it instead of three layers and eight moduli: about a third of the L1 traffic of the
// one-node-per-thread gather, which is what bounds that gather on large grids. kb: a node layer <= k0 shared
// by the wave (the addressing base, as in ku_gather). Returns the column's x.y contribution.
template <typename T, int MODE, int KZ>
__device__ __forceinline__ double apply_k_column(const Dims &g, const KeSym<T> &ks, const T *__restrict__ ep,
                                                 const T *__restrict__ x, const uint8_t *__restrict__ flags,
                                                 T *__restrict__ y, int i, int j, int k0, int kb) {
    const int zb = kb > 0 ? kb - 1 : 0;
    const size_t layer = static_cast<size_t>(g.nx1) * g.ny1;
    const T *xb = x + 3 * layer * zb;
    const uint8_t *fb = MODE == kApplyAGeneral ? flags + 3 * layer * zb : nullptr;
    const T *eb = ep + static_cast<size_t>(g.px) * g.py * kb;
    constexpr uint32_t sT = sizeof(T);
    const uint32_t ex = sT, ey = sT * g.px, ez = ey * g.py;
    const uint32_t sx = 3 * sT, sy = sx * g.nx1, sz = sy * g.ny1;
    const uint32_t oc = sx * i;
    const uint32_t xs[3] = {i > 0 ? oc - sx : oc, oc, i < g.nx ? oc + sx : oc};
    const uint32_t oy[3] = {j > 0 ? sy * (j - 1) : 0u, sy * j, j < g.ny ? sy * (j + 1) : sy * j};
    // Byte offset of node layer k (clamped to [0, nz]) and of padded modulus layer kp, from the bases.
    auto zoff = [&](int k) { return sz * static_cast<uint32_t>((k < 0 ? 0 : k > g.nz ? g.nz : k) - zb); };
    const uint32_t e_ij = sT * ((i + 1) + g.px * (j + 1));
    auto eoff = [&](int kp) { return e_ij + ez * static_cast<uint32_t>(kp - kb); };
    auto load = [&](int k, T (&L)[9][3]) {
        const uint32_t z = zoff(k);
        const uint32_t rows[3] = {oy[0] + z, oy[1] + z, oy[2] + z};
        load_layer<T, MODE>(xb, fb, xs, rows, L);
    };
    // E of the four elements in padded modulus layer kp around (i, j): corners 0-3 (node at their bottom)
    // for kp = k + 1, corners 4-7 (node at their top) for kp = k.
    auto load_e = [&](int kp, T (&E4)[4]) {
        const uint32_t o = eoff(kp);
#pragma unroll
        for (int a = 0; a < 4; a++) E4[a] = ld_off(eb, o - (corner_x(a) ? ex : 0u) - (corner_y(a) ? ey : 0u));
    };
    T la[9][3], lb[9][3], lc[9][3], e_lo[4], e_hi[4];
    load(k0 - 1, la);
    load(k0, lb);
    load_e(k0, e_lo);
    double dotv = 0.0;
#pragma unroll
    for (int st = 0; st < KZ; st++) {
        const int k = k0 + st;
        if (k > g.nz) break;
        // Roles rotate with the step: lo / mid / hi are la / lb / lc for st % 3 == 0, then lb / lc / la, ...
        T(&lo)[9][3] = st % 3 == 0 ? la : st % 3 == 1 ? lb : lc;
        T(&mid)[9][3] = st % 3 == 0 ? lb : st % 3 == 1 ? lc : la;
        T(&hi)[9][3] = st % 3 == 0 ? lc : st % 3 == 1 ? la : lb;
        load(k + 1, hi);
        load_e(k + 1, e_hi);
        T acc0 = 0, acc1 = 0, acc2 = 0;
        corner_rows_slab<4>(ks, lo, mid, e_lo[0], acc0, acc1, acc2);
        corner_rows_slab<5>(ks, lo, mid, e_lo[1], acc0, acc1, acc2);
        corner_rows_slab<6>(ks, lo, mid, e_lo[2], acc0, acc1, acc2);
        corner_rows_slab<7>(ks, lo, mid, e_lo[3], acc0, acc1, acc2);
        corner_rows_slab<0>(ks, mid, hi, e_hi[0], acc0, acc1, acc2);
        corner_rows_slab<1>(ks, mid, hi, e_hi[1], acc0, acc1, acc2);
        corner_rows_slab<2>(ks, mid, hi, e_hi[2], acc0, acc1, acc2);
        corner_rows_slab<3>(ks, mid, hi, e_hi[3], acc0, acc1, acc2);
        T out[3] = {acc0, acc1, acc2};
        const T xc[3] = {mid[4][0], mid[4][1], mid[4][2]};
        dotv += finish_node<T, MODE>(x, flags, y, i + g.nx1 * (j + g.ny1 * k), out, xc);
#pragma unroll
        for (int a = 0; a < 4; a++) e_lo[a] = e_hi[a];
    }
    return dotv;
}

// Half of K x at node (i, j, k) for the slab-split gathers (two lanes per node): half 0 sums the four
// elements below the node (corners 4-7, node layers k-1 and k), half 1 the four above it. Half 1 runs the
// very same instructions as half 0 on z-mirrored data: the cube's stiffness satisfies
// Ke[A][B] = S Ke[A^4][B^4] S with S = diag(1, 1, -1), so the elements above are the lower-corner rows applied
// to layers k+1 (as the lower layer) and k with their z components negated, and the result's z row negated.
// Both halves therefore share a wave without diverging, and each loads two layers, in one round trip.
// kx receives this half's sum, xc the node's own x values (k0: as in ku_gather).
template <typename T>
__device__ __forceinline__ void ku_gather_half(const Dims &g, const KeSym<T> &ks, const T *__restrict__ ep,
                                               const T *__restrict__ x, int i, int j, int k, int k0, int half,
                                               T (&kx)[3], T (&xc)[3]) {
    const int zb = k0 > 0 ? k0 - 1 : 0;
    const size_t layer = static_cast<size_t>(g.nx1) * g.ny1;
    const T *xb = x + 3 * layer * zb;
    const T *eb = ep + static_cast<size_t>(g.px) * g.py * k0;
    constexpr uint32_t sT = sizeof(T);
    const uint32_t ex = sT, ey = sT * g.px;
    // Padded modulus index of the element at quadrant (0, 0) of this half: layer k (below) or k + 1 (above).
    const uint32_t e0 = sT * ((i + 1) + g.px * ((j + 1) + g.py * (k + half - k0)));
    T E[4];
#pragma unroll
    for (int q = 0; q < 4; q++) E[q] = ld_off(eb, e0 - (corner_x(q) ? ex : 0u) - (corner_y(q) ? ey : 0u));
    const uint32_t sx = 3 * sT, sy = sx * g.nx1, sz = sy * g.ny1;
    const uint32_t oc = sx * i;
    const uint32_t xs[3] = {i > 0 ? oc - sx : oc, oc, i < g.nx ? oc + sx : oc};
    const uint32_t oy[3] = {j > 0 ? sy * (j - 1) : 0u, sy * j, j < g.ny ? sy * (j + 1) : sy * j};
    const int kl = half ? (k < g.nz ? k + 1 : k) : (k > 0 ? k - 1 : k);  // the (mirrored) lower layer
    const uint32_t zl = sz * (kl - zb), zm = sz * (k - zb);
    uint32_t rows[3];
    T lo[9][3], up[9][3];
#pragma unroll
    for (int dy = 0; dy < 3; dy++) rows[dy] = oy[dy] + zl;
    load_layer<T, kApplyK>(xb, nullptr, xs, rows, lo);
#pragma unroll
    for (int dy = 0; dy < 3; dy++) rows[dy] = oy[dy] + zm;
    load_layer<T, kApplyK>(xb, nullptr, xs, rows, up);
#pragma unroll
    for (int d = 0; d < 3; d++) xc[d] = up[4][d];
    const T sgn = half ? T(-1) : T(1);
#pragma unroll
    for (int q = 0; q < 9; q++) {
        lo[q][2] *= sgn;
        up[q][2] *= sgn;
    }
    T acc0 = 0, acc1 = 0, acc2 = 0;
    corner_rows_slab<4>(ks, lo, up, E[0], acc0, acc1, acc2);
    corner_rows_slab<5>(ks, lo, up, E[1], acc0, acc1, acc2);
    corner_rows_slab<6>(ks, lo, up, E[2], acc0, acc1, acc2);
    corner_rows_slab<7>(ks, lo, up, E[3], acc0, acc1, acc2);
    kx[0] = acc0;
    kx[1] = acc1;
    kx[2] = acc2 * sgn;
}

// Returns this node's contribution x.y (FP64) for fused dot products; n >= n_nodes is a no-op. k0: as in
// ku_gather.
template <typename T, int MODE>
__device__ __forceinline__ double apply_k_node(const Dims &g, const KeSym<T> &ks, const T *__restrict__ ep,
                                               const T *__restrict__ x, const uint8_t *__restrict__ flags,
                                               T *__restrict__ y, int n, int k0) {
    double dotv = 0.0;
    if (n >= g.n_nodes) return dotv;
    const int i = n % g.nx1, t = n / g.nx1, j = t % g.ny1, k = t / g.ny1;
    uint8_t fl[3] = {0, 0, 0};
    if (MODE != kApplyK) {
#pragma unroll
        for (int d = 0; d < 3; d++) fl[d] = flags[3 * n + d];
    }
    T out[3], xc[3];
    ku_gather<T, MODE>(g, ks, ep, x, flags, i, j, k, k0, out, xc);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int dof = 3 * n + d;
        if (MODE == kApplyAProjected) {
            if (fl[d] & kFlagFixed) out[d] = xc[d];
        } else if (MODE == kApplyAGeneral) {
            if (fl[d] & kFlagConstrained) out[d] = x[dof];
        }
        y[dof] = out[d];
        dotv += static_cast<double>(MODE == kApplyAGeneral ? x[dof] : xc[d]) * static_cast<double>(out[d]);
    }
    return dotv;
}

}  // namespace mycel
