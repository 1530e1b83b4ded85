// Element-centric K*u: every element's product Ke u_e is formed once, in the monomial basis of the trilinear
// element (KeMono in ke_sym.hpp), and each node sums the corner values of its elements in a fixed order.
//
// One wavefront owns a tile of 8 x 8 element columns, elements (i0 - 1 + a, j0 - 1 + b) for lane a + 8 b, and
// marches them along z. Per element layer a lane loads only the element's upper face (4 nodes) and its modulus:
// the lower face is the face the lane loaded one step before. The Walsh-Hadamard transform factors by face
// (corner values -> face sums in x and y, then faces -> element coefficients in z), so the face sums carry from
// step to step as well; the inverse transform is linear, so the upper-face part of one element and the
// lower-face part of the element above are added before the x/y stage, once per node layer. The node owned by
// lane (a, b), (i0 - 1 + a, j0 - 1 + b), collects the corner values of the four element columns around it from
// lanes a - 1 (a DPP row shift) and b - 1 (a lane permute), in the fixed order
//     (own x0y0 + left x1y0) + below(own x0y1 + left x1y1).
// Lanes with a = 0 or b = 0 own the tile's halo columns: they form elements for their neighbours and write no
// node, so a tile writes 7 x 7 node columns. No atomics and no barriers; the result is deterministic.
//
// Work per element: one face transform (24 add/sub), the element coefficients (21), the sparse product with
// the modulus (43) and the select on E != 0 (18), the inverse (54): about 160 operations instead of the 576
// multiply-adds of the dense product, plus 9 exchanged values per node.
#pragma once

#include "apply_k.hpp"
#include "ke_sym.hpp"

namespace mycel {

constexpr int kEcTile = 8;              // element columns per tile edge
constexpr int kEcOut = kEcTile - 1;     // node columns per tile edge that a tile writes

template <typename T>
struct FaceSums {
    T s, x, y, xy;  // sums of a face's 4 corner values weighted by 1, xi, eta, xi eta
};

// Corner values c0..c3 at (x0,y0), (x1,y0), (x1,y1), (x0,y1) of one face.
template <typename T>
__device__ __forceinline__ FaceSums<T> face_sums(T c0, T c1, T c2, T c3) {
    const T r0 = c1 + c0, r0x = c1 - c0, r1 = c2 + c3, r1x = c2 - c3;
    return {r1 + r0, r1x + r0x, r1 - r0, r1x - r0x};
}

// Element coefficients per component from the lower (z0) and upper (z1) face sums: monomials xi, eta, zeta,
// xi eta, xi zeta, eta zeta, xi eta zeta (the constant one has no stiffness and is not formed).
template <typename T>
__device__ __forceinline__ void mono_coefficients(const FaceSums<T> (&lo)[3], const FaceSums<T> (&hi)[3],
                                                  T (&c)[3][7]) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
        c[d][0] = hi[d].x + lo[d].x;
        c[d][1] = hi[d].y + lo[d].y;
        c[d][2] = hi[d].s - lo[d].s;
        c[d][3] = hi[d].xy + lo[d].xy;
        c[d][4] = hi[d].x - lo[d].x;
        c[d][5] = hi[d].y - lo[d].y;
        c[d][6] = hi[d].xy - lo[d].xy;
    }
}

// gm = E M c (same layout as c). Modulus-0 elements give exactly 0 (a select, so non-finite values at their
// nodes stay out).
template <typename T>
__device__ __forceinline__ void mono_product(const KeMono<T> &km, T E, const T (&c)[3][7], T (&gm)[3][7]) {
    const bool on = E != T(0);
    const T amb = km.a - km.b, fmh = km.f - km.h;
    const T tr = (c[0][0] + c[1][1]) + c[2][2];
    const T bt = (E * km.b) * tr, ab = E * amb;
    gm[0][0] = on ? fma_t(ab, c[0][0], bt) : T(0);
    gm[1][1] = on ? fma_t(ab, c[1][1], bt) : T(0);
    gm[2][2] = on ? fma_t(ab, c[2][2], bt) : T(0);
    const T ce = E * km.c;
    const T s_uv = on ? ce * (c[0][1] + c[1][0]) : T(0);  // (u_eta, v_xi)
    const T s_uw = on ? ce * (c[0][2] + c[2][0]) : T(0);  // (u_zeta, w_xi)
    const T s_vw = on ? ce * (c[1][2] + c[2][1]) : T(0);  // (v_zeta, w_eta)
    gm[0][1] = s_uv;
    gm[1][0] = s_uv;
    gm[0][2] = s_uw;
    gm[2][0] = s_uw;
    gm[1][2] = s_vw;
    gm[2][1] = s_vw;
    const T de = E * km.d, ee = E * km.e;
    gm[0][3] = on ? fma_t(de, c[0][3], ee * c[2][5]) : T(0);  // (u_xi.eta, w_eta.zeta)
    gm[2][5] = on ? fma_t(ee, c[0][3], de * c[2][5]) : T(0);
    gm[0][4] = on ? fma_t(de, c[0][4], ee * c[1][5]) : T(0);  // (u_xi.zeta, v_eta.zeta)
    gm[1][5] = on ? fma_t(ee, c[0][4], de * c[1][5]) : T(0);
    gm[1][3] = on ? fma_t(de, c[1][3], ee * c[2][4]) : T(0);  // (v_xi.eta, w_xi.zeta)
    gm[2][4] = on ? fma_t(ee, c[1][3], de * c[2][4]) : T(0);
    const T tw = (c[0][5] + c[1][4]) + c[2][3];  // (u_eta.zeta, v_xi.zeta, w_xi.eta)
    const T ht = (E * km.h) * tw, fh = E * fmh;
    gm[0][5] = on ? fma_t(fh, c[0][5], ht) : T(0);
    gm[1][4] = on ? fma_t(fh, c[1][4], ht) : T(0);
    gm[2][3] = on ? fma_t(fh, c[2][3], ht) : T(0);
    const T qe = E * km.q;
    gm[0][6] = on ? qe * c[0][6] : T(0);
    gm[1][6] = on ? qe * c[1][6] : T(0);
    gm[2][6] = on ? qe * c[2][6] : T(0);
}

// Inverse transform of one face from its sums: the corner values f0..f3 at (x0,y0), (x1,y0), (x1,y1), (x0,y1).
template <typename T>
__device__ __forceinline__ void face_corners(const FaceSums<T> &F, T &f0, T &f1, T &f2, T &f3) {
    const T r0 = F.s - F.y, r1 = F.s + F.y, r0x = F.x - F.xy, r1x = F.x + F.xy;
    f0 = r0 - r0x;
    f1 = r0 + r0x;
    f3 = r1 - r1x;
    f2 = r1 + r1x;
}

// Value of lane - 1 within a row of 16 lanes (lanes with a = 0 read a value they do not use).
__device__ __forceinline__ float from_left(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, false));
}
__device__ __forceinline__ double from_left(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x111, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x111, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// Value of lane - 8 (lanes with b = 0 read a value they do not use).
template <typename T>
__device__ __forceinline__ T from_below(T v) {
    return __shfl_up(v, kEcTile, 64);
}

// Tiles of a grid: cdiv(nx1, 7) x cdiv(ny1, 7) x cdiv(nz1, KZ), x fastest.
__host__ __device__ inline int ec_tiles(const Dims &g, int kz) {
    return ((g.nx1 + kEcOut - 1) / kEcOut) * ((g.ny1 + kEcOut - 1) / kEcOut) * ((g.nz1 + kz - 1) / kz);
}

// Where a tile reads x: the vector itself, with constrained DOFs read as 0 when MASK. A source has load (the raw
// words of one DOF at vector offset at + byte offset o, component d), value (raw words -> x) and ready, which
// the march calls once, after its first loads are issued and before any value is used (false: the march ends
// there and writes nothing).
template <typename T, bool MASK>
struct EcVector {
    const T *x;
    const uint8_t *flags;
    using Raw = T;
    __device__ Raw load(size_t at, uint32_t o, int d) const {
        T v = ld_off(x + at, o, d * static_cast<int>(sizeof(T)));
        if (MASK && (ld_off(flags + at, o / static_cast<uint32_t>(sizeof(T)), d) & kFlagConstrained)) v = T(0);
        return v;
    }
    __device__ T value(Raw r) const { return r; }
    template <typename Fin>
    __device__ bool ready(Fin &) {
        return true;
    }
};

// One wavefront's tile: node layers KZ*tz .. KZ*tz + KZ - 1 of its 7 x 7 node columns, x read through src. The
// epilogue fin has two stages: fin.pre(n) at the start of the step that completes node n (its loads then
// overlap the step's face loads and arithmetic; halo lanes pass a valid clamped node), and fin(n, kx, xc, pre)
// for every node the tile writes, with the raw (K x) at node n and the node's own x values as read. Called by
// all 64 lanes of the wave (tile is wave-uniform).
template <typename T, int KZ, typename Src, typename Fin>
__device__ __forceinline__ void ec_tile_march(const Dims &g, const KeMono<T> &km, const T *__restrict__ ep,
                                              Src &src, int tile, Fin &&fin) {
    const int lane = threadIdx.x & 63, a = lane & (kEcTile - 1), b = lane >> 3;
    const int ntx = (g.nx1 + kEcOut - 1) / kEcOut, nty = (g.ny1 + kEcOut - 1) / kEcOut;
    const int tx = tile % ntx, tr = tile / ntx, ty = tr % nty, tz = tr / nty;
    const int ie = kEcOut * tx - 1 + a, je = kEcOut * ty - 1 + b, k0 = KZ * tz;
    const bool col_ok = ie <= g.nx && je <= g.ny;  // the padded moduli hold this element column
    const bool out_ok = a > 0 && b > 0 && col_ok;  // (ie, je) is a node column of this tile's output
    constexpr uint32_t sT = sizeof(T), sx = 3 * sT;
    const int i0 = min(max(ie, 0), g.nx), i1 = min(ie + 1, g.nx), j0 = min(max(je, 0), g.ny), j1 = min(je + 1, g.ny);
    const uint32_t o[4] = {sx * (i0 + g.nx1 * j0), sx * (i1 + g.nx1 * j0), sx * (i1 + g.nx1 * j1),
                           sx * (i0 + g.nx1 * j1)};
    const uint32_t eo = col_ok ? sT * ((ie + 1) + g.px * (je + 1)) : 0u;
    const size_t layer = static_cast<size_t>(g.nx1) * g.ny1, elayer = static_cast<size_t>(g.px) * g.py;

    using Raw = typename Src::Raw;
    // Node layer kn (clamped to the grid: out-of-grid layers belong to elements of modulus 0).
    auto load_face = [&](int kn, Raw (&v)[4][3]) {
        const size_t at = 3 * layer * static_cast<size_t>(min(max(kn, 0), g.nz));
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int d = 0; d < 3; d++) v[q][d] = src.load(at, o[q], d);
    };

    Raw cur[4][3];
    load_face(k0 - 1, cur);
    const int kend = min(k0 + KZ - 1, g.nz);  // last node layer written
    // Software pipeline: step s issues the loads of step s + 1 (its upper face, its modulus and the epilogue's
    // loads) before its own arithmetic and stores, so a tile waits for memory about once, not once per step.
    // Every step runs (straight-line code); on the last z-block the steps past the grid read clamped layers,
    // take modulus 0 and write nothing. eo = 0 (a valid word) outside the padded columns.
    Raw up[4][3];
    load_face(k0, up);
    T Ev = ld_off(ep + elayer * (min(k0 - 1, g.nz) + 1), eo);
    if (!src.ready(fin)) return;
    FaceSums<T> lo[3], carry[3];
    T xc[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        lo[d] = face_sums(src.value(cur[0][d]), src.value(cur[1][d]), src.value(cur[2][d]), src.value(cur[3][d]));
        carry[d] = {T(0), T(0), T(0), T(0)};
        xc[d] = src.value(cur[0][d]);
    }
    decltype(fin.pre(0)) pre{};
#pragma unroll
    for (int s = 0; s <= KZ; s++) {
        const int ke = k0 - 1 + s;  // element layer of this step; node layer ke is completed by it
        const bool live = ke <= kend;
        Raw upn[4][3];
        T Evn = T(0);
        decltype(fin.pre(0)) pren{};
        if (s < KZ) {
            const int kn = min(ke + 1, g.nz);
            load_face(ke + 2, upn);
            Evn = ld_off(ep + elayer * (kn + 1), eo);
            pren = fin.pre(i0 + g.nx1 * (j0 + g.ny1 * kn));
        }
        const T E = col_ok && live ? Ev : T(0);
        FaceSums<T> hi[3];
        T c[3][7], gm[3][7];
#pragma unroll
        for (int d = 0; d < 3; d++)
            hi[d] = face_sums(src.value(up[0][d]), src.value(up[1][d]), src.value(up[2][d]), src.value(up[3][d]));
        mono_coefficients(lo, hi, c);
        mono_product(km, E, c, gm);
        // Inverse transform, z stage: the element's lower-face part joins the carried upper-face part of the
        // element below (node layer ke); its own upper-face part is carried to node layer ke + 1.
        FaceSums<T> F[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const T gx = gm[d][0], gy = gm[d][1], gz = gm[d][2], gxy = gm[d][3], gxz = gm[d][4], gyz = gm[d][5],
                    gxyz = gm[d][6];
            F[d].s = carry[d].s - gz;
            F[d].x = (carry[d].x + gx) - gxz;
            F[d].y = (carry[d].y + gy) - gyz;
            F[d].xy = (carry[d].xy + gxy) - gxyz;
            carry[d] = {gz, gx + gxz, gy + gyz, gxy + gxyz};
            lo[d] = hi[d];
        }
        if (s > 0) {
            T out[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                // x/y stage: the face's corner values, then the node's four element columns.
                T f0, f1, f2, f3;
                face_corners(F[d], f0, f1, f2, f3);
                const T q0 = f0 + from_left(f1), q3 = f3 + from_left(f2);
                out[d] = q0 + from_below(q3);
            }
            if (out_ok && live) fin(ie + g.nx1 * (je + g.ny1 * ke), out, xc, pre);
        }
#pragma unroll
        for (int d = 0; d < 3; d++) xc[d] = src.value(up[0][d]);
        if (s < KZ) {
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int d = 0; d < 3; d++) up[q][d] = upn[q][d];
            Ev = Evn;
            pre = pren;
        }
    }
}

// The march on a plain vector x (constrained DOFs read as 0 when MASK).
template <typename T, bool MASK, int KZ, typename Fin>
__device__ __forceinline__ void ec_tile_run(const Dims &g, const KeMono<T> &km, const T *__restrict__ ep,
                                            const T *__restrict__ x, const uint8_t *__restrict__ flags, int tile,
                                            Fin &&fin) {
    EcVector<T, MASK> src{x, flags};
    ec_tile_march<T, KZ>(g, km, ep, src, tile, fin);
}

// Epilogue of K x (MODE: as ku_gather): mode handling, store, and the x.y contribution.
template <typename T, int MODE>
struct KuFin {
    const T *x;
    const uint8_t *flags;
    T *y;
    double dot = 0.0;
    struct Pre {
        uint8_t fl[3];
        T xd[3];
    };
    __device__ Pre pre(int n) const {
        Pre p{};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (MODE != kApplyK) p.fl[d] = flags[3 * n + d];
            if (MODE == kApplyAGeneral) p.xd[d] = x[3 * n + d];
        }
        return p;
    }
    __device__ void operator()(int n, T (&out)[3], const T (&xc)[3], const Pre &p) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (MODE == kApplyAProjected) {
                if (p.fl[d] & kFlagFixed) out[d] = xc[d];
            } else if (MODE == kApplyAGeneral) {
                if (p.fl[d] & kFlagConstrained) out[d] = p.xd[d];
            }
            y[3 * n + d] = out[d];
            dot += static_cast<double>(MODE == kApplyAGeneral ? p.xd[d] : xc[d]) * static_cast<double>(out[d]);
        }
    }
};

// K x (MODE: as ku_gather) for one tile; returns the lane's x.y contribution.
template <typename T, int MODE, int KZ>
__device__ __forceinline__ double ec_tile(const Dims &g, const KeMono<T> &km, const T *__restrict__ ep,
                                          const T *__restrict__ x, const uint8_t *__restrict__ flags,
                                          T *__restrict__ y, int tile) {
    KuFin<T, MODE> fin{x, flags, y};
    ec_tile_run<T, MODE == kApplyAGeneral, KZ>(g, km, ep, x, flags, tile, fin);
    return fin.dot;
}

}  // namespace mycel
