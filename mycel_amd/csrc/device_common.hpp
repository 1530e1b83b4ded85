// Device-side helpers shared by the kernels: deterministic reductions, XCD-aware block mapping,
// SIMP stiffness interpolation.
#pragma once

#include <hip/hip_runtime.h>

namespace mycel {

// Hex8 local corner coordinates (Fem/Hex8Element.cs:8-16).
__host__ __device__ constexpr int corner_x(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0; }
__host__ __device__ constexpr int corner_y(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0; }
__host__ __device__ constexpr int corner_z(int a) { return a >= 4 ? 1 : 0; }

// p, known to address global memory. A pointer read from memory (a table of arguments in LDS or in a device
// buffer) is otherwise a generic one, and the compiler accesses it with flat instructions; a flat access also
// counts against the LDS/scalar wait counter, so every LDS access, lane exchange or argument read behind it waits
// for its memory round trip. The pointer returned is generic, but the compiler sees that it came from a global
// one and emits global instructions for every access through it (also in the inlined functions it is passed to).
template <typename T>
__device__ __forceinline__ T *gmem(T *p) {
    return (T *)__builtin_bit_cast(__attribute__((address_space(1))) T *, p);
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each with a private L2.
// Remap so that each XCD receives a contiguous range of logical tiles: neighbouring node tiles then
// share halo lines in the same L2. Bijective on [0, G).
__device__ __forceinline__ int xcd_block(int b, int G) {
    const int x = b & 7, q = G >> 3, r = G & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Butterfly sum over a 64-lane wavefront. lane l adds v[l] and v[l^off]; both partners compute the
// same commutative sum, so every lane ends with the bit-identical total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Fixed-order sum over a 256-thread block (4 wavefronts). Deterministic for a fixed launch shape.
__device__ __forceinline__ double block_sum256(double v, double *lds4) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds4[w] = v;
    __syncthreads();
    const double t = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    __syncthreads();
    return t;
}

// Every block that calls this computes the same value from the same partials in the same order.
__device__ __forceinline__ double sum_partials256(const double *__restrict__ part, int n, double *lds4) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) v += part[i];
    return block_sum256(v, lds4);
}

// The same fixed-order sum, formed by each wavefront on its own (no LDS, no barrier): lane l adds the
// partials l, l + 64, ... and the lanes are then combined by the butterfly. Every wave of every block
// gets the bit-identical value.
// Four independent accumulators (lane l, stride 256, offsets 0/64/128/192) keep the dependent-add chain
// short; they are combined as (a0 + a1) + (a2 + a3) before the butterfly. The loads of a round (16 per lane
// and array, 1024 partials) are all issued before the first add, so up to 1024 partials cost one memory
// round trip instead of one per 256; K arrays are summed together, each in exactly that order.
template <int K>
__device__ __forceinline__ void wave_sum_partials_n(const double *const (&part)[K], const int (&n)[K],
                                                    double (&out)[K]) {
    constexpr int U = 16;
    double acc[K][4];
    int nmax = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        acc[k][0] = acc[k][1] = acc[k][2] = acc[k][3] = 0.0;
        nmax = n[k] > nmax ? n[k] : nmax;
    }
    for (int b = threadIdx.x & 63; b < nmax; b += 64 * U) {
        double v[K][U];
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int i = b + 64 * q;
                v[k][q] = i < n[k] ? part[k][i] : 0.0;
            }
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int q = 0; q < U; q++) acc[k][q & 3] += v[k][q];
    }
#pragma unroll
    for (int k = 0; k < K; k++) out[k] = wave_sum((acc[k][0] + acc[k][1]) + (acc[k][2] + acc[k][3]));
}

__device__ __forceinline__ double wave_sum_partials(const double *__restrict__ part, int n) {
    const double *const p[1] = {part};
    const int m[1] = {n};
    double s[1];
    wave_sum_partials_n<1>(p, m, s);
    return s[0];
}

// Two interleaved partial arrays (part[2i], part[2i + 1]) summed per wave: lane l adds entries l, l + 128, ...
// into one accumulator and l + 64, l + 192, ... into another, then a0 + a1 and the butterfly. A round's
// loads (8 per lane and array) are issued together.
__device__ __forceinline__ void wave_sum_partials2(const double *__restrict__ part, int n, double &s0, double &s1) {
    constexpr int U = 8;
    double a[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
    for (int b = threadIdx.x & 63; b < n; b += 64 * U) {
        double va[U], vb[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int i = b + 64 * q;
            va[q] = i < n ? part[2 * i] : 0.0;
            vb[q] = i < n ? part[2 * i + 1] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            a[q & 1] += va[q];
            c[q & 1] += vb[q];
        }
    }
    s0 = wave_sum(a[0] + a[1]);
    s1 = wave_sum(c[0] + c[1]);
}

// Two fixed-order block sums with one LDS round trip.
__device__ __forceinline__ void block_sum256x2(double &a, double &b, double *lds8) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds8[w] = a;
        lds8[4 + w] = b;
    }
    __syncthreads();
    a = (lds8[0] + lds8[1]) + (lds8[2] + lds8[3]);
    b = (lds8[4] + lds8[5]) + (lds8[6] + lds8[7]);
    __syncthreads();
}

// rho^3 rounded once to nearest (up to double-rounding ties): rho^2 is split exactly into hi + lo
// with an FMA, then (hi + lo) * rho is formed with one more error-free product.
__device__ __forceinline__ double cube_rn(double r) {
    const double hi = r * r;
    const double lo = __fma_rn(r, r, -hi);
    const double p1 = hi * r;
    const double e1 = __fma_rn(hi, r, -p1);
    return p1 + __fma_rn(lo, r, e1);
}

// rho^p as Math.Pow would give it for the penalties in use (p = 1, 2, 3 exactly; pow otherwise).
__device__ __forceinline__ double pow_penal(double r, double p) {
    if (p == 3.0) return cube_rn(r);
    if (p == 2.0) return r * r;
    if (p == 1.0) return r;
    return pow(r, p);
}

}  // namespace mycel
