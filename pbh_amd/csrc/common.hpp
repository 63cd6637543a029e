// Shared device helpers of the pm3d HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pm3d {

constexpr int kThreads = 256;  // 4 wave64 per workgroup in every kernel

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
    R re, im;
};

template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
// a*b with explicit fused multiply-adds (FFT interior only; the parity-critical particle and
// Green's-function arithmetic never goes through here).
template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}
template <typename R>
__device__ __forceinline__ cplx<R> cconj(cplx<R> a) { return {a.re, -a.im}; }

// numpy.remainder for floats: fmod with the sign moved to the divisor's; may return b itself when
// a is a tiny negative number (numpy npy_divmod semantics; pm3d.py:124 `% self.L`).
__device__ __forceinline__ double npy_rem(double a, double b) {
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

// One axis of the cell-centred CIC stencil in FP64 (pm3d.py:76-79): g = X/dx - 0.5 with IEEE
// division, i0 = floor(g), f = g - i0; returns (i0 mod N, (i0+1) mod N) with Python's
// non-negative modulo.
__device__ __forceinline__ void cic_axis(double X, double dx, int N, int& lo, int& hi, double& f) {
    const double g = X / dx - 0.5;
    const double fl = floor(g);
    const long long i0 = (long long)fl;
    f = g - (double)i0;
    long long m = i0 % N;
    if (m < 0) m += N;
    lo = (int)m;
    hi = (m + 1 == N) ? 0 : (int)(m + 1);
}

// Wave64 max / min reductions through DPP-free shuffles.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}
__device__ __forceinline__ long long wave_min_i64(long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        long long w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
// Block-wide max of a non-negative double over kThreads threads (all threads must call).
__device__ __forceinline__ double block_max_nonneg(double v) {
    __shared__ unsigned long long wmax[kThreads / 64];
    const unsigned long long bits = wave_max_u64((unsigned long long)__double_as_longlong(v));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = bits;
    __syncthreads();
    unsigned long long m = wmax[0];
    for (int w = 1; w < kThreads / 64; ++w) m = wmax[w] > m ? wmax[w] : m;
    __syncthreads();
    return __longlong_as_double((long long)m);
}

// Order-preserving map double -> int64 (for min/max of doubles of any sign as integers).
__device__ __forceinline__ long long ordered_bits(double v) {
    long long i = __double_as_longlong(v);
    return i >= 0 ? i : (i ^ 0x7fffffffffffffffLL);
}
__device__ __forceinline__ double from_ordered_bits(long long o) {
    return __longlong_as_double(o >= 0 ? o : (o ^ 0x7fffffffffffffffLL));
}

// Block-wide min of a double over kThreads threads (all threads must call).
__device__ __forceinline__ double block_min(double v) {
    __shared__ long long wmin[kThreads / 64];
    const long long o = wave_min_i64(ordered_bits(v));
    if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = o;
    __syncthreads();
    long long m = wmin[0];
    for (int w = 1; w < kThreads / 64; ++w) m = wmin[w] < m ? wmin[w] : m;
    __syncthreads();
    return from_ordered_bits(m);
}

// ----------------------------------------------------------------- deterministic fixed point
// Cell accumulators of the deposit.  Integer addition commutes, so every launch order gives the
// same bits.
//   narrow: one uint64 word per cell, value = q * 2^-FB
//   wide  : two uint64 words {lo, hi} per cell, value = (hi*2^64 + lo) * 2^-60; the carry out of lo
//           is detected from the value atomicAdd returns, and the number of carries equals
//           floor(sum/2^64) whatever the order.  Quantum 2^-60 ~ 8.7e-19 of a particle, so the FP64
//           mode is exact to FP64 rounding even where delta = rho - 1 is tiny.
constexpr int kWideFB = 60;

template <bool WIDE>
__device__ __forceinline__ void fx_atomic_add(unsigned long long* mesh, size_t cell, unsigned long long q) {
    if constexpr (WIDE) {
        unsigned long long* c = mesh + 2 * cell;
        const unsigned long long old = atomicAdd(c, q);
        if (old + q < old) atomicAdd(c + 1, 1ull);
    } else {
        atomicAdd(mesh + cell, q);
    }
}

template <bool WIDE>
__device__ __forceinline__ double fx_value(const unsigned long long* mesh, size_t cell, double qscale) {
    if constexpr (WIDE) {
        const ulonglong2 v = *(const ulonglong2*)(mesh + 2 * cell);
        return (double)v.y * 16.0 + (double)v.x * 0x1p-60;
    } else {
        return (double)mesh[cell] * qscale;
    }
}

}  // namespace pm3d
