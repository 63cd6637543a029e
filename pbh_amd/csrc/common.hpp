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
// Order-preserving map double -> int64 (for atomicMin on doubles of any sign).
__device__ __forceinline__ long long ordered_bits(double v) {
    long long i = __double_as_longlong(v);
    return i >= 0 ? i : (i ^ 0x7fffffffffffffffLL);
}

}  // namespace pm3d
