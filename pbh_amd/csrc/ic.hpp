// Initial conditions on the device (gfx950): a triaxial Gaussian density peak plus an optional
// Gaussian random field, displaced to the Zel'dovich growing mode (the construction of
// pm3d.py:59-68, generalised to an explicit linear density field at a_i).
#pragma once
#include "common.hpp"

namespace pm3d {

// SplitMix64 finaliser (Steele, Lea & Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014): a counter-based stream, so cell i of a field has the same value on the host
// (numpy uint64 arithmetic) and on the device, whatever the launch shape.
__host__ __device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Standard normal deviate of cell i: Box-Muller on u1 in (0, 1], u2 in [0, 1) (53-bit).
__device__ __forceinline__ double cell_normal(unsigned long long key, long long i) {
    const unsigned long long a = splitmix64(key + 2ull * (unsigned long long)i);
    const unsigned long long b = splitmix64(key + 2ull * (unsigned long long)i + 1ull);
    const double u1 = (double)((a >> 11) + 1ull) * 0x1p-53;
    const double u2 = (double)(b >> 11) * 0x1p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__global__ __launch_bounds__(kThreads) void k_white(double* __restrict__ w, long long M, unsigned long long key) {
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < M; i += (long long)gridDim.x * kThreads)
        w[i] = cell_normal(key, i);
}

// delta_i = mu * exp(-0.5*((X/sx)^2 + (Y/sy)^2 + (Z/sz)^2)) at the cell centres, coordinates
// relative to the box-centre vertex, X = (i + 0.5)*dx - L/2 (pm3d.py:42-45).
__global__ __launch_bounds__(kThreads) void k_peak_field(double* __restrict__ out, int N, double dx, double L, double mu,
                                                         double sx, double sy, double sz) {
    const long long M = (long long)N * N * N;
    for (long long c = (long long)blockIdx.x * kThreads + threadIdx.x; c < M; c += (long long)gridDim.x * kThreads) {
        const int i = (int)(c / ((long long)N * N)), j = (int)((c / N) % N), k = (int)(c % N);
        const double X = (i + 0.5) * dx - L / 2, Y = (j + 0.5) * dx - L / 2, Z = (k + 0.5) * dx - L / 2;
        const double X2 = (X / sx) * (X / sx), Y2 = (Y / sy) * (Y / sy), Z2 = (Z / sz) * (Z / sz);
        out[c] = mu * exp(-0.5 * ((X2 + Y2) + Z2));
    }
}

// Half spectrum (device layout) times sqrt(P(k)) = 1/|k| for 0 < |k| <= kcut, else 0.
template <typename R>
__global__ __launch_bounds__(kThreads) void k_spec_amp(cplx<R>* __restrict__ spec, cplx<R>* __restrict__ nyq, int N,
                                                       double kval, double kcut) {
    const int NH = N / 2;
    const long long main = (long long)N * N * NH, tot = main + (long long)N * N;
    for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < tot; t += (long long)gridDim.x * kThreads) {
        int ix, iy, iz;
        cplx<R>* z;
        if (t < main) {
            ix = (int)(t / ((long long)N * NH));
            iy = (int)((t / NH) % N);
            iz = (int)(t % NH);
            z = spec + t;
        } else {
            const long long r = t - main;
            ix = (int)(r / N);
            iy = (int)(r % N);
            iz = NH;
            z = nyq + r;
        }
        const double kx = 6.283185307179586 * ((double)(ix < N / 2 ? ix : ix - N) * kval);
        const double ky = 6.283185307179586 * ((double)(iy < N / 2 ? iy : iy - N) * kval);
        const double kz = 6.283185307179586 * ((double)iz * kval);
        const double kk = sqrt((kx * kx + ky * ky) + kz * kz);
        const double amp = (kk > 0.0 && kk <= kcut) ? 1.0 / fmax(kk, 1e-30) : 0.0;
        *z = cplx<R>{(R)((double)z->re * amp), (R)((double)z->im * amp)};
    }
}

// Per-block sums of f and f^2 (host adds the blocks: deterministic).
__global__ __launch_bounds__(kThreads) void k_moments(const double* __restrict__ f, long long M, double* __restrict__ out) {
    __shared__ double s1[kThreads], s2[kThreads];
    double a = 0.0, b = 0.0;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < M; i += (long long)gridDim.x * kThreads) {
        a += f[i];
        b += f[i] * f[i];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = s1[0];
        out[2 * blockIdx.x + 1] = s2[0];
    }
}

__global__ __launch_bounds__(kThreads) void k_axpy(double* __restrict__ y, const double* __restrict__ x, long long M,
                                                   double a) {
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < M; i += (long long)gridDim.x * kThreads)
        y[i] = y[i] + a * x[i];
}

// Lattice placement: particle i at Q = (i_c + 0.5)*(L/Np) - L/2 (pm3d.py:64-65); the displacement
// is the CIC gather of the component-major mesh disp[c*N^3 + cell] at (Q + L/2) % L (pm3d.py:66), and
// x = ((Q + d) + L/2) % L, p = coef * d (pm3d.py:67-68 with coef = a_i^2 H(a_i)).
__global__ __launch_bounds__(kThreads) void k_lattice(const double* __restrict__ disp, int N, int Np, double L,
                                                      double coef, double* __restrict__ x, double* __restrict__ p) {
    const long long n = (long long)Np * Np * Np;
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int s[3] = {(int)(i / ((long long)Np * Np)), (int)((i / Np) % Np), (int)(i % Np)};
    const double dq = L / Np, dx = L / N;
    const size_t M = (size_t)N * N * N;  // displacement mesh disp[d*M + cell]
    double Q[3];
    int lo[3], hi[3];
    double f[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        Q[d] = (s[d] + 0.5) * dq - L / 2;
        cic_axis_fast(npy_rem(Q[d] + L / 2, L), dx, N, lo[d], hi[d], f[d]);
    }
    const double wx[2] = {1.0 - f[0], f[0]}, wy[2] = {1.0 - f[1], f[1]}, wz[2] = {1.0 - f[2], f[2]};
    double a[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ix = (c >> 2) ? hi[0] : lo[0], iy = ((c >> 1) & 1) ? hi[1] : lo[1], iz = (c & 1) ? hi[2] : lo[2];
        const double* g = disp + (((size_t)ix * N + iy) * N + iz);
#pragma unroll
        for (int d = 0; d < 3; ++d) a[d] = a[d] + ((g[d * M] * wx[c >> 2]) * wy[(c >> 1) & 1]) * wz[c & 1];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x[d * n + i] = npy_rem((Q[d] + a[d]) + L / 2, L);
        p[d * n + i] = coef * a[d];
    }
}

// FP64 SoA state -> mixed state (device to device).
__global__ __launch_bounds__(kThreads) void k_soa_to_mx(const double* __restrict__ xs, const double* __restrict__ ps,
                                                        long long n, int N, double dx, int* __restrict__ cell,
                                                        float* __restrict__ off, float* __restrict__ pf) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int lo, hi;
        double f;
        cic_axis_fast(xs[d * n + i], dx, N, lo, hi, f);
        float ff = (float)f;
        if (ff >= 1.0f) {
            ff = 0.0f;
            lo = hi;
        }
        cell[d * n + i] = lo;
        off[d * n + i] = ff;
        pf[d * n + i] = (float)ps[d * n + i];
    }
}

}  // namespace pm3d
