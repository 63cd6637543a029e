// Diagnostics of the inner Lagrangian sets on the device (pm3d.py:132-159, `analyse`).
//
// The host supplies index lists once per run (they depend only on the Lagrangian radii r_ell_q):
// set 0 is the core whose centre of mass is the centre, sets 1.. are the inner Lagrangian sets.
// The device computes, from the particle state in HBM:
//   centre   (k_an_centre): numpy's ((xs - xs[0] + L/2) % L - L/2).mean(axis=0) over the core in
//            index order, summed sequentially as numpy's axis-0 reduction does, then
//            (xs[0] + mean) % L: bit-identical to the reference's expression;
//   per set  (k_an_sum, k_an_moments, k_an_finish): the mean of the minimum-image positions
//            rel = (x - centre + L/2) % L - L/2 and the sums of the outer products of rel - mean
//            (the reference's r.T @ r before the division by len(r)); fixed-shape tree reductions,
//            so the result is the same on every run;
//   radii    (k_an_radii): counts of sqrt((rx^2 + ry^2) + rz^2) < R_k over all particles (exact);
//   rho max  (k_an_mesh_max): maximum of a fresh deposit of the state.
// The eigen-decomposition of the 3 x 3 inertia tensors stays on the host.
#pragma once
#include "common.hpp"

namespace pm3d {

constexpr int kAnBlocks = 64;  // blocks per set for the set sums
constexpr int kAnMaxRadii = 16;

// Position of particle i along axis d in FP64 box coordinates, as pm3d_get_state returns it.
struct AnPosF64 {
    const double* x;
    long long n;
    __device__ __forceinline__ double get(long long i, int d) const { return x[d * n + i]; }
};
struct AnPosMx {
    const int* cell;
    const float* off;
    long long n;
    double dx, L;
    __device__ __forceinline__ double get(long long i, int d) const {
        double v = ((double)cell[d * n + i] + 0.5 + (double)off[d * n + i]) * dx;
        if (v >= L) v -= L;
        return v;
    }
};

// (v - c + L/2) % L - L/2 with numpy's remainder, evaluated left to right as the reference does.
__device__ __forceinline__ double an_rel(double v, double c, double L, double hL) {
    return npy_rem((v - c) + hL, L) - hL;
}

// Block sum of `nv` doubles per thread over kThreads threads, fixed tree order; result in red[0..nv).
template <int NV>
__device__ __forceinline__ void an_block_sum(double (&v)[NV], double* red) {
    __shared__ double sh[NV][kThreads];
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < NV) red[threadIdx.x] = sh[threadIdx.x][0];
    __syncthreads();
}

// One thread: the centre from the core set (idx[0..ncore), ascending particle index).
template <class Pos>
__global__ void k_an_centre(Pos pos, const int* __restrict__ idx, int ncore, double L, double hL,
                            double* __restrict__ centre) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const long long i0 = idx[0];
    double x0[3], s[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) x0[d] = pos.get(i0, d);
    for (int j = 0; j < ncore; ++j) {
        const long long i = idx[j];
        for (int d = 0; d < 3; ++d) s[d] = s[d] + an_rel(pos.get(i, d), x0[d], L, hL);
    }
    for (int d = 0; d < 3; ++d) centre[d] = npy_rem(x0[d] + s[d] / (double)ncore, L);
}

// Partial sums of rel over block b's chunk of set s (grid kAnBlocks x nsets).
// part[(s * kAnBlocks + b) * 3 + d]
template <class Pos>
__global__ __launch_bounds__(kThreads) void k_an_sum(Pos pos, const int* __restrict__ idx,
                                                     const long long* __restrict__ offs,
                                                     const double* __restrict__ centre, double L, double hL,
                                                     double* __restrict__ part) {
    const int s = blockIdx.y, b = blockIdx.x;
    const long long o0 = offs[s], m = offs[s + 1] - o0;
    const long long chunk = (m + kAnBlocks - 1) / kAnBlocks;
    const long long j0 = o0 + b * chunk, j1 = o0 + min(m, (b + 1) * chunk);
    const double c[3] = {centre[0], centre[1], centre[2]};
    double v[3] = {0.0, 0.0, 0.0};
    for (long long j = j0 + threadIdx.x; j < j1; j += kThreads) {
        const long long i = idx[j];
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] += an_rel(pos.get(i, d), c[d], L, hL);
    }
    __shared__ double red[3];
    an_block_sum<3>(v, red);
    if (threadIdx.x < 3) part[(s * kAnBlocks + b) * 3 + threadIdx.x] = red[threadIdx.x];
}

// Mean of set s from the partial sums (same order in every block), then partial sums of the outer
// products of rel - mean: part2[(s * kAnBlocks + b) * 6 + k], k = xx, xy, xz, yy, yz, zz.
template <class Pos>
__global__ __launch_bounds__(kThreads) void k_an_moments(Pos pos, const int* __restrict__ idx,
                                                         const long long* __restrict__ offs,
                                                         const double* __restrict__ centre, double L, double hL,
                                                         const double* __restrict__ part, double* __restrict__ part2) {
    const int s = blockIdx.y, b = blockIdx.x;
    const long long o0 = offs[s], m = offs[s + 1] - o0;
    const long long chunk = (m + kAnBlocks - 1) / kAnBlocks;
    const long long j0 = o0 + b * chunk, j1 = o0 + min(m, (b + 1) * chunk);
    double mean[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double t = 0.0;
        for (int k = 0; k < kAnBlocks; ++k) t += part[(s * kAnBlocks + k) * 3 + d];
        mean[d] = t / (double)m;
    }
    const double c[3] = {centre[0], centre[1], centre[2]};
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long j = j0 + threadIdx.x; j < j1; j += kThreads) {
        const long long i = idx[j];
        double r[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) r[d] = an_rel(pos.get(i, d), c[d], L, hL) - mean[d];
        v[0] += r[0] * r[0];
        v[1] += r[0] * r[1];
        v[2] += r[0] * r[2];
        v[3] += r[1] * r[1];
        v[4] += r[1] * r[2];
        v[5] += r[2] * r[2];
    }
    __shared__ double red[6];
    an_block_sum<6>(v, red);
    if (threadIdx.x < 6) part2[(s * kAnBlocks + b) * 6 + threadIdx.x] = red[threadIdx.x];
}

// out[s * 9 + 0..2] = mean, out[s * 9 + 3..8] = outer-product sums of set s (one block per set).
__global__ void k_an_finish(const long long* __restrict__ offs, const double* __restrict__ part,
                            const double* __restrict__ part2, double* __restrict__ out) {
    const int s = blockIdx.x, k = threadIdx.x;
    const long long m = offs[s + 1] - offs[s];
    if (k < 3) {
        double t = 0.0;
        for (int b = 0; b < kAnBlocks; ++b) t += part[(s * kAnBlocks + b) * 3 + k];
        out[s * 9 + k] = t / (double)m;
    } else if (k < 9) {
        double t = 0.0;
        for (int b = 0; b < kAnBlocks; ++b) t += part2[(s * kAnBlocks + b) * 6 + (k - 3)];
        out[s * 9 + k] = t;
    }
}

struct AnRadii {
    double r[kAnMaxRadii];
    int nr;
};

// cnt[k] += #particles with sqrt((rx^2 + ry^2) + rz^2) < radii.r[k] (numpy's sum over axis 1 of
// rel**2 adds the three squares left to right; sqrt is correctly rounded on both sides).
template <class Pos>
__global__ __launch_bounds__(kThreads) void k_an_radii(Pos pos, long long n, const double* __restrict__ centre,
                                                       double L, double hL, AnRadii radii,
                                                       unsigned long long* __restrict__ cnt) {
    __shared__ unsigned int bc[kAnMaxRadii];
    if (threadIdx.x < kAnMaxRadii) bc[threadIdx.x] = 0u;
    __syncthreads();
    const double c[3] = {centre[0], centre[1], centre[2]};
    unsigned int mine[kAnMaxRadii] = {};
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
        double r2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double r = an_rel(pos.get(i, d), c[d], L, hL);
            r2 = d == 0 ? r * r : r2 + r * r;
        }
        const double rc = sqrt(r2);
#pragma unroll
        for (int k = 0; k < kAnMaxRadii; ++k) mine[k] += (k < radii.nr && rc < radii.r[k]) ? 1u : 0u;
    }
#pragma unroll
    for (int k = 0; k < kAnMaxRadii; ++k)
        if (mine[k]) atomicAdd(&bc[k], mine[k]);
    __syncthreads();
    if ((int)threadIdx.x < radii.nr && bc[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], (unsigned long long)bc[threadIdx.x]);
}

// Per-block maxima of a fixed-point mesh (grid-stride), then one block reduces them into *out.
template <bool WIDE>
__global__ __launch_bounds__(kThreads) void k_an_mesh_max(const unsigned long long* __restrict__ mesh, long long M,
                                                          double qscale, double* __restrict__ bmax) {
    double m = 0.0;
    for (long long c = (long long)blockIdx.x * kThreads + threadIdx.x; c < M; c += (long long)gridDim.x * kThreads)
        m = fmax(m, fx_value<WIDE>(mesh, (size_t)c, qscale));
    m = block_max_nonneg(m);
    if (threadIdx.x == 0) bmax[blockIdx.x] = m;
}
__global__ __launch_bounds__(kThreads) void k_an_max_reduce(const double* __restrict__ bmax, int nb,
                                                            double* __restrict__ out) {
    double m = 0.0;
    for (int t = threadIdx.x; t < nb; t += kThreads) m = fmax(m, bmax[t]);
    m = block_max_nonneg(m);
    if (threadIdx.x == 0) *out = m;
}

// ------------------------------------------------------------------------------- global sums
// The global sums that stay in FP64 (or exact) in every precision mode.  Mass: the total of the
// fixed-point mesh of a deposit, summed as a 128-bit integer (exact in any order); each block
// stores its partial total as {lo, hi}.  Momentum: sum of p_i per axis in FP64 over a fixed grid
// and a fixed tree, so the same state always gives the same bits.
constexpr int kTotBlocks = 1024;

template <bool WIDE>
__global__ __launch_bounds__(kThreads) void k_mesh_total(const unsigned long long* __restrict__ mesh, long long M,
                                                         unsigned long long* __restrict__ part) {
    unsigned __int128 s = 0;
    for (long long c = (long long)blockIdx.x * kThreads + threadIdx.x; c < M; c += (long long)gridDim.x * kThreads) {
        if constexpr (WIDE)
            s += ((unsigned __int128)mesh[2 * c + 1] << 64) | (unsigned __int128)mesh[2 * c];
        else
            s += (unsigned __int128)mesh[c];
    }
    __shared__ unsigned long long lo[kThreads], hi[kThreads];
    lo[threadIdx.x] = (unsigned long long)s;
    hi[threadIdx.x] = (unsigned long long)(s >> 64);
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            const unsigned __int128 a = ((unsigned __int128)hi[threadIdx.x] << 64) | lo[threadIdx.x];
            const unsigned __int128 b = ((unsigned __int128)hi[threadIdx.x + st] << 64) | lo[threadIdx.x + st];
            const unsigned __int128 t = a + b;
            lo[threadIdx.x] = (unsigned long long)t;
            hi[threadIdx.x] = (unsigned long long)(t >> 64);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = lo[0];
        part[2 * blockIdx.x + 1] = hi[0];
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_momentum_total(const T* __restrict__ p, long long n,
                                                             double* __restrict__ part) {
    double v[3] = {0.0, 0.0, 0.0};
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
#pragma unroll
        for (int d = 0; d < 3; ++d) v[d] += (double)p[d * n + i];
    }
    __shared__ double red[3];
    an_block_sum<3>(v, red);
    if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x];
}

}  // namespace pm3d
