// Owner-computes CIC deposit in column bricks (gfx950), shared by the FP64 and mixed modes.
//
// The mesh is cut into columns of BX x BY x N cells (the whole z axis).  The workgroup owning a
// column reads every particle whose Lagrangian site lies in the column or in the one-site ring
// around it in x and y ((BX+2) x (BY+2) rows of N sites: each row is N contiguous elements, so the
// reads are fully coalesced), accumulates the CIC corners that land inside its column in LDS with
// integer atomics, and writes the column once with plain stores (no global atomics, no clearing
// pass).  This relies on Np == N and on each particle's base cell being its site or the site just
// below it on every axis ("near").  A particle outside that range ("far") is recorded by the
// workgroup that owns its site, in that workgroup's own segment of `far` (no global counter), and is
// added afterwards by k_deposit_far_seg with global atomics (several workgroups per segment).
//
// Each corner weight ((wx*wy)*wz, the reference's order) is rounded to the fixed-point grid on its
// own, so every path that deposits a particle adds the same integers; integer accumulation
// commutes, so the result is bitwise identical for any schedule and to the atomic deposit.
#pragma once
#include "common.hpp"
#include "particles.hpp"
#include "particles_mixed.hpp"

namespace pm3d {

// Particle sources: raw loads of particle i (issued in batches, so that several sites' loads are in
// flight per thread), the CIC base/upper cells with the per-axis weights, and the fixed-point value
// of one corner weight (wx*wy)*wz, evaluated in the reference's order (pm3d.py:85-88).
struct SrcF64 {
    const double* x;
    long long n;
    double dx, qs;
    long long one;
    int N;
    using W = double;
    struct Raw {
        double v[3];
    };
    __device__ __forceinline__ Raw load(long long i) const { return Raw{{x[i], x[n + i], x[2 * n + i]}}; }
    __device__ __forceinline__ void axes(const Raw& r, int lo[3], int hi[3], double w0[3], double w1[3]) const {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double f;
            cic_axis_fast(r.v[d], dx, N, lo[d], hi[d], f);
            w0[d] = 1.0 - f;
            w1[d] = f;
        }
    }
    __device__ __forceinline__ unsigned long long quant(double w) const { return q52(w); }
};

struct SrcMx {
    const int* cell;
    const float* off;
    long long n;
    int N;
    using W = float;
    struct Raw {
        int c[3];
        float f[3];
    };
    __device__ __forceinline__ Raw load(long long i) const {
        return Raw{{cell[i], cell[n + i], cell[2 * n + i]}, {off[i], off[n + i], off[2 * n + i]}};
    }
    __device__ __forceinline__ void axes(const Raw& r, int lo[3], int hi[3], float w0[3], float w1[3]) const {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = r.c[d];
            hi[d] = lo[d] + 1 == N ? 0 : lo[d] + 1;
            w0[d] = 1.0f - r.f[d];
            w1[d] = r.f[d];
        }
    }
    __device__ __forceinline__ unsigned long long quant(float w) const { return q31(w); }
};

// Displacement of the base cell from the site along one axis (minimum image).
__device__ __forceinline__ int site_offset(int lo, int site, int N) {
    int dd = lo - site;
    if (dd > N / 2) dd -= N;
    if (dd < -N / 2) dd += N;
    return dd;
}

template <class Src, int BX, int BY, int N, bool WIDE, int NT>
__global__ __launch_bounds__(NT) void k_deposit_col(Src src, unsigned long long* __restrict__ mesh,
                                                    int* __restrict__ far, unsigned int* __restrict__ farcnt, XMap xm) {
    constexpr int NCELL = BX * BY * N;
    constexpr int EX = BX + 2, EY = BY + 2, NSITE = EX * EY * N;
    // WIDE: the 128-bit cell is accumulated as two 64-bit sums of the low and high 32-bit halves of
    // each contribution (no return value, no carry test per add) and recombined exactly at the end
    __shared__ unsigned long long acc[NCELL];
    __shared__ unsigned long long acch[WIDE ? NCELL : 1];
    __shared__ unsigned int nfar;
    for (int t = threadIdx.x; t < NCELL; t += NT) {
        acc[t] = 0ull;
        if (WIDE) acch[t] = 0ull;
    }
    if (threadIdx.x == 0) nfar = 0u;
    __syncthreads();
    constexpr int NBY = N / BY;
    const int col = (int)xcd_map(blockIdx.x, gridDim.x, xm);  // column id, also this column's far segment
    const int x0 = (col / NBY) * BX, y0 = (col % NBY) * BY;
    int* myfar = far + (size_t)col * (BX * BY * N);
    constexpr int U = 4;  // sites whose loads are issued together
    for (int t0 = threadIdx.x; t0 < NSITE; t0 += NT * U) {
        typename Src::Raw raw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * NT;
            if (t < NSITE) {
                const int sz = t % N, ey = (t / N) % EY, ex = t / (N * EY);
                const int sx = wrapN(x0 + ex - 1, N), sy = wrapN(y0 + ey - 1, N);
                raw[u] = src.load(((long long)sx * N + sy) * N + sz);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
        const int t = t0 + u * NT;
        if (t >= NSITE) break;
        const int sz = t % N, ey = (t / N) % EY, ex = t / (N * EY);
        const int sx = wrapN(x0 + ex - 1, N), sy = wrapN(y0 + ey - 1, N);
        const long long i = ((long long)sx * N + sy) * N + sz;
        int lo[3], hi[3];
        typename Src::W w0[3], w1[3];
        src.axes(raw[u], lo, hi, w0, w1);
        const int ox = site_offset(lo[0], sx, N), oy = site_offset(lo[1], sy, N), oz = site_offset(lo[2], sz, N);
        const bool is_far = ox < -1 || ox > 0 || oy < -1 || oy > 0 || oz < -1 || oz > 0;
        const bool own = ex >= 1 && ex <= BX && ey >= 1 && ey <= BY;
        const bool app = is_far && own;
        const unsigned long long bal = __ballot(app);
        if (bal) {  // one LDS atomic per wave, slots by prefix popcount
            const int lane = threadIdx.x & 63;
            const int leader = __ffsll((long long)bal) - 1;
            unsigned int base = 0;
            if (lane == leader) base = atomicAdd(&nfar, (unsigned int)__popcll(bal));
            base = __shfl(base, leader, 64);
            if (app) myfar[base + __popcll(bal & ((1ull << lane) - 1ull))] = (int)i;
        }
        if (is_far) continue;
        // column-local x/y of the two stencil planes; a plane outside the column maps to >= BX/BY
        const int rx[2] = {wrapN(lo[0] - x0, N), wrapN(hi[0] - x0, N)};
        const int ry[2] = {wrapN(lo[1] - y0, N), wrapN(hi[1] - y0, N)};
#pragma unroll
        for (int cxy = 0; cxy < 4; ++cxy) {
            const int cx = cxy >> 1, cy = cxy & 1;
            if (rx[cx] < BX && ry[cy] < BY) {  // only the corners that land in this column
                const int col = (rx[cx] * BY + ry[cy]) * N;
                const typename Src::W wxy = (cx ? w1[0] : w0[0]) * (cy ? w1[1] : w0[1]);
#pragma unroll
                for (int cz = 0; cz < 2; ++cz) {
                    const unsigned long long v = src.quant(wxy * (cz ? w1[2] : w0[2]));
                    if (v != 0ull) {
                        const int cell = col + (cz ? hi[2] : lo[2]);
                        if (WIDE) {
                            atomicAdd(&acc[cell], v & 0xffffffffull);
                            atomicAdd(&acch[cell], v >> 32);
                        } else {
                            atomicAdd(&acc[cell], v);
                        }
                    }
                }
            }
        }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) farcnt[col] = nfar;
    for (int t = threadIdx.x; t < NCELL; t += NT) {
        const int rz = t % N, ry = (t / N) % BY, rx = t / (N * BY);
        const size_t cell = ((size_t)(x0 + rx) * N + (y0 + ry)) * N + rz;
        if (WIDE) {
            const unsigned long long L = acc[t], H = acch[t];
            const unsigned long long lo = L + (H << 32);
            *(ulonglong2*)(mesh + 2 * cell) = make_ulonglong2(lo, (H >> 32) + (lo < L ? 1ull : 0ull));
        } else {
            mesh[cell] = acc[t];
        }
    }
}

// Far particles, added with global atomics: workgroup (b, y) takes every gridDim.y-th block of
// kThreads entries of column b's list (a list holds at most the column's own sites).
constexpr int kFarSub = 4;
template <class Src, bool WIDE>
__global__ __launch_bounds__(kThreads) void k_deposit_far_seg(Src src, unsigned long long* __restrict__ mesh,
                                                              const int* __restrict__ far,
                                                              const unsigned int* __restrict__ farcnt, int seg, int N) {
    const unsigned int cnt = farcnt[blockIdx.x];
    const int* list = far + (size_t)blockIdx.x * seg;
    for (unsigned int t = blockIdx.y * kThreads + threadIdx.x; t < cnt; t += gridDim.y * kThreads) {
        const long long i = list[t];
        int lo[3], hi[3];
        typename Src::W w0[3], w1[3];
        src.axes(src.load(i), lo, hi, w0, w1);
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const unsigned long long v = src.quant(((c >> 2 ? w1[0] : w0[0]) * ((c >> 1) & 1 ? w1[1] : w0[1])) *
                                                   (c & 1 ? w1[2] : w0[2]));
            if (v == 0ull) continue;
            const int ix = (c >> 2) ? hi[0] : lo[0], iy = ((c >> 1) & 1) ? hi[1] : lo[1], iz = (c & 1) ? hi[2] : lo[2];
            fx_atomic_add<WIDE>(mesh, ((size_t)ix * N + iy) * N + iz, v);
        }
    }
}

// Column shape per mesh size, cell width and LDS budget: the widest of 8x8, 4x8, 4x4, 2x4
// columns whose LDS accumulator fits in BUDGET bytes (gfx950 has 160 KB of LDS per CU).
template <int N, bool WIDE, int BUDGET = 131072> struct ColShape {
    static constexpr int CELLB = WIDE ? 16 : 8;
    static constexpr bool fits(int bx, int by) { return bx * by * N * CELLB <= BUDGET; }
    static constexpr int BX = fits(8, 8) ? 8 : fits(4, 8) ? 4 : fits(4, 4) ? 4 : 2;
    static constexpr int BY = fits(8, 8) ? 8 : fits(4, 8) ? 8 : fits(4, 4) ? 4 : 4;
    static constexpr int NBLK = (N / BX) * (N / BY);
    static_assert(BX * BY * N * CELLB <= 163840, "column does not fit in LDS");
};

}  // namespace pm3d
