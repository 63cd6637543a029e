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
// added afterwards with global atomics by k_deposit_far_bal (the lists of all columns dealt out evenly
// over the workgroups) or, for more columns than its prefix holds, k_deposit_far_seg (several
// workgroups per column).
//
// Each corner weight ((wx*wy)*wz, the reference's order) is rounded to the fixed-point grid on its
// own, so every path that deposits a particle adds the same integers; integer accumulation
// commutes, so the result is bitwise identical for any schedule and to the atomic deposit.
#pragma once
#include "common.hpp"
#include "particles.hpp"
#include "particles_mixed.hpp"

namespace pm3d {

// Particle sources.  `load` issues the raw loads of site i (32-bit index off per-axis base pointers,
// so every load is one instruction); `near_axis` is the column deposit's fast CIC of one axis, valid
// only when it returns true (base cell = the site or the site below it); `axes` is the general CIC
// of the far pass; `quant*` round one corner weight (wx*wy)*wz (the reference's order,
// pm3d.py:85-88) to the fixed-point grid of the mesh.
// FB: fraction bits of the mesh, kWideFB (128-bit cells, accumulated as 32-bit halves: quant2) or
// kCompactFB (64-bit cells: the scale folded into wx, quant_scaled).
template <int FB>
struct SrcF64T {
    const double* x0;
    const double* x1;
    const double* x2;
    double dx, inv_dx;
    int N;
    using W = double;
    static constexpr int kColThreads = 1024;
    struct Raw {
        double v[3];
    };
    __device__ __forceinline__ Raw load(unsigned i) const {
        const unsigned b = i * 8u;  // byte offset < 2^32: one load with a scalar base and a 32-bit offset
        return Raw{{ld_off(x0, b), ld_off(x1, b), ld_off(x2, b)}};
    }
    __device__ __forceinline__ Raw load(long long i) const { return load((unsigned)i); }
    // g = X/dx - 0.5, lo = floor(g) mod N, f = g - floor(g) (pm3d.py:76-79).  The quotient is the
    // Markstein-corrected product of div_dx without its range test: for |X| >= 1e290 the result is
    // huge or NaN and the particle is sent to the far pass, and for |X| <= 1e-290 any quotient
    // within a few ulps gives g == -0.5 exactly.  `ob` is the base cell's offset from site s (0 or -1).
    __device__ __forceinline__ bool near_axis(const Raw& r, int d, int s, int& ob, int& lo, double& f) const {
        const bool nr = near_f64(r.v[d], dx, inv_dx, N, s, lo, f);
        ob = lo == s ? 0 : -1;
        return nr;
    }
    // base cell and fraction of one axis (a particle taken from a row list: its row is known)
    __device__ __forceinline__ void cell_frac(const Raw& r, int d, int& lo, double& f) const {
        near_f64(r.v[d], dx, inv_dx, N, 0, lo, f);
    }
    __device__ __forceinline__ void axes(const Raw& r, int lo[3], int hi[3], double w0[3], double w1[3]) const {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double f;
            cic_axis_fast(r.v[d], dx, N, lo[d], hi[d], f);
            w0[d] = 1.0 - f;
            w1[d] = f;
        }
    }
    // weight scale folded into wx before the products (1 for 128-bit cells: q52 scales inside its
    // FMA).  2^FB (ws = (2^FB wx * wy) * wz) is exactly 2^FB ((wx * wy) * wz): a power-of-two scale
    // commutes with rounding, and a product that underflows rounds to 0 either way.
    static constexpr double kScale = FB == kWideFB ? 1.0 : (double)(1ull << FB);
    __device__ __forceinline__ unsigned long long quant(double w) const {
        if constexpr (FB == kWideFB) return q52(w);
        else return q_fixed(w, kScale);
    }
    __device__ __forceinline__ unsigned long long quant_scaled(double ws) const { return q_scaled(ws); }
    // q52 as its two 32-bit halves: 0x4330000000000000 has zero low bits, so no borrow crosses
    __device__ __forceinline__ void quant2(double w, unsigned& lo32, unsigned& hi32) const {
        const unsigned long long b = (unsigned long long)__double_as_longlong(fma(w, 0x1p52, 0x1p52));
        lo32 = (unsigned)b;
        hi32 = (unsigned)(b >> 32) - 0x43300000u;
    }
};
using SrcF64 = SrcF64T<kWideFB>;
using SrcF64C = SrcF64T<kCompactFB>;

struct SrcMx {
    const int* c0;
    const int* c1;
    const int* c2;
    const float* f0;
    const float* f1;
    const float* f2;
    int N;
    using W = float;
    static constexpr int kColThreads = 512;
    struct Raw {
        int c[3];
        float f[3];
    };
    __device__ __forceinline__ Raw load(unsigned i) const {
        const unsigned b = i * 4u;
        return Raw{{ld_off(c0, b), ld_off(c1, b), ld_off(c2, b)}, {ld_off(f0, b), ld_off(f1, b), ld_off(f2, b)}};
    }
    __device__ __forceinline__ Raw load(long long i) const { return load((unsigned)i); }
    // the state keeps the cell in [0, N); a cell outside that range is sent to the far pass, so the
    // column's LDS index stays in bounds whatever the state holds
    __device__ __forceinline__ bool near_axis(const Raw& r, int d, int s, int& ob, int& lo, float& f) const {
        lo = r.c[d];
        f = r.f[d];
        const bool nr = near_cell(lo, N, s);
        ob = lo == s ? 0 : -1;
        lo = (unsigned)lo < (unsigned)N ? lo : 0;
        return nr;
    }
    __device__ __forceinline__ void cell_frac(const Raw& r, int d, int& lo, float& f) const {
        lo = (unsigned)r.c[d] < (unsigned)N ? r.c[d] : 0;
        f = r.f[d];
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
    // 2^31 folded into wx: (2^31 wx * wy) * wz == 2^31 ((wx * wy) * wz) exactly (a power-of-two scale
    // commutes with rounding; an underflowed product rounds to 0 either way), so rint of the scaled
    // product is q31 of the corner weight
    static constexpr float kScale = 2147483648.0f;
    __device__ __forceinline__ unsigned long long quant(float w) const { return q31(w); }
    __device__ __forceinline__ unsigned quant_scaled(float ws) const { return (unsigned)rintf(ws); }
};

// The CIC corners of one particle that land in the column, added to the LDS accumulator: base cell
// at column-local (lx, ly) (-1 .. BX, -1 .. BY: the ring) and z cell lz (in [0, N) for every lane),
// fractions f.  A corner plane outside the column gets weight 0 and an in-column address; an (x, y)
// corner pair whose weight is 0 in every lane of the wave (a ring row's far plane, a zero fraction,
// an idle lane) is skipped, and within a pair only lanes with a non-zero weight issue the atomic.
// Each corner weight (wx*wy)*wz is rounded to the fixed-point grid on its own.  Called by every lane
// of the wave (`valid` false for idle lanes).  WIDE && SINGLE: the 128-bit cell is summed in one
// 64-bit accumulator (q52 whole: no cell of the column can reach 2^64, see deposit_col_body).
template <class Src, int BX, int BY, int NZ, bool WIDE, bool SINGLE>
__device__ __forceinline__ void col_add(const Src& src, unsigned long long* acc, unsigned long long* acch, bool valid,
                                        int lx, int ly, int lz, const typename Src::W f[3]) {
    using W = typename Src::W;
    const bool vx0 = valid && (unsigned)lx < (unsigned)BX, vx1 = valid && (unsigned)(lx + 1) < (unsigned)BX;
    const bool vy0 = (unsigned)ly < (unsigned)BY, vy1 = (unsigned)(ly + 1) < (unsigned)BY;
    const W one = W(1);
    const W wx[2] = {vx0 ? (one - f[0]) * Src::kScale : W(0), vx1 ? f[0] * Src::kScale : W(0)};
    const W wy[2] = {vy0 ? one - f[1] : W(0), vy1 ? f[1] : W(0)};
    const W wz[2] = {one - f[2], f[2]};
    const int px[2] = {vx0 ? lx : 0, vx1 ? lx + 1 : 0};
    const int py[2] = {vy0 ? ly : 0, vy1 ? ly + 1 : 0};
#pragma unroll
    for (int cxy = 0; cxy < 4; ++cxy) {
        const int cx = cxy >> 1, cy = cxy & 1;
        const W wxy = wx[cx] * wy[cy];
        if (__ballot(wxy != W(0)) == 0ull) continue;
        const int a = (px[cx] * BY + py[cy]) * NZ + lz;
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
            const W w = wxy * wz[cz];
            // only lanes with a contribution add: idle lanes (far particles at their sites, corners
            // outside the column) would otherwise pile zero adds onto shared LDS addresses
            if (w == W(0)) continue;
            if constexpr (WIDE && SINGLE) {
                atomicAdd(&acc[a + cz], src.quant(w));
            } else if constexpr (WIDE) {
                unsigned l32, h32;
                src.quant2(w, l32, h32);
                atomicAdd(&acc[a + cz], (unsigned long long)l32);
                atomicAdd(&acch[a + cz], (unsigned long long)h32);
            } else {
                atomicAdd(&acc[a + cz], (unsigned long long)src.quant_scaled(w));
            }
        }
    }
}

// Threads per column workgroup.  Mixed state (4-byte cells and offsets, 8-byte mesh cells): 512, so
// up to four 4x4 columns share a CU (peak256 at its crossing: deposit 213 -> 190 us and the tuner's
// choice 8x8 -> 4x4 columns; early steps 153 -> 141 us).  FP64: 1024 (512 made the 4x4 columns at
// the 128^3 crossing 31.5 -> 33.2 us and the 8x8 ones 43 -> 60 us).  256 was slower for both
// (FP64 4x4 at the crossing 44.5 us, mixed 214 us).
template <class Src>
constexpr int col_threads() { return Src::kColThreads; }

// Column deposit.  LDS holds the column as BX*BY rows of N+1 cells: entry N of a row stands for
// z = 0 (the upper z corner of a particle whose base cell is N-1) and is folded into entry 0 before
// the store, so the upper z corner is always the next entry.  Threads walk the (BX+2) x (BY+2) x N
// sites of the column and its ring; when 64 | N a wave covers 64 sites of one row, so the row, the
// site coordinates and the ownership test are wave-uniform.  Far particles: with row lists
// (fr.cnt != nullptr, built by this step's gather-kick-drift) the column skips them at their sites
// and takes them from the lists of the (BX+1) x (BY+1) base-cell rows whose corners can reach it
// (the column and its low ring); without, the owner of a far particle's site records it in its
// segment of `far` for the global-atomic pass.
template <class Src, int BX, int BY, int N, bool WIDE, int NT>
__device__ __forceinline__ void deposit_col_body(Src src, unsigned long long* __restrict__ mesh,
                                                 int* __restrict__ far, unsigned int* __restrict__ farcnt, FarRows fr,
                                                 XMap xm) {
    using W = typename Src::W;
    constexpr int NZ = N + 1;
    constexpr int NCELL = BX * BY * NZ;
    constexpr int EX = BX + 2, EY = BY + 2, NSITE = EX * EY * N;
    constexpr bool ROWS = N % 64 == 0;
    constexpr int RY = BY + 1, NR = (BX + 1) * RY;  // base-cell rows whose corners reach the column
    static_assert(NR < NT, "one thread per row list");
    // WIDE: a cell c takes near particles whose base cell is c or c-1 on every axis, i.e. sites c-2 ..
    // c+1 (base = site + {-1, 0, 1}): at most 64, and one contribution per entry of the four row lists
    // at and below it; each contribution is at most 2^52 units.  Rows hold at most 6N entries (the
    // acceleration buffer they alias), so for 64 + 24N < 4096, N <= 167, no cell reaches 2^64 and the
    // cell's 64-bit sum is exact in any order: SINGLE, one accumulator per cell and one LDS atomic
    // per corner.  Larger meshes accumulate the 128-bit cell as two 64-bit sums of the low and high
    // 32-bit halves of each contribution (no return value, no carry test per add), recombined
    // exactly at the end.
    constexpr bool SINGLE = WIDE && 64 + 24 * N < 4096;
    __shared__ unsigned long long acc[NCELL];
    __shared__ unsigned long long acch[WIDE && !SINGLE ? NCELL : 1];
    __shared__ unsigned int nfar;
    __shared__ unsigned int rpre[NR + 1];
    if (sched_ended(fr.sch)) return;  // a replay past the end of the table: the mesh stays as it is
    const bool lists = fr.cnt != nullptr;
    const int col = (int)xcd_map(blockIdx.x, gridDim.x, xm);  // column id, also this column's far segment
    constexpr int NBY = N / BY;
    const int x0 = (col / NBY) * BX, y0 = (col % NBY) * BY;
    unsigned int rcount = 0;
    if (lists && (int)threadIdx.x < NR) {  // issued first: the loads complete under the site loop
        const int ex = threadIdx.x / RY, ey = threadIdx.x % RY;
        const unsigned c = fr.cnt[wrapN(x0 + ex - 1, N) * N + wrapN(y0 + ey - 1, N)];
        rcount = c < (unsigned)fr.cap ? c : (unsigned)fr.cap;
    }
    for (int t = threadIdx.x; t < NCELL; t += NT) {
        acc[t] = 0ull;
        if (WIDE && !SINGLE) acch[t] = 0ull;
    }
    if (threadIdx.x == 0) nfar = 0u;
    __syncthreads();
    int* myfar = far + (size_t)col * (BX * BY * N);
    const int lane = threadIdx.x & 63;
    const int wbase = __builtin_amdgcn_readfirstlane((int)(threadIdx.x & ~63u));
    constexpr int U = 4;  // sites whose loads are issued together
    for (int t0 = 0; t0 < NSITE; t0 += NT * U) {
        typename Src::Raw raw[U];
        int exs[U], eys[U], szs[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int tw = t0 + u * NT + wbase;  // wave-uniform
            const int t = tw + lane;
            int row, sz;
            if (ROWS) {
                row = tw / N;
                sz = tw % N + lane;
            } else {
                row = t / N;
                sz = t % N;
            }
            exs[u] = row / EY;
            eys[u] = row % EY;
            szs[u] = sz;
            if ((ROWS ? tw : t) < NSITE) {
                const int sx = wrapN(x0 + exs[u] - 1, N), sy = wrapN(y0 + eys[u] - 1, N);
                raw[u] = src.load((unsigned)((sx * N + sy) * N + sz));
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int tw = t0 + u * NT + wbase;
            if ((ROWS ? tw : tw + lane) >= NSITE) break;
            const int ex = exs[u], ey = eys[u], sz = szs[u];
            const int sx = wrapN(x0 + ex - 1, N), sy = wrapN(y0 + ey - 1, N);
            int ob[3], lo[3];
            W f[3];
            const bool nx = src.near_axis(raw[u], 0, sx, ob[0], lo[0], f[0]);
            const bool ny = src.near_axis(raw[u], 1, sy, ob[1], lo[1], f[1]);
            const bool nz = src.near_axis(raw[u], 2, sz, ob[2], lo[2], f[2]);
            const bool is_near = nx && ny && nz;
            const bool own = ex >= 1 && ex <= BX && ey >= 1 && ey <= BY;
            const bool app = !is_near && own && !lists;
            const unsigned long long bal = __ballot(app);
            if (bal) {  // one LDS atomic per wave, slots by prefix popcount
                const int leader = __ffsll((long long)bal) - 1;
                unsigned int base = 0;
                if (lane == leader) base = atomicAdd(&nfar, (unsigned int)__popcll(bal));
                base = __shfl(base, leader, 64);
                if (app) myfar[base + __popcll(bal & ((1ull << lane) - 1ull))] = (int)((sx * N + sy) * N + sz);
            }
            // column-local planes of the base cell: ex-1+ob in [-2, BX]
            col_add<Src, BX, BY, NZ, WIDE, SINGLE>(src, acc, acch, is_near, ex - 1 + ob[0], ey - 1 + ob[1], lo[2], f);
        }
    }
    if (lists) {
        // exclusive prefix of the row counts (loaded above), then the entries of all rows dealt
        // out over the threads; an entry's row by binary search of the prefix
        if ((int)threadIdx.x < NR) rpre[threadIdx.x] = rcount;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned run = 0;
            for (int r = 0; r < NR; ++r) {
                const unsigned c = rpre[r];
                rpre[r] = run;
                run += c;
            }
            rpre[NR] = run;
        }
        __syncthreads();
        const unsigned total = rpre[NR];
        constexpr int UL = 4;  // entries per thread whose loads are in flight together
        for (unsigned t0 = 0; t0 < total; t0 += NT * UL) {
            int ra[UL], idx[UL];
            bool valid[UL];
#pragma unroll
            for (int u = 0; u < UL; ++u) {
                const unsigned t = t0 + u * NT + threadIdx.x;
                valid[u] = t < total;
                int a = 0;
                if (valid[u]) {
                    int b = NR - 1;
                    while (a < b) {
                        const int m = (a + b + 1) >> 1;
                        if (rpre[m] <= t) a = m;
                        else b = m - 1;
                    }
                }
                ra[u] = a;
                const int row = wrapN(x0 + a / RY - 1, N) * N + wrapN(y0 + a % RY - 1, N);
                idx[u] = valid[u] ? fr.list[(size_t)row * fr.cap + (t - rpre[a])] : 0;
            }
            typename Src::Raw raw[UL];
#pragma unroll
            for (int u = 0; u < UL; ++u)
                if (valid[u]) raw[u] = src.load((unsigned)idx[u]);
#pragma unroll
            for (int u = 0; u < UL; ++u) {
                if (t0 + u * NT >= total) break;  // wave-uniform: no lane has an entry
                int lz = 0;
                W f[3] = {W(0), W(0), W(0)};
                if (valid[u]) {
                    int lxy;
                    src.cell_frac(raw[u], 0, lxy, f[0]);
                    src.cell_frac(raw[u], 1, lxy, f[1]);
                    src.cell_frac(raw[u], 2, lz, f[2]);
                }
                col_add<Src, BX, BY, NZ, WIDE, SINGLE>(src, acc, acch, valid[u], ra[u] / RY - 1, ra[u] % RY - 1, lz, f);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) farcnt[col] = nfar;
    if (threadIdx.x < BX * BY) {  // fold z = N into z = 0
        const int r = threadIdx.x * NZ;
        acc[r] += acc[r + N];
        if (WIDE && !SINGLE) acch[r] += acch[r + N];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < BX * BY * N; t += NT) {
        const int rz = t % N, rxy = t / N;
        const int ry = rxy % BY, rx = rxy / BY;
        const unsigned cell = (unsigned)(((x0 + rx) * N + (y0 + ry)) * N + rz);
        const int l = rxy * NZ + rz;
        if (WIDE && SINGLE) {
            *(ulonglong2*)(mesh + 2 * (size_t)cell) = make_ulonglong2(acc[l], 0ull);
        } else if (WIDE) {
            const unsigned long long L = acc[l], H = acch[l];
            const unsigned long long lo = L + (H << 32);
            *(ulonglong2*)(mesh + 2 * (size_t)cell) = make_ulonglong2(lo, (H >> 32) + (lo < L ? 1ull : 0ull));
        } else {
            mesh[cell] = acc[l];
        }
    }
}

template <class Src, int BX, int BY, int N, bool WIDE, int NT>
__global__ __launch_bounds__(NT) void k_deposit_col(Src src, unsigned long long* __restrict__ mesh,
                                                    int* __restrict__ far, unsigned int* __restrict__ farcnt,
                                                    FarRows fr, XMap xm) {
    deposit_col_body<Src, BX, BY, N, WIDE, NT>(src, mesh, far, farcnt, fr, xm);
}

// One far particle's eight corners, added with global atomics.
template <class Src, bool WIDE>
__device__ __forceinline__ void far_add(const Src& src, unsigned long long* __restrict__ mesh, long long i, int N) {
    int lo[3], hi[3];
    typename Src::W w0[3], w1[3];
    src.axes(src.load(i), lo, hi, w0, w1);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const unsigned long long v =
            src.quant(((c >> 2 ? w1[0] : w0[0]) * ((c >> 1) & 1 ? w1[1] : w0[1])) * (c & 1 ? w1[2] : w0[2]));
        if (v == 0ull) continue;
        const int ix = (c >> 2) ? hi[0] : lo[0], iy = ((c >> 1) & 1) ? hi[1] : lo[1], iz = (c & 1) ? hi[2] : lo[2];
        fx_atomic_add<WIDE>(mesh, ((size_t)ix * N + iy) * N + iz, v);
    }
}

// Far particles, added with global atomics: workgroup (b, y) takes every gridDim.y-th block of
// kThreads entries of column b's list (a list holds at most the column's own sites).
constexpr int kFarSub = 4;
template <class Src, bool WIDE>
__device__ __forceinline__ void deposit_far_seg_body(Src src, unsigned long long* __restrict__ mesh,
                                                     const int* __restrict__ far,
                                                     const unsigned int* __restrict__ farcnt, int seg, int N,
                                                     const DevSched* __restrict__ sch) {
    if (sched_ended(sch)) return;  // the columns left the mesh as it was: nothing to add
    const unsigned int cnt = farcnt[blockIdx.x];
    const int* list = far + (size_t)blockIdx.x * seg;
    for (unsigned int t = blockIdx.y * kThreads + threadIdx.x; t < cnt; t += gridDim.y * kThreads)
        far_add<Src, WIDE>(src, mesh, list[t], N);
}

template <class Src, bool WIDE>
__global__ __launch_bounds__(kThreads) void k_deposit_far_seg(Src src, unsigned long long* __restrict__ mesh,
                                                              const int* __restrict__ far,
                                                              const unsigned int* __restrict__ farcnt, int seg, int N,
                                                              const DevSched* __restrict__ sch) {
    deposit_far_seg_body<Src, WIDE>(src, mesh, far, farcnt, seg, N, sch);
}

// Far particles, load-balanced: the far particles of a collapsing peak crowd into the few columns
// around it, so splitting per column leaves most workgroups idle while a handful walk long lists.
// Here every workgroup scans the ncol column counts into an exclusive prefix in LDS, and the
// concatenation of all lists is dealt out in blocks of kThreads entries, block k to workgroup
// k mod gridDim.x; an entry's column is found by binary search of the prefix.
constexpr int kFarColsMax = 8192;  // columns the prefix holds (32 KB of LDS)
constexpr int kFarBalGrid = 1024;  // workgroups of the balanced pass (at most)
//
// With row lists (fr.cnt != nullptr) the pass also deposits the lists' overflow (fr.ovl, when
// `lists` says this deposit consumed them) and clears the row counts for the next step; the overflow
// count is cleared by the forward transform that follows (every workgroup here reads it).
template <class Src, bool WIDE>
__device__ __forceinline__ void deposit_far_bal_body(Src src, unsigned long long* __restrict__ mesh,
                                                     const int* __restrict__ far,
                                                     const unsigned int* __restrict__ farcnt, int ncol, int seg,
                                                     int N, unsigned blk, unsigned nblk, FarRows fr, int lists) {
    if (sched_ended(fr.sch)) return;
    if (lists && fr.cnt) {
        // the columns took the far particles from the row lists (and recorded none themselves): only
        // the lists' overflow and the clearing of the row counts are left
        const unsigned nov = min(*fr.ovcnt, fr.ovcap);
        for (unsigned t = blk * kThreads + threadIdx.x; t < nov; t += nblk * kThreads) far_add<Src, WIDE>(src, mesh, fr.ovl[t], N);
        for (unsigned r = blk * kThreads + threadIdx.x; r < (unsigned)N * (unsigned)N; r += nblk * kThreads) fr.cnt[r] = 0u;
        return;
    }
    __shared__ unsigned int pre[kFarColsMax + 1];
    __shared__ unsigned int wsum[kThreads / 64];
    const int R = (ncol + kThreads - 1) / kThreads;  // columns per thread, contiguous
    const int c0 = (int)threadIdx.x * R;
    unsigned int sum = 0;
    for (int k = 0; k < R; ++k)
        if (c0 + k < ncol) sum += farcnt[c0 + k];
    // exclusive block scan of the per-thread sums
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned int inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    unsigned int run = inc - sum;
    for (int w = 0; w < wv; ++w) run += wsum[w];
    for (int k = 0; k < R; ++k)
        if (c0 + k < ncol) {
            pre[c0 + k] = run;
            run += farcnt[c0 + k];
        }
    if (threadIdx.x == kThreads - 1) pre[ncol] = run;  // the last thread's run ends at the total
    __syncthreads();
    const unsigned int total = pre[ncol];
    for (unsigned int t = blk * kThreads + threadIdx.x; t < total; t += nblk * kThreads) {
        int a = 0, b = ncol - 1;  // last column with pre[c] <= t (columns with no entries are skipped)
        while (a < b) {
            const int m = (a + b + 1) >> 1;
            if (pre[m] <= t) a = m;
            else b = m - 1;
        }
        far_add<Src, WIDE>(src, mesh, far[(size_t)a * seg + (t - pre[a])], N);
    }
    if (fr.cnt)  // row lists that this deposit did not take: cleared all the same
        for (unsigned r = blk * kThreads + threadIdx.x; r < (unsigned)N * (unsigned)N; r += nblk * kThreads) fr.cnt[r] = 0u;
}

template <class Src, bool WIDE>
__global__ __launch_bounds__(kThreads) void k_deposit_far_bal(Src src, unsigned long long* __restrict__ mesh,
                                                              const int* __restrict__ far,
                                                              const unsigned int* __restrict__ farcnt, int ncol,
                                                              int seg, int N, FarRows fr, int lists) {
    deposit_far_bal_body<Src, WIDE>(src, mesh, far, farcnt, ncol, seg, N, blockIdx.x, gridDim.x, fr, lists);
}

// Column shape per mesh size, cell width and LDS budget.  The candidates are the 8x8, 4x8, 4x4 and
// 2x4 columns, widest first; those whose LDS accumulator (rows of N+1 cells) fits in BUDGET bytes
// (gfx950 has 160 KB of LDS per CU) are variants 0, 1, ... (NV of them; a larger V selects the last).
// Variant 0 has the least ring traffic per owned site ((BX+2)(BY+2)/(BX*BY) sites read per site);
// the narrower ones launch more workgroups and fit more of them per CU.  All variants give the same
// mesh bits (integer sums), so the choice is made by timing.
template <int N, bool WIDE, int V = 0, int BUDGET = 135168> struct ColShape {
    static constexpr int CELLB = WIDE ? 16 : 8;
    static constexpr bool fits(int bx, int by) { return bx * by * (N + 1) * CELLB <= BUDGET; }
    static constexpr int kBX[4] = {8, 4, 4, 2}, kBY[4] = {8, 8, 4, 4};
    static constexpr int first() {
        for (int k = 0; k < 3; ++k)
            if (fits(kBX[k], kBY[k])) return k;
        return 3;
    }
    static constexpr int NV = 4 - first();
    static constexpr int K = first() + (V < NV ? V : NV - 1);
    static constexpr int BX = kBX[K], BY = kBY[K];
    static constexpr int NBLK = (N / BX) * (N / BY);
    static_assert(BX * BY * (N + 1) * CELLB <= 163840, "column does not fit in LDS");
    static_assert(N % BX == 0 && N % BY == 0, "columns must tile the mesh");
};

// Timing state for choosing the column shape: the particle of lattice site (s0, s1, s2) at cell
// coordinate g_d = s_d + j_d (g = x/dx - 0.5), j_d in [-0.45, 0.45) from a hash of (i, d), so every
// particle is near its own column and the columns carry the lattice's full load.  FP64 positions
// x = (g + 0.5) dx stay inside (0, L); the mixed state is the cell floor(g) mod N and offset g - floor(g).
__device__ __forceinline__ double tune_jitter(unsigned long long i, int d) {
    unsigned long long z = i * 3ull + (unsigned long long)d + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return ((double)(z >> 40) * 0x1p-24 - 0.5) * 0.9;
}
__global__ __launch_bounds__(kThreads) void k_tune_lattice(int N, double dx, double* __restrict__ x,
                                                           int* __restrict__ cell, float* __restrict__ off) {
    const long long n = (long long)N * N * N;
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int s[3] = {(int)(i / ((long long)N * N)), (int)((i / N) % N), (int)(i % N)};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double g = s[d] + tune_jitter((unsigned long long)i, d);
        if (x) {
            x[d * n + i] = (g + 0.5) * dx;
        } else {
            const double fl = floor(g);
            cell[d * n + i] = fl < 0.0 ? N - 1 : (int)fl;
            off[d * n + i] = (float)(g - fl);
        }
    }
}

}  // namespace pm3d
