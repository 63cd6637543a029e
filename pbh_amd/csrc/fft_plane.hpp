// Plane-fused z/y passes of the Poisson FFT (gfx950).
//
// The 3D transform of fft.hpp runs five HBM passes per solve (z r2c, y, x + Poisson + 3 inverse x,
// 3 inverse y, 3 z c2r).  When half an x plane of the half spectrum fits in LDS, the z and y
// passes of each direction run in one kernel per x plane, so a solve is three passes:
//
//   k_fwd_plane : real rows of plane x -> spec[x][ky][kz] (z r2c, then y), one read, one write
//   k_x_pass    : unchanged (fft.hpp)
//   k_inv_plane : gradient spectra of plane x, one component -> real rows (y inverse, then z c2r)
//
// Each plane is split between two workgroups by the parity h of the output index along y, with
// the first radix-2 stage of the y transform (decimation in frequency) done while loading:
//   forward: ky = 2m + h needs only  b_h[y'] = (row[y'] + (-1)^h row[y'+N/2]) * exp(-2 pi i h y'/N)
//            (the combination is linear, so it is applied to the real rows before the z transform
//            and the twiddle after it), followed by an N/2-point y transform;
//   inverse: y = 2m + h needs only   B_h[k'] = (A[k'] + (-1)^h A[k'+N/2]) * exp(+2 pi i h k'/N)
//            followed by an N/2-point inverse y transform and the z c2r of those rows.
// The workgroup's tile is (N/2+1) x (N/2+1) complex: kz = 0..N/2 (Nyquist included) by N/2 rows.
// Tile element (a, b) is at a*S + b with S = N/2+1; the z transforms treat rows as pencils
// (element kz at kz*S + row), the y transforms treat kz columns as pencils (element row at
// kz*S + row).  Consecutive threads take consecutive pencils in both, so a wave's LDS accesses are
// a contiguous run (z) or a run of stride S (y); S is odd, so neither conflicts on the banks.  (With
// consecutive butterflies of one pencil per thread instead, the first radix-R pass of the y
// transform writes at a stride of R elements, 128 bytes, one bank for the whole wave.)
#pragma once
#include "common.hpp"
#include "fft.hpp"

namespace pm3d {

template <typename R, int N>
struct PlaneTile {
    static constexpr int H = N / 2, S = H + 1;
    static constexpr int BYTES = S * S * (int)sizeof(cplx<R>) + N * (int)sizeof(cplx<R>);  // tile + twiddles
    // gfx950: 160 KB of LDS per CU; the block reductions need a few words more
    static constexpr bool FITS = N >= 8 && BYTES <= 160 * 1024 - 512;
    // threads per workgroup (two workgroups per CU below 80 KB of LDS)
    static constexpr int NT = N >= 256 ? 1024 : N >= 128 ? 512 : 256;
    static constexpr int ML2 = sizeof(R) == 8 ? 3 : 4;  // largest radix: 8 (FP64) / 16 (FP32)
};

// One in-place Stockham pass of radix RADIX over NP pencils of NC points in an LDS tile:
// element j of pencil c at buf[j*ES + c*PS].  PFAST: consecutive threads take consecutive pencils
// (else consecutive butterflies of one pencil).  tw[k] = exp(-2 pi i k/(NC*TWS)).
template <typename R, int NC, int NP, int ES, int PS, bool PFAST, bool INV, int TWS, int RADIX, int NT>
__device__ __forceinline__ void tile_pass(cplx<R>* buf, int Ns, const cplx<R>* __restrict__ tw) {
    constexpr int M = NC / RADIX, T = NP * M, IT = (T + NT - 1) / NT;
    const int s = TWS * (NC / (RADIX * Ns));
    cplx<R> v[IT][RADIX];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int t = threadIdx.x + it * NT;
        if (T % NT == 0 || t < T) {
            const int c = PFAST ? t % NP : t / M, j = PFAST ? t / NP : t % M, k = j % Ns;
#pragma unroll
            for (int r = 0; r < RADIX; ++r) v[it][r] = buf[(j + r * M) * ES + c * PS];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < RADIX; ++r) {
                    cplx<R> w = tw[r * k * s];
                    if (INV) w = cconj(w);
                    v[it][r] = cmul(v[it][r], w);
                }
            }
            dft_reg<R, RADIX, INV>(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int t = threadIdx.x + it * NT;
        if (T % NT == 0 || t < T) {
            const int c = PFAST ? t % NP : t / M, j = PFAST ? t / NP : t % M, k = j % Ns;
            const int dd = (j / Ns) * Ns * RADIX + k;
#pragma unroll
            for (int r = 0; r < RADIX; ++r) buf[(dd + r * Ns) * ES + c * PS] = v[it][r];
        }
    }
    __syncthreads();
}

// In-place FFT of NP pencils of NC = 2^a 3^b points in an LDS tile (power-of-two passes of radix
// 2^ML2, then the remaining power of two, then radix 3).  Ends with a barrier.
template <typename R, int NC, int NP, int ES, int PS, bool PFAST, bool INV, int TWS, int ML2, int NT, int NS = 1,
          int REM2 = count_factor(NC, 2), int REM3 = count_factor(NC, 3)>
__device__ __forceinline__ void tile_fft(cplx<R>* buf, const cplx<R>* __restrict__ tw) {
    if constexpr (REM2 > 0 || REM3 > 0) {
        constexpr int L2 = REM2 >= ML2 ? ML2 : REM2;
        constexpr int RADIX = REM2 > 0 ? (1 << L2) : 3;
        constexpr int R2 = REM2 > 0 ? REM2 - L2 : 0, R3 = REM2 > 0 ? REM3 : REM3 - 1;
        tile_pass<R, NC, NP, ES, PS, PFAST, INV, TWS, RADIX, NT>(buf, NS, tw);
        tile_fft<R, NC, NP, ES, PS, PFAST, INV, TWS, ML2, NT, NS * RADIX, R2, R3>(buf, tw);
    }
}

// Block-wide max of a non-negative double over NT threads (all threads must call).
template <int NT>
__device__ __forceinline__ double block_max_nonneg_nt(double v) {
    __shared__ unsigned long long wmax[NT / 64];
    const unsigned long long bits = wave_max_u64((unsigned long long)__double_as_longlong(v));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = bits;
    __syncthreads();
    unsigned long long m = wmax[0];
    for (int w = 1; w < NT / 64; ++w) m = wmax[w] > m ? wmax[w] : m;
    __syncthreads();
    return __longlong_as_double((long long)m);
}

// Half-spectrum bin k of a real row from the N/2-point transform Z of its even/odd pairs:
// X[k] = E + exp(-2 pi i k/N) O with E = (Z_k + conj Z_{N/2-k})/2, O = (Z_k - conj Z_{N/2-k})/(2i)
// (the same expressions as k_r2c_z).
template <typename R>
__device__ __forceinline__ cplx<R> r2c_bin(cplx<R> zk, cplx<R> zm, cplx<R> w) {
    const R er = (zk.re + zm.re) * (R)0.5, ei = (zk.im - zm.im) * (R)0.5;
    const R orr = (zk.im + zm.im) * (R)0.5, oi = (zm.re - zk.re) * (R)0.5;
    return cplx<R>{er + (w.re * orr - w.im * oi), ei + (w.re * oi + w.im * orr)};
}

// Input point k of the N/2-point inverse transform of a c2r row from bins X_k and X_{N/2-k}
// (the same expressions as k_c2r_z; w = exp(-2 pi i k/N), the odd part takes its conjugate).
template <typename R>
__device__ __forceinline__ cplx<R> c2r_point(cplx<R> xk, cplx<R> xm, cplx<R> w) {
    const R er = xk.re + xm.re, ei = xk.im - xm.im;
    const R dr = xk.re - xm.re, di = xk.im + xm.im;
    const R orr = dr * w.re + di * w.im, oi = di * w.re - dr * w.im;
    return cplx<R>{er - oi, ei + orr};
}

// Raw loads of two consecutive cells of a real row: fixed-point cells (WIDE: {lo, hi} words) or
// reals.  B = pairs whose loads a thread keeps in flight (about 32 VGPRs of loaded data).
template <typename R, bool FIXED, bool WIDE>
struct PlaneLoadBatch;
template <typename R, bool WIDE>
struct PlaneLoadBatch<R, true, WIDE> {
    struct Pair {
        unsigned long long w[WIDE ? 4 : 2];
    };
    static constexpr int B = WIDE ? 4 : 8;
    __device__ __forceinline__ static Pair load(const void* in, size_t g) {
        const ulonglong2* p = (const ulonglong2*)((const unsigned long long*)in + (WIDE ? 2 * g : g));
        Pair r;
        const ulonglong2 a = p[0];
        r.w[0] = a.x;
        r.w[1] = a.y;
        if constexpr (WIDE) {
            const ulonglong2 b = p[1];
            r.w[2] = b.x;
            r.w[3] = b.y;
        }
        return r;
    }
    // the same values as fx_value<WIDE>
    __device__ __forceinline__ static void values(const Pair& r, double qscale, double& v0, double& v1) {
        if constexpr (WIDE) {
            v0 = (double)r.w[1] * 4096.0 + (double)r.w[0] * 0x1p-52;
            v1 = (double)r.w[3] * 4096.0 + (double)r.w[2] * 0x1p-52;
        } else {
            v0 = (double)r.w[0] * qscale;
            v1 = (double)r.w[1] * qscale;
        }
    }
};
template <typename R>
struct PlaneLoadBatch<R, false, false> {
    struct Pair {
        R v[2];
    };
    static constexpr int B = 8;
    __device__ __forceinline__ static Pair load(const void* in, size_t g) {
        const cplx<R> c = *(const cplx<R>*)((const R*)in + g);
        return Pair{{c.re, c.im}};
    }
};

// ---------------------------------------------------------------------------- forward, plane x
// Workgroup (x, h), logical id 2x + h.  FIXED: the input is the fixed-point deposit (WIDE: 128-bit
// cells), converted to delta = rho/n_mean - 1 on load (pm3d.py:103-104); both workgroups of a plane
// store the plane's maximum rho to bmax[2x + h] (reduced by the x pass).  `clear` (nullable): a
// counter the deposit just consumed (the far-particle overflow count), zeroed by workgroup 0.
template <typename R, int N, bool FIXED, bool WIDE>
__device__ __forceinline__ void fwd_plane_body(const void* __restrict__ in, cplx<R>* __restrict__ spec,
                                               cplx<R>* __restrict__ nyq, const cplx<R>* __restrict__ tw_g,
                                               double* __restrict__ bmax,
                                               double qscale,
                                               double nmean, XMap xm, unsigned int* __restrict__ clear) {
    if (clear && blockIdx.x == 0 && threadIdx.x == 0) *clear = 0u;
    using PT = PlaneTile<R, N>;
    using LD = PlaneLoadBatch<R, FIXED, WIDE>;
    constexpr int H = PT::H, S = PT::S, NT = PT::NT;
    __shared__ cplx<R> T[S * S];
    // the twiddle table staged in LDS: the FFT passes read up to 15 twiddles per butterfly, from LDS
    // faster than through the vector memory path (mixed 256^3: fwd_zy 61 -> 58 us); first read after
    // the barrier that ends the load phase
    __shared__ cplx<R> TW[N];
    for (int t = threadIdx.x; t < N; t += NT) TW[t] = tw_g[t];
    const cplx<R>* tw = TW;
    const unsigned lin = xcd_map(blockIdx.x, gridDim.x, xm);
    const int x = (int)(lin >> 1), h = (int)(lin & 1);
    double lmax = 0.0;
    // delta = rho/n_mean - 1 as rho*(1/n_mean) - 1: the same bits when n_mean is a power of two (1
    // for a particle per cell), within an ulp otherwise, without a division per cell
    const double inv_nmean = 1.0 / nmean;
    // combined rows r[y'] = row(y') +/- row(y'+H), packed as (r[2m], r[2m+1]) at T[m*S + y'];
    // each thread issues the loads of LB cell pairs before it converts and stores any of them
    constexpr int IT = (H * H + NT - 1) / NT, LB = LD::B;
    for (int i0 = 0; i0 < IT; i0 += LB) {
        typename LD::Pair ra[LB], rb[LB];
#pragma unroll
        for (int b = 0; b < LB; ++b) {
            const int t = threadIdx.x + (i0 + b) * NT;
            if (i0 + b < IT && ((H * H) % NT == 0 || t < H * H)) {
                const int m = t % H, yp = t / H;
                const size_t g0 = ((size_t)x * N + yp) * N + 2 * m;
                ra[b] = LD::load(in, g0);
                rb[b] = LD::load(in, g0 + (size_t)H * N);
            }
        }
#pragma unroll
        for (int b = 0; b < LB; ++b) {
            const int t = threadIdx.x + (i0 + b) * NT;
            if (i0 + b < IT && ((H * H) % NT == 0 || t < H * H)) {
                const int m = t % H, yp = t / H;
                R a0, a1, b0, b1;
                if constexpr (FIXED) {
                    double r[4];
                    LD::values(ra[b], qscale, r[0], r[1]);
                    LD::values(rb[b], qscale, r[2], r[3]);
                    lmax = fmax(lmax, fmax(fmax(r[0], r[1]), fmax(r[2], r[3])));
                    a0 = (R)(r[0] * inv_nmean - 1.0);
                    a1 = (R)(r[1] * inv_nmean - 1.0);
                    b0 = (R)(r[2] * inv_nmean - 1.0);
                    b1 = (R)(r[3] * inv_nmean - 1.0);
                } else {
                    a0 = ra[b].v[0];
                    a1 = ra[b].v[1];
                    b0 = rb[b].v[0];
                    b1 = rb[b].v[1];
                }
                T[m * S + yp] = h ? cplx<R>{a0 - b0, a1 - b1} : cplx<R>{a0 + b0, a1 + b1};
            }
        }
    }
    if constexpr (FIXED) {
        const double mx = block_max_nonneg_nt<NT>(lmax);
        if (threadIdx.x == 0) bmax[lin] = mx;
    }
    __syncthreads();
    // z: H-point transform of every combined row (pencil y', element m)
    tile_fft<R, H, H, S, 1, true, false, 2, PT::ML2, NT>(T, tw);
    // even/odd split into bins 0..H (bin H, the Nyquist bin, goes to tile row H); parity 1 takes
    // the y twiddle exp(-2 pi i y'/N)
    for (int t = threadIdx.x; t < (H / 2 + 1) * H; t += NT) {
        const int yp = t % H, k = t / H;
        const cplx<R> zk = T[k * S + yp];
        const int km = k == 0 ? 0 : H - k;
        const cplx<R> zm = T[km * S + yp];
        cplx<R> xk = r2c_bin(zk, zm, tw[k]);
        cplx<R> xo = k == 0 ? cplx<R>{zk.re - zk.im, (R)0} : r2c_bin(zm, zk, tw[H - k]);
        if (h) {
            const cplx<R> w = tw[yp];
            xk = cmul(xk, w);
            xo = cmul(xo, w);
        }
        T[k * S + yp] = xk;
        if (k == 0) T[H * S + yp] = xo;
        else if (k != H - k) T[(H - k) * S + yp] = xo;
    }
    __syncthreads();
    // y: H-point transform of every kz column, kz = 0..H (pencil kz, element y')
    tile_fft<R, H, S, 1, S, true, false, 2, PT::ML2, NT>(T, tw);
    // ky = 2m + h
    for (int t = threadIdx.x; t < H * H; t += NT) {
        const int kz = t % H, m = t / H;
        spec[((size_t)x * N + 2 * m + h) * H + kz] = T[kz * S + m];
    }
    for (int t = threadIdx.x; t < H; t += NT) nyq[(size_t)x * N + 2 * t + h] = T[H * S + t];
}

template <typename R, int N, bool FIXED, bool WIDE>
__global__ __launch_bounds__((PlaneTile<R, N>::NT)) void k_fwd_plane(const void* __restrict__ in, cplx<R>* __restrict__ spec,
                                                                     cplx<R>* __restrict__ nyq, const cplx<R>* __restrict__ tw,
                                                                     double* __restrict__ bmax,
                                                                     double qscale,
                                                                     double nmean, XMap xm, unsigned int* __restrict__ clear) {
    fwd_plane_body<R, N, FIXED, WIDE>(in, spec, nyq, tw, bmax, qscale, nmean, xm, clear);
}

// ---------------------------------------------------------------------------- inverse, plane x
// Tile of the inverse plane kernel split Q ways along y: rows y = Q*m + q (m < N/Q) of one
// component's plane; T[kz*RS + m], kz = 0..N/2, RS = N/Q + 1.  A smaller tile (Q = 4, 8) leaves
// room for more workgroups per CU, whose load, LDS and store phases then overlap; each workgroup
// reads the whole component plane, the Q workgroups of a plane run on one XCD (the repeats come
// from L2).
template <typename R, int N, int Q>
struct InvTile {
    static constexpr int H = N / 2, S = H + 1, RW = N / Q, RS = RW + 1;
    static constexpr int BYTES = S * RS * (int)sizeof(cplx<R>) + N * (int)sizeof(cplx<R>);  // tile + twiddles
    static constexpr bool FITS = N >= 2 * Q && BYTES <= 160 * 1024 - 512;
    static constexpr int ML2 = PlaneTile<R, N>::ML2;
    // threads: about one per butterfly of the z pass (H/radix per row, RW rows), 256..1024
    static constexpr int WORK = RW * (H >> ML2);
    static constexpr int NT = WORK >= 1024 ? 1024 : WORK >= 512 ? 512 : 256;
};

// a * exp(+2 pi i e/8) for e in 0..7: exact for the quarter turns, twiddle product otherwise
// (tw[k] = exp(-2 pi i k/N), so exp(+2 pi i e/8) = conj(tw[e*N/8])).
template <typename R, int N>
__device__ __forceinline__ cplx<R> turn8(cplx<R> a, int e, const cplx<R>* __restrict__ tw) {
    switch (e & 7) {
        case 0: return a;
        case 2: return cplx<R>{-a.im, a.re};
        case 4: return cplx<R>{-a.re, -a.im};
        case 6: return cplx<R>{a.im, -a.re};
        default: return cmul(a, cconj(tw[(e & 7) * (N / 8)]));
    }
}

// Workgroup (comp, x, q), logical id (comp*N + x)*Q + q: inverse y then z c2r of component comp's
// spectrum on plane x, rows y = Q*m + q, written to out[comp*ocs + (x*N + y)*N + z] scaled by
// `scale` (1/N^3, numpy's backward normalisation).  The first radix-Q stage of the inverse y
// transform (decimation in frequency) is done while loading:
//   B_q[k'] = (sum_j A[k' + j*N/Q] exp(+2 pi i q j/Q)) exp(+2 pi i q k'/N),  k' < N/Q,
// then an N/Q-point inverse y transform gives rows m.  numpy's c2r uses only the real parts of the
// kz = 0 and kz = N/2 bins (numpy/fft/_pocketfft.py irfft), and so does this kernel.
template <typename R, int N, int Q>
__device__ __forceinline__ void inv_plane_body(const cplx<R>* __restrict__ gspec,
                                               const cplx<R>* __restrict__ gnyq, size_t gcs,
                                               size_t gcn, R* __restrict__ out, size_t ocs,
                                               const cplx<R>* __restrict__ tw_g, R scale, XMap xm) {
    using IT_ = InvTile<R, N, Q>;
    constexpr int H = IT_::H, S = IT_::S, RW = IT_::RW, RS = IT_::RS, NT = IT_::NT;
    static_assert(Q == 2 || Q == 4 || Q == 8, "split 2, 4 or 8 ways");
    __shared__ cplx<R> T[S * RS];
    // the twiddle table staged in LDS (the load phase's turns and the FFT passes read it; mixed
    // 256^3: inv_yz 157 -> 139 us)
    __shared__ cplx<R> TW[N];
    for (int t = threadIdx.x; t < N; t += NT) TW[t] = tw_g[t];
    __syncthreads();
    const cplx<R>* tw = TW;
    const unsigned lin = xcd_map(blockIdx.x, gridDim.x, xm);
    const int comp = (int)(lin / (Q * N)), x = (int)((lin / Q) % N), q = (int)(lin % Q);
    const cplx<R>* sp = gspec + comp * gcs + (size_t)x * N * H;
    const cplx<R>* np = gnyq + comp * gcn + (size_t)x * N;
    // each thread issues the loads of LB elements (Q spectrum values each) before it stores any: as
    // few batches as keep the loaded data within VB VGPRs, balanced (one memory latency each)
    constexpr int IT = (RW * S + NT - 1) / NT, VPE = Q * (int)sizeof(cplx<R>) / 4, VB = sizeof(R) == 8 ? 72 : 40;
    constexpr int NB = (IT * VPE + VB - 1) / VB, LB = (IT + NB - 1) / NB;
    for (int i0 = 0; i0 < IT; i0 += LB) {
        cplx<R> a[LB][Q], w[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int t = threadIdx.x + (i0 + u) * NT;
            if (i0 + u < IT && t < RW * S) {
                const int kz = t % S, kp = t / S;
                const cplx<R>* pa = kz < H ? sp + (size_t)kp * H + kz : np + kp;
                const size_t ob = kz < H ? (size_t)RW * H : (size_t)RW;
#pragma unroll
                for (int jj = 0; jj < Q; ++jj) a[u][jj] = pa[jj * ob];
                w[u] = tw[(q * kp) % N];
            }
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
            const int t = threadIdx.x + (i0 + u) * NT;
            if (i0 + u < IT && t < RW * S) {
                const int kz = t % S, kp = t / S;
                cplx<R> v = a[u][0];
#pragma unroll
                for (int jj = 1; jj < Q; ++jj) v = cadd(v, turn8<R, N>(a[u][jj], (q * jj * (8 / Q)) & 7, tw));
                T[kz * RS + kp] = q ? cmul(v, cconj(w[u])) : v;
            }
        }
    }
    __syncthreads();
    // c2r uses only the real parts of the kz = 0 and kz = H columns' inverse y transforms, which are
    // the transforms of their Hermitian parts (Herm(b)[k'] = (b[k'] + conj b[-k'])/2): column 0
    // takes Herm(b_0) + i Herm(b_H), so one transform yields both (real and imaginary part) and the
    // y pass runs over H columns, one per butterfly slot, instead of H + 1
    for (int k = threadIdx.x; k <= RW / 2; k += NT) {
        const int km = k == 0 ? 0 : RW - k;
        const cplx<R> a = T[k], am = T[km], b = T[H * RS + k], bm = T[H * RS + km];
        const R har = (a.re + am.re) * (R)0.5, hai = (a.im - am.im) * (R)0.5;
        const R hbr = (b.re + bm.re) * (R)0.5, hbi = (b.im - bm.im) * (R)0.5;
        T[k] = cplx<R>{har - hbi, hai + hbr};
        if (km != k) T[km] = cplx<R>{har + hbi, hbr - hai};
    }
    __syncthreads();
    // inverse y: RW-point transform of the kz columns 0 .. H-1 (pencil kz, element k')
    tile_fft<R, RW, S - 1, 1, RS, true, true, Q, IT_::ML2, NT>(T, tw);
    // c2r input points of every row m (pencil m, element k): point k from bins k and H-k (bin 0
    // pairs with the Nyquist bin, whose real part column 0 carries as its imaginary part)
    for (int t = threadIdx.x; t < (H / 2 + 1) * RW; t += NT) {
        const int m = t % RW, k = t / RW;
        cplx<R> xk = T[k * RS + m];
        const cplx<R> xm = k == 0 ? cplx<R>{xk.im, (R)0} : T[(H - k) * RS + m];
        if (k == 0) xk.im = (R)0;
        const cplx<R> pk = c2r_point(xk, xm, tw[k]);
        if (k != 0 && k != H - k) T[(H - k) * RS + m] = c2r_point(xm, xk, tw[H - k]);
        T[k * RS + m] = pk;
    }
    __syncthreads();
    // z: H-point inverse transform of every row
    tile_fft<R, H, RW, RS, 1, true, true, 2, IT_::ML2, NT>(T, tw);
    // cells 2j and 2j+1 of row y = Q*m + q are the real and imaginary parts of point j
    R* dst = out + comp * ocs + ((size_t)x * N + q) * N;
    for (int t = threadIdx.x; t < RW * N; t += NT) {
        const int z = t % N, m = t / N;
        const cplx<R> v = T[(z >> 1) * RS + m];
        dst[(size_t)(Q * m) * N + z] = ((z & 1) ? v.im : v.re) * scale;
    }
}

template <typename R, int N, int Q>
__global__ __launch_bounds__((InvTile<R, N, Q>::NT)) void k_inv_plane(const cplx<R>* __restrict__ gspec,
                                                                      const cplx<R>* __restrict__ gnyq, size_t gcs,
                                                                      size_t gcn, R* __restrict__ out, size_t ocs,
                                                                      const cplx<R>* __restrict__ tw, R scale, XMap xm) {
    inv_plane_body<R, N, Q>(gspec, gnyq, gcs, gcn, out, ocs, tw, scale, xm);
}

}  // namespace pm3d
