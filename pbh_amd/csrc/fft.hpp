// Hand-written Stockham FFT for the periodic Poisson solve (gfx950).
//
// The 3D real-to-complex transform of an N^3 mesh is done in three pencil passes, each one a
// kernel that stages P pencils in LDS, runs a mixed radix-4/2/3 Stockham FFT there and
// writes back coalesced rows:
//   z pass  (contiguous rows): real N -> N/2 complex FFT + even/odd split  -> half spectrum
//   y pass  (stride N/2)     : complex N, in place
//   x pass  (stride N*N/2)   : complex N forward, then the Green's function, the optional Gaussian
//                              filter and -i k_c fused in registers, then three inverse x FFTs
// and the inverse runs y and z passes back for the three gradient components.
//
// Half-spectrum layout: `spec` holds kz = 0..N/2-1 as [kx][ky][kz] (power-of-two rows) and the
// kz = N/2 Nyquist plane lives in a separate [kx][ky] array `nyq`.  numpy's layout (N,N,N/2+1) is
// assembled only on the host side of the test entry points.
//
// LDS layout of a pencil block: element j of pencil c at buf[lds_at(j, c)], rows of P elements
// (j*PP + c with PP = lds_pad(P): P for rows of 128 bytes or more, else P + 1; FP32 with P = 8
// in rows of 8 in padded groups of 8);
// threads map pencil-fastest in the butterflies, so every LDS access of a wave is a run of
// contiguous pencils.
#pragma once
#include "common.hpp"

namespace pm3d {

// Pencils per workgroup for each pass: every LDS ping-pong buffer holds about 32 KB, and the count
// is a power of two that divides the number of pencils (N is 2^k or 3*2^k).
constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int pow2_floor(int v) { return v >= 2 ? 2 * pow2_floor(v / 2) : 1; }
constexpr int pow2_dividing(int cap, int total) {  // largest power of two <= cap dividing total
    return (cap <= 1) ? 1 : (total % cap == 0 ? cap : pow2_dividing(cap / 2, total));
}
// Pass kinds, for the radix table.
enum FftPassKind { PK_R2C = 0, PK_YFWD, PK_X, PK_YINV, PK_C2R };

// log2 of the largest radix of each pass (MI355X timings, us per launch at 128^3 FP64 / 256^3 FP32,
// radix 4 -> radix 8 (FP64) or 16 (FP32)): FP64 x 25.9 -> 24.9, y inv 24.2 -> 23.7, the other
// passes equal within 1 %; FP32 r2c 56.9 -> 51.9, y fwd 45.0 -> 44.2, y inv 122 -> 114, c2r 171 ->
// 148, but x 99 -> 116 (the Poisson x pass keeps its per-thread spectrum in registers, which radix
// 16 crowds).  The FP32 x pass takes radix 8 (256 = 8 * 8 * 4: three LDS round trips instead of
// four, and every write of the 8-pencil grouped tile on its own bank offsets, where radix 4's write
// with Ns = 4 overlapped): mixed 256^3 x_poisson 80.4 -> 71 us.  The radix fixes the rounding
// sequence, so it is a constant, not a tuning choice.
template <typename R, int PK>
constexpr int radix_l2() {
    if constexpr (sizeof(R) == 8) return (PK == PK_X || PK == PK_YINV) ? 3 : 2;
    else return PK == PK_X ? 3 : 4;
}

// Tile configuration V of a pass (every V computes the same bits, so pm3d.hip picks by timing):
//   bit 0: pencils per workgroup sized for ~16 KB (0) or ~32 KB (1) per LDS buffer;
//   bit 1: two ping-pong buffers (0) or one buffer transformed in place (1: half the LDS, so twice
//          the resident workgroups, at the cost of a second barrier per radix pass).
template <typename R, int N, int V> struct Plan {
    static constexpr int EB = ((V & 1) ? 32768 : 16384) / (2 * (int)sizeof(R));  // complex elements per buffer
    static constexpr bool INPLACE = (V & 2) != 0;
    static constexpr int PZ = pow2_dividing(pow2_floor(cmin(2 * EB / N, N * N)), N * N);  // r2c rows
    static constexpr int PY = pow2_dividing(pow2_floor(cmin(EB / N, N / 2)), N / 2);      // y columns
    static constexpr int PX = pow2_dividing(pow2_floor(cmin(EB / N, N / 2)), N / 2);      // x columns
    // c2r rows: NCOMP components of PC rows share one tile of NCOMP*PC pencils of N/2 points
    template <int NCOMP>
    static constexpr int pc() { return pow2_dividing(pow2_floor(cmin(EB / ((N / 2) * NCOMP), N * N)), N * N); }
    static constexpr int NBUF = INPLACE ? 1 : 2;
};

// In-register DFT of size R in {2, 3, 4, 8, 16} (forward: exp(-2 pi i km/R); INV: conjugate).
// Sizes 8 and 16 split into even/odd halves (decimation in time) with the constant twiddles
// exp(-/+ 2 pi i k/R).
template <typename R, bool INV>
__device__ __forceinline__ cplx<R> rot_mi(cplx<R> a) {  // multiply by -i (forward) or +i (inverse)
    return INV ? cplx<R>{-a.im, a.re} : cplx<R>{a.im, -a.re};
}
template <typename R, bool INV>
__device__ __forceinline__ cplx<R> twc(cplx<R> a, double c, double sn) {  // a * (c -/+ i sn)
    const R cc = (R)c, ss = (R)(INV ? -sn : sn);
    return cplx<R>{a.re * cc + a.im * ss, a.im * cc - a.re * ss};
}
template <typename R, int RADIX, bool INV>
__device__ __forceinline__ void dft_reg(cplx<R>* v) {
    if constexpr (RADIX == 2) {
        const cplx<R> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (RADIX == 3) {
        const R h = (R)0.86602540378443864676;  // sin(2 pi / 3)
        const cplx<R> t1 = cadd(v[1], v[2]);
        const cplx<R> t2 = cplx<R>{v[0].re - (R)0.5 * t1.re, v[0].im - (R)0.5 * t1.im};
        const cplx<R> d = csub(v[1], v[2]);
        const cplx<R> t3 = INV ? cplx<R>{-h * d.im, h * d.re} : cplx<R>{h * d.im, -h * d.re};
        v[0] = cadd(v[0], t1);
        v[1] = cadd(t2, t3);
        v[2] = csub(t2, t3);
    } else if constexpr (RADIX == 4) {
        const cplx<R> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]), a2 = cadd(v[1], v[3]);
        const cplx<R> a3 = rot_mi<R, INV>(csub(v[1], v[3]));
        v[0] = cadd(a0, a2);
        v[1] = cadd(a1, a3);
        v[2] = csub(a0, a2);
        v[3] = csub(a1, a3);
    } else {
        constexpr int H = RADIX / 2;
        cplx<R> e[H], o[H];
#pragma unroll
        for (int m = 0; m < H; ++m) {
            e[m] = v[2 * m];
            o[m] = v[2 * m + 1];
        }
        dft_reg<R, H, INV>(e);
        dft_reg<R, H, INV>(o);
        // cos/sin(2 pi k / 16), k = 0..7 (k/16 = 2k/32 for RADIX 8 uses even entries)
        constexpr double C16[8] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                                   0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613};
        constexpr double S16[8] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                                   1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173};
#pragma unroll
        for (int k = 0; k < H; ++k) {
            constexpr int STEP = 16 / RADIX;
            cplx<R> t;
            if (k == 0) t = o[0];
            else if (k * STEP == 4) t = rot_mi<R, INV>(o[k]);
            else t = twc<R, INV>(o[k], C16[k * STEP], S16[k * STEP]);
            v[k] = cadd(e[k], t);
            v[k + H] = csub(e[k], t);
        }
    }
}

// Row stride of a P-pencil LDS tile buf[j*PP + c].  Threads take consecutive pencils, so a wave
// reads 64/P rows of P elements.  Rows of 128 bytes or more (FP64 complex with P >= 8, FP32 complex
// with P >= 16) are unpadded: measured on MI355X, the FP64 x pass at the 128^3 crossing lost 0.6 of
// its active LDS cycles to bank conflicts with P + 1 (1.1 with P + 2, 0.6 with P + 4) and none with
// P (P = 16; 0.02 with P = 8), x_poisson 19.7 -> 18.9 us; the FP32 x passes of the 32^3 .. 128^3
// boxes (P = 16, 32) went from 0.25 - 0.48 to 0, their active LDS cycles -15 %.  Shorter rows keep
// P + 1 (FP32 with P = 8 takes the grouped layout of lds_at instead: a stride of 18 dwords started
// the 8 rows of a wave unevenly, about 2e7 bank-conflict cycles per 256^3 x-pass launch, and 20
// dwords still left conflicts of 1.6x the active LDS cycles).
template <typename R>
constexpr int lds_pad(int P) { return P * (int)sizeof(cplx<R>) >= 128 ? P : P + 1; }

// Element (row, pencil c) of a P-pencil LDS tile, and the elements a tile of `rows` rows takes.
// FP32 with P = 8: rows of 8 elements (16 dwords) in groups of 8 rows, each group (128 dwords)
// followed by 8 unused elements (16 dwords).  A wave's accesses in the pencil passes are rows
// j .. j+7 (the radix-4 Stockham reads j + r*M and writes (j/Ns)*Ns*4 + r*Ns + j%Ns), which then
// start on distinct 16-dword bank offsets in every pass but the write with Ns = 4 (two runs of 4
// rows, 288 dwords apart).  Measured at the 256^3 crossing: conflict cycles 1.62 -> 0.33 of the
// active LDS cycles against a stride of P + 2.  The padding is additive, so the accesses of one
// thread keep constant distances (LDS instruction pairing).  Other tiles keep lds_pad.
template <typename R, int P>
constexpr bool lds_grouped() { return sizeof(R) == 4 && P == 8; }
template <typename R, int P>
__host__ __device__ __forceinline__ constexpr int lds_at(int row, int c) {
    if constexpr (lds_grouped<R, P>()) return row * 8 + (row >> 3) * 8 + c;
    else return row * lds_pad<R>(P) + c;
}
template <typename R, int P>
constexpr int lds_size(int rows) {
    return lds_grouped<R, P>() ? rows * 8 + ((rows + 7) >> 3) * 8 : rows * lds_pad<R>(P);
}
// Element distance from (row, c) to (row + d, c) when (row % 8) + (d % 8) < 8, so that the rows
// between cross d / 8 group ends: a constant for a constant d, which keeps the accesses of one
// thread at immediate offsets of one address (the compiler cannot derive that from lds_at).
template <typename R, int P>
__host__ __device__ __forceinline__ constexpr int lds_adv(int d) {
    return lds_grouped<R, P>() ? d * 8 + (d >> 3) * 8 : d * lds_pad<R>(P);
}
// A radix-RADIX pass reads rows j + r*M and writes rows dd + r*NS with dd = (j/NS)*NS*RADIX + k,
// k < NS.  lds_adv holds for the reads when M % 8 == 0, and for the writes when NS % 8 == 0 or
// NS and RADIX are powers of two (the NS*RADIX rows of dd's block then lie in whole groups, or in
// one group, and dd % 8 + (r*NS) % 8 stays below 8).
constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
template <typename R, int P, int M>
constexpr bool lds_rd_adv() { return !lds_grouped<R, P>() || M % 8 == 0; }
template <typename R, int P, int NS, int RADIX>
constexpr bool lds_wr_adv() { return !lds_grouped<R, P>() || NS % 8 == 0 || (is_pow2(NS) && is_pow2(RADIX)); }

// Stockham pass of radix RADIX on P pencils (see lds_fft).
template <typename R, int NC, int P, bool INV, int TWS, int RADIX, int Ns>
__device__ __forceinline__ void stockham_pass(const cplx<R>* src, cplx<R>* dst, const cplx<R>* __restrict__ tw) {
    constexpr int M = NC / RADIX;
    constexpr int s = TWS * (NC / (RADIX * Ns));
    for (int t = threadIdx.x; t < P * M; t += kThreads) {
        const int c = t % P, j = t / P;
        const int k = j % Ns;
        cplx<R> v[RADIX];
        const int j0 = lds_at<R, P>(j, c);
#pragma unroll
        for (int r = 0; r < RADIX; ++r)
            v[r] = src[lds_rd_adv<R, P, M>() ? j0 + lds_adv<R, P>(r * M) : lds_at<R, P>(j + r * M, c)];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < RADIX; ++r) {
                cplx<R> w = tw[r * k * s];
                if (INV) w = cconj(w);
                v[r] = cmul(v[r], w);
            }
        }
        dft_reg<R, RADIX, INV>(v);
        const int dd = (j / Ns) * Ns * RADIX + k, d0 = lds_at<R, P>(dd, c);
#pragma unroll
        for (int r = 0; r < RADIX; ++r)
            dst[lds_wr_adv<R, P, Ns, RADIX>() ? d0 + lds_adv<R, P>(r * Ns) : lds_at<R, P>(dd + r * Ns, c)] = v[r];
    }
    __syncthreads();
}

// The same pass on one buffer: every thread reads its RADIX-point inputs, applies the twiddles and
// the DFT in registers, and writes back after a barrier.
template <typename R, int NC, int P, bool INV, int TWS, int RADIX, int Ns>
__device__ __forceinline__ void stockham_pass_ip(cplx<R>* buf, const cplx<R>* __restrict__ tw) {
    constexpr int M = NC / RADIX, T = P * M, IT = (T + kThreads - 1) / kThreads;
    constexpr int s = TWS * (NC / (RADIX * Ns));
    cplx<R> v[IT][RADIX];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int t = threadIdx.x + it * kThreads;
        if (T % kThreads == 0 || t < T) {
            const int c = t % P, j = t / P, k = j % Ns, j0 = lds_at<R, P>(j, c);
#pragma unroll
            for (int r = 0; r < RADIX; ++r)
                v[it][r] = buf[lds_rd_adv<R, P, M>() ? j0 + lds_adv<R, P>(r * M) : lds_at<R, P>(j + r * M, c)];
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
        const int t = threadIdx.x + it * kThreads;
        if (T % kThreads == 0 || t < T) {
            const int c = t % P, j = t / P, k = j % Ns;
            const int dd = (j / Ns) * Ns * RADIX + k, d0 = lds_at<R, P>(dd, c);
#pragma unroll
            for (int r = 0; r < RADIX; ++r)
                buf[lds_wr_adv<R, P, Ns, RADIX>() ? d0 + lds_adv<R, P>(r * Ns) : lds_at<R, P>(dd + r * Ns, c)] =
                    v[it][r];
        }
    }
    __syncthreads();
}

constexpr int count_factor(int n, int f) { return (n % f == 0) ? 1 + count_factor(n / f, f) : 0; }

// In-LDS Stockham FFT of P pencils of NC points (NC = 2^a 3^b, unnormalised; forward =
// exp(-2 pi i jk/NC)).  Power-of-two passes of radix 2^ML2 (then one smaller power of two for the
// rest), then radix-3 passes; each pass is one LDS round trip.  Input in `a`, result in the
// returned buffer (always `a` when IP: in place, `b` unused).  Caller synchronises before the call;
// the call ends with a barrier.  tw[k] = exp(-2 pi i k / (NC*TWS)).
template <typename R, int NC, int P, bool INV, int TWS, int ML2, bool IP, int NS = 1, int REM2 = count_factor(NC, 2),
          int REM3 = count_factor(NC, 3)>
__device__ __forceinline__ cplx<R>* lds_fft(cplx<R>* a, cplx<R>* b, const cplx<R>* __restrict__ tw) {
    if constexpr (REM2 == 0 && REM3 == 0) {
        return a;
    } else {
        constexpr int L2 = REM2 >= ML2 ? ML2 : REM2;
        constexpr int RADIX = REM2 > 0 ? (1 << L2) : 3;
        constexpr int R2 = REM2 > 0 ? REM2 - L2 : 0, R3 = REM2 > 0 ? REM3 : REM3 - 1;
        if constexpr (IP) {
            stockham_pass_ip<R, NC, P, INV, TWS, RADIX, NS>(a, tw);
            return lds_fft<R, NC, P, INV, TWS, ML2, IP, NS * RADIX, R2, R3>(a, b, tw);
        } else {
            stockham_pass<R, NC, P, INV, TWS, RADIX, NS>(a, b, tw);
            return lds_fft<R, NC, P, INV, TWS, ML2, IP, NS * RADIX, R2, R3>(b, a, tw);
        }
    }
}

// ------------------------------------------------------------------------------------------ z r2c
// Rows (x,y) of the real mesh -> half spectrum along kz.  FIXED: the input is the fixed-point
// deposit (WIDE: 128-bit cells), converted on the fly to delta = rho/n_mean - 1 (pm3d.py:103-104),
// with each block's maximum rho stored to bmax[blockIdx.x] (reduced by the following y pass, so no
// two blocks contend for one address).  `clear` (nullable): a counter the deposit just consumed
// (the far-particle overflow count), zeroed by block 0.
template <typename R, int N, bool FIXED, bool WIDE, int V>
__global__ __launch_bounds__(kThreads) void k_r2c_z(const void* __restrict__ in, cplx<R>* __restrict__ spec,
                                                    cplx<R>* __restrict__ nyq, const cplx<R>* __restrict__ tw,
                                                    double* __restrict__ bmax, double qscale, double nmean,
                                                    unsigned int* __restrict__ clear) {
    if (clear && blockIdx.x == 0 && threadIdx.x == 0) *clear = 0u;
    constexpr int NC = N / 2, P = Plan<R, N, V>::PZ;
    __shared__ cplx<R> A[lds_size<R, P>(NC)], B[Plan<R, N, V>::INPLACE ? 1 : lds_size<R, P>(NC)];
    const size_t row0 = (size_t)blockIdx.x * P;
    double lmax = 0.0;
    const double inv_nmean = 1.0 / nmean;  // rho*(1/n_mean) - 1: exact for a power-of-two n_mean
    for (int t = threadIdx.x; t < P * NC; t += kThreads) {
        const int c = t / NC, m = t - c * NC;
        const size_t g = (row0 + c) * N + 2 * m;
        R v0, v1;
        if constexpr (FIXED) {
            const unsigned long long* mesh = (const unsigned long long*)in;
            const double r0 = fx_value<WIDE>(mesh, g, qscale), r1 = fx_value<WIDE>(mesh, g + 1, qscale);
            lmax = fmax(lmax, fmax(r0, r1));
            v0 = (R)(r0 * inv_nmean - 1.0);
            v1 = (R)(r1 * inv_nmean - 1.0);
        } else {
            v0 = ((const R*)in)[g];
            v1 = ((const R*)in)[g + 1];
        }
        A[lds_at<R, P>(m, c)] = cplx<R>{v0, v1};
    }
    if constexpr (FIXED) {
        const double m = block_max_nonneg(lmax);
        if (threadIdx.x == 0) bmax[blockIdx.x] = m;
    }
    __syncthreads();
    const cplx<R>* Z = lds_fft<R, NC, P, false, 2, radix_l2<R, PK_R2C>(), Plan<R, N, V>::INPLACE>(A, B, tw);
    for (int t = threadIdx.x; t < P * NC; t += kThreads) {
        const int c = t / NC, k = t - c * NC;
        const cplx<R> zk = Z[lds_at<R, P>(k, c)];
        const cplx<R> zm = Z[lds_at<R, P>(k == 0 ? 0 : NC - k, c)];
        // even part E = (Z_k + conj Z_{-k})/2, odd part O = (Z_k - conj Z_{-k})/(2i)
        const R er = (zk.re + zm.re) * (R)0.5, ei = (zk.im - zm.im) * (R)0.5;
        const R orr = (zk.im + zm.im) * (R)0.5, oi = (zm.re - zk.re) * (R)0.5;
        const cplx<R> w = tw[k];
        spec[(row0 + c) * NC + k] = cplx<R>{er + (w.re * orr - w.im * oi), ei + (w.re * oi + w.im * orr)};
        if (k == 0) nyq[row0 + c] = cplx<R>{zk.re - zk.im, (R)0};
    }
}

// ------------------------------------------------------------------------------------------ y c2c
// In-place complex FFT along y on `ncomp` = gridDim.y half spectra (component strides cs, cn).
// If nbmax > 0, the last block also reduces the r2c pass's per-block maxima bmax[0..nbmax) into
// *slot (the deposit's maximum rho, as the bits of a non-negative double).
template <typename R, int N, bool INV, int V>
__global__ __launch_bounds__(kThreads) void k_c2c_y(cplx<R>* __restrict__ spec, cplx<R>* __restrict__ nyq,
                                                    size_t cs, size_t cn, const cplx<R>* __restrict__ tw,
                                                    const double* __restrict__ bmax, int nbmax,
                                                    unsigned long long* __restrict__ slot, XMap xm) {
    constexpr int P = Plan<R, N, V>::PY, NH = N / 2, NBM = N * (NH / P);
    constexpr int ML2 = radix_l2<R, INV ? PK_YINV : PK_YFWD>();
    constexpr bool IP = Plan<R, N, V>::INPLACE;
    __shared__ cplx<R> A[lds_size<R, P>(N)], B[IP ? 1 : lds_size<R, P>(N)];
    if (nbmax > 0 && blockIdx.x == gridDim.x - 1 && blockIdx.y == 0) {
        double m = 0.0;
        for (int t = threadIdx.x; t < nbmax; t += kThreads) m = fmax(m, bmax[t]);
        m = block_max_nonneg(m);
        if (threadIdx.x == 0) *slot = (unsigned long long)__double_as_longlong(m);
    }
    const unsigned lin = xcd_map(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y, xm);
    const int b = (int)(lin % gridDim.x);
    spec += (lin / gridDim.x) * cs;
    nyq += (lin / gridDim.x) * cn;
    if (b < NBM) {
        const int x = b / (NH / P), col0 = (b % (NH / P)) * P;
        cplx<R>* base = spec + (size_t)x * N * NH + col0;
        for (int t = threadIdx.x; t < N * P; t += kThreads) {
            const int c = t % P, j = t / P;
            A[lds_at<R, P>(j, c)] = base[(size_t)j * NH + c];
        }
        __syncthreads();
        const cplx<R>* res = lds_fft<R, N, P, INV, 1, ML2, IP>(A, B, tw);
        for (int t = threadIdx.x; t < N * P; t += kThreads) {
            const int c = t % P, j = t / P;
            base[(size_t)j * NH + c] = res[lds_at<R, P>(j, c)];
        }
    } else {
        cplx<R>* base = nyq + (size_t)(b - NBM) * P * N;  // pencil c = row x0+c, contiguous in y
        for (int t = threadIdx.x; t < N * P; t += kThreads) {
            const int j = t % N, c = t / N;
            A[lds_at<R, P>(j, c)] = base[(size_t)c * N + j];
        }
        __syncthreads();
        const cplx<R>* res = lds_fft<R, N, P, INV, 1, ML2, IP>(A, B, tw);
        for (int t = threadIdx.x; t < N * P; t += kThreads) {
            const int j = t % N, c = t / N;
            base[(size_t)c * N + j] = res[lds_at<R, P>(j, c)];
        }
    }
}

// numpy fftfreq / rfftfreq: k = (2*pi) * (i * (1/(N*dx))) with the Nyquist index of the full axis
// negative (pm3d.py:31-32; numpy/fft/_helper.py fftfreq, rfftfreq).
__device__ __forceinline__ double kfull(int i, int N, double kval) {
    return 6.283185307179586 * ((double)(i < N / 2 ? i : i - N) * kval);
}
__device__ __forceinline__ double khalf(int i, double kval) { return 6.283185307179586 * ((double)i * kval); }

// ------------------------------------------------------------------------------------------ x pass
// MODE 0: forward FFT along x in place.  MODE 1: inverse in place.  MODE 2 (the Poisson pass):
// forward FFT along x, then per mode  phik = ((s*dk) * (1/k2)) [* exp((-0.5*k2)*sm2)], phik(0)=0
// (pm3d.py:105-109; numpy's complex/real division is a multiply by the reciprocal), then for
// c = x,y,z: -1j*k_c*phik = (k_c*Im, -k_c*Re) (pm3d.py:112) and the inverse FFT along x, written to
// the three gradient spectra (k_y and k_z are applied after the transform of phik, which is linear).  If nbmax > 0, the last block also reduces per-block deposit maxima
// bmax[0..nbmax) into *slot (left there by the plane forward pass, fft_plane.hpp).
template <typename R, int N, int MODE, int V>
__device__ __forceinline__ void x_pass_body(cplx<R>* __restrict__ spec, cplx<R>* __restrict__ nyq,
                                            cplx<R>* __restrict__ gspec, cplx<R>* __restrict__ gnyq,
                                            size_t gcs, size_t gcn, const cplx<R>* __restrict__ tw_g,
                                            double kval, double s, double sm2, int smooth_on,
                                            const DevSched* __restrict__ sch, const DevStep* __restrict__ tab,
                                            const double* __restrict__ bmax, int nbmax,
                                            unsigned long long* __restrict__ slot, XMap xm) {
    if (nbmax > 0 && blockIdx.x == gridDim.x - 1) {
        double m = 0.0;
        for (int t = threadIdx.x; t < nbmax; t += kThreads) m = fmax(m, bmax[t]);
        m = block_max_nonneg(m);
        if (threadIdx.x == 0) *slot = (unsigned long long)__double_as_longlong(m);
    }
    if (MODE == 2 && sch) {  // Green's function factor of the scheduled step (last entry past the end)
        const int k = sch->counter < sch->nsteps ? sch->counter : sch->nsteps - 1;
        s = tab[k].s;
        sm2 = tab[k].sm2;
        smooth_on = sm2 > 0.0 ? 1 : 0;
    }
    constexpr int P = Plan<R, N, V>::PX, NH = N / 2, NBM = N * (NH / P);
    constexpr int EPT = (N * P + kThreads - 1) / kThreads;
    constexpr int ML2 = radix_l2<R, PK_X>();
    constexpr bool IP = Plan<R, N, V>::INPLACE;
    __shared__ cplx<R> A[lds_size<R, P>(N)], B[IP ? 1 : lds_size<R, P>(N)];
    // the twiddle table staged in LDS for the pencil FFTs (FP64 128^3: 22.1 -> 21.1 us); first read
    // after the barrier that ends the load phase
    __shared__ cplx<R> TW[N];
    for (int t = threadIdx.x; t < N; t += kThreads) TW[t] = tw_g[t];
    const cplx<R>* tw = TW;
    const int b = (int)xcd_map(blockIdx.x, gridDim.x, xm);
    const bool main = b < NBM;
    int y = 0, col0 = 0;
    if (main) {
        y = b / (NH / P);
        col0 = (b % (NH / P)) * P;
    } else {
        col0 = (b - NBM) * P;  // first ky of this Nyquist block
    }
    // element (x = j, pencil c) offset within a spectrum
    auto off = [&](int j, int c) -> size_t {
        return main ? ((size_t)j * N + y) * NH + col0 + c : (size_t)j * N + col0 + c;
    };
    cplx<R>* src = main ? spec : nyq;
    for (int t = threadIdx.x; t < N * P; t += kThreads) {
        const int c = t % P, j = t / P;
        A[lds_at<R, P>(j, c)] = src[off(j, c)];
    }
    __syncthreads();
    if constexpr (MODE == 3) {
        // Zel'dovich displacement: psi_c = 1j*k_c*dk/k2 with dk(0) = 0, evaluated as numpy does
        // ((0 + 1j*k_c) * dk, then complex/real as a multiply by 1/k2), then inverse x FFTs
        const cplx<R>* res = lds_fft<R, N, P, false, 1, ML2, IP>(A, B, tw);
        cplx<R> dkr[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int t = threadIdx.x + e * kThreads;
            dkr[e] = cplx<R>{(R)0, (R)0};
            if (t < N * P) dkr[e] = res[lds_at<R, P>(t / P, t % P)];
        }
        __syncthreads();
        for (int comp = 0; comp < 3; ++comp) {
#pragma unroll
            for (int e = 0; e < EPT; ++e) {
                const int t = threadIdx.x + e * kThreads;
                if (t < N * P) {
                    const int c = t % P, j = t / P;
                    const int iy = main ? y : col0 + c, iz = main ? col0 + c : NH;
                    const double kx = kfull(j, N, kval), ky = kfull(iy, N, kval), kz = khalf(iz, kval);
                    const bool dc = (j == 0) && (iy == 0) && (iz == 0);
                    double k2 = (kx * kx + ky * ky) + kz * kz;
                    if (dc) k2 = 1.0;
                    const double rk = 1.0 / k2;
                    const double kc = comp == 0 ? kx : comp == 1 ? ky : kz;
                    double pr = (-(kc * (double)dkr[e].im)) * rk, pi = (kc * (double)dkr[e].re) * rk;
                    if (dc) { pr = 0.0; pi = 0.0; }
                    A[lds_at<R, P>(j, c)] = cplx<R>{(R)pr, (R)pi};
                }
            }
            __syncthreads();
            const cplx<R>* r2 = lds_fft<R, N, P, true, 1, ML2, IP>(A, B, tw);
            cplx<R>* dst = main ? gspec + comp * gcs : gnyq + comp * gcn;
            for (int t = threadIdx.x; t < N * P; t += kThreads) {
                const int c = t % P, j = t / P;
                dst[off(j, c)] = r2[lds_at<R, P>(j, c)];
            }
            __syncthreads();
        }
        return;
    } else if constexpr (MODE < 2) {
        const cplx<R>* res = lds_fft<R, N, P, MODE == 1, 1, ML2, IP>(A, B, tw);
        for (int t = threadIdx.x; t < N * P; t += kThreads) {
            const int c = t % P, j = t / P;
            src[off(j, c)] = res[lds_at<R, P>(j, c)];
        }
        return;
    } else {
        const cplx<R>* res = lds_fft<R, N, P, false, 1, ML2, IP>(A, B, tw);
        cplx<R> ph[EPT];
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int t = threadIdx.x + e * kThreads;
            ph[e] = cplx<R>{(R)0, (R)0};
            if (t < N * P) {
                const int c = t % P, j = t / P;
                const int iy = main ? y : col0 + c, iz = main ? col0 + c : NH;
                const double kx = kfull(j, N, kval), ky = kfull(iy, N, kval), kz = khalf(iz, kval);
                const bool dc = (j == 0) && (iy == 0) && (iz == 0);
                double k2 = (kx * kx + ky * ky) + kz * kz;
                if (dc) k2 = 1.0;
                const double rk = 1.0 / k2;
                const cplx<R> z = res[lds_at<R, P>(j, c)];
                double pr = (s * (double)z.re) * rk, pi = (s * (double)z.im) * rk;
                if (smooth_on) {
                    const double f = exp((-0.5 * k2) * sm2);
                    pr = pr * f;
                    pi = pi * f;
                }
                if (dc) { pr = 0.0; pi = 0.0; }
                ph[e] = cplx<R>{(R)pr, (R)pi};
            }
        }
        __syncthreads();
        // -1j*k_c for c = y, z is constant along an x pencil, so their inverse x transforms are
        // -1j*k_c times the inverse transform of phik: pass 0 transforms phik once and writes both
        // (the product rounded after the transform instead of before it), pass 1 transforms
        // -1j*k_x*phik for the x component
        for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
            for (int e = 0; e < EPT; ++e) {
                const int t = threadIdx.x + e * kThreads;
                if (t < N * P) {
                    const int c = t % P, j = t / P;
                    if (pass == 0) {
                        A[lds_at<R, P>(j, c)] = ph[e];
                    } else {
                        const double kx = kfull(j, N, kval);
                        A[lds_at<R, P>(j, c)] = cplx<R>{(R)(kx * (double)ph[e].im), (R)(-(kx * (double)ph[e].re))};
                    }
                }
            }
            __syncthreads();
            const cplx<R>* r2 = lds_fft<R, N, P, true, 1, ML2, IP>(A, B, tw);
            if (pass == 0) {
                cplx<R>* d1 = main ? gspec + gcs : gnyq + gcn;
                cplx<R>* d2 = main ? gspec + 2 * gcs : gnyq + 2 * gcn;
                for (int t = threadIdx.x; t < N * P; t += kThreads) {
                    const int c = t % P, j = t / P;
                    const int iy = main ? y : col0 + c, iz = main ? col0 + c : NH;
                    const double ky = kfull(iy, N, kval), kz = khalf(iz, kval);
                    const cplx<R> z = r2[lds_at<R, P>(j, c)];
                    d1[off(j, c)] = cplx<R>{(R)(ky * (double)z.im), (R)(-(ky * (double)z.re))};
                    d2[off(j, c)] = cplx<R>{(R)(kz * (double)z.im), (R)(-(kz * (double)z.re))};
                }
            } else {
                cplx<R>* d0 = main ? gspec : gnyq;
                for (int t = threadIdx.x; t < N * P; t += kThreads) {
                    const int c = t % P, j = t / P;
                    d0[off(j, c)] = r2[lds_at<R, P>(j, c)];
                }
            }
            __syncthreads();
        }
    }
}

template <typename R, int N, int MODE, int V>
__global__ __launch_bounds__(kThreads) void k_x_pass(cplx<R>* __restrict__ spec, cplx<R>* __restrict__ nyq,
                                                     cplx<R>* __restrict__ gspec, cplx<R>* __restrict__ gnyq,
                                                     size_t gcs, size_t gcn, const cplx<R>* __restrict__ tw,
                                                     double kval, double s, double sm2, int smooth_on,
                                                     const DevSched* __restrict__ sch, const DevStep* __restrict__ tab,
                                                     const double* __restrict__ bmax, int nbmax,
                                                     unsigned long long* __restrict__ slot, XMap xm) {
    x_pass_body<R, N, MODE, V>(spec, nyq, gspec, gnyq, gcs, gcn, tw, kval, s, sm2, smooth_on, sch, tab, bmax, nbmax, slot, xm);
}

// ------------------------------------------------------------------------------------------ z c2r
// Half spectra along kz -> real rows, NCOMP components written component-major
// out[comp*ocs + cell], scaled by `scale` (1/N^3, numpy's backward normalisation).  numpy's c2r uses only the real parts
// of the kz = 0 and kz = N/2 bins (numpy/fft/_pocketfft.py irfft), and so does this kernel.
// The NCOMP components of P rows are transformed together as NCOMP*P pencils of one LDS tile
// (pencil comp*P + c), and every row is written straight from that tile.
template <typename R, int N, int NCOMP, int V>
__global__ __launch_bounds__(kThreads) void k_c2r_z(const cplx<R>* __restrict__ gspec, const cplx<R>* __restrict__ gnyq,
                                                    size_t gcs, size_t gcn, R* __restrict__ out, size_t ocs,
                                                    const cplx<R>* __restrict__ tw, R scale) {
    constexpr int NC = N / 2, P = Plan<R, N, V>::template pc<NCOMP>(), NP = NCOMP * P;
    __shared__ cplx<R> A[lds_size<R, NP>(NC)], B[Plan<R, N, V>::INPLACE ? 1 : lds_size<R, NP>(NC)];
    const size_t row0 = (size_t)blockIdx.x * P;
    for (int t = threadIdx.x; t < NP * NC; t += kThreads) {
        const int q = t / NC, k = t - q * NC;  // q = comp*P + c
        const int comp = q / P, c = q - comp * P;
        const cplx<R>* sp = gspec + comp * gcs + (row0 + c) * NC;
        cplx<R> xk = sp[k];
        cplx<R> xm = (k == 0) ? gnyq[comp * gcn + row0 + c] : sp[NC - k];
        if (k == 0) { xk.im = (R)0; xm.im = (R)0; }
        const R er = xk.re + xm.re, ei = xk.im - xm.im;
        const R dr = xk.re - xm.re, di = xk.im + xm.im;
        const cplx<R> w = tw[k];  // exp(-2 pi i k/N); the odd part takes exp(+2 pi i k/N)
        const R orr = dr * w.re + di * w.im, oi = di * w.re - dr * w.im;
        A[lds_at<R, NP>(k, q)] = cplx<R>{er - oi, ei + orr};
    }
    __syncthreads();
    const cplx<R>* z = lds_fft<R, NC, NP, true, 2, radix_l2<R, PK_C2R>(), Plan<R, N, V>::INPLACE>(A, B, tw);
    // cells 2m and 2m+1 of a row are the real and imaginary parts of point m; consecutive threads
    // write consecutive cells of one row
    for (int t = threadIdx.x; t < NP * N; t += kThreads) {
        const int q = t / N, e = t - q * N;  // q = comp*P + c
        const int comp = q / P, c = q - comp * P;
        const cplx<R> v = z[lds_at<R, NP>(e >> 1, q)];
        out[comp * ocs + (row0 + c) * N + e] = ((e & 1) ? v.im : v.re) * scale;
    }
}

}  // namespace pm3d
