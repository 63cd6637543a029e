// Shared device helpers of the pm3d HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pm3d {

constexpr int kThreads = 256;  // 4 wave64 per workgroup in every kernel

// Device-resident step schedule: entry `counter` of the table holds the scalars of the next step,
// so one captured graph of a step replays for every step.  The FP32 factors are folded once on the
// host for the mixed mode.
struct DevStep {
    double c1p, denp;   // pending closing kick of the previous step: p += (c1p*acc)/denp
    double c1, den1;    // opening kick: p += (c1*acc)/den1
    double da, den_d;   // drift: x += (da*p)/den_d
    double den2;        // closing kick of this step (applied with the next opening kick)
    double s, sm2;      // Green's function factor -1.5*H_I^2/a_new and (smooth*dx)^2
    double rdenp, rden1, rden_d;  // 1/denp, 1/den1, 1/den_d, correctly rounded (qdiv)
    float kp, k1, sd, pad;  // mixed mode: c1p/denp, c1/den1, da/(den_d*dx)
};

// Load / store of base[byte_offset / sizeof(T)] with a 32-bit byte offset from a wave-uniform base
// (arrays under 4 GiB): one instruction with the base in scalar registers and the offset in one
// vector register, instead of a 64-bit address computed per lane and per array.
template <class T>
__device__ __forceinline__ T ld_off(const T* base, unsigned byte_offset) {
    return *(const T*)((const char*)base + byte_offset);
}
template <class T>
__device__ __forceinline__ void st_off(T* base, unsigned byte_offset, T v) {
    *(T*)((char*)base + byte_offset) = v;
}

// x / y for a divisor y fixed over the launch, with r = 1/y correctly rounded on the host: q0 = x*r
// is within about an ulp of x/y, and one correction with the exact FMA residual x - q0*y (Markstein,
// IBM J. Res. Dev. 34 (1990) 111) gives the IEEE quotient outside the under/overflow range (checked
// bit for bit against x/y on the CPU for the divisors of the KDK schedule, tests/test_div_exact.py).
// A zero x keeps its sign (-0/y = -0), as the division does.
__device__ __forceinline__ double qdiv(double x, double y, double r) {
    const double q0 = x * r;
    const double q = fma(fma(-q0, y, x), r, q0);
    return x == 0.0 ? q0 : q;
}

// Unsigned division by a divisor fixed for the launch (Granlund & Montgomery, PLDI 1994, the
// round-up method): q = (umulhi(n, m) + n) >> s with m = floor(2^32 (2^s - d) / d) + 1,
// s = ceil(log2 d); exact for n < 2^31 (umulhi(n, m) <= n, so the sum does not wrap).
struct Div32 {
    unsigned m, s;
};
inline Div32 div32_make(unsigned d) {
    unsigned s = 0;
    while ((1ull << s) < d) ++s;
    return Div32{(unsigned)(((1ull << 32) * ((1ull << s) - d)) / d + 1), s};
}
__device__ __forceinline__ unsigned div32(unsigned n, Div32 v) { return (__umulhi(n, v.m) + n) >> v.s; }

// Lattice site (a, b, c) of particle i = (a*Np + b)*Np + c, with dNp = div32_make(Np) and
// dNp2 = div32_make(Np*Np).
struct SiteDiv {
    Div32 dNp, dNp2;
    unsigned Np;
};
inline SiteDiv site_div(int Np) { return SiteDiv{div32_make((unsigned)Np), div32_make((unsigned)Np * Np), (unsigned)Np}; }
__device__ __forceinline__ void site_of(unsigned i, const SiteDiv& sd, int s[3]) {
    const unsigned a = div32(i, sd.dNp2), r = i - a * sd.Np * sd.Np, b = div32(r, sd.dNp);
    s[0] = (int)a;
    s[1] = (int)b;
    s[2] = (int)(r - b * sd.Np);
}

// c in (-2^30, 2^30) wrapped into [0, N): one compare for a step of less than a box, the remainder
// otherwise.
__device__ __forceinline__ int wrap_cell(int c, int N) {
    if ((unsigned)c >= (unsigned)N) {
        c = c < 0 ? c + N : c - N;
        if ((unsigned)c >= (unsigned)N) {
            c %= N;
            if (c < 0) c += N;
        }
    }
    return c;
}
struct DevSched {
    int counter;  // next entry; advanced by the last kernel of each step
    int nsteps;   // entries in the table
    int pad0, pad1;
};

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

// npy_rem for the drift, where v = x + dx_step stays within one box of [0, L): fmod is exact
// (v itself, or v - L by Sterbenz), and only v < 0 takes the rounded `mod += L` of numpy.
__device__ __forceinline__ double npy_rem_near(double v, double L) {
    double m;
    if (v >= 0.0 && v < L) m = v;
    else if (v >= L && v < 2.0 * L) m = v - L;
    else if (v < 0.0 && v > -L) m = v + L;
    else return npy_rem(v, L);
    return m == 0.0 ? 0.0 : m;  // numpy returns +0.0 for a zero remainder
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

// Hot-path form of cic_axis with 32-bit integer arithmetic: identical results (fl is integral, so
// g - fl == g - (double)(int64)fl), periodic wrap by conditional add/subtract, and the general
// 64-bit modulo only for a position more than one box away from [0, L).
__device__ __forceinline__ double div_dx(double X, double dx, double inv_dx) {
    // X/dx correctly rounded: q0 = X*(1/dx) is within 1 ulp and one FMA correction step with the
    // correctly rounded reciprocal gives the IEEE quotient (Markstein), outside under/overflow
    // range; checked bit for bit against X/dx on 2.2e8 positions, cell centres and cell edges.
    const double ax = fabs(X);
    if (ax < 1e290 && (ax > 1e-290 || X == 0.0)) {
        const double q0 = X * inv_dx;
        return fma(fma(-q0, dx, X), inv_dx, q0);
    }
    return X / dx;
}

__device__ __forceinline__ void cic_axis_fast(double X, double dx, int N, int& lo, int& hi, double& f) {
    const double g = div_dx(X, dx, 1.0 / dx) - 0.5;
    const double fl = floor(g);
    f = g - fl;
    int m;
    if (fabs(fl) < 1e9) {
        m = (int)fl;
        if (m < 0) m += N;
        else if (m >= N) m -= N;
        if ((unsigned)m >= (unsigned)N) {
            m %= N;
            if (m < 0) m += N;
        }
    } else {
        long long t = (long long)fl % N;
        m = (int)(t < 0 ? t + N : t);
    }
    lo = m;
    hi = (m + 1 == N) ? 0 : m + 1;
}

// v in [-N, 2N) wrapped into [0, N) without an integer division.
__device__ __forceinline__ int wrapN(int v, int N) { return v < 0 ? v + N : (v >= N ? v - N : v); }

// Workgroups are dispatched to the 8 XCDs of gfx950 round-robin by linear workgroup id, and each XCD
// has its own L2.  Neighbouring tiles share cache lines (pencil segments shorter than a line, CIC
// corners of adjacent particle planes, the one-site ring of deposit columns); these maps send them
// to one XCD so that they meet in one L2:
//   xcd_block: every XCD runs one contiguous range of logical ids;
//   xcd_map mode 2: runs of c consecutive logical ids are dealt round-robin to the XCDs (needs
//   G % (8c) == 0), so all XCDs stream from neighbouring addresses at any moment.
// Both are bijections of [0, G), so results never depend on them.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G) {
    const unsigned q = G / kXcds, r = G % kXcds;
    const unsigned x = b % kXcds, k = b / kXcds;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}
struct XMap {
    int mode;  // 0 identity, 1 xcd_block, 2 runs of c round-robin over the XCDs
    unsigned c;
};
__device__ __forceinline__ unsigned xcd_map(unsigned b, unsigned G, XMap m) {
    if (m.mode == 0) return b;
    if (m.mode == 1) return xcd_block(b, G);
    const unsigned x = b % kXcds, k = b / kXcds;
    return (k / m.c) * (kXcds * m.c) + x * m.c + (k % m.c);
}

// Exact conversion of an integral double in [0, 2^63) to uint64 with two 32-bit conversions.
__device__ __forceinline__ unsigned long long u64_of_integral(double q) {
    const unsigned int hi = (unsigned int)(q * 0x1p-32);
    const unsigned int lo = (unsigned int)(q - (double)hi * 0x1p32);
    return ((unsigned long long)hi << 32) | lo;
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

// ------------------------------------------------------------------ far-particle row lists
// The column deposit takes a particle from its Lagrangian site when its CIC base cell is the site
// or the site below it on every axis ("near").  The others ("far": 19-42 % of the particles at the
// first shell crossing of a collapsing peak, most 2-10 cells from their sites) are sorted by the
// high part of their cell key, the row r = ix0*N + iy0, by the kernel that moves them (the
// gather-kick-drift of the same step): each is appended to the list of its row, and the column
// deposit reads the lists of the rows its column and low ring cover, so far particles are
// accumulated in LDS like near ones.  A row holds `cap` entries; the rest go to a flat overflow
// list deposited with global atomics.  The row counts are cleared by the far-particle pass of the
// deposit that consumes them, the overflow count by the forward transform that follows it.
struct FarRows {
    unsigned int* cnt;    // [N*N] entries appended per row (may exceed cap); nullptr: no lists
    int* list;            // [N*N][cap] particle indices
    int cap;
    int* ovl;             // overflow entries, flat
    unsigned int* ovcnt;  // their count (appends past ovcap are counted but not stored)
    unsigned int ovcap;   // entries ovl holds
    // the device schedule of a scheduled (graph-replayed) deposit, nullptr otherwise: once its table
    // has ended, the deposit kernels leave the mesh, the lists and their counts as they are, so a
    // replay past the end recomputes the same force meshes (the gather-kick-drift, which fills the
    // lists, does nothing then either)
    const DevSched* sch;
};
__device__ __forceinline__ bool sched_ended(const DevSched* sch) { return sch && sch->counter >= sch->nsteps; }

// Appends particle i to its row's list for the lanes with `far` set.  The lanes of a wave are
// consecutive lattice sites, whose far particles come in runs of lanes with one row: each run takes
// its slots with one atomic (all runs of the wave in one instruction), and a lane's slot is its
// run's base plus its place in the run.  Must be called by every lane of the wave.
__device__ __forceinline__ void far_rows_append(const FarRows& fr, bool far, unsigned row, int i) {
    const int lane = threadIdx.x & 63;
    const unsigned long long fm = __ballot(far);
    if (fm == 0ull) return;
    const unsigned prow = (unsigned)__shfl_up((int)row, 1u, 64);
    const bool head = far && (lane == 0 || !((fm >> ((lane + 63) & 63)) & 1ull) || prow != row);
    const unsigned long long hm = __ballot(head);
    const unsigned long long below = (2ull << lane) - 1ull;  // lanes 0..lane (all 64 for lane 63)
    unsigned base = 0;
    if (head) {  // the run ends at the next head or the next lane that is not far
        const unsigned long long after = (hm | ~fm) & ~below;
        const int end = after ? __ffsll((long long)after) - 1 : 64;
        base = atomicAdd(fr.cnt + row, (unsigned)(end - lane));
    }
    const int hl = far ? 63 - __clzll((long long)(hm & below)) : lane;  // this lane's run head
    const unsigned hbase = (unsigned)__shfl((int)base, hl, 64);
    bool ov = false;
    if (far) {
        const unsigned slot = hbase + (unsigned)(lane - hl);
        if (slot < (unsigned)fr.cap) fr.list[(size_t)row * fr.cap + slot] = i;
        else ov = true;
    }
    const unsigned long long ovm = __ballot(ov);
    if (ovm) {
        const int leader = __ffsll((long long)ovm) - 1;
        unsigned obase = 0;
        if (lane == leader) obase = atomicAdd(fr.ovcnt, (unsigned)__popcll(ovm));
        obase = (unsigned)__shfl((int)obase, leader, 64);
        const unsigned slot = obase + (unsigned)__popcll(ovm & ((1ull << lane) - 1ull));
        if (ov && slot < fr.ovcap) fr.ovl[slot] = i;
    }
}

// The column deposit's "near" test on one axis for an FP64 coordinate (the shared definition of
// the deposit and of the kernel that sorts the far particles): g = X/dx - 0.5 with the
// Markstein-corrected quotient of div_dx, lo = floor(g) mod N, f = g - floor(g); near when the base
// cell is site s or the cell below it (periodic).  |X| >= 1e290 or NaN gives not near.
__device__ __forceinline__ bool near_f64(double X, double dx, double inv_dx, int N, int s, int& lo, double& f) {
    const double q0 = X * inv_dx;
    const double q = fma(fma(-q0, dx, X), inv_dx, q0);
    const double g = q - 0.5;
    const double fl = floor(g);
    f = g - fl;
    const bool inr = fl >= -1.0 && fl <= (double)(N - 1);
    const int l = (int)(inr ? fl : 0.0);
    lo = l < 0 ? l + N : l;
    const int o = lo - s;
    return inr && (o == 0 || o == -1 || o == N - 1);
}
__device__ __forceinline__ bool near_cell(int lo, int N, int s) {
    const int o = lo - s;
    return (unsigned)lo < (unsigned)N && (o == 0 || o == -1 || o == N - 1);
}

// ----------------------------------------------------------------- deterministic fixed point
// Cell accumulators of the deposit.  Integer addition commutes, so every launch order gives the
// same bits.
//   narrow: one uint64 word per cell, value = q * 2^-FB
//   wide  : two uint64 words {lo, hi} per cell, value = (hi*2^64 + lo) * 2^-52; the carry out of lo
//           is detected from the value atomicAdd returns, and the number of carries equals
//           floor(sum/2^64) whatever the order.  Each corner weight is rounded to 2^-52 (the spacing
//           of doubles just below 1), so the FP64 mode stays at FP64 rounding level.
constexpr int kWideFB = 52;

// Round-to-nearest-even of w*2^52 as an integer, for w in [0, 1]: the FMA lands in [2^52, 2^53],
// where doubles are the integers, and the bit pattern minus that of 2^52 is the integer.
__device__ __forceinline__ unsigned long long q52(double w) {
    return (unsigned long long)__double_as_longlong(fma(w, 0x1p52, 0x1p52)) - 0x4330000000000000ull;
}

// FP64 mode on 64-bit cells: 2^-40 of a particle per corner weight.  A cell holds at most the
// total n * (2^40 + 4) (each of a particle's 8 rounded weights is at most 1/2 unit high), below 2^64
// while n <= 2^23; larger FP64 boxes keep the 128-bit cells.
constexpr int kCompactFB = 40;
constexpr long long kCompactMaxN = 1LL << 23;

// Round-to-nearest-even of ws as an integer, for ws in [0, 2^52]: ws + 2^52 lands in [2^52, 2^53],
// where the doubles are the integers (the same rounding as rint).
__device__ __forceinline__ unsigned long long q_scaled(double ws) {
    return (unsigned long long)__double_as_longlong(ws + 0x1p52) - 0x4330000000000000ull;
}
// The same of w * qs (qs a power of two, w * qs <= 2^52), in one FMA.
__device__ __forceinline__ unsigned long long q_fixed(double w, double qs) {
    return (unsigned long long)__double_as_longlong(fma(w, qs, 0x1p52)) - 0x4330000000000000ull;
}

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
        return (double)v.y * 4096.0 + (double)v.x * 0x1p-52;
    } else {
        return (double)mesh[cell] * qscale;
    }
}

}  // namespace pm3d
