// Mixed-precision particle kernels (gfx950).  A position is stored per axis as the CIC base cell
// and fraction of pm3d.py:76-79 directly:
//     x = (cell + 1/2 + off) * dx,   cell int32 in [0, N),   off float in [0, 1)
// so the deposit and the gather need no division and the weights are (1-off, off).  Momenta and
// accelerations are FP32; the scalars of the step arrive from the host in FP64 and are folded into
// FP32 factors once per launch.  The deposit mesh is one uint64 per cell at 2^-31: the whole box
// holds at most n * 2^31 < 2^63 units for n < 2^32, so it cannot overflow.
#pragma once
#include "common.hpp"

namespace pm3d {

constexpr int kNarrowFB = 31;

// Narrow fixed point: a weight w in [0, 1] as round(w * 2^31), one FP32 -> u32 conversion.
__device__ __forceinline__ unsigned int q31(float w) { return (unsigned int)rintf(w * 2147483648.0f); }

__device__ __forceinline__ void mx_weights(const float f[3], float w[8]) {
    const float wx[2] = {1.0f - f[0], f[0]}, wy[2] = {1.0f - f[1], f[1]}, wz[2] = {1.0f - f[2], f[2]};
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = (wx[c >> 2] * wy[(c >> 1) & 1]) * wz[c & 1];
}

// The 8 corner weights, each rounded to nearest at 2^-31 independently.
__device__ __forceinline__ void mx_quantise8(const float f[3], const float w[8], unsigned long long q[8]) {
    (void)f;
#pragma unroll
    for (int c = 0; c < 8; ++c) q[c] = q31(w[c]);
}

struct MxPos {
    const int* cell;
    const float* off;
    long long n;
    __device__ __forceinline__ void load(long long i, int lo[3], float f[3]) const {
        const unsigned o = (unsigned)i * 4u;  // n * 4 < 2^32
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = ld_off(cell + d * n, o);
            f[d] = ld_off(off + d * n, o);
        }
    }
};

__device__ __forceinline__ void mx_scatter(const int lo[3], const float f[3], int N, unsigned long long* mesh) {
    float w[8];
    unsigned long long q[8];
    mx_weights(f, w);
    mx_quantise8(f, w, q);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (q[c] == 0) continue;
        const int ix = (c >> 2) ? (lo[0] + 1 == N ? 0 : lo[0] + 1) : lo[0];
        const int iy = ((c >> 1) & 1) ? (lo[1] + 1 == N ? 0 : lo[1] + 1) : lo[1];
        const int iz = (c & 1) ? (lo[2] + 1 == N ? 0 : lo[2] + 1) : lo[2];
        atomicAdd(mesh + ((size_t)ix * N + iy) * N + iz, q[c]);
    }
}

__global__ __launch_bounds__(kThreads) void k_deposit_atomic_mx(MxPos P, int N, unsigned long long* __restrict__ mesh) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= P.n) return;
    int lo[3];
    float f[3];
    P.load(i, lo, f);
    mx_scatter(lo, f, N, mesh);
}

// CIC gather of the component-major FP32 gradient mesh grad[c*N^3 + cell] at base cell lo, fraction f.
__device__ __forceinline__ void gather3_mx_at(const int lo[3], const float f[3], int N, const float* __restrict__ grad,
                                              float a[3]) {
    const int hx = lo[0] + 1 == N ? 0 : lo[0] + 1, hy = lo[1] + 1 == N ? 0 : lo[1] + 1, hz = lo[2] + 1 == N ? 0 : lo[2] + 1;
    const float wx[2] = {1.0f - f[0], f[0]}, wy[2] = {1.0f - f[1], f[1]}, wz[2] = {1.0f - f[2], f[2]};
    const size_t M = (size_t)N * N * N;
    const float *g0 = grad, *g1 = grad + M, *g2 = grad + 2 * M;
    const unsigned uN = (unsigned)N;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ix = (c >> 2) ? hx : lo[0], iy = ((c >> 1) & 1) ? hy : lo[1], iz = (c & 1) ? hz : lo[2];
        const unsigned o = (((unsigned)ix * uN + (unsigned)iy) * uN + (unsigned)iz) * 4u;
        const float w = (wx[c >> 2] * wy[(c >> 1) & 1]) * wz[c & 1];
        a0 = fmaf(ld_off(g0, o), w, a0);
        a1 = fmaf(ld_off(g1, o), w, a1);
        a2 = fmaf(ld_off(g2, o), w, a2);
    }
    a[0] = a0;
    a[1] = a1;
    a[2] = a2;
}

// Gather of every particle into acc (SoA).
__device__ __forceinline__ void gather3_mx_body(MxPos P, int N, const float* __restrict__ grad,
                                                float* __restrict__ acc, XMap xm) {
    const long long i = (long long)xcd_map(blockIdx.x, gridDim.x, xm) * kThreads + threadIdx.x;
    if (i >= P.n) return;
    int lo[3];
    float f[3], a[3];
    P.load(i, lo, f);
    gather3_mx_at(lo, f, N, grad, a);
#pragma unroll
    for (int d = 0; d < 3; ++d) st_off(acc + d * P.n, (unsigned)i * 4u, a[d]);
}

__global__ __launch_bounds__(kThreads) void k_gather3_mx(MxPos P, int N, const float* __restrict__ grad,
                                                         float* __restrict__ acc, XMap xm) {
    gather3_mx_body(P, N, grad, acc, xm);
}

// Pending closing kick (kp = c1p/denp), opening kick (k1 = c1/den1), drift by p*sd cells
// (sd = da/(den_d*dx)) with the carry of the offset into the cell index, periodic in [0, N).
// The cell array is written only where it changes.
__device__ __forceinline__ void kick_drift_mx_body(int* __restrict__ cell, float* __restrict__ off,
                                                   float* __restrict__ p, const float* __restrict__ acc,
                                                   long long n, int N, int pend, float kp, float k1, float sd,
                                                   unsigned long long* __restrict__ bad,
                                                   const DevSched* __restrict__ sch,
                                                   const DevStep* __restrict__ tab) {
    if (sch) {  // scalars from the device schedule (see k_kick_drift)
        const int k = sch->counter;
        if (k >= sch->nsteps) return;
        pend = 1;
        kp = tab[k].kp;
        k1 = tab[k].k1;
        sd = tab[k].sd;
    }
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    bool ok = true;
    if (i < n) {
        const unsigned o = (unsigned)i * 4u;  // n * 4 < 2^32
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float a = ld_off(acc + d * n, o);
            float pd = ld_off(p + d * n, o);
            if (pend) pd = fmaf(kp, a, pd);
            pd = fmaf(k1, a, pd);
            float f = ld_off(off + d * n, o) + pd * sd;
            float fl = floorf(f);
            f -= fl;
            if (f >= 1.0f) {  // a tiny negative f rounds up to 1 after the subtraction: next cell, offset 0
                f = 0.0f;
                fl += 1.0f;
            }
            st_off(p + d * n, o, pd);
            st_off(off + d * n, o, f);
            ok &= isfinite(pd) && isfinite(f) && fabsf(fl) < 1e6f;
            if (fl != 0.0f && ok) st_off(cell + d * n, o, wrap_cell(ld_off(cell + d * n, o) + (int)fl, N));
        }
    }
    const unsigned long long nb = __ballot(!ok);
    if ((threadIdx.x & 63) == 0 && nb) atomicAdd(bad, (unsigned long long)__popcll(nb));
}

__global__ __launch_bounds__(kThreads) void k_kick_drift_mx(int* __restrict__ cell, float* __restrict__ off,
                                                            float* __restrict__ p, const float* __restrict__ acc,
                                                            long long n, int N, int pend, float kp, float k1, float sd,
                                                            unsigned long long* __restrict__ bad,
                                                            const DevSched* __restrict__ sch,
                                                            const DevStep* __restrict__ tab) {
    kick_drift_mx_body(cell, off, p, acc, n, N, pend, kp, k1, sd, bad, sch, tab);
}

// The gather of the previous step's forces fused with this step's kicks and drift (the mixed form
// of k_gather_kick_drift): the cell and offset are loaded once, acc = gather3_mx_at, then exactly
// kick_drift_mx_body's arithmetic.  The acceleration is never stored.
// With row lists (fr.cnt != nullptr; Np == N) a particle whose new cell is not near its lattice site
// is appended to its row's list (see k_gather_kick_drift).
__device__ __forceinline__ void gather_kick_drift_mx_body(int* __restrict__ cell, float* __restrict__ off,
                                                          float* __restrict__ p, long long n, int N,
                                                          const float* __restrict__ grad, int pend, float kp, float k1,
                                                          float sd, unsigned long long* __restrict__ bad,
                                                          const DevSched* __restrict__ sch,
                                                          const DevStep* __restrict__ tab, XMap xm, FarRows fr,
                                                          SiteDiv sdv) {
    if (sch) {
        const int k = sch->counter;
        if (k >= sch->nsteps) return;
        pend = 1;
        kp = tab[k].kp;
        k1 = tab[k].k1;
        sd = tab[k].sd;
    }
    const long long i = (long long)xcd_map(blockIdx.x, gridDim.x, xm) * kThreads + threadIdx.x;
    bool ok = true, far = false;
    unsigned row = 0;
    if (i < n) {
        const unsigned o = (unsigned)i * 4u;  // n * 4 < 2^32
        int lo[3];
        float f[3], a[3];
        MxPos{cell, off, n}.load(i, lo, f);
        gather3_mx_at(lo, f, N, grad, a);
        int nc[3] = {lo[0], lo[1], lo[2]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float* pk = (float*)((char*)(p + d * n) + o);
            // p and the new offsets are streamed with non-temporal loads/stores: they are not read
            // again before the next step, and leaving them out of the caches keeps more of the
            // positions there for the deposit that follows (mixed 256^3: deposit 152 -> 142 us).
            float pd = __builtin_nontemporal_load(pk);
            if (pend) pd = fmaf(kp, a[d], pd);
            pd = fmaf(k1, a[d], pd);
            float fo = f[d] + pd * sd;
            float fl = floorf(fo);
            fo -= fl;
            if (fo >= 1.0f) {  // a tiny negative offset rounds up to 1 after the subtraction: next cell, offset 0
                fo = 0.0f;
                fl += 1.0f;
            }
            __builtin_nontemporal_store(pd, pk);
            __builtin_nontemporal_store(fo, (float*)((char*)(off + d * n) + o));
            ok &= isfinite(pd) && isfinite(fo) && fabsf(fl) < 1e6f;
            if (fl != 0.0f && ok) {
                const int c = wrap_cell(lo[d] + (int)fl, N);
                st_off(cell + d * n, o, c);
                nc[d] = c;
            }
        }
        if (fr.cnt) {
            int st[3];
            site_of((unsigned)i, sdv, st);
            far = !(near_cell(nc[0], N, st[0]) && near_cell(nc[1], N, st[1]) && near_cell(nc[2], N, st[2]));
            const bool inr = (unsigned)nc[0] < (unsigned)N && (unsigned)nc[1] < (unsigned)N;
            row = inr ? (unsigned)(nc[0] * N + nc[1]) : 0u;
        }
    }
    if (fr.cnt) far_rows_append(fr, far, row, (int)i);
    const unsigned long long nb = __ballot(!ok);
    if ((threadIdx.x & 63) == 0 && nb) atomicAdd(bad, (unsigned long long)__popcll(nb));
}

__global__ __launch_bounds__(kThreads) void k_gather_kick_drift_mx(int* __restrict__ cell, float* __restrict__ off,
                                                                   float* __restrict__ p, long long n, int N,
                                                                   const float* __restrict__ grad, int pend, float kp,
                                                                   float k1, float sd,
                                                                   unsigned long long* __restrict__ bad,
                                                                   const DevSched* __restrict__ sch,
                                                                   const DevStep* __restrict__ tab, XMap xm, FarRows fr,
                                                                   SiteDiv sdv) {
    gather_kick_drift_mx_body(cell, off, p, n, N, grad, pend, kp, k1, sd, bad, sch, tab, xm, fr, sdv);
}

__global__ __launch_bounds__(kThreads) void k_kick_mx(float* __restrict__ p, const float* __restrict__ acc, long long n3,
                                                      float kp) {
    const long long k = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (k < n3) p[k] = fmaf(kp, acc[k], p[k]);
}

// Host FP64 AoS (staged on the device) -> mixed state, and back.
__global__ __launch_bounds__(kThreads) void k_set_pos_mx(const double* __restrict__ xa, int* __restrict__ cell,
                                                         float* __restrict__ off, long long n, int N, double dx) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int lo, hi;
        double f;
        cic_axis(xa[i * 3 + d], dx, N, lo, hi, f);
        float ff = (float)f;
        if (ff >= 1.0f) {
            ff = 0.0f;
            lo = hi;
        }
        cell[d * n + i] = lo;
        off[d * n + i] = ff;
    }
}
__global__ __launch_bounds__(kThreads) void k_get_pos_mx(const int* __restrict__ cell, const float* __restrict__ off,
                                                         double* __restrict__ xa, long long n, double dx, double L) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double v = ((double)cell[d * n + i] + 0.5 + (double)off[d * n + i]) * dx;
        if (v >= L) v -= L;
        xa[i * 3 + d] = v;
    }
}
__global__ __launch_bounds__(kThreads) void k_f64_to_f32_soa(const double* __restrict__ a, float* __restrict__ s, long long n) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d * n + i] = (float)a[i * 3 + d];
}
__global__ __launch_bounds__(kThreads) void k_f32_soa_to_f64(const float* __restrict__ s, double* __restrict__ a, long long n) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) a[i * 3 + d] = (double)s[d * n + i];
}

// Last kernel of a scheduled step: reduce the Jacobian kernel's per-block minima into
// jhist[counter] and advance the schedule.  A minimum <= 0 (the first shell crossing) ends the
// table there (nsteps = counter), so the member's later steps do nothing; so does a non-zero count
// of non-finite particles (`bad`, the handle's sticky counter), so a broken member stops at once.
__device__ __forceinline__ void sched_advance_body(DevSched* __restrict__ sch, const double* __restrict__ bmin,
                                                   int nb, double* __restrict__ jhist,
                                                   const unsigned long long* __restrict__ bad) {
    double m = 1e300;
    for (int t = threadIdx.x; t < nb; t += kThreads) m = fmin(m, bmin[t]);
    m = block_min(m);
    if (threadIdx.x == 0) {
        const int k = sch->counter;
        if (k < sch->nsteps) {
            jhist[k] = m;
            sch->counter = k + 1;
            if (m <= 0.0 || *bad != 0ull) sch->nsteps = k + 1;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_sched_advance(DevSched* __restrict__ sch, const double* __restrict__ bmin,
                                                            int nb, double* __restrict__ jhist,
                                                            const unsigned long long* __restrict__ bad) {
    sched_advance_body(sch, bmin, nb, jhist, bad);
}

// Minimum of nb per-block minima into *out (one block), so a host read of min det(dx/dq) copies 8 bytes.
__global__ __launch_bounds__(kThreads) void k_min_reduce(const double* __restrict__ bmin, int nb, double* __restrict__ out) {
    long long o = ordered_bits(1e300);
    for (int t = threadIdx.x; t < nb; t += kThreads) {
        const long long v = ordered_bits(bmin[t]);
        o = (t == (int)threadIdx.x || v < o) ? v : o;
    }
    const double m = block_min(from_ordered_bits(o));
    if (threadIdx.x == 0) *out = m;
}

// min det(dx/dq) over the Lagrangian lattice from cell + offset: the forward differences in cells
// (exact integer cell differences, minimum image, plus the offset differences) form J*c with c the
// cells per lattice site, so det(J) = det(J*c) * inv_c3, inv_c3 = 1/c^3 (one multiply instead of nine
// divisions).  32-bit indices (3n < 2^31).
__device__ __forceinline__ void min_jacobian_mx_body(MxPos P, int Np, int N, double inv_c3,
                                                     double* __restrict__ bmin) {
    const unsigned n = (unsigned)P.n;
    const unsigned i = blockIdx.x * kThreads + threadIdx.x;
    double det = 1e300;
    if (i < n) {
        const unsigned np = (unsigned)Np, a = i / (np * np), r = i - a * (np * np), b = r / np, c = r - b * np;
        const unsigned nb[3] = {((a + 1 == np ? 0u : a + 1) * np + b) * np + c, (a * np + (b + 1 == np ? 0u : b + 1)) * np + c,
                                (a * np + b) * np + (c + 1 == np ? 0u : c + 1)};
        int ci[3];
        float fi[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ci[d] = P.cell[d * n + i];
            fi[d] = P.off[d * n + i];
        }
        double J[3][3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                int dc = P.cell[d * n + nb[e]] - ci[d];
                if (dc > N / 2) dc -= N;
                if (dc < -N / 2) dc += N;
                J[e][d] = (double)dc + ((double)P.off[d * n + nb[e]] - (double)fi[d]);
            }
        }
        det = (J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) * inv_c3;
    }
    const double m = block_min(det);
    if (threadIdx.x == 0) bmin[blockIdx.x] = m;
}

__global__ __launch_bounds__(kThreads) void k_min_jacobian_mx(MxPos P, int Np, int N, double inv_c3,
                                                              double* __restrict__ bmin) {
    min_jacobian_mx_body(P, Np, N, inv_c3, bmin);
}

// The determinant of one site's forward differences (the arithmetic of min_jacobian_mx_body).
// POW2 (N = 2^k, k >= 1, sh = 32 - k): the minimum image of a cell difference dc in (-N, N) is the
// sign-extended low k bits of dc, which lie in [-N/2, N/2), except dc = +N/2, which the compare form
// keeps as +N/2: one bit-field extract, one compare and one select instead of two compares, two adds
// and two selects, and the same integer for every dc in (-N, N) (cells of the state are in [0, N)).
// Compiled, the four-site loop takes 412 VALU instructions instead of 522.
template <bool POW2>
__device__ __forceinline__ double det_mx(const int ci[3], const float fi[3], const int cn[3][3], const float fn[3][3],
                                         int N, int sh, double inv_c3) {
    double J[3][3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            int dc = cn[e][d] - ci[d];
            if constexpr (POW2) {
                const int w = __builtin_amdgcn_sbfe(dc, 0u, 32u - (unsigned)sh);  // v_bfe_i32
                dc = dc == N / 2 ? dc : w;
            } else {
                if (dc > N / 2) dc -= N;
                if (dc < -N / 2) dc += N;
            }
            J[e][d] = (double)dc + ((double)fn[e][d] - (double)fi[d]);
        }
    }
    return (J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) * inv_c3;
}

// The same minimum with four consecutive lattice sites (one z row, Np % 4 == 0) per thread: the
// cell and offset words of the four sites and of their +x and +y neighbours are one 16-byte load
// per array (18 loads for four sites instead of 96 4-byte loads), the +z neighbour of a site is
// the next site of the four, and only the fourth one's is loaded on its own.  Each site's
// determinant is the arithmetic of min_jacobian_mx_body; the per-thread minimum is taken on the
// ordered bits, like block_min.  Block b holds the minimum of sites [1024 b, 1024 b + 1024).
template <bool POW2>
__device__ __forceinline__ void min_jacobian_mx4_sites(MxPos P, int N, double inv_c3, SiteDiv sdv,
                                                       double* __restrict__ bmin) {
    const int sh = POW2 ? 32 - __builtin_ctz((unsigned)N) : 0;
    const unsigned n = (unsigned)P.n, np = sdv.Np;
    const unsigned i = 4u * (blockIdx.x * kThreads + threadIdx.x);
    long long ob = ordered_bits(1e300);
    if (i < n) {
        int st[3];
        site_of(i, sdv, st);
        const unsigned a = (unsigned)st[0], b = (unsigned)st[1], c = (unsigned)st[2];
        const unsigned ixn = i + ((a + 1 == np ? 0u : a + 1) - a) * np * np;  // +x neighbour of site i
        const unsigned iyn = i + ((b + 1 == np ? 0u : b + 1) - b) * np;       // +y neighbour
        const unsigned izn = i - c + (c + 4 == np ? 0u : c + 4);              // +z neighbour of site i + 3
        int4 C[3], CX[3], CY[3];
        float4 F[3], FX[3], FY[3];
        int cz[3];
        float fz[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int4* cb = (const int4*)(P.cell + (size_t)d * n);
            const float4* fb = (const float4*)(P.off + (size_t)d * n);
            C[d] = ld_off(cb, i * 4u);
            F[d] = ld_off(fb, i * 4u);
            CX[d] = ld_off(cb, ixn * 4u);
            FX[d] = ld_off(fb, ixn * 4u);
            CY[d] = ld_off(cb, iyn * 4u);
            FY[d] = ld_off(fb, iyn * 4u);
            cz[d] = ld_off(P.cell + (size_t)d * n, izn * 4u);
            fz[d] = ld_off(P.off + (size_t)d * n, izn * 4u);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int ci[3], cn[3][3];
            float fi[3], fn[3][3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int cv[4] = {C[d].x, C[d].y, C[d].z, C[d].w}, cxv[4] = {CX[d].x, CX[d].y, CX[d].z, CX[d].w},
                          cyv[4] = {CY[d].x, CY[d].y, CY[d].z, CY[d].w};
                const float fv[4] = {F[d].x, F[d].y, F[d].z, F[d].w}, fxv[4] = {FX[d].x, FX[d].y, FX[d].z, FX[d].w},
                            fyv[4] = {FY[d].x, FY[d].y, FY[d].z, FY[d].w};
                ci[d] = cv[k];
                fi[d] = fv[k];
                cn[0][d] = cxv[k];
                fn[0][d] = fxv[k];
                cn[1][d] = cyv[k];
                fn[1][d] = fyv[k];
                cn[2][d] = k < 3 ? cv[k + 1 < 4 ? k + 1 : 3] : cz[d];
                fn[2][d] = k < 3 ? fv[k + 1 < 4 ? k + 1 : 3] : fz[d];
            }
            const long long o = ordered_bits(det_mx<POW2>(ci, fi, cn, fn, N, sh, inv_c3));
            ob = (k == 0 || o < ob) ? o : ob;
        }
    }
    const double m = block_min(from_ordered_bits(ob));
    if (threadIdx.x == 0) bmin[blockIdx.x] = m;
}

// N is uniform over the launch (over a box's blocks in the batch): one branch picks the form.
__device__ __forceinline__ void min_jacobian_mx4_body(MxPos P, int N, double inv_c3, SiteDiv sdv,
                                                      double* __restrict__ bmin) {
    if (N >= 2 && (N & (N - 1)) == 0) min_jacobian_mx4_sites<true>(P, N, inv_c3, sdv, bmin);
    else min_jacobian_mx4_sites<false>(P, N, inv_c3, sdv, bmin);
}

__global__ __launch_bounds__(kThreads) void k_min_jacobian_mx4(MxPos P, int N, double inv_c3, SiteDiv sdv,
                                                               double* __restrict__ bmin) {
    min_jacobian_mx4_body(P, N, inv_c3, sdv, bmin);
}

}  // namespace pm3d
