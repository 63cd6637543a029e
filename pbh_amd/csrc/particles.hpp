// Particle-side kernels of the pm3d step in FP64 (gfx950): CIC key/fraction, the reference-shaped
// fixed-point atomic deposit, CIC gather, kick+drift, and the Lagrangian-lattice Jacobian.  (The
// column deposit of the step is in deposit.hpp.)
//
// Particle arrays are addressed as x[d*sd + i*si]: SoA device state uses (sd, si) = (n, 1), host
// (n,3) AoS arrays staged for the reference-shaped entry points use (1, 3).
#pragma once
#include "common.hpp"

namespace pm3d {

struct Strided {
    const double* p;
    long long sd, si;
    __device__ __forceinline__ double operator()(int d, long long i) const { return p[d * sd + i * si]; }
};

// ---------------------------------------------------------------------------------------- CIC key
// _cic_index on the three axes (pm3d.py:76-79): key = (ix0*N+iy0)*N+iz0, frac (n,3).
__global__ __launch_bounds__(kThreads) void k_cic_index(Strided X, long long n, int N, double dx,
                                                        long long* __restrict__ key, double* __restrict__ frac) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    int lo[3], hi[3];
    double f[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) cic_axis_fast(X(d, i), dx, N, lo[d], hi[d], f[d]);
    key[i] = ((long long)lo[0] * N + lo[1]) * N + lo[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) frac[i * 3 + d] = f[d];
}

// ------------------------------------------------------------------------------------ CIC weights
// The 8 corner weights ((wx*wy)*wz) in the reference's loop order (pm3d.py:85-88): corner
// c = 4*cx + 2*cy + cz, side 0 weight 1-f, side 1 weight f.
__device__ __forceinline__ void cic_weights(const double f[3], double w[8]) {
    const double wx[2] = {1.0 - f[0], f[0]}, wy[2] = {1.0 - f[1], f[1]}, wz[2] = {1.0 - f[2], f[2]};
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = (wx[c >> 2] * wy[(c >> 1) & 1]) * wz[c & 1];
}

// ------------------------------------------------------------------------------- deposit (atomic)
// Direct scatter: each particle adds its 8 fixed-point weights with 64-bit integer atomics.
// Integer addition commutes, so the mesh is bitwise identical whatever the launch order.
template <bool WIDE>
__global__ __launch_bounds__(kThreads) void k_deposit_atomic(Strided X, long long n, int N, double dx,
                                                             unsigned long long* __restrict__ mesh, double qs,
                                                             long long one) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    int lo[3], hi[3];
    double f[3], w[8];
    long long q[8];
#pragma unroll
    for (int d = 0; d < 3; ++d) cic_axis_fast(X(d, i), dx, N, lo[d], hi[d], f[d]);
    cic_weights(f, w);
    // the grid of the mesh: 2^-52 on 128-bit cells; 1/qs on 64-bit cells (2^-40 in FP64 mode, 2^-31
    // in mixed mode)
#pragma unroll
    for (int c = 0; c < 8; ++c) q[c] = WIDE ? (long long)q52(w[c]) : (long long)q_fixed(w[c], qs);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (q[c] == 0) continue;
        const int ix = (c >> 2) ? hi[0] : lo[0], iy = ((c >> 1) & 1) ? hi[1] : lo[1], iz = (c & 1) ? hi[2] : lo[2];
        fx_atomic_add<WIDE>(mesh, ((size_t)ix * N + iy) * N + iz, (unsigned long long)q[c]);
    }
}

// ------------------------------------------------------------------------------------- gather
// acc_c = sum over corners ((g_c*wx)*wy)*wz in the reference's order (pm3d.py:91-99,113), from the
// component-major gradient mesh grad[c*N^3 + cell], at position X.
__device__ __forceinline__ void gather3_at(const double X[3], int N, double dx, const double* __restrict__ grad,
                                           double a[3]) {
    int lo[3], hi[3];
    double f[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) cic_axis_fast(X[d], dx, N, lo[d], hi[d], f[d]);
    const double wx[2] = {1.0 - f[0], f[0]}, wy[2] = {1.0 - f[1], f[1]}, wz[2] = {1.0 - f[2], f[2]};
    const size_t M = (size_t)N * N * N;
    const double *g0 = grad, *g1 = grad + M, *g2 = grad + 2 * M;  // one component mesh < 4 GiB
    const unsigned uN = (unsigned)N;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ix = (c >> 2) ? hi[0] : lo[0], iy = ((c >> 1) & 1) ? hi[1] : lo[1], iz = (c & 1) ? hi[2] : lo[2];
        const unsigned o = (((unsigned)ix * uN + (unsigned)iy) * uN + (unsigned)iz) * 8u;
        const double u = wx[c >> 2], v = wy[(c >> 1) & 1], w = wz[c & 1];
        a0 = a0 + ((ld_off(g0, o) * u) * v) * w;
        a1 = a1 + ((ld_off(g1, o) * u) * v) * w;
        a2 = a2 + ((ld_off(g2, o) * u) * v) * w;
    }
    a[0] = a0;
    a[1] = a1;
    a[2] = a2;
}

// Gather of every particle into acc (SoA).
__global__ __launch_bounds__(kThreads) void k_gather3(const double* __restrict__ x, long long n, int N, double dx,
                                                      const double* __restrict__ grad, double* __restrict__ acc,
                                                      XMap xm) {
    const long long i = (long long)xcd_map(blockIdx.x, gridDim.x, xm) * kThreads + threadIdx.x;
    if (i >= n) return;
    const unsigned o = (unsigned)i * 8u;  // 3n doubles < 4 GiB
    const double X[3] = {ld_off(x, o), ld_off(x + n, o), ld_off(x + 2 * n, o)};
    double a[3];
    gather3_at(X, N, dx, grad, a);
#pragma unroll
    for (int d = 0; d < 3; ++d) st_off(acc + d * n, o, a[d]);
}

// Scalar-field gather for the interp(F, X) entry point (pm3d.py:91-99).
__global__ __launch_bounds__(kThreads) void k_gather1(Strided X, long long n, int N, double dx,
                                                      const double* __restrict__ F, double* __restrict__ out) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    int lo[3], hi[3];
    double f[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) cic_axis_fast(X(d, i), dx, N, lo[d], hi[d], f[d]);
    const double wx[2] = {1.0 - f[0], f[0]}, wy[2] = {1.0 - f[1], f[1]}, wz[2] = {1.0 - f[2], f[2]};
    double a = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ix = (c >> 2) ? hi[0] : lo[0], iy = ((c >> 1) & 1) ? hi[1] : lo[1], iz = (c & 1) ? hi[2] : lo[2];
        a = a + ((F[((size_t)ix * N + iy) * N + iz] * wx[c >> 2]) * wy[(c >> 1) & 1]) * wz[c & 1];
    }
    out[i] = a;
}

// ---------------------------------------------------------------------------------- kick + drift
// Host scalars of one KDK step (pm3d.py:118-127), with the correctly rounded reciprocals of the
// three divisors: the device forms each quotient with qdiv, the IEEE quotient of the reference.
struct KickScalars {
    double c1p, denp, rdenp;  // pending closing kick of the previous step: p += (c1p*acc)/denp
    double c1, den1, rden1;   // opening kick: p += (c1*acc)/den1
    double da, den_d, rden_d; // drift: x += (da*p)/den_d
};
__device__ __forceinline__ KickScalars kick_from_tab(const DevStep& e) {
    return KickScalars{e.c1p, e.denp, e.rdenp, e.c1, e.den1, e.rden1, e.da, e.den_d, e.rden_d};
}

// Pending closing kick (with `pend`), opening kick and drift of one coordinate, each its own
// rounding as in pm3d.py:122-127; returns the new position, updates p.
__device__ __forceinline__ double kick_drift_1(double xk, double& pd, double ak, int pend, const KickScalars& ks,
                                               double L) {
    if (pend) pd = pd + qdiv(ks.c1p * ak, ks.denp, ks.rdenp);
    pd = pd + qdiv(ks.c1 * ak, ks.den1, ks.rden1);
    return npy_rem_near(xk + qdiv(ks.da * pd, ks.den_d, ks.rden_d), L);
}

// Opening half kick with the cached acceleration, p += (c1*acc)/den1 (pm3d.py:122), then the drift
// x = (x + (da*p)/den_d) % L (pm3d.py:123-124).  Non-finite results are counted in *bad.
// With `pend`, the previous step's deferred closing kick p += (c1p*acc)/denp (pm3d.py:127) is
// applied first, as a separate rounding.
// With a schedule (sch != nullptr) the scalars come from tab[sch->counter] and the pending kick is
// always applied; past the end of the table the step does nothing.
__global__ __launch_bounds__(kThreads) void k_kick_drift(double* __restrict__ x, double* __restrict__ p,
                                                         const double* __restrict__ acc, long long n, int pend,
                                                         KickScalars ks, double L, unsigned long long* __restrict__ bad,
                                                         const DevSched* __restrict__ sch, const DevStep* __restrict__ tab) {
    if (sch) {
        const int k = sch->counter;
        if (k >= sch->nsteps) return;
        ks = kick_from_tab(tab[k]);
        pend = 1;
    }
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    bool ok = true;
    if (i < n) {
        const unsigned o = (unsigned)i * 8u;  // n * 8 < 2^32
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double pd = ld_off(p + d * n, o);
            const double xd = kick_drift_1(ld_off(x + d * n, o), pd, ld_off(acc + d * n, o), pend, ks, L);
            st_off(p + d * n, o, pd);
            st_off(x + d * n, o, xd);
            ok &= isfinite(pd) && isfinite(xd);
        }
    }
    const unsigned long long nb = __ballot(!ok);
    if ((threadIdx.x & 63) == 0 && nb) atomicAdd(bad, (unsigned long long)__popcll(nb));
}

// The gather of the previous step's forces fused with this step's kicks and drift.  A step ends
// with the gradient mesh of the forces at its new positions; the gather that turns it into the
// acceleration is deferred to here, where the same positions are loaded once for the gather and
// the drift: acc = gather3_at(x), then exactly k_kick_drift's arithmetic on it (pending closing
// kick, opening kick, drift, each its own rounding).  The acceleration is never stored.  With row
// lists (fr.cnt != nullptr; particles on the mesh lattice, Np == N) every particle whose new base
// cell is not near its lattice site (the column deposit's test) is appended to its row's list.
__global__ __launch_bounds__(kThreads) void k_gather_kick_drift(double* __restrict__ x, double* __restrict__ p,
                                                                long long n, int N, double dx,
                                                                const double* __restrict__ grad, int pend,
                                                                KickScalars ks, double L,
                                                                unsigned long long* __restrict__ bad,
                                                                const DevSched* __restrict__ sch,
                                                                const DevStep* __restrict__ tab, XMap xm, FarRows fr,
                                                                SiteDiv sdv) {
    if (sch) {
        const int k = sch->counter;
        if (k >= sch->nsteps) return;
        ks = kick_from_tab(tab[k]);
        pend = 1;
    }
    const long long i = (long long)xcd_map(blockIdx.x, gridDim.x, xm) * kThreads + threadIdx.x;
    bool ok = true, far = false;
    unsigned row = 0;
    if (i < n) {
        const unsigned o = (unsigned)i * 8u;  // n * 8 < 2^32
        const double X[3] = {ld_off(x, o), ld_off(x + n, o), ld_off(x + 2 * n, o)};
        double a[3], xn[3];
        gather3_at(X, N, dx, grad, a);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double pd = ld_off(p + d * n, o);
            const double xd = kick_drift_1(X[d], pd, a[d], pend, ks, L);
            st_off(p + d * n, o, pd);
            st_off(x + d * n, o, xd);
            xn[d] = xd;
            ok &= isfinite(pd) && isfinite(xd);
        }
        if (fr.cnt) {
            int s[3];
            site_of((unsigned)i, sdv, s);
            const double inv_dx = 1.0 / dx;
            int lo[3];
            bool nr = true;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double f;
                nr &= near_f64(xn[d], dx, inv_dx, N, s[d], lo[d], f);
            }
            far = !nr;
            row = (unsigned)(lo[0] * N + lo[1]);
        }
    }
    if (fr.cnt) far_rows_append(fr, far, row, (int)i);
    const unsigned long long nb = __ballot(!ok);
    if ((threadIdx.x & 63) == 0 && nb) atomicAdd(bad, (unsigned long long)__popcll(nb));
}

// ----------------------------------------------------------------------------------- utilities
__global__ __launch_bounds__(kThreads) void k_fill_u64(unsigned long long* __restrict__ a, long long n,
                                                       unsigned long long v) {
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        a[i] = v;
}

// rho of a fixed-point mesh; with `as_delta`, delta = rho/n_mean - 1 (pm3d.py:104).
template <bool WIDE>
__global__ __launch_bounds__(kThreads) void k_fixed_to_real(const unsigned long long* __restrict__ q, long long m,
                                                            double qscale, int as_delta, double nmean,
                                                            double* __restrict__ out) {
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < m; i += (long long)gridDim.x * kThreads) {
        const double r = fx_value<WIDE>(q, i, qscale);
        out[i] = as_delta ? r / nmean - 1.0 : r;
    }
}

// (n,3) AoS <-> SoA copies of FP64 particle arrays.
__global__ __launch_bounds__(kThreads) void k_aos_to_soa(const double* __restrict__ a, double* __restrict__ s, long long n) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) s[d * n + i] = a[i * 3 + d];
}
__global__ __launch_bounds__(kThreads) void k_soa_to_aos(const double* __restrict__ s, double* __restrict__ a, long long n) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) a[i * 3 + d] = s[d * n + i];
}

// Closing half kick applied lazily before the host reads p: p += (c1*acc)/den (rden = 1/den).
__global__ __launch_bounds__(kThreads) void k_kick(double* __restrict__ p, const double* __restrict__ acc, long long n3,
                                                   double c1, double den, double rden) {
    const long long k = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (k < n3) p[k] = p[k] + qdiv(c1 * acc[k], den, rden);
}

// ---------------------------------------------------------------------------- shell crossing
// min over Lagrangian lattice cells of det(dx/dq): forward differences to the +x, +y, +z lattice
// neighbours, minimum-image unwrapped, divided by the lattice spacing dq.  Each block stores its
// minimum to bmin[blockIdx.x]; the host takes the minimum of those.
__global__ __launch_bounds__(kThreads) void k_min_jacobian(const double* __restrict__ x, int Np, double L, double dq,
                                                           double* __restrict__ bmin) {
    const long long n = (long long)Np * Np * Np;
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    double det = 1e300;
    if (i < n) {
        const int ii = (int)i, a = ii / (Np * Np), b = (ii / Np) - a * Np, c = ii - (ii / Np) * Np;
        const long long nb[3] = {((long long)(a + 1 == Np ? 0 : a + 1) * Np + b) * Np + c,
                                 ((long long)a * Np + (b + 1 == Np ? 0 : b + 1)) * Np + c,
                                 ((long long)a * Np + b) * Np + (c + 1 == Np ? 0 : c + 1)};
        double J[3][3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double v = x[d * n + nb[e]] - x[d * n + i];
                v = v - L * rint(v / L);
                J[e][d] = v / dq;
            }
        }
        det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    }
    const double m = block_min(det);
    if (threadIdx.x == 0) bmin[blockIdx.x] = m;
}

// v - L*rint(v/L), the minimum image of a difference of two positions in [0, L), without the
// division where the quotient's rounding cannot matter: |v| < lo = 0.49 L gives rint = +-0 (the
// sign of v), lo' = 0.51 L < |v| < hi = 1.49 L gives +-1; the bands in between and beyond take the
// division.  L*(+-0) and L*(+-1) are exact, so the result is the expression's bit for bit.
struct MinImage {
    double L, lo, mid, hi;
};
__device__ __forceinline__ double min_image(double v, const MinImage& m) {
    const double a = fabs(v);
    double k;
    if (a < m.lo) k = 0.0;
    else if (a > m.mid && a < m.hi) k = 1.0;
    else k = fabs(rint(v / m.L));
    return v - m.L * copysign(k, v);
}

// min det(dx/dq) with four consecutive lattice sites of one z row per thread (Np % 4 == 0): the
// positions of the four sites and of their +x and +y neighbours are two 16-byte loads per axis and
// array, the +z neighbour of a site is the next site of the four (only the fourth's is loaded on its
// own).  Entries: min_image, then the quotient by dq as qdiv with the host's reciprocal rdq (the
// IEEE quotient on the CPU checks of tests/test_div_exact.py).  Block b holds the minimum of sites
// [1024 b, 1024 b + 1024).
__global__ __launch_bounds__(kThreads) void k_min_jacobian4(const double* __restrict__ x, unsigned n, MinImage mi,
                                                            double dq, double rdq, SiteDiv sdv,
                                                            double* __restrict__ bmin) {
    const unsigned np = sdv.Np;
    const unsigned i = 4u * (blockIdx.x * kThreads + threadIdx.x);
    long long ob = ordered_bits(1e300);
    if (i < n) {
        int st[3];
        site_of(i, sdv, st);
        const unsigned a = (unsigned)st[0], b = (unsigned)st[1], c = (unsigned)st[2];
        const unsigned ixn = i + ((a + 1 == np ? 0u : a + 1) - a) * np * np;  // +x neighbour of site i
        const unsigned iyn = i + ((b + 1 == np ? 0u : b + 1) - b) * np;       // +y neighbour
        const unsigned izn = i - c + (c + 4 == np ? 0u : c + 4);              // +z neighbour of site i + 3
        double X[3][4], XX[3][4], XY[3][4], xz[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double* xd = x + (size_t)d * n;
            const double2 p0 = ld_off((const double2*)xd, i * 8u), p1 = ld_off((const double2*)xd, i * 8u + 16u);
            const double2 q0 = ld_off((const double2*)xd, ixn * 8u), q1 = ld_off((const double2*)xd, ixn * 8u + 16u);
            const double2 r0 = ld_off((const double2*)xd, iyn * 8u), r1 = ld_off((const double2*)xd, iyn * 8u + 16u);
            X[d][0] = p0.x, X[d][1] = p0.y, X[d][2] = p1.x, X[d][3] = p1.y;
            XX[d][0] = q0.x, XX[d][1] = q0.y, XX[d][2] = q1.x, XX[d][3] = q1.y;
            XY[d][0] = r0.x, XY[d][1] = r0.y, XY[d][2] = r1.x, XY[d][3] = r1.y;
            xz[d] = ld_off(xd, izn * 8u);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double J[3][3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double xi = X[d][k];
                const double nb[3] = {XX[d][k], XY[d][k], k < 3 ? X[d][k + 1 < 4 ? k + 1 : 3] : xz[d]};
#pragma unroll
                for (int e = 0; e < 3; ++e) J[e][d] = qdiv(min_image(nb[e] - xi, mi), dq, rdq);
            }
            const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            const long long o = ordered_bits(det);
            ob = (k == 0 || o < ob) ? o : ob;
        }
    }
    const double m = block_min(from_ordered_bits(ob));
    if (threadIdx.x == 0) bmin[blockIdx.x] = m;
}

}  // namespace pm3d
