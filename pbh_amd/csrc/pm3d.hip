// pm3d particle-mesh library for MI355X (gfx950): C-ABI entry points of include/pm3d_hip.h.
//
// One handle = one periodic box stepped on one HIP stream.  Device state is SoA in Lagrangian
// particle order (particle i is lattice site i, pm3d.py:64-65), so the x/y/z components of the
// positions, momenta and cached accelerations are each n contiguous doubles.
//
// A step (pm3d.py:117-129) is 8 launches:
//   kick+drift -> fill(mesh) -> deposit -> r2c_z (int64 -> delta fused) -> y fwd
//   -> x (fwd FFT, Green's function, -ik, 3 inverse FFTs) -> y inv (3 comps) -> c2r_z (3 comps)
//   -> gather (+ closing kick)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pm3d_hip.h"
#include "common.hpp"
#include "fft.hpp"
#include "particles.hpp"

using namespace pm3d;

namespace {

std::string g_create_error;
constexpr int kRing = 4096;  // deposit maxima kept on the device for lazy delta.max()

enum KernelClass {
    KC_KICK_DRIFT = 0,
    KC_FILL,
    KC_DEPOSIT,
    KC_DEPOSIT_FAR,
    KC_R2C_Z,
    KC_Y_FWD,
    KC_X_POISSON,
    KC_Y_INV,
    KC_C2R_Z,
    KC_GATHER,
    KC_OTHER,
    KC_COUNT
};
const char* kKernelNames =
    "kick_drift,fill,deposit,deposit_far,r2c_z,y_fwd,x_poisson,y_inv,c2r_z,gather,other";

struct Handle {
    pm3d_cfg cfg{};
    int N = 0, Np = 0, log2N = 0;
    long long n = 0, M = 0;
    double L = 0, dx = 0, qscale = 0, qs = 0;
    long long qone = 0;
    hipStream_t st = nullptr;
    // particle state (FP64 SoA)
    double *x = nullptr, *p = nullptr, *acc = nullptr;
    // meshes
    unsigned long long* mesh = nullptr;  // fixed-point deposit: FP64 mode {lo,hi} per cell (2^-60)
    bool wide = true;
    cplx<double>* spec = nullptr;        // [N][N][N/2] + nyq [N][N]
    cplx<double>* gspec = nullptr;       // 3 x the same
    double* grad = nullptr;              // [M][3]
    double* delta = nullptr;             // [M] (written on demand)
    cplx<double>* tw = nullptr;          // exp(-2 pi i k/N), k < N
    // device scalars: [0] nfar, [1] bad count (sticky until pm3d_synchronize), [2] jmin bits,
    // [8 + s % kRing] max fixed-point cell value of deposit serial s
    unsigned long long* dscal = nullptr;
    long long serial = 0;                // number of deposits of the state so far
    int* far = nullptr;                  // far-particle list of the tiled deposit
    double* bmax = nullptr;              // per-block maxima of the r2c pass
    // staging for reference-shaped entry points
    double* sbuf = nullptr;
    size_t sbuf_bytes = 0;
    bool state_set = false;
    bool acc_valid = false;
    // deferred closing kick (pm3d.py:127) -- applied before the next kick or a read of p
    bool kick_pending = false;
    double pend_c1 = 0, pend_den = 0;
    int deposit_mode = 1;  // 0 atomic, 1 tiled
    // timing
    bool timing = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
    std::vector<hipEvent_t> ev_pool;
    long long t_count[KC_COUNT] = {};
    double t_ms[KC_COUNT] = {};
    std::string err;
    size_t bytes = 0;
};

#define HCHK(h, call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);                     \
            return PM3D_EHIP;                                                                 \
        }                                                                                     \
    } while (0)

int fail(Handle* h, int code, const std::string& msg) {
    h->err = msg;
    return code;
}

hipEvent_t take_event(Handle* h) {
    if (h->ev_pool.empty()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    hipEvent_t e = h->ev_pool.back();
    h->ev_pool.pop_back();
    return e;
}

// Brackets one launch with events when timing is on.
struct Timed {
    Handle* h;
    int kc;
    hipEvent_t e1 = nullptr;
    Timed(Handle* h_, int kc_) : h(h_), kc(kc_) {
        if (h->timing) {
            hipEvent_t e0 = take_event(h);
            e1 = take_event(h);
            if (e0 && e1) {
                hipEventRecord(e0, h->st);
                h->ev_used.push_back({kc, {e0, e1}});
            }
        }
    }
    ~Timed() {
        if (h->timing && e1) hipEventRecord(e1, h->st);
    }
};

inline unsigned grid_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline unsigned long long* qmax_slot(Handle* h) { return h->dscal + 8 + (h->serial % kRing); }

// ---------------------------------------------------------------- dispatch over power-of-two N
template <int N>
int launch_fft_forward_tmpl(Handle* h, const void* in, bool fixed, double* delta_out, double nmean) {
    using R = double;
    const size_t NH = N / 2, specN = (size_t)N * N * NH;
    cplx<R>* spec = h->spec;
    cplx<R>* nyq = h->spec + specN;
    {
        Timed t(h, KC_R2C_Z);
        const unsigned g = (unsigned)((size_t)N * N / Plan<R, N>::PZ);
        if (fixed)
            k_r2c_z<R, N, true, true><<<g, kThreads, 0, h->st>>>(in, spec, nyq, h->tw, delta_out, h->bmax, h->dscal,
                                                                 h->qscale, nmean);
        else
            k_r2c_z<R, N, false, true><<<g, kThreads, 0, h->st>>>(in, spec, nyq, h->tw, nullptr, nullptr, nullptr, 0.0,
                                                                  1.0);
    }
    {
        Timed t(h, KC_Y_FWD);
        constexpr int P = Plan<R, N>::PY;
        dim3 g(N * (N / 2 / P) + N / P, 1);
        const int nb = fixed ? (int)((size_t)N * N / Plan<R, N>::PZ) : 0;
        k_c2c_y<R, N, false><<<g, kThreads, 0, h->st>>>(spec, nyq, 0, 0, h->tw, h->bmax, nb, qmax_slot(h));
    }
    return 0;
}

template <int N>
int launch_x_tmpl(Handle* h, int mode, double s, double sm2) {
    using R = double;
    const size_t NH = N / 2, specN = (size_t)N * N * NH, tot = specN + (size_t)N * N;
    constexpr int P = Plan<R, N>::PX;
    const unsigned g = N * (N / 2 / P) + N / P;
    const double kval = 1.0 / ((double)N * h->dx);
    Timed t(h, KC_X_POISSON);
    if (mode == 0)
        k_x_pass<R, N, 0><<<g, kThreads, 0, h->st>>>(h->spec, h->spec + specN, nullptr, nullptr, 0, 0, h->tw, kval, 0, 0, 0);
    else if (mode == 1)
        k_x_pass<R, N, 1><<<g, kThreads, 0, h->st>>>(h->spec, h->spec + specN, nullptr, nullptr, 0, 0, h->tw, kval, 0, 0, 0);
    else
        k_x_pass<R, N, 2><<<g, kThreads, 0, h->st>>>(h->spec, h->spec + specN, h->gspec, h->gspec + specN, tot, tot, h->tw,
                                                     kval, s, sm2, sm2 > 0.0 ? 1 : 0);
    return 0;
}

// Inverse y (ncomp spectra starting at base) and c2r z into out (ncomp interleaved).
template <int N>
int launch_inverse_tmpl(Handle* h, cplx<double>* base, int ncomp, double* out) {
    using R = double;
    const size_t NH = N / 2, specN = (size_t)N * N * NH, tot = specN + (size_t)N * N;
    {
        Timed t(h, KC_Y_INV);
        constexpr int P = Plan<R, N>::PY;
        dim3 g(N * (N / 2 / P) + N / P, ncomp);
        k_c2c_y<R, N, true><<<g, kThreads, 0, h->st>>>(base, base + specN, tot, tot, h->tw, nullptr, 0, nullptr);
    }
    {
        Timed t(h, KC_C2R_Z);
        const unsigned g = (unsigned)((size_t)N * N / Plan<R, N>::PC);
        const double scale = 1.0 / ((double)N * N * N);
        if (ncomp == 3)
            k_c2r_z<R, N, 3><<<g, kThreads, 0, h->st>>>(base, base + specN, tot, tot, out, h->tw, scale);
        else
            k_c2r_z<R, N, 1><<<g, kThreads, 0, h->st>>>(base, base + specN, tot, tot, out, h->tw, scale);
    }
    return 0;
}

#define PM3D_DISPATCH(h, FN, ...)                    \
    switch ((h)->log2N) {                            \
        case 3: FN<8>(h, __VA_ARGS__); break;        \
        case 4: FN<16>(h, __VA_ARGS__); break;       \
        case 5: FN<32>(h, __VA_ARGS__); break;       \
        case 6: FN<64>(h, __VA_ARGS__); break;       \
        case 7: FN<128>(h, __VA_ARGS__); break;      \
        case 8: FN<256>(h, __VA_ARGS__); break;      \
        case 9: FN<512>(h, __VA_ARGS__); break;      \
        default: return fail(h, PM3D_EUNSUPPORTED, "N must be a power of two in [8,512]"); \
    }

int check_launch(Handle* h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(h, PM3D_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

int ensure_sbuf(Handle* h, size_t bytes) {
    if (h->sbuf_bytes >= bytes) return 0;
    HCHK(h, hipStreamSynchronize(h->st));  // queued work may still read the old buffer
    if (h->sbuf) HCHK(h, hipFree(h->sbuf));
    h->sbuf = nullptr;
    h->sbuf_bytes = 0;
    HCHK(h, hipMalloc(&h->sbuf, bytes));
    h->sbuf_bytes = bytes;
    return 0;
}

// Deposit of the device SoA state into h->mesh (int64 fixed point).  The tiled path writes every
// cell, so only the atomic path clears the mesh first.
int deposit_state(Handle* h) {
    // bricks of 16^3 cells with a one-site halo: the halo must not wrap onto the brick itself
    const bool tiled = h->deposit_mode == 1 && h->Np == h->N && h->N >= 32;
    // nfar is cleared by the r2c pass of the previous deposit (and at creation); the maximum of
    // this deposit is stored (not accumulated) into its ring slot by the y pass
    h->serial += 1;
    if (!tiled) {
        Timed t(h, KC_FILL);
        k_fill_u64<<<std::min<long long>(grid_for(2 * h->M), 4096), kThreads, 0, h->st>>>(h->mesh, 2 * h->M, 0ull);
    }
    if (tiled) {
        {
            Timed t(h, KC_DEPOSIT);
            const int B = 16;
            const unsigned g = (unsigned)((h->N / B) * (h->N / B) * (h->N / B));
            k_deposit_brick<16, 16, 16, 1, true, 1024><<<g, 1024, 0, h->st>>>(h->x, h->n, h->N, h->dx, h->mesh, h->qs,
                                                                               h->qone, h->far, (unsigned int*)(h->dscal + 0));
        }
        {
            Timed t(h, KC_DEPOSIT_FAR);
            k_deposit_far<true><<<1024, kThreads, 0, h->st>>>(h->x, h->n, h->N, h->dx, h->mesh, h->qs, h->qone, h->far,
                                                        (const unsigned int*)(h->dscal + 0));
        }
    } else {
        Timed t(h, KC_DEPOSIT);
        Strided X{h->x, h->n, 1};
        k_deposit_atomic<true><<<grid_for(h->n), kThreads, 0, h->st>>>(X, h->n, h->N, h->dx, h->mesh, h->qs, h->qone);
    }
    return check_launch(h, "deposit");
}

// forces at the current positions: deposit, Poisson, gather into acc_out (SoA, 3n).
int forces_state(Handle* h, double s, double sm2, double* delta_out, double* acc_out) {
    int rc = deposit_state(h);
    if (rc) return rc;
    const double nmean = (double)h->Np * h->Np * h->Np / ((double)h->N * h->N * h->N);
    PM3D_DISPATCH(h, launch_fft_forward_tmpl, (const void*)h->mesh, true, delta_out, nmean);
    PM3D_DISPATCH(h, launch_x_tmpl, 2, s, sm2);
    PM3D_DISPATCH(h, launch_inverse_tmpl, h->gspec, 3, h->grad);
    {
        Timed t(h, KC_GATHER);
        k_gather3<false><<<grid_for(h->n), kThreads, 0, h->st>>>(h->x, h->n, h->N, h->dx, h->grad, acc_out, h->p, 0, 1);
    }
    return check_launch(h, "forces");
}

int flush_kick(Handle* h) {
    if (!h->kick_pending) return 0;
    Timed t(h, KC_OTHER);
    k_kick<<<grid_for(3 * h->n), kThreads, 0, h->st>>>(h->p, h->acc, 3 * h->n, h->pend_c1, h->pend_den);
    h->kick_pending = false;
    return check_launch(h, "kick");
}

int read_bad(Handle* h, int* bad) {
    unsigned long long v = 0;
    HCHK(h, hipMemcpyAsync(&v, h->dscal + 1, sizeof(v), hipMemcpyDeviceToHost, h->st));
    k_fill_u64<<<1, kThreads, 0, h->st>>>(h->dscal + 1, 1, 0ull);
    HCHK(h, hipStreamSynchronize(h->st));
    *bad = (int)(v > 0x7fffffffull ? 0x7fffffff : v);
    return 0;
}

void free_all(Handle* h) {
    double* bufs[] = {h->x, h->p, h->acc, h->grad, h->delta, h->sbuf};
    for (double* b : bufs)
        if (b) hipFree(b);
    if (h->mesh) hipFree(h->mesh);
    if (h->spec) hipFree(h->spec);
    if (h->gspec) hipFree(h->gspec);
    if (h->tw) hipFree(h->tw);
    if (h->dscal) hipFree(h->dscal);
    if (h->far) hipFree(h->far);
    if (h->bmax) hipFree(h->bmax);
    for (auto& u : h->ev_used) {
        hipEventDestroy(u.second.first);
        hipEventDestroy(u.second.second);
    }
    for (auto e : h->ev_pool) hipEventDestroy(e);
    if (h->st) hipStreamDestroy(h->st);
}

template <typename T>
int dalloc(Handle* h, T** ptr, size_t count) {
    const size_t b = count * sizeof(T);
    hipError_t e = hipMalloc((void**)ptr, b);
    if (e != hipSuccess) {
        h->err = std::string("hipMalloc(") + std::to_string(b) + " B): " + hipGetErrorString(e);
        return PM3D_ENOMEM;
    }
    h->bytes += b;
    return 0;
}

}  // namespace

// ================================================================================ C-ABI
extern "C" {

int pm3d_create(const pm3d_cfg* cfg, void** handle) {
    if (!cfg || !handle) {
        g_create_error = "null argument";
        return PM3D_EINVAL;
    }
    *handle = nullptr;
    const int N = cfg->N;
    int lg = 0;
    while ((1 << lg) < N) ++lg;
    if (N < 8 || N > 512 || (1 << lg) != N) {
        g_create_error = "N must be a power of two in [8,512]";
        return PM3D_EUNSUPPORTED;
    }
    if (cfg->Np < 1 || !(cfg->L > 0.0) || cfg->smooth < 0.0) {
        g_create_error = "invalid Np, L or smooth";
        return PM3D_EINVAL;
    }
    if (cfg->precision != PM3D_FP64) {
        g_create_error = "precision mode not built";
        return PM3D_EUNSUPPORTED;
    }
    // FP64 mode: 128-bit cells at 2^-60; frac_bits applies to the 64-bit cells of the mixed mode
    const int fb = cfg->precision == PM3D_FP64 ? kWideFB : (cfg->frac_bits ? cfg->frac_bits : 40);
    if (fb < 16 || fb > 60) {
        g_create_error = "frac_bits must be in [16,60]";
        return PM3D_EINVAL;
    }
    Handle* h = new Handle();
    h->cfg = *cfg;
    h->cfg.frac_bits = fb;
    h->N = N;
    h->log2N = lg;
    h->Np = cfg->Np;
    h->n = (long long)cfg->Np * cfg->Np * cfg->Np;
    h->M = (long long)N * N * N;
    h->L = cfg->L;
    h->dx = cfg->L / N;  // pm3d.py:30
    h->qs = std::ldexp(1.0, fb);
    h->qscale = std::ldexp(1.0, -fb);
    h->qone = 1LL << fb;
    int rc = 0;
    auto bail = [&](int code) {
        g_create_error = h->err;
        free_all(h);
        delete h;
        return code;
    };
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        h->err = "hipStreamCreate failed";
        return bail(PM3D_EHIP);
    }
    const size_t specN = (size_t)N * N * (N / 2) + (size_t)N * N;
    if ((rc = dalloc(h, &h->x, 3 * h->n)) || (rc = dalloc(h, &h->p, 3 * h->n)) || (rc = dalloc(h, &h->acc, 3 * h->n)) ||
        (rc = dalloc(h, &h->mesh, 2 * h->M)) || (rc = dalloc(h, &h->spec, specN)) || (rc = dalloc(h, &h->gspec, 3 * specN)) ||
        (rc = dalloc(h, &h->grad, 3 * h->M)) || (rc = dalloc(h, &h->tw, (size_t)N)) ||
        (rc = dalloc(h, &h->dscal, 8 + kRing)) || (rc = dalloc(h, &h->bmax, (size_t)N * N)) ||
        (rc = dalloc(h, &h->far, (size_t)h->n)))
        return bail(rc);
    // twiddles exp(-2 pi i k/N) from long-double cos/sin of an exactly reduced angle
    std::vector<cplx<double>> tw(N);
    for (int k = 0; k < N; ++k) {
        const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)N;
        tw[k] = cplx<double>{(double)cosl(ang), (double)sinl(ang)};
    }
    if (hipMemcpy(h->tw, tw.data(), N * sizeof(cplx<double>), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(h->dscal, 0, (8 + kRing) * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(h->acc, 0, 3 * h->n * sizeof(double)) != hipSuccess) {
        h->err = "initial copies failed";
        return bail(PM3D_EHIP);
    }
    *handle = h;
    return PM3D_OK;
}

int pm3d_destroy(void* handle) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    hipStreamSynchronize(h->st);
    free_all(h);
    delete h;
    return PM3D_OK;
}

const char* pm3d_last_error(void* handle) {
    if (!handle) return g_create_error.c_str();
    return ((Handle*)handle)->err.c_str();
}

int64_t pm3d_device_bytes(void* handle) { return handle ? (int64_t)((Handle*)handle)->bytes : 0; }

int pm3d_synchronize(void* handle) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    HCHK(h, hipStreamSynchronize(h->st));
    int bad = 0;
    int rc = read_bad(h, &bad);
    return rc ? rc : bad;
}

int pm3d_set_state(void* handle, const double* x, const double* p, int64_t n) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (n != h->n) return fail(h, PM3D_EINVAL, "n must equal Np^3");
    if (!x && !p) return 0;
    int rc = ensure_sbuf(h, 3 * n * sizeof(double));
    if (rc) return rc;
    if (x) {
        HCHK(h, hipMemcpyAsync(h->sbuf, x, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->st));
        k_aos_to_soa<<<grid_for(n), kThreads, 0, h->st>>>(h->sbuf, h->x, n);
    }
    if (p) {
        // the host value of p already contains every kick: drop a pending one
        h->kick_pending = false;
        HCHK(h, hipMemcpyAsync(h->sbuf, p, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->st));
        k_aos_to_soa<<<grid_for(n), kThreads, 0, h->st>>>(h->sbuf, h->p, n);
    }
    HCHK(h, hipStreamSynchronize(h->st));
    h->state_set = true;
    return check_launch(h, "set_state");
}

int pm3d_get_state(void* handle, double* x, double* p) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->state_set) return fail(h, PM3D_ESTATE, "state not set");
    int rc = ensure_sbuf(h, 3 * h->n * sizeof(double));
    if (rc) return rc;
    if (x) {
        k_soa_to_aos<<<grid_for(h->n), kThreads, 0, h->st>>>(h->x, h->sbuf, h->n);
        HCHK(h, hipMemcpyAsync(x, h->sbuf, 3 * h->n * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HCHK(h, hipStreamSynchronize(h->st));
    }
    if (p) {
        if ((rc = flush_kick(h))) return rc;
        k_soa_to_aos<<<grid_for(h->n), kThreads, 0, h->st>>>(h->p, h->sbuf, h->n);
        HCHK(h, hipMemcpyAsync(p, h->sbuf, 3 * h->n * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HCHK(h, hipStreamSynchronize(h->st));
    }
    return check_launch(h, "get_state");
}

int pm3d_cic_index(void* handle, const double* X, int64_t n, int64_t* key, double* frac) {
    if (!handle || !X || !key || !frac || n < 0) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (n == 0) return 0;
    const size_t xb = 3 * n * sizeof(double), kb = n * sizeof(long long);
    int rc = ensure_sbuf(h, 2 * xb + kb);
    if (rc) return rc;
    double* dX = h->sbuf;
    double* dF = h->sbuf + 3 * n;
    long long* dK = (long long*)(h->sbuf + 6 * n);
    HCHK(h, hipMemcpyAsync(dX, X, xb, hipMemcpyHostToDevice, h->st));
    k_cic_index<<<grid_for(n), kThreads, 0, h->st>>>(Strided{dX, 1, 3}, n, h->N, h->dx, dK, dF);
    HCHK(h, hipMemcpyAsync(key, dK, kb, hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipMemcpyAsync(frac, dF, xb, hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    return check_launch(h, "cic_index");
}

int pm3d_deposit(void* handle, const double* X, int64_t n, double* rho) {
    if (!handle || !rho || (n > 0 && !X) || n < 0) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    const size_t xb = 3 * (size_t)n * sizeof(double);
    int rc = ensure_sbuf(h, std::max(xb, (size_t)h->M * sizeof(double)));
    if (rc) return rc;
    k_fill_u64<<<std::min<long long>(grid_for(2 * h->M), 4096), kThreads, 0, h->st>>>(h->mesh, 2 * h->M, 0ull);
    if (n > 0) {
        HCHK(h, hipMemcpyAsync(h->sbuf, X, xb, hipMemcpyHostToDevice, h->st));
        k_deposit_atomic<true><<<grid_for(n), kThreads, 0, h->st>>>(Strided{h->sbuf, 1, 3}, n, h->N, h->dx, h->mesh, h->qs,
                                                                    h->qone);
    }
    k_fixed_to_real<true><<<std::min<long long>(grid_for(h->M), 4096), kThreads, 0, h->st>>>(h->mesh, h->M, h->qscale, 0,
                                                                                             1.0, h->sbuf);
    HCHK(h, hipMemcpyAsync(rho, h->sbuf, h->M * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    return check_launch(h, "deposit");
}

int pm3d_interp(void* handle, const double* F, const double* X, int64_t n, double* out) {
    if (!handle || !F || !out || (n > 0 && !X) || n < 0) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (n == 0) return 0;
    const size_t xb = 3 * (size_t)n * sizeof(double), fb = (size_t)h->M * sizeof(double);
    int rc = ensure_sbuf(h, xb + fb + n * sizeof(double));
    if (rc) return rc;
    double* dX = h->sbuf;
    double* dF = h->sbuf + 3 * n;
    double* dO = dF + h->M;
    HCHK(h, hipMemcpyAsync(dX, X, xb, hipMemcpyHostToDevice, h->st));
    HCHK(h, hipMemcpyAsync(dF, F, fb, hipMemcpyHostToDevice, h->st));
    k_gather1<<<grid_for(n), kThreads, 0, h->st>>>(Strided{dX, 1, 3}, n, h->N, h->dx, dF, dO);
    HCHK(h, hipMemcpyAsync(out, dO, n * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    return check_launch(h, "interp");
}

static int fetch_delta(Handle* h, double* delta, double* dmax) {
    if (delta) {
        HCHK(h, hipMemcpyAsync(delta, h->delta, h->M * sizeof(double), hipMemcpyDeviceToHost, h->st));
    }
    if (dmax) {
        int rc = pm3d_delta_max_at(h, h->serial, dmax);
        if (rc) return rc;
    }
    return 0;
}

int64_t pm3d_deposit_serial(void* handle) { return handle ? (int64_t)((Handle*)handle)->serial : -1; }

int pm3d_delta_max_at(void* handle, int64_t serial, double* dmax) {
    if (!handle || !dmax) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (serial < 1 || serial > h->serial || serial <= h->serial - kRing)
        return fail(h, PM3D_ESTATE, "deposit serial no longer held on the device");
    unsigned long long q = 0;
    HCHK(h, hipMemcpyAsync(&q, h->dscal + 8 + (serial % kRing), sizeof(q), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    const double nmean = (double)h->Np * h->Np * h->Np / ((double)h->N * h->N * h->N);
    double rmax;
    std::memcpy(&rmax, &q, sizeof(rmax));
    *dmax = rmax / nmean - 1.0;
    return 0;
}

int pm3d_get_delta(void* handle, double* delta) {
    if (!handle || !delta) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (h->serial < 1) return fail(h, PM3D_ESTATE, "no deposit yet");
    int rc = ensure_sbuf(h, h->M * sizeof(double));
    if (rc) return rc;
    const double nmean = (double)h->Np * h->Np * h->Np / ((double)h->N * h->N * h->N);
    k_fixed_to_real<true><<<std::min<long long>(grid_for(h->M), 4096), kThreads, 0, h->st>>>(h->mesh, h->M, h->qscale, 1,
                                                                                             nmean, h->sbuf);
    HCHK(h, hipMemcpyAsync(delta, h->sbuf, h->M * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    return check_launch(h, "get_delta");
}

int pm3d_forces(void* handle, double a, double s, double sm2, double* acc, double* delta) {
    (void)a;
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->state_set) return fail(h, PM3D_ESTATE, "state not set");
    // the step's cached acceleration is left alone (pm3d.py:101-114 does not touch _acc):
    // this call gathers into staging memory
    int rc = ensure_sbuf(h, 6 * h->n * sizeof(double));
    if (rc) return rc;
    if (delta && !h->delta && (rc = dalloc(h, &h->delta, h->M))) return rc;
    if ((rc = forces_state(h, s, sm2, delta ? h->delta : nullptr, h->sbuf))) return rc;
    if (acc) {
        k_soa_to_aos<<<grid_for(h->n), kThreads, 0, h->st>>>(h->sbuf, h->sbuf + 3 * h->n, h->n);
        HCHK(h, hipMemcpyAsync(acc, h->sbuf + 3 * h->n, 3 * h->n * sizeof(double), hipMemcpyDeviceToHost, h->st));
    }
    if ((rc = fetch_delta(h, delta, nullptr))) return rc;
    HCHK(h, hipStreamSynchronize(h->st));
    int bad = 0;
    if ((rc = read_bad(h, &bad))) return rc;
    return bad;
}

int pm3d_step(void* handle, const pm3d_step_scalars* sc, double* delta, double* delta_max) {
    if (!handle || !sc) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->state_set) return fail(h, PM3D_ESTATE, "state not set");
    int rc;
    if (delta && !h->delta && (rc = dalloc(h, &h->delta, h->M))) return rc;
    if (sc->need_forces || !h->acc_valid) {
        // the previous closing kick used the old acceleration: apply it before replacing acc
        if ((rc = flush_kick(h))) return rc;
        if ((rc = forces_state(h, sc->s_old, sc->sm2, nullptr, h->acc))) return rc;
        h->acc_valid = true;
    }
    {
        // pending closing kick of the previous step (pm3d.py:127) + opening kick (:122) + drift
        // (:123-124) in one pass, with the reference's two separate roundings of p
        Timed t(h, KC_KICK_DRIFT);
        k_kick_drift<<<grid_for(h->n), kThreads, 0, h->st>>>(h->x, h->p, h->acc, h->n, h->kick_pending ? 1 : 0,
                                                             h->pend_c1, h->pend_den, sc->c1, sc->den1, sc->da,
                                                             sc->den_d, h->L, h->dscal + 1);
    }
    h->kick_pending = false;
    if ((rc = check_launch(h, "kick_drift"))) return rc;
    // the closing kick is deferred: the gather stores acc, and the kick is applied with the next
    // opening kick or before p is read
    if ((rc = forces_state(h, sc->s_new, sc->sm2, delta ? h->delta : nullptr, h->acc))) return rc;
    h->acc_valid = true;
    h->kick_pending = true;
    h->pend_c1 = sc->c1;
    h->pend_den = sc->den2;
    if (delta || delta_max) {
        if ((rc = fetch_delta(h, delta, delta_max))) return rc;
        HCHK(h, hipStreamSynchronize(h->st));
        int bad = 0;
        if ((rc = read_bad(h, &bad))) return rc;
        return bad;
    }
    return 0;
}

int pm3d_rfftn(void* handle, const double* in, double* out_c) {
    if (!handle || !in || !out_c) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    const int N = h->N, NH = N / 2;
    int rc = ensure_sbuf(h, h->M * sizeof(double));
    if (rc) return rc;
    HCHK(h, hipMemcpyAsync(h->sbuf, in, h->M * sizeof(double), hipMemcpyHostToDevice, h->st));
    PM3D_DISPATCH(h, launch_fft_forward_tmpl, (const void*)h->sbuf, false, nullptr, 1.0);
    PM3D_DISPATCH(h, launch_x_tmpl, 0, 0.0, 0.0);
    if ((rc = check_launch(h, "rfftn"))) return rc;
    const size_t specN = (size_t)N * N * NH;
    std::vector<cplx<double>> hs(specN + (size_t)N * N);
    HCHK(h, hipMemcpyAsync(hs.data(), h->spec, hs.size() * sizeof(cplx<double>), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    for (size_t r = 0; r < (size_t)N * N; ++r) {
        std::memcpy(out_c + 2 * r * (NH + 1), &hs[r * NH], NH * sizeof(cplx<double>));
        out_c[2 * (r * (NH + 1) + NH)] = hs[specN + r].re;
        out_c[2 * (r * (NH + 1) + NH) + 1] = hs[specN + r].im;
    }
    return 0;
}

int pm3d_irfftn(void* handle, const double* in_c, double* out) {
    if (!handle || !in_c || !out) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    const int N = h->N, NH = N / 2;
    const size_t specN = (size_t)N * N * NH;
    std::vector<cplx<double>> hs(specN + (size_t)N * N);
    for (size_t r = 0; r < (size_t)N * N; ++r) {
        std::memcpy(&hs[r * NH], in_c + 2 * r * (NH + 1), NH * sizeof(cplx<double>));
        hs[specN + r] = cplx<double>{in_c[2 * (r * (NH + 1) + NH)], in_c[2 * (r * (NH + 1) + NH) + 1]};
    }
    int rc = ensure_sbuf(h, h->M * sizeof(double));
    if (rc) return rc;
    HCHK(h, hipMemcpyAsync(h->spec, hs.data(), hs.size() * sizeof(cplx<double>), hipMemcpyHostToDevice, h->st));
    PM3D_DISPATCH(h, launch_x_tmpl, 1, 0.0, 0.0);
    PM3D_DISPATCH(h, launch_inverse_tmpl, h->spec, 1, h->sbuf);
    if ((rc = check_launch(h, "irfftn"))) return rc;
    HCHK(h, hipMemcpyAsync(out, h->sbuf, h->M * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

int pm3d_poisson_gradient(void* handle, const double* delta, double s, double sm2, double* g) {
    if (!handle || !delta || !g) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    int rc = ensure_sbuf(h, h->M * sizeof(double));
    if (rc) return rc;
    HCHK(h, hipMemcpyAsync(h->sbuf, delta, h->M * sizeof(double), hipMemcpyHostToDevice, h->st));
    PM3D_DISPATCH(h, launch_fft_forward_tmpl, (const void*)h->sbuf, false, nullptr, 1.0);
    PM3D_DISPATCH(h, launch_x_tmpl, 2, s, sm2);
    PM3D_DISPATCH(h, launch_inverse_tmpl, h->gspec, 3, h->grad);
    if ((rc = check_launch(h, "poisson_gradient"))) return rc;
    HCHK(h, hipMemcpyAsync(g, h->grad, 3 * h->M * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    return 0;
}

int pm3d_min_jacobian(void* handle, double* jmin) {
    if (!handle || !jmin) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->state_set) return fail(h, PM3D_ESTATE, "state not set");
    k_fill_u64<<<1, kThreads, 0, h->st>>>(h->dscal + 2, 1, 0x7fffffffffffffffull);
    k_min_jacobian<<<grid_for(h->n), kThreads, 0, h->st>>>(h->x, h->Np, h->L, h->L / h->Np, (long long*)(h->dscal + 2));
    long long o = 0;
    HCHK(h, hipMemcpyAsync(&o, h->dscal + 2, sizeof(o), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    const long long bits = o >= 0 ? o : (o ^ 0x7fffffffffffffffLL);
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *jmin = v;
    return check_launch(h, "min_jacobian");
}

int pm3d_set_timing(void* handle, int enable) {
    if (!handle) return PM3D_EINVAL;
    ((Handle*)handle)->timing = enable != 0;
    return 0;
}

int pm3d_get_timings(void* handle, int64_t* counts, double* total_ms, int nmax) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    HCHK(h, hipStreamSynchronize(h->st));
    for (auto& u : h->ev_used) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, u.second.first, u.second.second) == hipSuccess) {
            h->t_count[u.first] += 1;
            h->t_ms[u.first] += ms;
        }
        h->ev_pool.push_back(u.second.first);
        h->ev_pool.push_back(u.second.second);
    }
    h->ev_used.clear();
    const int m = nmax < KC_COUNT ? nmax : KC_COUNT;
    for (int k = 0; k < m; ++k) {
        if (counts) counts[k] = h->t_count[k];
        if (total_ms) total_ms[k] = h->t_ms[k];
    }
    for (int k = 0; k < KC_COUNT; ++k) {
        h->t_count[k] = 0;
        h->t_ms[k] = 0;
    }
    return KC_COUNT;
}

const char* pm3d_kernel_names(void) { return kKernelNames; }

}  // extern "C"
