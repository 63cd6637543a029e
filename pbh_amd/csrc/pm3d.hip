// pm3d particle-mesh library for MI355X (gfx950): C-ABI entry points of include/pm3d_hip.h.
//
// One handle = one periodic box stepped on one HIP stream.  Device state is SoA in Lagrangian
// particle order (particle i is lattice site i, pm3d.py:64-65).
//   FP64 mode : x, p, acc as 3 x n doubles; deposit mesh 128-bit fixed point {lo,hi} per cell at
//               2^-52.
//   mixed mode: cell (3 x n int32) + off (3 x n float), p and acc 3 x n float; deposit mesh one
//               uint64 per cell at 2^-31; FFT, Green's function and gradient mesh in FP32.
//
// A step (pm3d.py:117-129) is:
//   gather of the previous step's force meshes + its closing kick + this step's opening kick and
//   drift (one kernel) -> deposit (+ far pass) -> z and y forward per x plane (fixed point -> delta
//   fused) -> x (forward FFT, Green's function, -ik, inverse FFTs) -> y and z inverse per x plane
//   into the three force meshes, which the next step gathers.  Sizes whose half plane does not fit
//   in LDS run the z, y, x, y, z passes as separate pencil kernels.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pm3d_hip.h"
#include "common.hpp"
#include "fft.hpp"
#include "particles.hpp"
#include "particles_mixed.hpp"
#include "deposit.hpp"
#include "fft_plane.hpp"
#include "batch_kernels.hpp"
#include "ic.hpp"
#include "analyse.hpp"

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
    KC_FWD_ZY,
    KC_INV_YZ,
    KC_GATHER_KICK,
    KC_COUNT
};
const char* kKernelNames =
    "kick_drift,fill,deposit,deposit_far,r2c_z,y_fwd,x_poisson,y_inv,c2r_z,gather,other,fwd_zy,inv_yz,gather_kick";

// Trials of the step deposit's column shape during a run (live_deposit).  Every column shape writes
// the same mesh bits, so any deposit may use any of them; as particles cluster, the workgroups of
// the dense columns take longer and a narrower shape (more, smaller workgroups) can overtake the
// one chosen at creation on a lattice.
constexpr int kLivePeriod = 64;  // step deposits between the starts of two trials
constexpr int kLiveReps = 2;     // deposits per candidate in a trial (its faster one counts)
constexpr int kLiveSlots = 4 * kLiveReps;
// A candidate slower than the trial's winner by more than this factor sits out the next
// kLiveRestTrials trials: its deposits are the trial's cost (FP64 128^3 at the crossing: 8x8
// columns 44 us against 32 for 4x4, about 1.5 us per step over a window that holds a trial).
constexpr float kLiveRestMargin = 1.2f;
constexpr int kLiveRestTrials = 3;
struct LiveShape {
    unsigned cand = 0;       // bit v: shape variant v is a candidate (0 or one bit: no trials)
    long long deposits = 0;  // step deposits made
    long long next = 0;      // step deposit at which the next trial starts (a new state: its first)
    int plan[kLiveSlots] = {};  // variants of the running trial, in launch order (the current one last)
    int nplan = 0, pos = 0;     // trial length (0: none running), deposits of it launched
    int unread = 0;             // samples of the trial not yet read
    float best[4] = {};         // fastest sample per variant in the running trial (ms)
    int rest[4] = {};           // trials a candidate sits out after losing by more than kLiveRestMargin
    hipEvent_t ev[2 * kLiveSlots] = {};  // event pair of each sample, in plan order
    int trials = 0, changes = 0;         // finished trials, shape changes
};

struct Handle {
    pm3d_cfg cfg{};
    int N = 0, Np = 0, log2N = 0;
    long long n = 0, M = 0;
    double L = 0, dx = 0, qscale = 0, qs = 0, nmean = 1;
    long long qone = 0;
    bool mixed = false;
    bool wide = true;  // 128-bit mesh cells (FP64 mode)
    hipStream_t st = nullptr;
    // particle state, FP64 mode
    double *x = nullptr, *p = nullptr, *acc = nullptr;
    // particle state, mixed mode
    int* cell = nullptr;
    float *off = nullptr, *pf = nullptr, *accf = nullptr;
    // meshes
    unsigned long long* mesh = nullptr;
    void* spec = nullptr;   // [N][N][N/2] + nyq [N][N], complex R
    void* gspec = nullptr;  // 3 x the same
    void* grad = nullptr;   // [3][M] R (component-major)
    void* tw = nullptr;     // exp(-2 pi i k/N), complex R
    // device scalars: [0] unused, [1] bad count (sticky until pm3d_synchronize), [2] overflow count
    // of the row lists, [3..7] unused,
    // [8 + s % kRing] max rho (double bits) of deposit serial s
    unsigned long long* dscal = nullptr;
    long long serial = 0;
    int* far = nullptr;                // far particles, one segment per deposit column (or the row lists' overflow)
    unsigned int* farcnt = nullptr;    // far count per column
    // far-particle row lists (FarRows): counts per base-cell row; the lists alias acc / accf, which
    // only a state with acc_valid uses (a step that builds the lists clears acc_valid)
    unsigned int* rowcnt = nullptr;
    int row_cap = 0;
    bool lists_fresh = false;  // the last gather-kick-drift built lists that no deposit has consumed
    double* bmax = nullptr;  // per-block maxima / minima scratch
    size_t bmax_len = 0;
    double* sbuf = nullptr;  // staging for reference-shaped entry points
    size_t sbuf_bytes = 0;
    // device step schedule for graph replay
    DevSched* sched = nullptr;
    DevStep* tab = nullptr;
    double* jhist = nullptr;
    int tab_cap = 0;
    DevStep* htab = nullptr;    // pinned host copy of the schedule table (htab_cap entries)
    int htab_cap = 0;
    DevSched* hsched = nullptr; // pinned host copy of the counter the last pm3d_set_schedule uploaded
    double* hjmin = nullptr;    // pinned host slot pm3d_min_jacobian's reduction writes
    bool sched_upload = false;  // its copies may still read htab / hsched
    bool use_sched = false;  // set while a batch graph is being captured
    bool state_set = false;
    bool acc_valid = false;   // acc holds the acceleration at the current positions
    bool grad_valid = false;  // grad holds the force meshes of the last solve at the current positions
    bool kick_pending = false;
    double pend_c1 = 0, pend_den = 0;
    // timing
    bool timing = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
    std::vector<hipEvent_t> ev_pool;
    long long t_count[KC_COUNT] = {};
    double t_ms[KC_COUNT] = {};
    std::string err;
    size_t bytes = 0;
    int fftv[5] = {};  // FFT tile variant per pass (r2c_z, y fwd, x, y inv, c2r_z)
    int depv = 0;      // column shape variant of the deposit (ColShape<N, WIDE, V>)
    LiveShape live;    // trials of the step deposit's column shape during a run
    bool live_arm = false;  // the next deposit_state is a step's own deposit (pm3d_step)
    int bmax_pending = 0;  // per-block deposit maxima left in bmax by the plane forward pass
    // analysis sets (pm3d_set_analysis_sets): indices, offsets, centre/moment/partial buffers
    int* an_idx = nullptr;
    long long* an_off = nullptr;
    double* an_buf = nullptr;  // [0..3) centre, [3..4) rho max, [4 + 9*s) per set, then partial sums
    unsigned long long* an_cnt = nullptr;
    int an_nsets = 0;
    long long an_off_host1 = 0;  // size of the core set
};

#define HCHK(h, call)                                                     \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) {                                           \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_); \
            return PM3D_EHIP;                                             \
        }                                                                 \
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

// Times the launches of one scope when timing is on.  A scope of one launch made through
// `launch` (`one` = true) has the dispatch itself stamp the event pair (hipExtLaunchKernelGGL: the
// kernel's own start and end, the duration a profiler's kernel trace reports); any other scope is
// bracketed by two event records, whose time includes the gaps around its launches.
struct Timed {
    Handle* h;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool one;
    Timed(Handle* h_, int kc, bool one_launch = false) : h(h_), one(one_launch) {
        if (h->timing) {
            e0 = take_event(h);
            e1 = take_event(h);
            if (e0 && e1) {
                if (!one) hipEventRecord(e0, h->st);
                h->ev_used.push_back({kc, {e0, e1}});
            } else {
                e1 = nullptr;
            }
        }
    }
    ~Timed() {
        if (e1 && !one) hipEventRecord(e1, h->st);
    }
    // kernel<<<g, b, 0, h->st>>>(args...), every argument converted to the kernel's parameter type
    template <typename... P, typename... A>
    void launch(void (*kernel)(P...), dim3 g, dim3 b, A... args) {
        static_assert(sizeof...(P) == sizeof...(A), "argument count");
        if (e1 && one)
            hipExtLaunchKernelGGL(kernel, g, b, 0, h->st, e0, e1, 0, static_cast<P>(args)...);
        else
            kernel<<<g, b, 0, h->st>>>(static_cast<P>(args)...);
    }
};

inline unsigned grid_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline unsigned grid_cap(long long n) { return (unsigned)std::min<long long>(grid_for(n), 4096); }
inline unsigned long long* qmax_slot(Handle* h) { return h->dscal + 8 + (h->serial % kRing); }
inline MxPos mxpos(Handle* h) { return MxPos{h->cell, h->off, h->n}; }
// XCD mapping of a launch of G blocks whose runs of `share` consecutive logical blocks share data:
// such runs are dealt round-robin to the XCDs, so the XCDs work on neighbouring addresses at any
// moment (no two XCDs stream from regions a power of two apart).  Falls back to one contiguous
// range per XCD when G does not split into whole rounds.  On MI355X (128^3 FP64 / 256^3 FP32 step,
// us): gather 34.0 / 162 unmapped, 33.0 / 175 contiguous ranges, 32.2 / 164 runs; y inv 24.5 / 116,
// 24.7 / 106, 24.8 / 102; the column deposit prefers contiguous ranges (178 / 166 / 172 in FP32).
inline XMap xmap(unsigned G, unsigned share) {
    if (share >= 1 && G % (kXcds * share) == 0) return XMap{2, share};
    return XMap{1, 1};
}

int check_launch(Handle* h, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(h, PM3D_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

// Far-particle row lists of the handle (FarRows), or none (cnt == nullptr) when its deposit does
// not use columns or uses the per-column far pass (N = 512).
FarRows far_rows(const Handle* h) {
    const DevSched* sch = h->use_sched ? h->sched : nullptr;
    if (!h->rowcnt) return FarRows{nullptr, nullptr, 0, nullptr, nullptr, 0u, sch};
    return FarRows{h->rowcnt, (int*)(h->mixed ? (void*)h->accf : (void*)h->acc), h->row_cap, h->far,
                   (unsigned int*)(h->dscal + 2), (unsigned int)h->n, sch};
}
// the far-particle overflow count, cleared by the forward transform after the deposit consumed it
unsigned int* ovf_clear(const Handle* h) { return h->rowcnt ? (unsigned int*)(h->dscal + 2) : nullptr; }

// ---------------------------------------------------------------- FFT launches, R and N templated
// Each pass runs in tile variant h->fftv[pass] (fft.hpp Plan<R, N, V>); every variant computes
// the same bits, so the choice is made by timing (tune_fft).
enum FftPass { FP_R2C = 0, FP_YFWD, FP_X, FP_YINV, FP_C2R, FP_COUNT };

#define PM3D_V4(v, fn, ...)                   \
    switch (v) {                              \
        case 1: fn<1>(__VA_ARGS__); break;    \
        case 2: fn<2>(__VA_ARGS__); break;    \
        case 3: fn<3>(__VA_ARGS__); break;    \
        default: fn<0>(__VA_ARGS__); break;   \
    }

template <typename R, int N>
struct Fft {
    static constexpr size_t NH = N / 2, specN = (size_t)N * N * NH, tot = specN + (size_t)N * N;
    // z and y passes fused per x plane (fft_plane.hpp) when half a plane fits in LDS
    static constexpr bool PLANE = PlaneTile<R, N>::FITS;

    static void fwd_plane(Handle* h, const void* in, bool fixed) {
        constexpr int NT = PlaneTile<R, N>::NT;
        cplx<R>* spec = (cplx<R>*)h->spec;
        const cplx<R>* tw = (const cplx<R>*)h->tw;
        const unsigned g = 2 * N;
        const XMap xm = xmap(g, 2);  // the two workgroups of a plane read the same rows: one XCD
        Timed t(h, KC_FWD_ZY, true);
        if (fixed && h->wide)
            t.launch(k_fwd_plane<R, N, true, true>, g, NT, in, spec, spec + specN, tw, h->bmax, h->qscale, h->nmean, xm,
                     ovf_clear(h));
        else if (fixed)
            t.launch(k_fwd_plane<R, N, true, false>, g, NT, in, spec, spec + specN, tw, h->bmax, h->qscale, h->nmean, xm,
                     ovf_clear(h));
        else
            t.launch(k_fwd_plane<R, N, false, false>, g, NT, in, spec, spec + specN, tw, nullptr, 0.0, 1.0, xm, nullptr);
        h->bmax_pending = fixed ? (int)g : 0;
    }

    template <int Q>
    static void inv_plane_q(Handle* h, cplx<R>* base, int ncomp, void* out) {
        constexpr int NT = InvTile<R, N, Q>::NT;
        const R scale = (R)(1.0 / ((double)N * N * N));
        const unsigned g = Q * N * ncomp;
        Timed t(h, KC_INV_YZ, true);
        t.launch(k_inv_plane<R, N, Q>, g, NT, base, base + specN, tot, tot, (R*)out, (size_t)N * N * N,
                 (const cplx<R>*)h->tw, scale, xmap(g, Q));
    }
    // split 2 ways along y: measured on MI355X against 4 and 8 (smaller tiles, more workgroups per
    // CU, the plane read 4 or 8 times from L2), 128^3 FP64 / 256^3 FP32 inverse pass: 30.2 / 159 us
    // (2), 32.7 / 189 us (4), 40.1 / 216 us (8, before the bank-conflict fix of the y phase); with the
    // twiddles in LDS: 26.8 / 137.5 us (2), 27.7 / 156.4 us (4)
    static constexpr int kInvQ = 2;
    static void inv_plane(Handle* h, cplx<R>* base, int ncomp, void* out) { inv_plane_q<kInvQ>(h, base, ncomp, out); }

    // consecutive pencil blocks whose P-element row segments share a 128-byte line
    template <int P>
    static unsigned pencil_share() {
        const unsigned seg = (unsigned)(P * sizeof(cplx<R>));
        return seg >= 128 ? 1u : 128u / seg;
    }

    template <int V>
    static unsigned r2c_grid() { return (unsigned)((size_t)N * N / Plan<R, N, V>::PZ); }

    template <int V>
    static void r2c(Handle* h, const void* in, bool fixed) {
        cplx<R>* spec = (cplx<R>*)h->spec;
        const cplx<R>* tw = (const cplx<R>*)h->tw;
        const unsigned gz = r2c_grid<V>();
        Timed t(h, KC_R2C_Z);
        if (fixed && h->wide)
            k_r2c_z<R, N, true, true, V><<<gz, kThreads, 0, h->st>>>(in, spec, spec + specN, tw, h->bmax,
                                                                    h->qscale, h->nmean, ovf_clear(h));
        else if (fixed)
            k_r2c_z<R, N, true, false, V><<<gz, kThreads, 0, h->st>>>(in, spec, spec + specN, tw, h->bmax,
                                                                     h->qscale, h->nmean, ovf_clear(h));
        else
            k_r2c_z<R, N, false, false, V><<<gz, kThreads, 0, h->st>>>(in, spec, spec + specN, tw, nullptr, 0.0, 1.0,
                                                                      nullptr);
    }

    // nbmax: number of per-block maxima the r2c launch left in h->bmax (0 when the input was real)
    template <int V>
    static void yfwd(Handle* h, int nbmax) {
        cplx<R>* spec = (cplx<R>*)h->spec;
        constexpr int P = Plan<R, N, V>::PY;
        dim3 g(N * (N / 2 / P) + N / P, 1);
        Timed t(h, KC_Y_FWD);
        k_c2c_y<R, N, false, V><<<g, kThreads, 0, h->st>>>(spec, spec + specN, 0, 0, (const cplx<R>*)h->tw, h->bmax,
                                                           nbmax, qmax_slot(h), xmap(g.x, pencil_share<P>()));
    }

    // x pass: mode 0 forward, 1 inverse (in place on h->spec), 2 Poisson (h->spec -> h->gspec),
    // 3 Zel'dovich displacement spectra (h->spec -> h->gspec)
    template <int V>
    static void x(Handle* h, int mode, double s, double sm2) {
        cplx<R>* spec = (cplx<R>*)h->spec;
        cplx<R>* gs = (cplx<R>*)h->gspec;
        const cplx<R>* tw = (const cplx<R>*)h->tw;
        constexpr int P = Plan<R, N, V>::PX;
        const unsigned g = N * (N / 2 / P) + N / P;
        const double kval = 1.0 / ((double)N * h->dx);
        Timed t(h, KC_X_POISSON, true);
        const DevSched* sch = h->use_sched ? h->sched : nullptr;
        const XMap xm = xmap(g, pencil_share<P>());
        // the plane forward pass leaves the deposit maximum in per-block pieces: reduced here
        const int nb = h->bmax_pending;
        unsigned long long* slot = nb > 0 ? qmax_slot(h) : nullptr;
        h->bmax_pending = 0;
        if (mode == 0)
            t.launch(k_x_pass<R, N, 0, V>, g, kThreads, spec, spec + specN, nullptr, nullptr, 0, 0, tw, kval, 0, 0, 0,
                     nullptr, nullptr, h->bmax, nb, slot, xm);
        else if (mode == 1)
            t.launch(k_x_pass<R, N, 1, V>, g, kThreads, spec, spec + specN, nullptr, nullptr, 0, 0, tw, kval, 0, 0, 0,
                     nullptr, nullptr, h->bmax, nb, slot, xm);
        else if (mode == 2)
            t.launch(k_x_pass<R, N, 2, V>, g, kThreads, spec, spec + specN, gs, gs + specN, tot, tot, tw, kval, s, sm2,
                     sm2 > 0.0 ? 1 : 0, sch, h->tab, h->bmax, nb, slot, xm);
        else
            t.launch(k_x_pass<R, N, 3, V>, g, kThreads, spec, spec + specN, gs, gs + specN, tot, tot, tw, kval, 0, 0, 0,
                     nullptr, nullptr, h->bmax, nb, slot, xm);
    }

    template <int V>
    static void yinv(Handle* h, cplx<R>* base, int ncomp) {
        constexpr int P = Plan<R, N, V>::PY;
        dim3 g(N * (N / 2 / P) + N / P, ncomp);
        Timed t(h, KC_Y_INV);
        k_c2c_y<R, N, true, V><<<g, kThreads, 0, h->st>>>(base, base + specN, tot, tot, (const cplx<R>*)h->tw, nullptr,
                                                          0, nullptr, xmap(g.x * g.y, pencil_share<P>()));
    }

    template <int V>
    static void c2r(Handle* h, cplx<R>* base, int ncomp, void* out) {
        const R scale = (R)(1.0 / ((double)N * N * N));
        const cplx<R>* tw = (const cplx<R>*)h->tw;
        Timed t(h, KC_C2R_Z);
        if (ncomp == 3) {
            const unsigned g = (unsigned)((size_t)N * N / Plan<R, N, V>::template pc<3>());
            k_c2r_z<R, N, 3, V><<<g, kThreads, 0, h->st>>>(base, base + specN, tot, tot, (R*)out, (size_t)N * N * N, tw,
                                                           scale);
        } else {
            const unsigned g = (unsigned)((size_t)N * N / Plan<R, N, V>::template pc<1>());
            k_c2r_z<R, N, 1, V><<<g, kThreads, 0, h->st>>>(base, base + specN, tot, tot, (R*)out, (size_t)N * N * N, tw,
                                                           scale);
        }
    }

    // r2c z + forward y on h->spec.  fixed: input is the deposit mesh.
    static void forward(Handle* h, const void* in, bool fixed) {
        if constexpr (PLANE) {
            fwd_plane(h, in, fixed);
            return;
        }
        PM3D_V4(h->fftv[FP_R2C], r2c, h, in, fixed);
        const int nb = fixed ? (int)((h->fftv[FP_R2C] & 1) ? r2c_grid<1>() : r2c_grid<0>()) : 0;  // PZ depends on bit 0 only
        PM3D_V4(h->fftv[FP_YFWD], yfwd, h, nb);
    }

    static void xpass(Handle* h, int mode, double s, double sm2) {
        PM3D_V4(h->fftv[FP_X], x, h, mode, s, sm2);
    }

    // inverse y (ncomp spectra from base) + c2r z into out (component c at out + c*N^3)
    static void inverse(Handle* h, void* base_, int ncomp, void* out) {
        cplx<R>* base = (cplx<R>*)base_;
        if constexpr (PLANE) {
            inv_plane(h, base, ncomp, out);
            return;
        }
        PM3D_V4(h->fftv[FP_YINV], yinv, h, base, ncomp);
        PM3D_V4(h->fftv[FP_C2R], c2r, h, base, ncomp, out);
    }

    // Launches pass `pass` in variant V as the step does (fixed-point r2c input, Poisson x pass,
    // three components inverse).
    template <int V>
    static void one(Handle* h, int pass) {
        switch (pass) {
            case FP_R2C: r2c<V>(h, h->mesh, true); break;
            case FP_YFWD: yfwd<V>(h, (int)r2c_grid<V>()); break;
            case FP_X: x<V>(h, 2, 1.0, 0.0); break;
            case FP_YINV: yinv<V>(h, (cplx<R>*)h->gspec, 3); break;
            default: c2r<V>(h, (cplx<R>*)h->gspec, 3, h->grad); break;
        }
    }

    static void launch_variant(Handle* h, int pass, int v) { PM3D_V4(v, one, h, pass); }
};

template <typename R>
int fft_call(Handle* h, int op, const void* in, bool fixed, int mode, double s, double sm2, void* base, int ncomp,
             void* out) {
#define PM3D_FFT_CASE(NN)                                     \
    case NN:                                                  \
        if (op == 0)                                          \
            Fft<R, NN>::forward(h, in, fixed);                \
        else if (op == 1)                                     \
            Fft<R, NN>::xpass(h, mode, s, sm2);               \
        else if (op == 2)                                     \
            Fft<R, NN>::inverse(h, base, ncomp, out);         \
        else                                                  \
            Fft<R, NN>::launch_variant(h, mode, ncomp);       \
        break;
    switch (h->N) {
        PM3D_FFT_CASE(8)
        PM3D_FFT_CASE(16)
        PM3D_FFT_CASE(32)
        PM3D_FFT_CASE(48)
        PM3D_FFT_CASE(64)
        PM3D_FFT_CASE(96)
        PM3D_FFT_CASE(128)
        PM3D_FFT_CASE(256)
        PM3D_FFT_CASE(512)
        default:
            return fail(h, PM3D_EUNSUPPORTED, "unsupported mesh size");
    }
#undef PM3D_FFT_CASE
    return 0;
}

int fft_forward(Handle* h, const void* in, bool fixed) {
    return h->mixed ? fft_call<float>(h, 0, in, fixed, 0, 0, 0, nullptr, 0, nullptr)
                    : fft_call<double>(h, 0, in, fixed, 0, 0, 0, nullptr, 0, nullptr);
}
int fft_xpass(Handle* h, int mode, double s, double sm2) {
    return h->mixed ? fft_call<float>(h, 1, nullptr, false, mode, s, sm2, nullptr, 0, nullptr)
                    : fft_call<double>(h, 1, nullptr, false, mode, s, sm2, nullptr, 0, nullptr);
}
int fft_inverse(Handle* h, void* base, int ncomp, void* out) {
    return h->mixed ? fft_call<float>(h, 2, nullptr, false, 0, 0, 0, base, ncomp, out)
                    : fft_call<double>(h, 2, nullptr, false, 0, 0, 0, base, ncomp, out);
}
// true when the z and y passes run fused per x plane (fft_plane.hpp) for this handle's N
template <typename R>
bool plane_for(int N) {
    switch (N) {
        case 8: return Fft<R, 8>::PLANE;
        case 16: return Fft<R, 16>::PLANE;
        case 32: return Fft<R, 32>::PLANE;
        case 48: return Fft<R, 48>::PLANE;
        case 64: return Fft<R, 64>::PLANE;
        case 96: return Fft<R, 96>::PLANE;
        case 128: return Fft<R, 128>::PLANE;
        case 256: return Fft<R, 256>::PLANE;
        case 512: return Fft<R, 512>::PLANE;
        default: return false;
    }
}
bool fft_plane_used(const Handle* h) { return h->mixed ? plane_for<float>(h->N) : plane_for<double>(h->N); }

int fft_launch_variant(Handle* h, int pass, int v) {
    return h->mixed ? fft_call<float>(h, 3, nullptr, false, pass, 0, 0, nullptr, v, nullptr)
                    : fft_call<double>(h, 3, nullptr, false, pass, 0, 0, nullptr, v, nullptr);
}

// Picks the fastest tile variant of each FFT pass (fft.hpp Plan: pencils per workgroup x one or two
// LDS buffers) by timing all four on this handle's buffers (5 launches each after one warm-up, on
// zeroed buffers).  Every variant computes the same bits, so the choice never changes a result.
// The choice is cached per (N, precision) for the process.  PM3D_FFT_TILES = one digit 0-3 forces
// that variant for every pass, five digits one per pass (r2c_z, y fwd, x, y inv, c2r_z).
std::mutex g_tune_mu;
std::map<std::pair<int, int>, std::array<int, FP_COUNT>> g_tuned;
// key of the tuning caches: mesh size and kind (0: FP64 on 128-bit cells, 1: mixed, 2: FP64 on
// 64-bit cells)
std::pair<int, int> tune_key(const Handle* h) { return {h->N, h->mixed ? 1 : h->wide ? 0 : 2}; }

int tune_fft(Handle* h) {
    const char* env = std::getenv("PM3D_FFT_TILES");
    if (env && *env && std::strcmp(env, "auto") != 0) {
        const size_t len = std::strlen(env);
        if (len != 1 && len != FP_COUNT) return fail(h, PM3D_EINVAL, "PM3D_FFT_TILES must be auto, one digit or 5 digits");
        for (int p = 0; p < FP_COUNT; ++p) {
            const char c = env[len == 1 ? 0 : p];
            if (c < '0' || c > '3') return fail(h, PM3D_EINVAL, "PM3D_FFT_TILES digits must be 0-3");
            h->fftv[p] = c - '0';
        }
        return 0;
    }
    const std::pair<int, int> key = tune_key(h);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            for (int p = 0; p < FP_COUNT; ++p) h->fftv[p] = it->second[p];
            return 0;
        }
    }
    hipEvent_t e0, e1;
    HCHK(h, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return fail(h, PM3D_EHIP, "hipEventCreate failed");
    }
    int rc = 0;
    std::array<int, FP_COUNT> best{};
    const bool plane = fft_plane_used(h);
    for (int p = 0; p < FP_COUNT && !rc; ++p) {
        if (plane && p != FP_X) continue;  // z and y run in the plane kernels (no tile variants)
        float ms[4] = {0.f, 0.f, 0.f, 0.f};
        for (int v = 0; v < 4 && !rc; ++v) {
            if ((rc = fft_launch_variant(h, p, v))) break;
            if (hipEventRecord(e0, h->st) != hipSuccess) { rc = fail(h, PM3D_EHIP, "hipEventRecord failed"); break; }
            for (int r = 0; r < 5 && !rc; ++r) rc = fft_launch_variant(h, p, v);
            if (rc) break;
            if (hipEventRecord(e1, h->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms[v], e0, e1) != hipSuccess)
                rc = fail(h, PM3D_EHIP, "FFT tile timing failed");
            if (!rc) rc = check_launch(h, "FFT tile timing");
        }
        int b = 0;
        for (int v = 1; v < 4; ++v)
            if (ms[v] < ms[b]) b = v;
        best[p] = b;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    for (int p = 0; p < FP_COUNT; ++p) h->fftv[p] = best[p];
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = best;
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

// Particle sources of the column deposit: per-axis base pointers of the SoA state.
SrcF64 src_f64(const Handle* h) {
    return SrcF64{h->x, h->x + h->n, h->x + 2 * h->n, h->dx, 1.0 / h->dx, h->N};
}
SrcF64C src_f64c(const Handle* h) {
    return SrcF64C{h->x, h->x + h->n, h->x + 2 * h->n, h->dx, 1.0 / h->dx, h->N};
}
SrcMx src_mx(const Handle* h) {
    return SrcMx{h->cell, h->cell + h->n, h->cell + 2 * h->n, h->off, h->off + h->n, h->off + 2 * h->n, h->N};
}

// Workgroups of the balanced far pass: enough for every particle to be far in one sweep of up to
// kFarBalGrid workgroups.
unsigned far_bal_grid(long long n) {
    return (unsigned)std::min<long long>(kFarBalGrid, (n + kThreads - 1) / kThreads);
}
// With row lists the pass only deposits their overflow (rare) and clears the N^2 row counts: one
// count per thread.
unsigned far_lists_grid(int N) {
    return (unsigned)std::max<long long>(1, std::min<long long>(kFarBalGrid, (long long)N * N / kThreads));
}

// XCD mapping of the column deposit: one x row of columns (N / BY of them) after another dealt
// round-robin to the XCDs.  Neighbouring columns of a row share ring lines in one L2, and the heavy
// columns of a collapsing core are spread over all XCDs instead of landing in one XCD's range.
// Measured at the first shell crossing, mixed 256^3 (8x8 columns) / FP64 128^3 (4x4), us: one
// contiguous range per XCD 271 / 42, runs of 1-8 columns 249-256 / 37, 16: 228 / 35.6, 32 (a row):
// 214 / 35.5, 64: 218 / 37.5; early steps within 1 us of each other (151 / 27).
template <int N, int BY>
XMap column_map(unsigned ncol) {
    return xmap(ncol, (unsigned)(N / BY));
}

// Column deposit (deposit.hpp) for one mesh size and column shape variant; `lists`: take the far
// particles from the row lists the step's gather-kick-drift built.
template <int N, bool WIDE, int V, class Src>
void deposit_columns(Handle* h, const Src& src, unsigned long long* mesh, bool lists) {
    using S = ColShape<N, WIDE, V>;
    const FarRows fr = far_rows(h);
    {
        Timed t(h, KC_DEPOSIT, true);
        constexpr int NT = col_threads<Src>();
        t.launch(k_deposit_col<Src, S::BX, S::BY, N, WIDE, NT>, S::NBLK, NT, src, mesh, h->far, h->farcnt,
                 lists ? fr : FarRows{nullptr, nullptr, 0, nullptr, nullptr, 0u, fr.sch}, column_map<N, S::BY>(S::NBLK));
    }
    Timed t(h, KC_DEPOSIT_FAR, true);
    if constexpr (S::NBLK <= kFarColsMax) {
        t.launch(k_deposit_far_bal<Src, WIDE>, lists ? far_lists_grid(N) : far_bal_grid(h->n), kThreads, src, mesh,
                 h->far, h->farcnt, S::NBLK, S::BX * S::BY * N, N, fr, lists ? 1 : 0);
    } else {
        // a column lists at most its own BX*BY*N sites: kFarSub workgroups per column bound the work
        // of the fullest one
        t.launch(k_deposit_far_seg<Src, WIDE>, dim3(S::NBLK, kFarSub), kThreads, src, mesh, h->far, h->farcnt,
                 S::BX * S::BY * N, N, fr.sch);
    }
}

template <int N, bool WIDE, class Src>
void deposit_columns_v(Handle* h, const Src& src, unsigned long long* mesh, int v, bool lists) {
    switch (v) {
        case 1: deposit_columns<N, WIDE, 1>(h, src, mesh, lists); break;
        case 2: deposit_columns<N, WIDE, 2>(h, src, mesh, lists); break;
        case 3: deposit_columns<N, WIDE, 3>(h, src, mesh, lists); break;
        default: deposit_columns<N, WIDE, 0>(h, src, mesh, lists); break;
    }
}

template <bool WIDE, class Src>
int deposit_columns_n(Handle* h, const Src& src, unsigned long long* mesh, int v, bool lists) {
    switch (h->N) {
        case 16: deposit_columns_v<16, WIDE>(h, src, mesh, v, lists); break;
        case 32: deposit_columns_v<32, WIDE>(h, src, mesh, v, lists); break;
        case 48: deposit_columns_v<48, WIDE>(h, src, mesh, v, lists); break;
        case 64: deposit_columns_v<64, WIDE>(h, src, mesh, v, lists); break;
        case 96: deposit_columns_v<96, WIDE>(h, src, mesh, v, lists); break;
        case 128: deposit_columns_v<128, WIDE>(h, src, mesh, v, lists); break;
        case 256: deposit_columns_v<256, WIDE>(h, src, mesh, v, lists); break;
        case 512: deposit_columns_v<512, WIDE>(h, src, mesh, v, lists); break;
        default: return fail(h, PM3D_EUNSUPPORTED, "column deposit needs N in [16,512]");
    }
    return 0;
}

// Column shape (BX, BY) of variant v for this handle's mesh size and cell width.
template <int N, bool WIDE>
void col_shape_v(int v, int& bx, int& by) {
    switch (v) {
        case 1: bx = ColShape<N, WIDE, 1>::BX, by = ColShape<N, WIDE, 1>::BY; break;
        case 2: bx = ColShape<N, WIDE, 2>::BX, by = ColShape<N, WIDE, 2>::BY; break;
        case 3: bx = ColShape<N, WIDE, 3>::BX, by = ColShape<N, WIDE, 3>::BY; break;
        default: bx = ColShape<N, WIDE, 0>::BX, by = ColShape<N, WIDE, 0>::BY; break;
    }
}
template <bool WIDE>
void col_shape_n(int N, int v, int& bx, int& by) {
    bx = by = 0;
    switch (N) {
        case 16: col_shape_v<16, WIDE>(v, bx, by); break;
        case 32: col_shape_v<32, WIDE>(v, bx, by); break;
        case 48: col_shape_v<48, WIDE>(v, bx, by); break;
        case 64: col_shape_v<64, WIDE>(v, bx, by); break;
        case 96: col_shape_v<96, WIDE>(v, bx, by); break;
        case 128: col_shape_v<128, WIDE>(v, bx, by); break;
        case 256: col_shape_v<256, WIDE>(v, bx, by); break;
        case 512: col_shape_v<512, WIDE>(v, bx, by); break;
        default: break;
    }
}
void col_shape(const Handle* h, int v, int& bx, int& by) {
    if (h->wide) col_shape_n<true>(h->N, v, bx, by);
    else col_shape_n<false>(h->N, v, bx, by);
}
bool columns_used(const Handle* h) { return h->Np == h->N && h->N >= 16; }

// Deposit of the device state into h->mesh.  The column path (Np == N, N >= 16) writes every cell;
// the atomic path clears the mesh first.  This deposit's maximum is stored into its ring slot by the
// y pass of the FFT.
// Deposit of the device state into `mesh` (h->mesh, or a scratch mesh of the same size).  `lists`:
// the far particles are in the row lists of the step's gather-kick-drift (only the step's own
// deposit of the state).  Every column deposit's far pass clears the row counts, so lists not
// consumed by this deposit are gone after it.
// `v`: column shape variant (h->depv when negative).
int deposit_into(Handle* h, unsigned long long* mesh, bool lists = false, int v = -1) {
    h->lists_fresh = false;
    if (v < 0) v = h->depv;
    const bool tiled = columns_used(h);  // a column plus its one-site ring must not wrap onto itself
    const int W = h->wide ? 2 : 1;
    if (!tiled) {
        Timed t(h, KC_FILL);
        k_fill_u64<<<grid_cap(W * h->M), kThreads, 0, h->st>>>(mesh, W * h->M, 0ull);
    }
    int rc = 0;
    if (!h->mixed) {
        if (tiled) {
            rc = h->wide ? deposit_columns_n<true>(h, src_f64(h), mesh, v, lists)
                         : deposit_columns_n<false>(h, src_f64c(h), mesh, v, lists);
        } else {
            Timed t(h, KC_DEPOSIT);
            if (h->wide)
                k_deposit_atomic<true><<<grid_for(h->n), kThreads, 0, h->st>>>(Strided{h->x, h->n, 1}, h->n, h->N, h->dx,
                                                                               mesh, h->qs, h->qone);
            else
                k_deposit_atomic<false><<<grid_for(h->n), kThreads, 0, h->st>>>(Strided{h->x, h->n, 1}, h->n, h->N,
                                                                                h->dx, mesh, h->qs, h->qone);
        }
    } else {
        if (tiled) {
            rc = deposit_columns_n<false>(h, src_mx(h), mesh, v, lists);
        } else {
            Timed t(h, KC_DEPOSIT);
            k_deposit_atomic_mx<<<grid_for(h->n), kThreads, 0, h->st>>>(mxpos(h), h->N, mesh);
        }
    }
    return rc ? rc : check_launch(h, "deposit");
}

// Reads the finished samples of the running trial (in launch order, without waiting) and, once all
// are in, takes the fastest candidate when it beats the current shape by 2 %.
void live_read(Handle* h) {
    LiveShape& t = h->live;
    while (t.unread > 0) {
        const int s = t.pos - t.unread;
        const hipError_t q = hipEventQuery(t.ev[2 * s + 1]);
        if (q != hipSuccess) {
            if (q == hipErrorNotReady) (void)hipGetLastError();  // not an error: leave no trace of it
            return;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.ev[2 * s], t.ev[2 * s + 1]) == hipSuccess && ms < t.best[t.plan[s]])
            t.best[t.plan[s]] = ms;
        --t.unread;
    }
    if (t.nplan && t.pos == t.nplan) {
        int best = h->depv;
        for (int v = 0; v < 4; ++v)
            if ((t.cand >> v & 1u) && t.best[v] < 0.98f * t.best[h->depv] && t.best[v] < t.best[best]) best = v;
        for (int v = 0; v < 4; ++v)  // timed this trial (best < 1e30) and far behind: rest
            if ((t.cand >> v & 1u) && v != best && t.best[v] < 1e30f && t.best[v] > kLiveRestMargin * t.best[best])
                t.rest[v] = kLiveRestTrials;
        t.changes += best != h->depv;
        h->depv = best;
        t.nplan = t.pos = 0;
        ++t.trials;
    }
}

// A step's own deposit of the state.  Every kLivePeriod of them (and at the first after a new state
// is set) a trial starts: kLiveReps deposits
// in each candidate shape, the current one last, each between an event pair on the handle's
// stream; live_read takes the result at later steps.
int live_deposit(Handle* h) {
    LiveShape& t = h->live;
    live_read(h);
    if (!t.nplan && t.deposits >= t.next) {
        t.next = t.deposits + kLivePeriod;
        for (int v = 0; v < 4; ++v) {
            t.best[v] = 1e30f;
            if (!(t.cand >> v & 1u) || v == h->depv) continue;
            if (t.rest[v] > 0) {
                --t.rest[v];
                continue;
            }
            for (int r = 0; r < kLiveReps; ++r) t.plan[t.nplan++] = v;
        }
        for (int r = 0; r < kLiveReps; ++r) t.plan[t.nplan++] = h->depv;
        t.pos = t.unread = 0;
    }
    ++t.deposits;
    if (!t.nplan || t.pos == t.nplan) return deposit_into(h, h->mesh, h->lists_fresh);
    const int s = t.pos;
    if (hipEventRecord(t.ev[2 * s], h->st) != hipSuccess) return fail(h, PM3D_EHIP, "hipEventRecord failed");
    int rc = deposit_into(h, h->mesh, h->lists_fresh, t.plan[s]);
    if (rc) return rc;
    if (hipEventRecord(t.ev[2 * s + 1], h->st) != hipSuccess) return fail(h, PM3D_EHIP, "hipEventRecord failed");
    ++t.pos;
    ++t.unread;
    return 0;
}

int deposit_state(Handle* h) {
    h->serial += 1;
    if (h->live_arm && __builtin_popcount(h->live.cand) > 1 && !h->use_sched) return live_deposit(h);
    return deposit_into(h, h->mesh, h->lists_fresh);
}

// Picks the column shape of the deposit (ColShape variants: 8x8, 4x8, 4x4, 2x4 columns, those that
// fit in LDS) by timing the column and far passes of each (3 interleaved rounds of 4 launches) on a
// jittered lattice written into this handle's position buffers (k_tune_lattice; pm3d_set_state or
// pm3d_make_peak_ic replaces it).  Every shape gives the same mesh bits, so the choice never changes
// a result.  The shapes within 25 % of the fastest stay candidates of the trials during a run
// (live_deposit).  Cached per (N, precision) for the process.  PM3D_DEPOSIT_SHAPE: "auto" (default),
// "static" (no trials during a run), "all" (every fitting shape a candidate), or one digit 0-3
// forcing the variant.
struct DepTuned {
    int v;
    unsigned cand;
};
std::map<std::pair<int, int>, DepTuned> g_dep_tuned;

// Event pairs of the live trials, when there are at least two candidates.
int live_init(Handle* h, unsigned cand) {
    h->live.cand = cand;
    if (__builtin_popcount(cand) < 2) return 0;
    for (auto& e : h->live.ev)
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            h->live.cand = 0;
            return fail(h, PM3D_EHIP, "hipEventCreate failed");
        }
    return 0;
}

int tune_deposit(Handle* h) {
    h->depv = 0;
    if (!columns_used(h)) return 0;
    const char* env = std::getenv("PM3D_DEPOSIT_SHAPE");
    const bool live = !(env && std::strcmp(env, "static") == 0);
    const bool all = env && std::strcmp(env, "all") == 0;
    if (env && *env && std::strcmp(env, "auto") != 0 && live && !all) {
        if (std::strlen(env) != 1 || env[0] < '0' || env[0] > '3')
            return fail(h, PM3D_EINVAL, "PM3D_DEPOSIT_SHAPE must be auto, static, all or one digit 0-3");
        h->depv = env[0] - '0';
        return 0;
    }
    // distinct shapes (variants past the last fitting shape repeat it)
    int nv = 0;
    for (int v = 0, pbx = -1, pby = -1; v < 4; ++v) {
        int bx, by;
        col_shape(h, v, bx, by);
        if (bx == pbx && by == pby) break;
        pbx = bx, pby = by, nv = v + 1;
    }
    const unsigned every = (1u << nv) - 1u;
    const std::pair<int, int> key = tune_key(h);
    {
        std::unique_lock<std::mutex> lk(g_tune_mu);
        auto it = g_dep_tuned.find(key);
        if (it != g_dep_tuned.end()) {
            const DepTuned t = it->second;
            lk.unlock();
            h->depv = t.v;
            return live_init(h, all ? every : live ? t.cand : 0u);
        }
    }
    k_tune_lattice<<<grid_for(h->n), kThreads, 0, h->st>>>(h->N, h->dx, h->mixed ? nullptr : h->x, h->cell, h->off);
    int rc = check_launch(h, "deposit shape timing state");
    if (rc) return rc;
    hipEvent_t e0, e1;
    HCHK(h, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return fail(h, PM3D_EHIP, "hipEventCreate failed");
    }
    auto launch = [&](int v) {
        return h->mixed  ? deposit_columns_n<false>(h, src_mx(h), h->mesh, v, false)
               : h->wide ? deposit_columns_n<true>(h, src_f64(h), h->mesh, v, false)
                         : deposit_columns_n<false>(h, src_f64c(h), h->mesh, v, false);
    };
    // three interleaved rounds of 4 launches per shape (after one warm-up each); a shape's time is
    // its fastest round, and a narrower shape replaces the widest only when 3 % faster
    float best_ms[4] = {1e30f, 1e30f, 1e30f, 1e30f};
    for (int v = 0; v < nv && !rc; ++v) rc = launch(v);
    for (int round = 0; round < 3 && !rc; ++round) {
        for (int v = 0; v < nv && !rc; ++v) {
            float ms = 0.f;
            if (hipEventRecord(e0, h->st) != hipSuccess) { rc = fail(h, PM3D_EHIP, "hipEventRecord failed"); break; }
            for (int r = 0; r < 4 && !rc; ++r) rc = launch(v);
            if (rc) break;
            if (hipEventRecord(e1, h->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                rc = fail(h, PM3D_EHIP, "deposit shape timing failed");
            if (!rc) rc = check_launch(h, "deposit shape timing");
            if (!rc && ms < best_ms[v]) best_ms[v] = ms;
        }
    }
    int best = 0;
    for (int v = 1; v < nv; ++v)
        if (best_ms[v] < 0.97f * best_ms[0] && best_ms[v] < best_ms[best]) best = v;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    unsigned cand = 0;
    for (int v = 0; v < nv; ++v)
        if (best_ms[v] <= 1.25f * best_ms[best]) cand |= 1u << v;
    h->depv = best;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_dep_tuned[key] = DepTuned{best, cand};
    }
    return live_init(h, all ? every : live ? cand : 0u);
}

// The FP64 step scalars with the reciprocals of their divisors (qdiv), from the host scalars of
// pm3d.py:118-127 and the pending closing kick of the previous step.
KickScalars kick_scalars(const Handle* h, const pm3d_step_scalars* sc) {
    const double dp = h->kick_pending ? h->pend_den : 1.0;
    return KickScalars{h->pend_c1, dp, 1.0 / dp, sc->c1, sc->den1, 1.0 / sc->den1, sc->da, sc->den_d, 1.0 / sc->den_d};
}

// 1/c^3 for the mixed Jacobian, c = N/Np cells per lattice site
double inv_c3(const Handle* h) {
    const double c = (double)h->N / h->Np;
    return 1.0 / (c * c * c);
}

// Gather blocks of kThreads consecutive particles, dealt to the XCDs in runs of 4 blocks (1024
// particles: a few rows of a Lagrangian x plane, whose corner rows then share an XCD's L2).  Longer
// runs hold up the collapse: with an eighth of a plane per run (32 blocks at 256^3) the waves of the
// collapsed core queue on its few gradient lines (L2 -> L1 return stalls 60x the early steps').
// Measured, mixed 256^3 / FP64 128^3 gather-kick at the crossing (us): run of 1 block 284 / 56.3,
// 2: 285 / 51.8, 4: 280 / 49.3, 8: 284 / 51.0, 16: 293 / 51.7, 32: 340 / -; early steps 228-234 /
// 47-50 for runs of 2-8.
XMap gather_map(Handle* h) {
    const unsigned G = grid_for(h->n);
    return xmap(G, G % (kXcds * 4) == 0 ? 4u : 1u);
}

// Force meshes at the current positions: deposit, Poisson, the three gradient meshes in h->grad.
int solve_state(Handle* h, double s, double sm2) {
    int rc;
    if ((rc = deposit_state(h))) return rc;
    if ((rc = fft_forward(h, h->mesh, true))) return rc;
    if ((rc = fft_xpass(h, 2, s, sm2))) return rc;
    return fft_inverse(h, h->gspec, 3, h->grad);
}

// Gather of h->grad at the current positions into acc_out (SoA 3n, state precision).
int gather_state(Handle* h, void* acc_out) {
    {
        Timed t(h, KC_GATHER);
        if (!h->mixed)
            k_gather3<<<grid_for(h->n), kThreads, 0, h->st>>>(h->x, h->n, h->N, h->dx, (const double*)h->grad,
                                                              (double*)acc_out, gather_map(h));
        else
            k_gather3_mx<<<grid_for(h->n), kThreads, 0, h->st>>>(mxpos(h), h->N, (const float*)h->grad, (float*)acc_out,
                                                               gather_map(h));
    }
    return check_launch(h, "gather");
}

// forces at the current positions: deposit, Poisson, gather into acc_out.
int forces_state(Handle* h, double s, double sm2, void* acc_out) {
    int rc = solve_state(h, s, sm2);
    return rc ? rc : gather_state(h, acc_out);
}

// A step leaves the acceleration as force meshes (grad_valid, !acc_valid); the next step gathers
// them in its first kernel.  Anything else that needs acc, or that replaces the force meshes or the
// positions, gathers them into acc first.
int settle_acc(Handle* h) {
    if (!h->grad_valid || h->acc_valid) return 0;
    int rc = gather_state(h, h->mixed ? (void*)h->accf : (void*)h->acc);
    if (!rc) h->acc_valid = true;
    return rc;
}

int flush_kick(Handle* h) {
    if (!h->kick_pending) return 0;
    if (int rc = settle_acc(h)) return rc;
    Timed t(h, KC_OTHER);
    if (!h->mixed)
        k_kick<<<grid_for(3 * h->n), kThreads, 0, h->st>>>(h->p, h->acc, 3 * h->n, h->pend_c1, h->pend_den,
                                                            1.0 / h->pend_den);
    else
        k_kick_mx<<<grid_for(3 * h->n), kThreads, 0, h->st>>>(h->pf, h->accf, 3 * h->n, (float)(h->pend_c1 / h->pend_den));
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
    void* bufs[] = {h->x, h->p, h->acc, h->cell, h->off, h->pf, h->accf, h->mesh, h->spec, h->gspec,
                    h->grad, h->tw, h->dscal, h->far, h->farcnt, h->rowcnt, h->bmax, h->sbuf, h->sched, h->tab, h->jhist,
                    h->an_idx, h->an_off, h->an_buf, h->an_cnt};
    for (void* b : bufs)
        if (b) hipFree(b);
    for (auto& u : h->ev_used) {
        hipEventDestroy(u.second.first);
        hipEventDestroy(u.second.second);
    }
    for (auto e : h->ev_pool) hipEventDestroy(e);
    for (auto e : h->live.ev)
        if (e) hipEventDestroy(e);
    if (h->htab) hipHostFree(h->htab);
    if (h->hsched) hipHostFree(h->hsched);
    if (h->hjmin) hipHostFree(h->hjmin);
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
int dalloc_bytes(Handle* h, void** ptr, size_t b) { return dalloc(h, (char**)ptr, b); }

// delta = rho/n_mean - 1 of the most recent deposit into FP64 device memory `out`.
void mesh_to_delta(Handle* h, double* out) {
    if (h->wide)
        k_fixed_to_real<true><<<grid_cap(h->M), kThreads, 0, h->st>>>(h->mesh, h->M, h->qscale, 1, h->nmean, out);
    else
        k_fixed_to_real<false><<<grid_cap(h->M), kThreads, 0, h->st>>>(h->mesh, h->M, h->qscale, 1, h->nmean, out);
}

// numpy-layout half spectrum (N,N,N/2+1) <-> device layout
template <typename R>
void spec_to_numpy(const std::vector<cplx<R>>& hs, int N, double* out_c) {
    const size_t NH = N / 2, specN = (size_t)N * N * NH;
    for (size_t r = 0; r < (size_t)N * N; ++r) {
        for (size_t k = 0; k < NH; ++k) {
            out_c[2 * (r * (NH + 1) + k)] = hs[r * NH + k].re;
            out_c[2 * (r * (NH + 1) + k) + 1] = hs[r * NH + k].im;
        }
        out_c[2 * (r * (NH + 1) + NH)] = hs[specN + r].re;
        out_c[2 * (r * (NH + 1) + NH) + 1] = hs[specN + r].im;
    }
}
template <typename R>
void numpy_to_spec(const double* in_c, int N, std::vector<cplx<R>>& hs) {
    const size_t NH = N / 2, specN = (size_t)N * N * NH;
    hs.resize(specN + (size_t)N * N);
    for (size_t r = 0; r < (size_t)N * N; ++r) {
        for (size_t k = 0; k < NH; ++k)
            hs[r * NH + k] = cplx<R>{(R)in_c[2 * (r * (NH + 1) + k)], (R)in_c[2 * (r * (NH + 1) + k) + 1]};
        hs[specN + r] = cplx<R>{(R)in_c[2 * (r * (NH + 1) + NH)], (R)in_c[2 * (r * (NH + 1) + NH) + 1]};
    }
}

template <typename R>
int rfftn_impl(Handle* h, const double* in, double* out_c) {
    std::vector<R> hin(h->M);
    for (long long i = 0; i < h->M; ++i) hin[i] = (R)in[i];
    int rc = ensure_sbuf(h, h->M * sizeof(double));
    if (rc) return rc;
    HCHK(h, hipMemcpyAsync(h->sbuf, hin.data(), h->M * sizeof(R), hipMemcpyHostToDevice, h->st));
    if ((rc = fft_forward(h, h->sbuf, false)) || (rc = fft_xpass(h, 0, 0, 0))) return rc;
    if ((rc = check_launch(h, "rfftn"))) return rc;
    const size_t specN = (size_t)h->N * h->N * (h->N / 2);
    std::vector<cplx<R>> hs(specN + (size_t)h->N * h->N);
    HCHK(h, hipMemcpyAsync(hs.data(), h->spec, hs.size() * sizeof(cplx<R>), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    spec_to_numpy(hs, h->N, out_c);
    return 0;
}

template <typename R>
int irfftn_impl(Handle* h, const double* in_c, double* out) {
    std::vector<cplx<R>> hs;
    numpy_to_spec(in_c, h->N, hs);
    int rc = ensure_sbuf(h, h->M * sizeof(double));
    if (rc) return rc;
    HCHK(h, hipMemcpyAsync(h->spec, hs.data(), hs.size() * sizeof(cplx<R>), hipMemcpyHostToDevice, h->st));
    if ((rc = fft_xpass(h, 1, 0, 0)) || (rc = fft_inverse(h, h->spec, 1, h->sbuf))) return rc;
    if ((rc = check_launch(h, "irfftn"))) return rc;
    std::vector<R> ho(h->M);
    HCHK(h, hipMemcpyAsync(ho.data(), h->sbuf, h->M * sizeof(R), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    for (long long i = 0; i < h->M; ++i) out[i] = (double)ho[i];
    return 0;
}

template <typename R>
int poisson_impl(Handle* h, const double* delta, double s, double sm2, double* g) {
    std::vector<R> hin(h->M);
    for (long long i = 0; i < h->M; ++i) hin[i] = (R)delta[i];
    int rc = ensure_sbuf(h, h->M * sizeof(double));
    if (rc) return rc;
    HCHK(h, hipMemcpyAsync(h->sbuf, hin.data(), h->M * sizeof(R), hipMemcpyHostToDevice, h->st));
    if ((rc = fft_forward(h, h->sbuf, false)) || (rc = fft_xpass(h, 2, s, sm2)) ||
        (rc = fft_inverse(h, h->gspec, 3, h->grad)))
        return rc;
    if ((rc = check_launch(h, "poisson_gradient"))) return rc;
    std::vector<R> ho(3 * h->M);
    HCHK(h, hipMemcpyAsync(ho.data(), h->grad, 3 * h->M * sizeof(R), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    for (long long i = 0; i < h->M; ++i)  // device component-major -> (N,N,N,3)
        for (int c = 0; c < 3; ++c) g[3 * i + c] = (double)ho[c * h->M + i];
    return 0;
}

// spectrum amplitude filter for the random field (FP64 handles only)
template <int N>
void spec_amp_n(Handle* h, double kcut) {
    using R = double;
    const size_t specN = (size_t)N * N * (N / 2);
    cplx<R>* sp = (cplx<R>*)h->spec;
    k_spec_amp<R><<<grid_cap((long long)specN + (long long)N * N), kThreads, 0, h->st>>>(sp, sp + specN, N,
                                                                                       1.0 / ((double)N * h->dx), kcut);
}

// Per-block minima of det(dx/dq) into h->bmax; returns the number of blocks.  Lattices with
// Np % 4 == 0 take four sites per thread (k_min_jacobian4, k_min_jacobian_mx4): FP64 128^3 31.2 ->
// 18.1 us, 96^3 16.3 -> 12.3 us (the one-site kernel's two IEEE divisions per entry were most of it).
unsigned launch_min_jacobian(Handle* h) {
    if (!h->mixed) {
        const double dq = h->L / h->Np;
        if (h->Np % 4 == 0) {
            const unsigned g = grid_for(h->n / 4);
            const MinImage mi{h->L, 0.49 * h->L, 0.51 * h->L, 1.49 * h->L};
            k_min_jacobian4<<<g, kThreads, 0, h->st>>>(h->x, (unsigned)h->n, mi, dq, 1.0 / dq, site_div(h->Np),
                                                       h->bmax);
            return g;
        }
        const unsigned g = grid_for(h->n);
        k_min_jacobian<<<g, kThreads, 0, h->st>>>(h->x, h->Np, h->L, dq, h->bmax);
        return g;
    }
    if (h->Np % 4 == 0) {
        const unsigned g = grid_for(h->n / 4);
        k_min_jacobian_mx4<<<g, kThreads, 0, h->st>>>(mxpos(h), h->N, inv_c3(h), site_div(h->Np), h->bmax);
        return g;
    }
    const unsigned g = grid_for(h->n);
    k_min_jacobian_mx<<<g, kThreads, 0, h->st>>>(mxpos(h), h->Np, h->N, inv_c3(h), h->bmax);
    return g;
}

// One step driven by the device schedule (for graph capture): the previous step's gather fused with
// the kicks and drift of the scheduled scalars, force meshes, Jacobian minimum, schedule advance.
// Launches only; no host synchronisation.
int launch_sched_step(Handle* h) {
    const FarRows fr = far_rows(h);
    if (!h->mixed)
        k_gather_kick_drift<<<grid_for(h->n), kThreads, 0, h->st>>>(h->x, h->p, h->n, h->N, h->dx, (const double*)h->grad,
                                                                    1, KickScalars{0, 1, 1, 0, 1, 1, 0, 1, 1}, h->L,
                                                                    h->dscal + 1, h->sched, h->tab, gather_map(h), fr,
                                                                    site_div(h->Np));
    else
        k_gather_kick_drift_mx<<<grid_for(h->n), kThreads, 0, h->st>>>(h->cell, h->off, h->pf, h->n, h->N,
                                                                       (const float*)h->grad, 1, 0, 0, 0, h->dscal + 1,
                                                                       h->sched, h->tab, gather_map(h), fr,
                                                                       site_div(h->Np));
    h->lists_fresh = fr.cnt != nullptr;
    h->use_sched = true;
    int rc = solve_state(h, 0.0, 0.0);
    h->use_sched = false;
    if (rc) return rc;
    const unsigned g = launch_min_jacobian(h);
    k_sched_advance<<<1, kThreads, 0, h->st>>>(h->sched, h->bmax, (int)g, h->jhist, h->dscal + 1);
    return check_launch(h, "scheduled step");
}

// Per-member pointers of a batch's status pass (k_batch_status).
struct BatchRef {
    const DevSched* sched;
    const double* jhist;
};
__global__ __launch_bounds__(kThreads) void k_batch_status(const BatchRef* __restrict__ refs, int count,
                                                           pm3d_member_status* __restrict__ out) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    const DevSched s = *refs[i].sched;
    pm3d_member_status o;
    o.done = s.counter;
    o.nsteps = s.nsteps;
    o.jmin = s.counter > 0 ? refs[i].jhist[s.counter - 1] : 1e300;
    out[i] = o;
}

struct Batch {
    hipStream_t st = nullptr;
    int steps_per_graph = 1;              // scheduled steps one graph launch advances (pm3d_batch_set_steps)
    BatchRef* refs = nullptr;             // device, one per member
    pm3d_member_status* stat = nullptr;   // pinned host memory the status pass writes
    int ref_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
    std::vector<hipStream_t> cls_st;   // one capture stream per mesh-size class (graph branches)
    std::vector<hipEvent_t> cls_join;
    std::vector<Handle*> members;
    std::vector<MxBox*> dbox;  // device descriptor arrays of the box-grouped classes
    std::string err;
};

// Members that share the box-grouped launches (batch_kernels.hpp): mixed precision, particles on
// the mesh lattice (column deposit), plane FFT.
bool box_groupable(const Handle* h) {
    if (!h->mixed || h->Np != h->N || h->N < 16 || !fft_plane_used(h)) return false;
    switch (h->N) {
        case 16: case 32: case 48: case 64: case 96: case 128: case 256: return true;
        default: return false;
    }
}

MxBox box_of(Handle* h) {
    const size_t specN = (size_t)h->N * h->N * (h->N / 2);
    MxBox b{};
    b.cell = h->cell;
    b.off = h->off;
    b.p = h->pf;
    b.acc = h->accf;
    b.n = h->n;
    b.N = h->N;
    b.Np = h->Np;
    b.dscal = h->dscal;
    b.sched = h->sched;
    b.tab = h->tab;
    b.jhist = h->jhist;
    b.mesh = h->mesh;
    b.far = h->far;
    b.farcnt = h->farcnt;
    b.spec = (cplx<float>*)h->spec;
    b.gspec = (cplx<float>*)h->gspec;
    b.grad = (float*)h->grad;
    b.tw = (const cplx<float>*)h->tw;
    b.bmax = h->bmax;
    b.fr = far_rows(h);
    b.qslot = qmax_slot(h);
    b.kval = 1.0 / ((double)h->N * h->dx);
    b.inv_c3 = inv_c3(h);
    b.sdv = site_div(h->Np);
    b.qscale = h->qscale;
    b.nmean = h->nmean;
    (void)specN;
    return b;
}

// Column deposit and far pass of nb boxes of mesh size N in column shape variant V.
template <int N, int V>
void launch_boxes_deposit(Handle* h0, const MxBox* dev, int nb, hipStream_t st) {
    using S = ColShape<N, false, V>;
    k_deposit_col_boxes<S::BX, S::BY, N><<<dim3(S::NBLK, nb), col_threads<SrcMx>(), 0, st>>>(dev, column_map<N, S::BY>(S::NBLK));
    if constexpr (S::NBLK <= kFarColsMax)
        k_deposit_far_bal_boxes<<<dim3(h0->rowcnt ? far_lists_grid(N) : far_bal_grid(h0->n), nb), kThreads, 0, st>>>(
            dev, S::NBLK, S::BX * S::BY * N);
    // (the grouped gather-kick-drift of the same replay built every member's row lists)
    else
        k_deposit_far_seg_boxes<<<dim3(S::NBLK, kFarSub, nb), kThreads, 0, st>>>(dev, S::BX * S::BY * N);
}

// One scheduled step of nb boxes of mesh size N in eight launches (the kernels and grids of
// launch_sched_step per box, with the boxes along the grid's y dimension).  The deposit's column
// shape and the x pass's tile variant are the ones the tuner chose for this mesh size (every member
// of the class shares them).
template <int N>
void launch_boxes_n(Handle* h0, const MxBox* dev, int nb, hipStream_t st) {
    const unsigned g = grid_for(h0->n);
    k_gather_kick_drift_mx_boxes<<<dim3(g, nb), kThreads, 0, st>>>(dev, gather_map(h0));
    switch (h0->depv) {
        case 1: launch_boxes_deposit<N, 1>(h0, dev, nb, st); break;
        case 2: launch_boxes_deposit<N, 2>(h0, dev, nb, st); break;
        case 3: launch_boxes_deposit<N, 3>(h0, dev, nb, st); break;
        default: launch_boxes_deposit<N, 0>(h0, dev, nb, st); break;
    }
    k_fwd_plane_boxes<N><<<dim3(2 * N, nb), PlaneTile<float, N>::NT, 0, st>>>(dev, xmap(2 * N, 2));
    switch (h0->fftv[FP_X]) {
#define PM3D_XBOX(V)                                                                                    \
    case V: {                                                                                           \
        constexpr int P = Plan<float, N, V>::PX;                                                        \
        const unsigned gx = N * (N / 2 / P) + N / P;                                                    \
        k_x_poisson_boxes<N, V><<<dim3(gx, nb), kThreads, 0, st>>>(                                     \
            dev, xmap(gx, Fft<float, N>::template pencil_share<P>()));                                  \
        break;                                                                                          \
    }
        PM3D_XBOX(1)
        PM3D_XBOX(2)
        PM3D_XBOX(3)
        default:
        PM3D_XBOX(0)
#undef PM3D_XBOX
    }
    {
        constexpr int Q = Fft<float, N>::kInvQ;
        k_inv_plane_boxes<N, Q><<<dim3(3 * Q * N, nb), InvTile<float, N, Q>::NT, 0, st>>>(dev, xmap(3 * Q * N, Q));
    }
    const unsigned g4 = grid_for(h0->n / 4);  // Np = N, a multiple of 4
    k_min_jacobian_mx_boxes<<<dim3(g4, nb), kThreads, 0, st>>>(dev);
    k_sched_advance_boxes<<<dim3(1, nb), kThreads, 0, st>>>(dev, (int)g4);
}

void launch_boxes(Handle* h0, const MxBox* dev, int nb, hipStream_t st) {
    switch (h0->N) {
        case 16: launch_boxes_n<16>(h0, dev, nb, st); break;
        case 32: launch_boxes_n<32>(h0, dev, nb, st); break;
        case 48: launch_boxes_n<48>(h0, dev, nb, st); break;
        case 64: launch_boxes_n<64>(h0, dev, nb, st); break;
        case 96: launch_boxes_n<96>(h0, dev, nb, st); break;
        case 128: launch_boxes_n<128>(h0, dev, nb, st); break;
        default: launch_boxes_n<256>(h0, dev, nb, st); break;
    }
}

void batch_free_boxes(Batch* b) {
    for (MxBox* d : b->dbox) hipFree(d);
    b->dbox.clear();
}

void batch_clear_graph(Batch* b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

template <class Pos>
void launch_analyse(Handle* h, const Pos& pos, const AnRadii& radii) {
    const double L = h->L, hL = h->L / 2;
    const int ns = h->an_nsets - 1;
    double* centre = h->an_buf;
    double* out = h->an_buf + 4;
    double* part = out + 9 * ns;
    double* part2 = part + (size_t)kAnBlocks * 3 * ns;
    Timed t(h, KC_OTHER);
    k_an_centre<<<1, 64, 0, h->st>>>(pos, h->an_idx, (int)(h->an_off_host1), L, hL, centre);
    if (ns > 0) {
        const dim3 g(kAnBlocks, ns);
        k_an_sum<<<g, kThreads, 0, h->st>>>(pos, h->an_idx, h->an_off + 1, centre, L, hL, part);
        k_an_moments<<<g, kThreads, 0, h->st>>>(pos, h->an_idx, h->an_off + 1, centre, L, hL, part, part2);
        k_an_finish<<<ns, 64, 0, h->st>>>(h->an_off + 1, part, part2, out);
    }
    k_fill_u64<<<1, kThreads, 0, h->st>>>(h->an_cnt, kAnMaxRadii, 0ull);
    k_an_radii<<<grid_cap(h->n), kThreads, 0, h->st>>>(pos, h->n, centre, L, hL, radii, h->an_cnt);
}

// ------------------------------------------------------- allocation-free deposit (scratch array)
// pm3d_deposit_noalloc keeps its 128-bit fixed-point mesh in this file-scope device array instead
// of memory the library allocates: 2 x 128^3 words (32 MiB), held by every GPU the library runs on
// until the process ends.  In host code `scratch` is only a symbol: its device address on each GPU
// comes from hipGetSymbolAddress with that GPU current (scratch_on_device), and kernels and fills
// receive that address, never the array's name.
constexpr long long kScratchCells = 128LL * 128 * 128;  // largest mesh: N = 128
__device__ __align__(256) unsigned long long scratch[2 * kScratchCells];

std::mutex g_scratch_mu;
std::map<int, unsigned long long*> g_scratch_addr;  // device ordinal -> address of `scratch` there

int scratch_on_device(unsigned long long** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        g_create_error = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return PM3D_EHIP;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(dev);
    if (it == g_scratch_addr.end()) {
        unsigned long long* ptr = nullptr;
        e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            g_create_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return PM3D_EHIP;
        }
        it = g_scratch_addr.emplace(dev, ptr).first;
    }
    *out = it->second;
    return 0;
}

// ------------------------------------------------------------ attainable-bandwidth probe
// Streaming kernels of the probe: one pass over the buffer, every thread moving U 16-byte vectors
// (all loads issued before the stores) with non-temporal loads and stores.  Measured on MI355X
// (1 GiB copy, GB/s): grid-stride 8192 workgroups 5540 (U = 4), one pass 5860, one pass
// non-temporal 6360; U = 8 and U = 1 were slower.
typedef unsigned int probe_v4 __attribute__((ext_vector_type(4)));
constexpr int kProbeU = 4;
__global__ __launch_bounds__(kThreads) void k_probe_copy(const probe_v4* __restrict__ src, probe_v4* __restrict__ dst,
                                                         long long n16) {
    const long long b = (long long)blockIdx.x * kThreads * kProbeU + threadIdx.x;
    probe_v4 v[kProbeU];
#pragma unroll
    for (int u = 0; u < kProbeU; ++u)
        if (b + u * kThreads < n16) v[u] = __builtin_nontemporal_load(src + b + u * kThreads);
#pragma unroll
    for (int u = 0; u < kProbeU; ++u)
        if (b + u * kThreads < n16) __builtin_nontemporal_store(v[u], dst + b + u * kThreads);
}
__global__ __launch_bounds__(kThreads) void k_probe_read(const probe_v4* __restrict__ src, long long n16,
                                                         unsigned* __restrict__ out) {
    const long long b = (long long)blockIdx.x * kThreads * kProbeU + threadIdx.x;
    unsigned acc = 0;
#pragma unroll
    for (int u = 0; u < kProbeU; ++u)
        if (b + u * kThreads < n16) {
            const probe_v4 v = __builtin_nontemporal_load(src + b + u * kThreads);
            acc ^= v.x ^ v.y ^ v.z ^ v.w;
        }
    if (acc == 0x9e3779b9u) out[blockIdx.x & 8191] = acc;  // keeps the loads; practically never stores
}

}  // namespace

// ================================================================================ C-ABI
extern "C" {

constexpr int kMaxNp = 812;
constexpr int kSchedHostCap = 512;  // schedule steps the pinned host table of a new handle holds

int pm3d_create(const pm3d_cfg* cfg, void** handle) {
    if (!cfg || !handle) {
        g_create_error = "null argument";
        return PM3D_EINVAL;
    }
    *handle = nullptr;
    const int N = cfg->N;
    int lg = 0;
    while ((1 << lg) < N) ++lg;
    if (!((N >= 8 && N <= 512 && (1 << lg) == N) || N == 48 || N == 96)) {
        g_create_error = "N must be a power of two in [8,512], or 48 or 96";
        return PM3D_EUNSUPPORTED;
    }
    // Np <= 812: every per-particle array of 8-byte elements stays under 4 GiB, so the particle
    // kernels address it with 32-bit byte offsets (n * 8 < 2^32)
    if (cfg->Np < 1 || cfg->Np > kMaxNp || !(cfg->L > 0.0) || !(cfg->smooth >= 0.0)) {
        g_create_error = "invalid Np (1 .. 812), L or smooth";
        return PM3D_EINVAL;
    }
    if (cfg->precision != PM3D_FP64 && cfg->precision != PM3D_MIXED) {
        g_create_error = "unknown precision mode";
        return PM3D_EINVAL;
    }
    const bool mixed = cfg->precision == PM3D_MIXED;
    // Mixed mode: 64-bit cells at 2^-31 (cannot overflow).  FP64 mode: 128-bit cells at 2^-52, or on
    // request (frac_bits = 40, n <= 2^23 so that no cell can overflow) 64-bit cells at 2^-40, whose
    // absolute error (about 1e-12 of a particle per cell) is far above FP64 rounding against the
    // density contrast of a nearly uniform state.
    const long long np3 = (long long)cfg->Np * cfg->Np * cfg->Np;
    const bool compact_ok = np3 <= kCompactMaxN;
    int fb = mixed ? kNarrowFB : kWideFB;
    if (cfg->frac_bits != 0) {
        const bool ok = mixed ? cfg->frac_bits == kNarrowFB
                              : cfg->frac_bits == kWideFB || (cfg->frac_bits == kCompactFB && compact_ok);
        if (!ok) {
            g_create_error = "frac_bits: mixed 31; FP64 52 (128-bit cells) or 40 (64-bit cells, at most 2^23 "
                             "particles); 0 = the mode's default (31, 52)";
            return PM3D_EINVAL;
        }
        fb = cfg->frac_bits;
    }
    Handle* h = new Handle();
    h->cfg = *cfg;
    h->cfg.frac_bits = fb;
    h->mixed = mixed;
    h->wide = fb == kWideFB;
    h->N = N;
    h->log2N = lg;
    h->Np = cfg->Np;
    h->n = (long long)cfg->Np * cfg->Np * cfg->Np;
    h->M = (long long)N * N * N;
    h->L = cfg->L;
    h->dx = cfg->L / N;                      // pm3d.py:30
    h->nmean = (double)h->n / (double)h->M;  // pm3d.py:103, Np**3/N**3
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
    const size_t rb = mixed ? 4 : 8;  // bytes per real
    const long long n3 = 3 * h->n;
    if (!mixed) {
        if ((rc = dalloc(h, &h->x, n3)) || (rc = dalloc(h, &h->p, n3)) || (rc = dalloc(h, &h->acc, n3))) return bail(rc);
    } else {
        if ((rc = dalloc(h, &h->cell, n3)) || (rc = dalloc(h, &h->off, n3)) || (rc = dalloc(h, &h->pf, n3)) ||
            (rc = dalloc(h, &h->accf, n3)))
            return bail(rc);
    }
    // per-block maxima / minima
    h->bmax_len = std::max<size_t>((size_t)N * N, (size_t)grid_for(h->n));
    if ((rc = dalloc(h, &h->mesh, (h->wide ? 2 : 1) * h->M)) || (rc = dalloc_bytes(h, &h->spec, specN * 2 * rb)) ||
        (rc = dalloc_bytes(h, &h->gspec, 3 * specN * 2 * rb)) || (rc = dalloc_bytes(h, &h->grad, 3 * h->M * rb)) ||
        (rc = dalloc_bytes(h, &h->tw, (size_t)N * 2 * rb)) || (rc = dalloc(h, &h->dscal, 8 + kRing)) ||
        (rc = dalloc(h, &h->bmax, h->bmax_len)) || (rc = dalloc(h, &h->far, (size_t)h->n)) ||
        (rc = dalloc(h, &h->farcnt, (size_t)N * N)))  // far count per column
        return bail(rc);
    // pinned host copies of a schedule (pm3d_set_schedule), allocated here so that setting a
    // schedule allocates nothing for tables of up to kSchedHostCap steps
    if (hipHostMalloc((void**)&h->hsched, sizeof(DevSched), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->htab, kSchedHostCap * sizeof(DevStep), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&h->hjmin, sizeof(double), hipHostMallocDefault) != hipSuccess) {
        h->err = "pinned host schedule buffers: allocation failed";
        return bail(PM3D_ENOMEM);
    }
    h->htab_cap = kSchedHostCap;
    // twiddles exp(-2 pi i k/N) from long-double cos/sin
    std::vector<double> tw(2 * N);
    std::vector<float> twf(2 * N);
    for (int k = 0; k < N; ++k) {
        const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)N;
        tw[2 * k] = (double)cosl(ang);
        tw[2 * k + 1] = (double)sinl(ang);
        twf[2 * k] = (float)cosl(ang);
        twf[2 * k + 1] = (float)sinl(ang);
    }
    if (hipMemcpy(h->tw, mixed ? (void*)twf.data() : (void*)tw.data(), N * 2 * rb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(h->dscal, 0, (8 + kRing) * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(mixed ? (void*)h->accf : (void*)h->acc, 0, n3 * rb) != hipSuccess ||
        hipMemset(h->mesh, 0, (h->wide ? 16 : 8) * (size_t)h->M) != hipSuccess ||
        hipMemset(h->spec, 0, specN * 2 * rb) != hipSuccess || hipMemset(h->gspec, 0, 3 * specN * 2 * rb) != hipSuccess ||
        hipMemset(h->farcnt, 0, (size_t)N * N * sizeof(unsigned int)) != hipSuccess) {
        h->err = "initial copies failed";
        return bail(PM3D_EHIP);
    }
    // far-particle row lists: column deposit with the balanced far pass (N <= 256); capacity per row
    // = what the acceleration buffer they alias holds, 6N (FP64: 24 B per particle) or 3N (mixed: 12
    // B) int entries for Np = N.  PM3D_ROW_CAP = 1 .. that overrides it (to exercise the overflow
    // path), 0 turns the lists off (every far particle through the global-atomic pass); all give the
    // same mesh bits.
    const int cap_max = (int)(3 * (size_t)h->n * (mixed ? sizeof(float) : sizeof(double)) / sizeof(int) /
                              ((size_t)N * N));
    int cap = (columns_used(h) && N <= 256) ? cap_max : 0;
    if (const char* env = std::getenv("PM3D_ROW_CAP")) {
        const int c = std::atoi(env);
        if (c < 0 || c > cap_max) {
            h->err = "PM3D_ROW_CAP must be in [0, 6N] (FP64) or [0, 3N] (mixed)";
            return bail(PM3D_EINVAL);
        }
        if (cap) cap = c;
    }
    if (cap) {
        h->row_cap = cap;
        if ((rc = dalloc(h, &h->rowcnt, (size_t)N * N))) return bail(rc);
        if (hipMemset(h->rowcnt, 0, (size_t)N * N * sizeof(unsigned int)) != hipSuccess) {
            h->err = "initial copies failed";
            return bail(PM3D_EHIP);
        }
    }
    if ((rc = tune_fft(h)) || (rc = tune_deposit(h))) return bail(rc);
    // the timing launches wrote a deposit maximum into the ring
    if (hipMemsetAsync(h->dscal, 0, (8 + kRing) * sizeof(unsigned long long), h->st) != hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess) {
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

int pm3d_device_info(void* handle, char* name, int name_len, char* arch, int arch_len, int64_t* total_bytes) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    int dev = 0;
    HCHK(h, hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HCHK(h, hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
        std::strncpy(name, prop.name, (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (arch && arch_len > 0) {
        std::strncpy(arch, prop.gcnArchName, (size_t)arch_len - 1);
        arch[arch_len - 1] = 0;
    }
    if (total_bytes) *total_bytes = (int64_t)prop.totalGlobalMem;
    return 0;
}

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
        // the cached acceleration stays that of the old positions (as the reference's _acc)
        if (!p && (rc = settle_acc(h))) return rc;
        h->grad_valid = false;
        HCHK(h, hipMemcpyAsync(h->sbuf, x, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->st));
        if (!h->mixed)
            k_aos_to_soa<<<grid_for(n), kThreads, 0, h->st>>>(h->sbuf, h->x, n);
        else
            k_set_pos_mx<<<grid_for(n), kThreads, 0, h->st>>>(h->sbuf, h->cell, h->off, n, h->N, h->dx);
    }
    if (p) {
        h->kick_pending = false;  // the host value of p already contains every kick
        HCHK(h, hipMemcpyAsync(h->sbuf, p, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->st));
        if (!h->mixed)
            k_aos_to_soa<<<grid_for(n), kThreads, 0, h->st>>>(h->sbuf, h->p, n);
        else
            k_f64_to_f32_soa<<<grid_for(n), kThreads, 0, h->st>>>(h->sbuf, h->pf, n);
    }
    HCHK(h, hipStreamSynchronize(h->st));
    h->state_set = true;
    h->live.next = h->live.deposits;  // a trial at the new state's first step, every candidate in it
    std::fill(std::begin(h->live.rest), std::end(h->live.rest), 0);
    return check_launch(h, "set_state");
}

int pm3d_get_state(void* handle, double* x, double* p) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->state_set) return fail(h, PM3D_ESTATE, "state not set");
    int rc = ensure_sbuf(h, 3 * h->n * sizeof(double));
    if (rc) return rc;
    if (x) {
        if (!h->mixed)
            k_soa_to_aos<<<grid_for(h->n), kThreads, 0, h->st>>>(h->x, h->sbuf, h->n);
        else
            k_get_pos_mx<<<grid_for(h->n), kThreads, 0, h->st>>>(h->cell, h->off, h->sbuf, h->n, h->dx, h->L);
        HCHK(h, hipMemcpyAsync(x, h->sbuf, 3 * h->n * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HCHK(h, hipStreamSynchronize(h->st));
    }
    if (p) {
        if ((rc = flush_kick(h))) return rc;
        if (!h->mixed)
            k_soa_to_aos<<<grid_for(h->n), kThreads, 0, h->st>>>(h->p, h->sbuf, h->n);
        else
            k_f32_soa_to_f64<<<grid_for(h->n), kThreads, 0, h->st>>>(h->pf, h->sbuf, h->n);
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
    if (!h->mixed && !h->wide && n > kCompactMaxN)
        return fail(h, PM3D_EINVAL, "more than 2^23 particles on 64-bit cells at 2^-40 (frac_bits = 52 gives 128-bit cells)");
    const size_t xb = 3 * (size_t)n * sizeof(double);
    int rc = ensure_sbuf(h, std::max(xb, (size_t)h->M * sizeof(double)));
    if (rc) return rc;
    const int W = h->wide ? 2 : 1;
    k_fill_u64<<<grid_cap(W * h->M), kThreads, 0, h->st>>>(h->mesh, W * h->M, 0ull);
    if (n > 0) {
        HCHK(h, hipMemcpyAsync(h->sbuf, X, xb, hipMemcpyHostToDevice, h->st));
        if (h->wide)
            k_deposit_atomic<true><<<grid_for(n), kThreads, 0, h->st>>>(Strided{h->sbuf, 1, 3}, n, h->N, h->dx, h->mesh,
                                                                        h->qs, h->qone);
        else
            k_deposit_atomic<false><<<grid_for(n), kThreads, 0, h->st>>>(Strided{h->sbuf, 1, 3}, n, h->N, h->dx, h->mesh,
                                                                         h->qs, h->qone);
    }
    if (h->wide)
        k_fixed_to_real<true><<<grid_cap(h->M), kThreads, 0, h->st>>>(h->mesh, h->M, h->qscale, 0, 1.0, h->sbuf);
    else
        k_fixed_to_real<false><<<grid_cap(h->M), kThreads, 0, h->st>>>(h->mesh, h->M, h->qscale, 0, 1.0, h->sbuf);
    HCHK(h, hipMemcpyAsync(rho, h->sbuf, h->M * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    h->serial += 1;  // the state's last deposit is no longer in the mesh
    return check_launch(h, "deposit");
}

int pm3d_deposit_noalloc(const double* X, int64_t n, int32_t N, double L, double* rho, void* stream) {
    if (n < 0 || (n > 0 && !X) || !rho || !(L > 0.0) || !std::isfinite(L)) {
        g_create_error = "pm3d_deposit_noalloc: invalid argument";
        return PM3D_EINVAL;
    }
    int lg = 0;
    while ((1 << lg) < N && lg < 30) ++lg;
    if (!((N >= 8 && (1 << lg) == N) || N == 48 || N == 96) || (long long)N * N * N > kScratchCells) {
        g_create_error = "pm3d_deposit_noalloc: N must be a power of two in [8,128], or 48 or 96 (scratch holds 128^3 cells)";
        return PM3D_EINVAL;
    }
    if (n > (1LL << 38)) {
        g_create_error = "pm3d_deposit_noalloc: more than 2^38 particles";
        return PM3D_EINVAL;
    }
    unsigned long long* mesh = nullptr;
    if (int rc = scratch_on_device(&mesh)) return rc;
    hipStream_t st = (hipStream_t)stream;
    const long long M = (long long)N * N * N;
    // the same kernels and constants as pm3d_deposit on an FP64 handle (2^-52 on 128-bit cells)
    k_fill_u64<<<grid_cap(2 * M), kThreads, 0, st>>>(mesh, 2 * M, 0ull);
    if (n > 0)
        k_deposit_atomic<true><<<grid_for(n), kThreads, 0, st>>>(Strided{X, 1, 3}, n, N, L / N, mesh,
                                                                  std::ldexp(1.0, kWideFB), 1LL << kWideFB);
    k_fixed_to_real<true><<<grid_cap(M), kThreads, 0, st>>>(mesh, M, std::ldexp(1.0, -kWideFB), 0, 1.0, rho);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_create_error = std::string("pm3d_deposit_noalloc: ") + hipGetErrorString(e);
        return PM3D_EHIP;
    }
    return 0;
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

int64_t pm3d_deposit_serial(void* handle) { return handle ? (int64_t)((Handle*)handle)->serial : -1; }

int pm3d_delta_max_at(void* handle, int64_t serial, double* dmax) {
    if (!handle || !dmax) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (serial < 1 || serial > h->serial || serial <= h->serial - kRing)
        return fail(h, PM3D_ESTATE, "deposit serial no longer held on the device");
    unsigned long long q = 0;
    HCHK(h, hipMemcpyAsync(&q, h->dscal + 8 + (serial % kRing), sizeof(q), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    double rmax;
    std::memcpy(&rmax, &q, sizeof(rmax));
    *dmax = rmax / h->nmean - 1.0;
    return 0;
}

int pm3d_get_delta(void* handle, double* delta) {
    if (!handle || !delta) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (h->serial < 1) return fail(h, PM3D_ESTATE, "no deposit yet");
    int rc = ensure_sbuf(h, h->M * sizeof(double));
    if (rc) return rc;
    mesh_to_delta(h, h->sbuf);
    HCHK(h, hipMemcpyAsync(delta, h->sbuf, h->M * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    return check_launch(h, "get_delta");
}

int pm3d_forces(void* handle, double a, double s, double sm2, double* acc, double* delta) {
    (void)a;
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->state_set) return fail(h, PM3D_ESTATE, "state not set");
    // the step's cached acceleration is left alone (pm3d.py:101-114 does not touch _acc): this
    // call gathers into staging memory
    const size_t rb = h->mixed ? 4 : 8;
    int rc = ensure_sbuf(h, 3 * h->n * rb + std::max<size_t>(3 * h->n, h->M) * sizeof(double));
    if (rc) return rc;
    double* stage = (double*)((char*)h->sbuf + 3 * h->n * rb);
    if ((rc = settle_acc(h))) return rc;  // the force meshes are about to be replaced
    h->grad_valid = false;
    if ((rc = forces_state(h, s, sm2, h->sbuf))) return rc;
    if (acc) {
        if (!h->mixed)
            k_soa_to_aos<<<grid_for(h->n), kThreads, 0, h->st>>>(h->sbuf, stage, h->n);
        else
            k_f32_soa_to_f64<<<grid_for(h->n), kThreads, 0, h->st>>>((const float*)h->sbuf, stage, h->n);
        HCHK(h, hipMemcpyAsync(acc, stage, 3 * h->n * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HCHK(h, hipStreamSynchronize(h->st));
    }
    if (delta) {
        mesh_to_delta(h, stage);
        HCHK(h, hipMemcpyAsync(delta, stage, h->M * sizeof(double), hipMemcpyDeviceToHost, h->st));
    }
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
    if (sc->need_forces || !(h->acc_valid || h->grad_valid)) {
        // the previous closing kick used the old acceleration: apply it before replacing it
        if ((rc = flush_kick(h))) return rc;
        if ((rc = solve_state(h, sc->s_old, sc->sm2))) return rc;
        h->acc_valid = false;
        h->grad_valid = true;
    }
    if (h->grad_valid) {
        // gather of the force meshes (pm3d.py:113, the previous step's or the solve above) +
        // pending closing kick of the previous step (:127) + opening kick (:122) + drift
        // (:123-124) in one pass; FP64 keeps the reference's two separate roundings of p
        Timed t(h, KC_GATHER_KICK, true);
        const FarRows fr = far_rows(h);
        if (!h->mixed)
            t.launch(k_gather_kick_drift, grid_for(h->n), kThreads, h->x, h->p, h->n, h->N, h->dx,
                     (const double*)h->grad, h->kick_pending ? 1 : 0, kick_scalars(h, sc), h->L, h->dscal + 1, nullptr,
                     nullptr, gather_map(h), fr, site_div(h->Np));
        else
            t.launch(k_gather_kick_drift_mx, grid_for(h->n), kThreads, h->cell, h->off, h->pf, h->n, h->N,
                     (const float*)h->grad, h->kick_pending ? 1 : 0,
                     (float)(h->kick_pending ? h->pend_c1 / h->pend_den : 0.0), (float)(sc->c1 / sc->den1),
                     (float)(sc->da / (sc->den_d * h->dx)), h->dscal + 1, nullptr, nullptr, gather_map(h), fr,
                     site_div(h->Np));
        h->lists_fresh = fr.cnt != nullptr;
    } else {
        // the stored acceleration (a state whose force meshes were replaced after it was gathered)
        Timed t(h, KC_KICK_DRIFT);
        if (!h->mixed)
            k_kick_drift<<<grid_for(h->n), kThreads, 0, h->st>>>(h->x, h->p, h->acc, h->n, h->kick_pending ? 1 : 0,
                                                                 kick_scalars(h, sc), h->L, h->dscal + 1, nullptr,
                                                                 nullptr);
        else
            k_kick_drift_mx<<<grid_for(h->n), kThreads, 0, h->st>>>(
                h->cell, h->off, h->pf, h->accf, h->n, h->N, h->kick_pending ? 1 : 0,
                (float)(h->kick_pending ? h->pend_c1 / h->pend_den : 0.0), (float)(sc->c1 / sc->den1),
                (float)(sc->da / (sc->den_d * h->dx)), h->dscal + 1, nullptr, nullptr);
    }
    h->kick_pending = false;
    h->acc_valid = false;
    h->grad_valid = false;
    if ((rc = check_launch(h, "kick_drift"))) return rc;
    // the gather and the closing kick are deferred: the force meshes stay in h->grad, and the next
    // step gathers them with its kicks (or they are gathered and the kick applied before p is read)
    h->live_arm = true;
    rc = solve_state(h, sc->s_new, sc->sm2);
    h->live_arm = false;
    if (rc) return rc;
    h->grad_valid = true;
    h->kick_pending = true;
    h->pend_c1 = sc->c1;
    h->pend_den = sc->den2;
    if (delta || delta_max) {
        if (delta && (rc = pm3d_get_delta(h, delta))) return rc;
        if (delta_max && (rc = pm3d_delta_max_at(h, h->serial, delta_max))) return rc;
        int bad = 0;
        if ((rc = read_bad(h, &bad))) return rc;
        return bad;
    }
    return 0;
}

int pm3d_rfftn(void* handle, const double* in, double* out_c) {
    if (!handle || !in || !out_c) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    return h->mixed ? rfftn_impl<float>(h, in, out_c) : rfftn_impl<double>(h, in, out_c);
}

int pm3d_irfftn(void* handle, const double* in_c, double* out) {
    if (!handle || !in_c || !out) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    return h->mixed ? irfftn_impl<float>(h, in_c, out) : irfftn_impl<double>(h, in_c, out);
}

int pm3d_poisson_gradient(void* handle, const double* delta, double s, double sm2, double* g) {
    if (!handle || !delta || !g) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (int rc = settle_acc(h)) return rc;  // the force meshes are about to be replaced
    h->grad_valid = false;
    return h->mixed ? poisson_impl<float>(h, delta, s, sm2, g) : poisson_impl<double>(h, delta, s, sm2, g);
}

int pm3d_min_jacobian(void* handle, double* jmin) {
    if (!handle || !jmin) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->state_set) return fail(h, PM3D_ESTATE, "state not set");
    const unsigned g = launch_min_jacobian(h);
    // the reduction writes the minimum straight into pinned host memory: no copy after it (a call
    // took about 70 us with an 8-byte copy into pageable memory behind it, whatever the mesh size)
    k_min_reduce<<<1, kThreads, 0, h->st>>>(h->bmax, (int)g, h->hjmin);
    HCHK(h, hipGetLastError());
    HCHK(h, hipStreamSynchronize(h->st));
    *jmin = *h->hjmin;
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

int pm3d_set_analysis_sets(void* handle, const int32_t* idx, const int64_t* offsets, int nsets) {
    if (!handle || !idx || !offsets || nsets < 1) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (offsets[0] != 0) return fail(h, PM3D_EINVAL, "offsets[0] must be 0");
    for (int s = 0; s < nsets; ++s)
        if (offsets[s + 1] <= offsets[s]) return fail(h, PM3D_EINVAL, "every analysis set must be non-empty");
    const long long tot = offsets[nsets];
    for (long long j = 0; j < tot; ++j)
        if (idx[j] < 0 || idx[j] >= h->n) return fail(h, PM3D_EINVAL, "analysis set index out of range");
    HCHK(h, hipStreamSynchronize(h->st));  // queued analysis work may still read the old sets
    void* old[] = {h->an_idx, h->an_off, h->an_buf, h->an_cnt};
    for (void* b : old)
        if (b) HCHK(h, hipFree(b));
    h->an_idx = nullptr;
    h->an_off = nullptr;
    h->an_buf = nullptr;
    h->an_cnt = nullptr;
    h->an_nsets = 0;
    const int ns = nsets - 1;
    int rc;
    if ((rc = dalloc(h, &h->an_idx, (size_t)tot)) || (rc = dalloc(h, &h->an_off, (size_t)nsets + 1)) ||
        (rc = dalloc(h, &h->an_buf, 4 + 9 * (size_t)ns + (size_t)kAnBlocks * 9 * ns + 1)) ||
        (rc = dalloc(h, &h->an_cnt, (size_t)kAnMaxRadii)))
        return rc;
    HCHK(h, hipMemcpy(h->an_idx, idx, (size_t)tot * sizeof(int), hipMemcpyHostToDevice));
    static_assert(sizeof(long long) == sizeof(int64_t), "offsets");
    HCHK(h, hipMemcpy(h->an_off, offsets, ((size_t)nsets + 1) * sizeof(long long), hipMemcpyHostToDevice));
    h->an_nsets = nsets;
    h->an_off_host1 = offsets[1];
    return 0;
}


int pm3d_analyse(void* handle, const double* radii, int nradii, double* centre, double* moments, int64_t* counts,
                 double* rho_max) {
    if (!handle || nradii < 0 || nradii > kAnMaxRadii || (nradii > 0 && (!radii || !counts))) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (h->an_nsets < 1) return fail(h, PM3D_ESTATE, "no analysis sets (pm3d_set_analysis_sets)");
    if (!h->state_set) return fail(h, PM3D_ESTATE, "no particle state");
    AnRadii rr{};
    rr.nr = nradii;
    for (int k = 0; k < nradii; ++k) rr.r[k] = radii[k];
    if (h->mixed)
        launch_analyse(h, AnPosMx{h->cell, h->off, h->n, h->dx, h->L}, rr);
    else
        launch_analyse(h, AnPosF64{h->x, h->n}, rr);
    int rc = check_launch(h, "analyse");
    if (rc) return rc;
    if (rho_max) {
        // maximum of a fresh deposit of the state, into the gradient spectra (free between steps), so
        // the density mesh of the last step stays readable
        unsigned long long* fresh = (unsigned long long*)h->gspec;
        if ((rc = deposit_into(h, fresh))) return rc;
        const unsigned nb = (unsigned)std::min<size_t>(1024, h->bmax_len);
        if (h->wide)
            k_an_mesh_max<true><<<nb, kThreads, 0, h->st>>>(fresh, h->M, h->qscale, h->bmax);
        else
            k_an_mesh_max<false><<<nb, kThreads, 0, h->st>>>(fresh, h->M, h->qscale, h->bmax);
        k_an_max_reduce<<<1, kThreads, 0, h->st>>>(h->bmax, (int)nb, h->an_buf + 3);
        if ((rc = check_launch(h, "analyse rho max"))) return rc;
    }
    const int ns = h->an_nsets - 1;
    std::vector<double> hb(4 + 9 * (size_t)ns);
    unsigned long long hc[kAnMaxRadii];
    HCHK(h, hipMemcpyAsync(hb.data(), h->an_buf, hb.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipMemcpyAsync(hc, h->an_cnt, sizeof(hc), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    if (centre) std::memcpy(centre, hb.data(), 3 * sizeof(double));
    if (moments && ns > 0) std::memcpy(moments, hb.data() + 4, 9 * (size_t)ns * sizeof(double));
    for (int k = 0; k < nradii; ++k) counts[k] = (int64_t)hc[k];
    if (rho_max) *rho_max = hb[3];
    return 0;
}

int pm3d_totals(void* handle, double* mass, double* mass_dev, double* momentum) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->state_set) return fail(h, PM3D_ESTATE, "state not set");
    if (h->serial < 1 && (mass || mass_dev)) return fail(h, PM3D_ESTATE, "no deposit yet");
    int rc = flush_kick(h);  // the momenta after the pending closing kick (same bits later on)
    if (rc) return rc;
    rc = ensure_sbuf(h, (size_t)kTotBlocks * 5 * sizeof(double));
    if (rc) return rc;
    unsigned long long* mpart = (unsigned long long*)h->sbuf;
    double* ppart = h->sbuf + 2 * kTotBlocks;
    if (h->wide)
        k_mesh_total<true><<<kTotBlocks, kThreads, 0, h->st>>>(h->mesh, h->M, mpart);
    else
        k_mesh_total<false><<<kTotBlocks, kThreads, 0, h->st>>>(h->mesh, h->M, mpart);
    if (h->mixed)
        k_momentum_total<float><<<kTotBlocks, kThreads, 0, h->st>>>(h->pf, h->n, ppart);
    else
        k_momentum_total<double><<<kTotBlocks, kThreads, 0, h->st>>>(h->p, h->n, ppart);
    if ((rc = check_launch(h, "totals"))) return rc;
    std::vector<unsigned long long> hm(2 * kTotBlocks);
    std::vector<double> hp(3 * kTotBlocks);
    HCHK(h, hipMemcpyAsync(hm.data(), mpart, hm.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipMemcpyAsync(hp.data(), ppart, hp.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    unsigned __int128 tot = 0;
    for (int b = 0; b < kTotBlocks; ++b) tot += ((unsigned __int128)hm[2 * b + 1] << 64) | hm[2 * b];
    const int fb = h->cfg.frac_bits;
    // n particles of unit weight would total exactly n * 2^fb
    const unsigned __int128 ideal = (unsigned __int128)h->n << fb;
    const bool over = tot >= ideal;
    const unsigned __int128 dev = over ? tot - ideal : ideal - tot;
    auto to_double = [fb](unsigned __int128 v) {
        return std::ldexp((double)(unsigned long long)(v >> 64), 64 - fb) + std::ldexp((double)(unsigned long long)v, -fb);
    };
    if (mass) *mass = to_double(tot);
    if (mass_dev) *mass_dev = over ? to_double(dev) : -to_double(dev);
    if (momentum) {
        double s[3] = {0.0, 0.0, 0.0};
        for (int b = 0; b < kTotBlocks; ++b)
            for (int d = 0; d < 3; ++d) s[d] += hp[3 * b + d];
        for (int d = 0; d < 3; ++d) momentum[d] = s[d];
    }
    return 0;
}

int pm3d_stream_probe(int64_t bytes, int reps, double* copy_gbs, double* read_gbs) {
    if (bytes < (1 << 20) || bytes > (int64_t)1 << 36 || reps < 1 || !copy_gbs || !read_gbs) {
        g_create_error = "pm3d_stream_probe: bytes in [1 MiB, 64 GiB], reps >= 1";
        return PM3D_EINVAL;
    }
    const long long n16 = bytes / 16;
    probe_v4 *a = nullptr, *b = nullptr;
    unsigned* sink = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    const unsigned grid = (unsigned)((n16 + kThreads * kProbeU - 1) / (kThreads * kProbeU));  // one pass
    float ms_copy = 0.f, ms_read = 0.f;
    auto ok = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && !rc) {
            g_create_error = std::string("pm3d_stream_probe: ") + what + ": " + hipGetErrorString(e);
            rc = PM3D_EHIP;
        }
        return !rc;
    };
    if (ok(hipMalloc(&a, (size_t)n16 * 16), "hipMalloc") && ok(hipMalloc(&b, (size_t)n16 * 16), "hipMalloc") &&
        ok(hipMalloc(&sink, 8192 * sizeof(unsigned)), "hipMalloc") &&
        ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream") && ok(hipEventCreate(&e0), "event") &&
        ok(hipEventCreate(&e1), "event") && ok(hipMemsetAsync(a, 1, (size_t)n16 * 16, st), "memset")) {
        k_probe_copy<<<grid, kThreads, 0, st>>>(a, b, n16);  // warm-up
        k_probe_read<<<grid, kThreads, 0, st>>>(a, n16, sink);
        ok(hipEventRecord(e0, st), "record");
        for (int r = 0; r < reps; ++r) k_probe_copy<<<grid, kThreads, 0, st>>>(a, b, n16);
        ok(hipEventRecord(e1, st), "record");
        ok(hipEventSynchronize(e1), "sync");
        ok(hipEventElapsedTime(&ms_copy, e0, e1), "elapsed");
        ok(hipEventRecord(e0, st), "record");
        for (int r = 0; r < reps; ++r) k_probe_read<<<grid, kThreads, 0, st>>>(a, n16, sink);
        ok(hipEventRecord(e1, st), "record");
        ok(hipEventSynchronize(e1), "sync");
        ok(hipEventElapsedTime(&ms_read, e0, e1), "elapsed");
        ok(hipGetLastError(), "launch");
    }
    if (st) hipStreamSynchronize(st);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    if (a) hipFree(a);
    if (b) hipFree(b);
    if (sink) hipFree(sink);
    if (rc) return rc;
    *copy_gbs = 2.0 * (double)n16 * 16 * reps / (ms_copy * 1e-3) / 1e9;
    *read_gbs = (double)n16 * 16 * reps / (ms_read * 1e-3) / 1e9;
    return 0;
}

int pm3d_fft_tiles(void* handle, int32_t* variants, int nmax) {
    if (!handle || !variants || nmax < 0) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    for (int p = 0; p < FP_COUNT && p < nmax; ++p) variants[p] = h->fftv[p];
    return FP_COUNT;
}

int pm3d_deposit_tiles(void* handle, int32_t* shape) {
    if (!handle || !shape) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    shape[0] = shape[1] = shape[2] = 0;
    if (!columns_used(h)) return 0;
    int bx, by;
    col_shape(h, h->depv, bx, by);
    shape[0] = bx;
    shape[1] = by;
    shape[2] = h->N;
    return 1;
}

int pm3d_frac_bits(void* handle) { return handle ? ((Handle*)handle)->cfg.frac_bits : PM3D_EINVAL; }

int pm3d_deposit_trials(void* handle, int64_t* out) {
    if (!handle || !out) return PM3D_EINVAL;
    const LiveShape& t = ((Handle*)handle)->live;
    out[0] = t.trials;
    out[1] = t.changes;
    out[2] = t.cand;
    return 0;
}

int pm3d_make_peak_ic(void* handle, const pm3d_peak_ic* ic) {
    if (!handle || !ic) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (h->mixed) return fail(h, PM3D_EUNSUPPORTED, "initial conditions are built on an FP64 handle (copy with pm3d_copy_state)");
    const long long M = h->M;
    const int nb = (int)grid_cap(M);
    int rc = ensure_sbuf(h, (2 * M + 2 * (size_t)nb) * sizeof(double));
    if (rc) return rc;
    double* field = h->sbuf;
    double* noise = h->sbuf + M;
    double* mom = h->sbuf + 2 * M;
    k_peak_field<<<nb, kThreads, 0, h->st>>>(field, h->N, h->dx, h->L, ic->mu, ic->sx, ic->sy, ic->sz);
    if (ic->noise_rms > 0.0) {
        // white noise -> rfftn -> * |k|^-1 (0 < |k| <= kcut) -> irfftn -> scaled to rms noise_rms
        k_white<<<nb, kThreads, 0, h->st>>>(noise, M, splitmix64(ic->seed));
        if ((rc = fft_forward(h, noise, false)) || (rc = fft_xpass(h, 0, 0, 0))) return rc;
        switch (h->N) {
            case 8: spec_amp_n<8>(h, ic->noise_kcut); break;
            case 16: spec_amp_n<16>(h, ic->noise_kcut); break;
            case 32: spec_amp_n<32>(h, ic->noise_kcut); break;
            case 48: spec_amp_n<48>(h, ic->noise_kcut); break;
            case 64: spec_amp_n<64>(h, ic->noise_kcut); break;
            case 96: spec_amp_n<96>(h, ic->noise_kcut); break;
            case 128: spec_amp_n<128>(h, ic->noise_kcut); break;
            case 256: spec_amp_n<256>(h, ic->noise_kcut); break;
            case 512: spec_amp_n<512>(h, ic->noise_kcut); break;
        }
        if ((rc = fft_xpass(h, 1, 0, 0)) || (rc = fft_inverse(h, h->spec, 1, noise))) return rc;
        k_moments<<<nb, kThreads, 0, h->st>>>(noise, M, mom);
        std::vector<double> hm(2 * nb);
        HCHK(h, hipMemcpyAsync(hm.data(), mom, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HCHK(h, hipStreamSynchronize(h->st));
        double s1 = 0.0, s2 = 0.0;
        for (int b = 0; b < nb; ++b) {
            s1 += hm[2 * b];
            s2 += hm[2 * b + 1];
        }
        const double mean = s1 / M, var = s2 / M - mean * mean;
        if (!(var > 0.0)) return fail(h, PM3D_EINVAL, "random field has zero variance (empty band)");
        k_axpy<<<nb, kThreads, 0, h->st>>>(field, noise, M, ic->noise_rms / std::sqrt(var));
    }
    // Zel'dovich displacement psi = irfftn(1j*k*dk/k2) on the mesh, then the lattice placement
    if ((rc = fft_forward(h, field, false)) || (rc = fft_xpass(h, 3, 0, 0)) || (rc = fft_inverse(h, h->gspec, 3, h->grad)))
        return rc;
    k_lattice<<<grid_for(h->n), kThreads, 0, h->st>>>((const double*)h->grad, h->N, h->Np, h->L, ic->coef, h->x, h->p);
    HCHK(h, hipStreamSynchronize(h->st));
    h->state_set = true;
    h->live.next = h->live.deposits;  // a trial at the new state's first step, every candidate in it
    std::fill(std::begin(h->live.rest), std::end(h->live.rest), 0);
    h->acc_valid = false;
    h->grad_valid = false;
    h->kick_pending = false;
    return check_launch(h, "make_peak_ic");
}

int pm3d_copy_state(void* dst_handle, void* src_handle) {
    if (!dst_handle || !src_handle) return PM3D_EINVAL;
    Handle* d = (Handle*)dst_handle;
    Handle* s = (Handle*)src_handle;
    if (s->mixed || !s->state_set) return fail(d, PM3D_EINVAL, "source must be an FP64 handle with a state");
    if (s->n != d->n || s->N != d->N || s->L != d->L) return fail(d, PM3D_EINVAL, "source and destination differ in N, Np or L");
    int rc = 0;
    HCHK(d, hipStreamSynchronize(s->st));
    if (!d->mixed) {
        HCHK(d, hipMemcpyAsync(d->x, s->x, 3 * s->n * sizeof(double), hipMemcpyDeviceToDevice, d->st));
        HCHK(d, hipMemcpyAsync(d->p, s->p, 3 * s->n * sizeof(double), hipMemcpyDeviceToDevice, d->st));
    } else {
        k_soa_to_mx<<<grid_for(d->n), kThreads, 0, d->st>>>(s->x, s->p, d->n, d->N, d->dx, d->cell, d->off, d->pf);
    }
    HCHK(d, hipStreamSynchronize(d->st));
    d->state_set = true;
    d->live.next = d->live.deposits;  // a trial at the new state's first step, every candidate in it
    std::fill(std::begin(d->live.rest), std::end(d->live.rest), 0);
    d->acc_valid = false;
    d->grad_valid = false;
    d->kick_pending = false;
    (void)rc;
    return check_launch(d, "copy_state");
}

int pm3d_set_schedule(void* handle, const pm3d_sched_entry* table, int nsteps) {
    if (!handle || !table || nsteps < 1) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->state_set || !h->grad_valid || !h->kick_pending)
        return fail(h, PM3D_ESTATE, "a schedule continues a stepped state: run one pm3d_step first");
    if (h->sched_upload) {  // the previous table's copies read htab / hsched
        HCHK(h, hipStreamSynchronize(h->st));
        h->sched_upload = false;
    }
    if (nsteps > h->tab_cap) {
        HCHK(h, hipStreamSynchronize(h->st));
        if (h->tab) HCHK(h, hipFree(h->tab));
        if (h->jhist) HCHK(h, hipFree(h->jhist));
        h->tab = nullptr;
        h->jhist = nullptr;
        h->tab_cap = 0;
        int rc;
        if ((rc = dalloc(h, &h->tab, (size_t)nsteps)) || (rc = dalloc(h, &h->jhist, (size_t)nsteps))) return rc;
        h->tab_cap = nsteps;
    }
    if (nsteps > h->htab_cap) {  // a table longer than the pinned copy made at creation
        if (h->htab) HCHK(h, hipHostFree(h->htab));
        h->htab_cap = 0;
        if (hipHostMalloc((void**)&h->htab, (size_t)nsteps * sizeof(DevStep), hipHostMallocDefault) != hipSuccess) {
            h->htab = nullptr;
            return fail(h, PM3D_ENOMEM, "pinned host schedule table allocation failed");
        }
        h->htab_cap = nsteps;
    }
    if (!h->sched) {
        int rc;
        if ((rc = dalloc(h, &h->sched, 1))) return rc;
    }
    for (int k = 0; k < nsteps; ++k) {
        const pm3d_sched_entry& e = table[k];
        DevStep& d = h->htab[k];
        d.c1p = e.c1p; d.denp = e.denp; d.c1 = e.c1; d.den1 = e.den1; d.da = e.da; d.den_d = e.den_d;
        d.den2 = e.den2; d.s = e.s; d.sm2 = e.sm2;
        d.rdenp = 1.0 / e.denp; d.rden1 = 1.0 / e.den1; d.rden_d = 1.0 / e.den_d;
        d.kp = (float)(e.c1p / e.denp);
        d.k1 = (float)(e.c1 / e.den1);
        d.sd = (float)(e.da / (e.den_d * h->dx));
        d.pad = 0.f;
    }
    // the copies read the handle's host table and counter, which stay untouched until the next
    // pm3d_set_schedule waits for them; no synchronisation here, so the tables of many members
    // upload while their first steps run
    *h->hsched = DevSched{0, nsteps, 0, 0};
    HCHK(h, hipMemcpyAsync(h->tab, h->htab, nsteps * sizeof(DevStep), hipMemcpyHostToDevice, h->st));
    HCHK(h, hipMemcpyAsync(h->sched, h->hsched, sizeof(DevSched), hipMemcpyHostToDevice, h->st));
    h->sched_upload = true;
    return 0;
}

int pm3d_schedule_status(void* handle, int* done, double* jhist) {
    if (!handle || !done) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->sched) return fail(h, PM3D_ESTATE, "no schedule");
    // on the handle's stream: ordered after its schedule upload and its own scheduled steps (a
    // batch's replays are synchronised by pm3d_batch_status / pm3d_batch_synchronize)
    DevSched hs;
    HCHK(h, hipMemcpyAsync(&hs, h->sched, sizeof(hs), hipMemcpyDeviceToHost, h->st));
    HCHK(h, hipStreamSynchronize(h->st));
    *done = hs.counter;
    if (jhist && hs.counter > 0) {
        HCHK(h, hipMemcpyAsync(jhist, h->jhist, hs.counter * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HCHK(h, hipStreamSynchronize(h->st));
    }
    return 0;
}

int pm3d_schedule_end(void* handle, int* done) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    int k = 0;
    int rc = pm3d_schedule_status(handle, &k, nullptr);
    if (rc) return rc;
    if (k > 0) {  // the gather and closing kick of the last scheduled step are pending
        h->grad_valid = true;
        h->acc_valid = false;
        h->kick_pending = true;
        h->pend_c1 = h->htab[k - 1].c1;
        h->pend_den = h->htab[k - 1].den2;
    }
    if (done) *done = k;
    return 0;
}

int pm3d_schedule_step(void* handle) {
    if (!handle) return PM3D_EINVAL;
    Handle* h = (Handle*)handle;
    if (!h->sched) return fail(h, PM3D_ESTATE, "no schedule");
    return launch_sched_step(h);
}

int pm3d_batch_create(void** batch) {
    if (!batch) return PM3D_EINVAL;
    Batch* b = new Batch();
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
        g_create_error = "batch stream/event creation failed";
        delete b;
        return PM3D_EHIP;
    }
    *batch = b;
    return 0;
}

int pm3d_batch_destroy(void* batch) {
    if (!batch) return PM3D_EINVAL;
    Batch* b = (Batch*)batch;
    hipStreamSynchronize(b->st);
    batch_clear_graph(b);
    batch_free_boxes(b);
    if (b->refs) hipFree(b->refs);
    if (b->stat) hipHostFree(b->stat);
    for (auto e : b->joins) hipEventDestroy(e);
    for (auto e : b->cls_join) hipEventDestroy(e);
    for (auto cs : b->cls_st) hipStreamDestroy(cs);
    hipEventDestroy(b->fork);
    hipStreamDestroy(b->st);
    delete b;
    return 0;
}

const char* pm3d_batch_last_error(void* batch) { return batch ? ((Batch*)batch)->err.c_str() : ""; }

int pm3d_batch_set(void* batch, void* const* handles, int count) {
    if (!batch || count < 0 || (count > 0 && !handles)) return PM3D_EINVAL;
    Batch* b = (Batch*)batch;
    hipStreamSynchronize(b->st);
    batch_clear_graph(b);
    batch_free_boxes(b);
    b->members.clear();
    for (int i = 0; i < count; ++i) {
        Handle* h = (Handle*)handles[i];
        if (!h || !h->sched) {
            b->err = "every member needs a schedule (pm3d_set_schedule)";
            return PM3D_ESTATE;
        }
        if (hipStreamSynchronize(h->st) != hipSuccess) {
            b->err = "member stream failed";
            return PM3D_EHIP;
        }
        b->members.push_back(h);
    }
    if (count == 0) return 0;
    if (count > b->ref_cap) {
        if (b->refs) hipFree(b->refs);
        if (b->stat) hipHostFree(b->stat);
        b->refs = nullptr;
        b->stat = nullptr;
        b->ref_cap = 0;
        if (hipMalloc((void**)&b->refs, count * sizeof(BatchRef)) != hipSuccess ||
            hipHostMalloc((void**)&b->stat, count * sizeof(pm3d_member_status), hipHostMallocDefault) != hipSuccess) {
            b->err = "allocation of the batch status buffers failed";
            return PM3D_ENOMEM;
        }
        b->ref_cap = count;
    }
    {
        std::vector<BatchRef> hr(count);
        for (int i = 0; i < count; ++i) hr[i] = BatchRef{b->members[i]->sched, b->members[i]->jhist};
        if (hipMemcpy(b->refs, hr.data(), count * sizeof(BatchRef), hipMemcpyHostToDevice) != hipSuccess) {
            b->err = "copy of the batch status pointers failed";
            return PM3D_EHIP;
        }
    }
    while ((int)b->joins.size() < count) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            b->err = "event creation failed";
            return PM3D_EHIP;
        }
        b->joins.push_back(e);
    }
    // members of one mesh size that can share launches: their descriptors, uploaded before capture
    std::map<int, std::vector<int>> classes;
    std::vector<bool> grouped(count, false);
    for (int i = 0; i < count; ++i)
        if (box_groupable(b->members[i])) classes[b->members[i]->N].push_back(i);
    while (b->cls_st.size() < classes.size()) {
        hipStream_t cs;
        hipEvent_t ce;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
            b->err = "stream creation failed";
            return PM3D_EHIP;
        }
        b->cls_st.push_back(cs);
        if (hipEventCreateWithFlags(&ce, hipEventDisableTiming) != hipSuccess) {
            b->err = "event creation failed";
            return PM3D_EHIP;
        }
        b->cls_join.push_back(ce);
    }
    std::vector<std::pair<int, MxBox*>> cls_dev;  // (first member, descriptors)
    for (auto& kv : classes) {
        std::vector<MxBox> hb;
        for (int i : kv.second) {
            b->members[i]->serial += 1;  // as deposit_state does in a one-box capture: the replays' deposit
            hb.push_back(box_of(b->members[i]));
            grouped[i] = true;
        }
        MxBox* d = nullptr;
        if (hipMalloc((void**)&d, hb.size() * sizeof(MxBox)) != hipSuccess) {
            b->err = "hipMalloc of batch descriptors failed";
            return PM3D_ENOMEM;
        }
        b->dbox.push_back(d);
        if (hipMemcpy(d, hb.data(), hb.size() * sizeof(MxBox), hipMemcpyHostToDevice) != hipSuccess) {
            b->err = "copy of batch descriptors failed";
            return PM3D_EHIP;
        }
        cls_dev.push_back({kv.second[0], d});
    }
    // capture one step: one set of launches per class on the batch stream, and every other member as
    // an independent branch on its own stream
    if (hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed) != hipSuccess) {
        b->err = "hipStreamBeginCapture failed";
        return PM3D_EHIP;
    }
    hipError_t e = hipEventRecord(b->fork, b->st);
    int rc = 0;
    // each class is its own branch of the graph, so its kernels may run beside another class's
    // (64-peak set: 201 -> 209 solves/s); PM3D_BATCH_BRANCHES=0 runs the classes one after the other
    const char* br = std::getenv("PM3D_BATCH_BRANCHES");
    const bool branches = !(br && br[0] == '0');
    for (size_t c = 0; c < cls_dev.size() && e == hipSuccess; ++c) {
        Handle* h0 = b->members[cls_dev[c].first];
        if (branches) {
            hipStream_t cs = b->cls_st[c];
            e = hipStreamWaitEvent(cs, b->fork, 0);
            // a class runs its steps of the launch one after the other without waiting for the
            // other classes between them
            for (int k = 0; k < b->steps_per_graph && e == hipSuccess; ++k) {
                launch_boxes(h0, cls_dev[c].second, (int)classes[h0->N].size(), cs);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipEventRecord(b->cls_join[c], cs);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->st, b->cls_join[c], 0);
        } else {
            for (int k = 0; k < b->steps_per_graph && e == hipSuccess; ++k) {
                launch_boxes(h0, cls_dev[c].second, (int)classes[h0->N].size(), b->st);
                e = hipGetLastError();
            }
        }
    }
    for (int i = 0; i < count && e == hipSuccess && rc == 0; ++i) {
        if (grouped[i]) continue;
        Handle* h = b->members[i];
        const bool timing = h->timing;
        h->timing = false;
        e = hipStreamWaitEvent(h->st, b->fork, 0);
        for (int k = 0; k < b->steps_per_graph && e == hipSuccess && rc == 0; ++k) rc = launch_sched_step(h);
        if (e == hipSuccess && rc == 0) e = hipEventRecord(b->joins[i], h->st);
        if (e == hipSuccess && rc == 0) e = hipStreamWaitEvent(b->st, b->joins[i], 0);
        h->timing = timing;
        if (rc) b->err = h->err;
    }
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->st, &g);
    if (e != hipSuccess || rc || e2 != hipSuccess) {
        if (g) hipGraphDestroy(g);
        if (b->err.empty()) b->err = std::string("graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2);
        return rc ? rc : PM3D_EHIP;
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        b->err = "hipGraphInstantiate failed";
        batch_clear_graph(b);
        return PM3D_EHIP;
    }
    return 0;
}

int pm3d_batch_set_steps(void* batch, int steps) {
    if (!batch || steps < 1 || steps > 64) return PM3D_EINVAL;
    ((Batch*)batch)->steps_per_graph = steps;
    return 0;
}

int pm3d_batch_run(void* batch, int nsteps) {
    if (!batch || nsteps < 0) return PM3D_EINVAL;
    Batch* b = (Batch*)batch;
    if (!b->exec) return 0;
    const int launches = (nsteps + b->steps_per_graph - 1) / b->steps_per_graph;
    for (int k = 0; k < launches; ++k) {
        hipError_t e = hipGraphLaunch(b->exec, b->st);
        if (e != hipSuccess) {
            b->err = std::string("hipGraphLaunch: ") + hipGetErrorString(e);
            return PM3D_EHIP;
        }
    }
    return 0;
}

int pm3d_batch_status(void* batch, pm3d_member_status* out, int count) {
    if (!batch || !out || count < 0) return PM3D_EINVAL;
    Batch* b = (Batch*)batch;
    if (count != (int)b->members.size()) {
        b->err = "pm3d_batch_status: count differs from the members of pm3d_batch_set";
        return PM3D_EINVAL;
    }
    if (count == 0) return 0;
    k_batch_status<<<grid_for(count), kThreads, 0, b->st>>>(b->refs, count, b->stat);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(b->st);
    if (e != hipSuccess) {
        b->err = std::string("batch status: ") + hipGetErrorString(e);
        return PM3D_EHIP;
    }
    std::memcpy(out, b->stat, count * sizeof(pm3d_member_status));
    return 0;
}

int pm3d_batch_synchronize(void* batch) {
    if (!batch) return PM3D_EINVAL;
    Batch* b = (Batch*)batch;
    hipError_t e = hipStreamSynchronize(b->st);
    if (e != hipSuccess) {
        b->err = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
        return PM3D_EHIP;
    }
    return 0;
}

}  // extern "C"
