/* C-ABI of the MI355X (gfx950) particle-mesh library `pbh_amd/_pm3d_hip.so`.
 *
 * It replaces the numpy hot path of the reference module pbhgr/pbhgr/pm3d.py (uniquehuman/pbh):
 * CIC deposit, FFT Poisson solve with the spectral gradient, CIC gather and the KDK leapfrog in a.
 * The calling convention follows the reference's own native boundary (pbhgr/csrc/kernels.c with
 * its ctypes loader pbhgr/pbhgr/cosmo_ev.py:31-47): plain C functions, `int` return, scalars as
 * int/double, arrays as caller-owned host pointers plus explicit sizes.  One deviation: device state
 * (positions, momenta, meshes) lives behind an opaque handle so that it stays resident in HBM.
 *
 * Return values: 0 = OK; > 0 = count of flagged particles (non-finite position or momentum, the
 *   analogue of the reference kernels' "bad cell" counts, kernels.c:36,73,105,166);
 *   < 0 = error (see PM3D_E*), with a message from pm3d_last_error().
 * Host arrays are C-contiguous.  Particle arrays are (n,3) FP64 in Lagrangian order (particle i is
 * lattice site i in C order, pm3d.py:64-65).  Meshes are (N,N,N) C order.  Half spectra are
 * (N,N,N/2+1) complex128 stored as interleaved (re,im) doubles, numpy's rfftn layout.
 * A handle is bound to one HIP stream and is not thread-safe; distinct handles may be used from
 * distinct host threads.
 */
#ifndef PM3D_HIP_H
#define PM3D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM3D_FP64 0   /* FP64 positions/momenta/FFT; fixed-point deposit (128-bit cells at 2^-52, or 64-bit at 2^-40) */
#define PM3D_MIXED 1  /* int32 cell + FP32 offset, FP32 momenta/FFT/forces; int64 deposit; FP64 scalars */

#define PM3D_OK 0
#define PM3D_EINVAL (-1)
#define PM3D_EHIP (-2)
#define PM3D_ENOMEM (-3)
#define PM3D_EUNSUPPORTED (-4)
#define PM3D_ESTATE (-5)

typedef struct pm3d_cfg {
    int32_t N;          /* mesh cells per side: power of two in [8, 512], or 48 or 96 */
    int32_t Np;         /* particles per side (n = Np^3), 1 .. 812 */
    double L;           /* periodic box side */
    double smooth;      /* Gaussian force smoothing length in cells, 0 = none (pm3d.py:27,107-108) */
    int32_t precision;  /* PM3D_FP64 or PM3D_MIXED */
    int32_t frac_bits;  /* fixed-point fraction bits of the deposit: 0 = the mode's default (mixed 31 on
                         * 64-bit cells, FP64 52 on 128-bit cells); FP64 also accepts 40 (64-bit cells,
                         * Np^3 <= 2^23 so that no cell can overflow: half the mesh bytes, an absolute
                         * error of about 1e-12 of a particle per cell); others are rejected */
} pm3d_cfg;

/* Host scalars of one KDK step, evaluated in Python exactly as pm3d.py:118-127 does
 * (`**` is libm pow, so the device never recomputes them). */
typedef struct pm3d_step_scalars {
    double a;          /* a at the start of the step */
    double da;
    double c1;         /* 0.5*da */
    double den1;       /* a*H(a) */
    double den_d;      /* a_mid**3 * H(a_mid) */
    double den2;       /* a_new*H(a_new) */
    double s_old;      /* -1.5*H_I**2/a     (used only when the cached acceleration is stale) */
    double s_new;      /* -1.5*H_I**2/a_new */
    double sm2;        /* (smooth*dx)**2 */
    int32_t need_forces; /* 1: recompute the opening acceleration at a (pm3d.py:119-120) */
    int32_t pad;
} pm3d_step_scalars;

/* Lifecycle.  Replaces PM3D.__init__ (pm3d.py:27-37): the k-grids are regenerated on the device
 * from integer indices with numpy's fftfreq formula instead of being stored. */
int pm3d_create(const pm3d_cfg* cfg, void** handle);
int pm3d_destroy(void* handle);
const char* pm3d_last_error(void* handle);  /* handle may be NULL: the last create error */
int pm3d_synchronize(void* handle);         /* waits for the stream; returns the sticky bad count */
int64_t pm3d_device_bytes(void* handle);    /* HBM held by the handle */
/* Name, gcnArchName and total memory of the handle's device, for run records (the reference's
 * provenance record, provenance.py:52-64, names the host CPU and compiler). */
int pm3d_device_info(void* handle, char* name, int name_len, char* arch, int arch_len, int64_t* total_bytes);

/* State transfer (the reference keeps pm.x / pm.p as numpy attributes, pm3d.py:67-68,124). */
int pm3d_set_state(void* handle, const double* x, const double* p, int64_t n);
int pm3d_get_state(void* handle, double* x, double* p);  /* either pointer may be NULL */

/* PM3D._cic_index on the three axes (pm3d.py:76-79): key[i] = (ix0*N+iy0)*N+iz0 (int64, the
 * deposit's flat index of corner 0, pm3d.py:88) and frac[i,d] = g - floor(g). */
int pm3d_cic_index(void* handle, const double* X, int64_t n, int64_t* key, double* frac);

/* PM3D.deposit(X) (pm3d.py:81-89): particle counts per cell, (N,N,N) FP64, from X (n,3), on the
 * handle's fixed-point grid (at most 2^23 particles on 64-bit cells at 2^-40: PM3D_EINVAL beyond). */
int pm3d_deposit(void* handle, const double* X, int64_t n, double* rho);

/* Allocation-free variant of pm3d_deposit, with no handle: X (n,3) FP64 and rho (N,N,N) FP64 are
 * DEVICE pointers the caller owns; the call is asynchronous on `stream` (a hipStream_t, NULL = the
 * null stream).  No counterpart in the reference: it is for callers that cannot allocate (code under
 * hipGraph capture, tight serving loops, hosts that forbid hipMalloc).  It calls no hipMalloc,
 * hipFree or allocating library routine: the fixed-point mesh lives in a device array of the
 * library of 2 x 128^3 uint64 (32 MiB), held by every GPU the library runs on until the process
 * ends, so N is limited to 128 (power of two in [8,128], or 48, 96; otherwise PM3D_EINVAL).  The
 * result is bit for bit that of pm3d_deposit on an FP64 handle created with frac_bits = 52 (the same
 * 2^-52 grid on 128-bit cells).  Every call
 * on one GPU shares that array: calls that overlap on one GPU (other streams, other threads) race
 * and must be serialised by the caller.  Errors are read with pm3d_last_error(NULL). */
int pm3d_deposit_noalloc(const double* X, int64_t n, int32_t N, double L, double* rho, void* stream);

/* PM3D.interp(F, X) (pm3d.py:91-99): CIC gather of the (N,N,N) mesh F at X (n,3) into out (n,). */
int pm3d_interp(void* handle, const double* F, const double* X, int64_t n, double* out);

/* PM3D.forces(a) (pm3d.py:101-114) at the current state: acc (n,3) and delta (N,N,N); either
 * output may be NULL.  s = -1.5*H_I**2/a and sm2 = (smooth*dx)**2 come from the host.  Like the
 * reference, it leaves the step's cached opening acceleration untouched. */
int pm3d_forces(void* handle, double a, double s, double sm2, double* acc, double* delta);

/* PM3D.step(a, da) (pm3d.py:117-129).  delta (N,N,N) and delta_max may be NULL; when both are
 * NULL the call is asynchronous (no host synchronisation).  The step ends with the force meshes of
 * the new positions on the device; their gather and the closing kick run at the start of the next
 * step, fused with its opening kick and drift (same operations and order, same bits), or when p is
 * read. */
int pm3d_step(void* handle, const pm3d_step_scalars* sc, double* delta, double* delta_max);

/* Lazy access to the density contrast of a deposit (the `delta` that PM3D.step returns,
 * pm3d.py:104,129).  Every deposit of the state gets a serial number; the maximum of delta of the
 * last 4096 deposits stays on the device.  pm3d_get_delta converts the most recent deposit. */
int64_t pm3d_deposit_serial(void* handle);
int pm3d_delta_max_at(void* handle, int64_t serial, double* delta_max);
int pm3d_get_delta(void* handle, double* delta);

/* numpy.fft.rfftn / irfftn(s=(N,N,N)) on the device FFT (pm3d.py:105,112; numpy conventions:
 * unnormalised forward, 1/N^3 inverse, c2r ignores the imaginary parts of the kz=0 and kz=N/2 bins). */
int pm3d_rfftn(void* handle, const double* in, double* out_c);
int pm3d_irfftn(void* handle, const double* in_c, double* out);

/* Spectral gradient meshes g_c = irfftn(-1j*k_c*phik) of pm3d.py:105-112 for a given real mesh
 * delta: g (N,N,N,3) interleaved.  Used by the parity tests of the fused Poisson pass. */
int pm3d_poisson_gradient(void* handle, const double* delta, double s, double sm2, double* g);

/* First-shell-crossing diagnostic (not in the reference; defined by this library):
 * min over Lagrangian lattice cells of det(dx/dq) with forward differences along the q-lattice,
 * minimum-image unwrapped, in units of the lattice spacing.  Requires Np == particles per side. */
int pm3d_min_jacobian(void* handle, double* jmin);

/* Per-kernel timing with hipEvents on the handle's stream.  When enabled, each step kernel's
 * dispatch stamps its own start/stop event pair (hipExtLaunchKernelGGL), other launches are
 * bracketed by event records; pm3d_get_timings returns (count, total ms) per kernel class in the
 * order of pm3d_kernel_names(), and resets the accumulators. */
int pm3d_set_timing(void* handle, int enable);
int pm3d_get_timings(void* handle, int64_t* counts, double* total_ms, int nmax);
const char* pm3d_kernel_names(void);  /* comma-separated */
/* Diagnostics of inner Lagrangian sets on the device (analyse, pm3d.py:132-159).
 * pm3d_set_analysis_sets uploads nsets index lists (ascending particle indices, set s =
 * idx[offsets[s]..offsets[s+1])): set 0 is the core whose centre of mass is the centre
 * ((xs - xs[0] + L/2) % L - L/2 averaged, pm3d.py:137-141), sets 1.. the inner Lagrangian sets
 * (pm3d.py:144-146).  pm3d_analyse returns centre[3]; moments[(s-1)*9 + 0..8] for s >= 1 = the set
 * mean of rel = (x - centre + L/2) % L - L/2 and the sums xx,xy,xz,yy,yz,zz of (rel - mean)
 * outer products (the reference's r.T @ r before / len(r), pm3d.py:147-148); counts[k] = number of
 * particles with |rel| < radii[k] (pm3d.py:153-155, nradii <= 16); rho_max = maximum of a deposit
 * of the current state (pm3d.py:158; the last step's density mesh is left untouched). */
int pm3d_set_analysis_sets(void* handle, const int32_t* idx, const int64_t* offsets, int nsets);
int pm3d_analyse(void* handle, const double* radii, int nradii, double* centre, double* moments, int64_t* counts,
                 double* rho_max);

/* Global sums kept in FP64 (or exact) in every precision mode.  mass = total of the most recent
 * deposit in particle units, summed exactly as a 128-bit integer of fixed-point weights and then
 * rounded to FP64; mass_dev = that exact total minus n (each corner weight is rounded to the
 * fixed-point grid on its own, so the deposit's mass is n to within that rounding); momentum[3] =
 * sum over particles of p (after the pending closing kick), accumulated in FP64 over a fixed
 * reduction tree (the same bits for the same state).  Any output may be NULL.  The reference
 * conserves mass and momentum implicitly (np.bincount of unit weights, pm3d.py:88; the momentum
 * update pm3d.py:122,127) and has no such diagnostic. */
int pm3d_totals(void* handle, double* mass, double* mass_dev, double* momentum);

/* Attainable HBM bandwidth of this GPU, for reading kernel rooflines against: a hand-written
 * streaming copy (16-byte non-temporal vector loads, then stores, one pass; read + write bytes
 * counted) and a streaming read of `bytes` (1 MiB .. 64 GiB), each launched `reps` times on
 * buffers the call allocates and frees.
 * No reference counterpart (a diagnostic of this library). */
int pm3d_stream_probe(int64_t bytes, int reps, double* copy_gbs, double* read_gbs);

/* FFT tile variant per pass (r2c_z, y fwd, x, y inv, c2r_z), two bits: bit 0 = pencils per
 * workgroup sized for 32 KB rather than 16 KB LDS buffers, bit 1 = one buffer transformed in place
 * rather than two ping-pong buffers.  Every variant gives the same bits; pm3d_create times all four
 * per pass and keeps the fastest (cached per mesh size and precision).  Environment
 * PM3D_FFT_TILES = one digit 0-3, or five digits, forces the variants.  Returns the number of
 * passes.  No reference counterpart (numpy's FFT has no tiles). */
int pm3d_fft_tiles(void* handle, int32_t* variants, int nmax);

/* Current column shape of the deposit: shape[0..2] = BX, BY, N of the owner-computes columns
 * (BX x BY x N cells per workgroup), all 0 for the atomic path (Np != N or N < 16).  pm3d_create
 * times the 8x8, 4x8, 4x4 and 2x4 columns that fit in LDS on a jittered lattice and keeps the
 * widest unless a narrower one is at least 3 % faster (cached per mesh size and precision); the
 * shapes within 25 % of the fastest stay candidates, and every 64 step deposits pm3d_step times two
 * deposits in each (event pairs read back without waiting) and moves to one that is 2 % faster than
 * the current shape, as clustering changes which is best.  Every shape gives the same mesh bits.
 * Environment PM3D_DEPOSIT_SHAPE = auto (default), static (the creation choice throughout), all
 * (every fitting shape a trial candidate), or one digit 0-3 forcing the n-th fitting shape.
 * Returns 1 for columns, 0 for the atomic path.  No reference counterpart (pm3d.py:81-89 deposits
 * with one np.bincount per CIC corner). */
int pm3d_deposit_tiles(void* handle, int32_t* shape);

/* Fixed-point fraction bits of the handle's deposit (the frac_bits pm3d_create chose), or < 0. */
int pm3d_frac_bits(void* handle);

/* Shape trials of the step deposits so far: out[0] = finished trials, out[1] = shape changes they
 * made, out[2] = candidate mask (bit v: shape variant v; fewer than two bits: no trials).  No
 * reference counterpart (diagnostic of pm3d_deposit_tiles' trials). */
int pm3d_deposit_trials(void* handle, int64_t* out);

/* Initial conditions built on the device (the construction of PM3D.make_ic, pm3d.py:59-68,
 * generalised to an explicit linear density field at a_i): delta = mu*exp(-0.5*sum (x_c/s_c)^2)
 * around the box-centre vertex, plus optionally a Gaussian random field with sqrt(P(k)) = 1/|k| for
 * 0 < |k| <= noise_kcut, from the SplitMix64 counter stream `seed` (Box-Muller), scaled to rms
 * noise_rms; psi = irfftn(1j*k*dk/k2) is gathered at the particle lattice, x = (Q + psi + L/2) % L,
 * p = coef*psi (coef = a_i^2 H(a_i)).  The handle must be FP64; a mixed handle takes the state
 * with pm3d_copy_state. */
typedef struct pm3d_peak_ic {
    double mu, sx, sy, sz;        /* amplitude and Gaussian widths (box units, not fractions) */
    double noise_rms, noise_kcut; /* random field rms (0 = none) and band edge */
    double coef;                  /* a_i^2 H(a_i) */
    uint64_t seed;
} pm3d_peak_ic;
int pm3d_make_peak_ic(void* handle, const pm3d_peak_ic* ic);
/* Device-to-device copy of the state of an FP64 handle into a handle of the same N, Np and L
 * (FP64 or mixed: positions become cell + offset, momenta FP32). */
int pm3d_copy_state(void* dst_handle, void* src_handle);

/* ---------------------------------------------------------------------------------------------
 * Device-resident schedules and batched graph replay.  No counterpart in the reference (its driver
 * loops over PM3D.step from Python, pbhgr/scripts/pm3d_peak.py:26-33); these entry points run the
 * same step for many boxes at once.
 *
 * A schedule is a table of per-step host scalars (the same expressions as pm3d_step_scalars,
 * evaluated in Python); entry k also carries the previous step's closing kick (c1p, denp).  The
 * device keeps a step counter; every scheduled step reads entry [counter], records the minimum of
 * det(dx/dq) of its new positions in a history and advances the counter.  A schedule continues a
 * stepped state (the force meshes of the current positions and a pending closing kick must exist:
 * run one pm3d_step first).  A step whose Jacobian minimum is <= 0 (the first shell crossing) ends
 * the table there; steps past the end of the table leave the state unchanged. */
typedef struct pm3d_sched_entry {
    double c1p, denp;  /* closing kick of the previous step: p += (c1p*acc)/denp */
    double c1, den1;   /* opening kick */
    double da, den_d;  /* drift */
    double den2;       /* this step's closing kick denominator (kick applied by the next entry) */
    double s, sm2;     /* -1.5*H_I**2/a_new, (smooth*dx)**2 */
} pm3d_sched_entry;

int pm3d_set_schedule(void* handle, const pm3d_sched_entry* table, int nsteps);
int pm3d_schedule_step(void* handle);                             /* one scheduled step on the handle's stream */
int pm3d_schedule_status(void* handle, int* done, double* jhist); /* steps done, jmin history (done entries) */
int pm3d_schedule_end(void* handle, int* done);                   /* hand the state back to pm3d_step */

/* A batch replays one captured hipGraph holding `steps` scheduled steps (1 unless set with
 * pm3d_batch_set_steps) of every member as independent branches (members of any mesh size and
 * precision); a member's table ends on the device at its first shell crossing, after which its
 * kernels do nothing. */
int pm3d_batch_create(void** batch);
int pm3d_batch_destroy(void* batch);
const char* pm3d_batch_last_error(void* batch);
int pm3d_batch_set_steps(void* batch, int steps);                 /* steps per graph launch, 1..64; next capture */
int pm3d_batch_set(void* batch, void* const* handles, int count); /* (re)capture for these members */
int pm3d_batch_run(void* batch, int nsteps);                      /* advance nsteps steps (whole launches, rounded up), asynchronous */
int pm3d_batch_synchronize(void* batch);

/* Schedule status of every member of the batch, in the order of pm3d_batch_set, read in one
 * device pass after the replays queued so far (one synchronisation for the whole batch instead of
 * two copies per member).  No counterpart in the reference: the reference runs one box per
 * process (pbhgr/scripts/pm3d_peak.py:24-33) and its campaign driver polls process results
 * (pbhgr/pbhgr/scan.py:67-76). */
typedef struct {
    int32_t done;   /* scheduled steps taken */
    int32_t nsteps; /* entries of the table; set to done when a step ends at the first shell crossing */
    double jmin;    /* min det(dx/dq) after the last step taken (1e300 before the first) */
} pm3d_member_status;
int pm3d_batch_status(void* batch, pm3d_member_status* out, int count); /* count = members of the batch */

#ifdef __cplusplus
}
#endif
#endif /* PM3D_HIP_H */
