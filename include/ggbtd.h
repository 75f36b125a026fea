/*
 * ggbtd.h -- C ABI of the MI355X real-time TDDFT propagator (libggbtd.so).
 *
 * The reference (graham-g-brown/ggbTDDFT) has no library interface: its hot
 * path is the C++ function
 *     void h_timePropagation(ksoSetTDStruct, hartreeStruct, propagatorsStruct,
 *                            meshStruct, observableStruct, tempStruct,
 *                            handleStruct, int scandx)
 * (src/subFunctions/timePropagation/h_timePropagation.cuh:32-41), fed by
 * loaders that read workingData/ files and by compile-time macros in
 * params.cuh.  Each entry point below names the reference piece it replaces.
 *
 * Conventions: every function returns 0 on success and a negative GGB_E*
 * code on failure (never exit()); ggb_last_error() gives the message.  Host
 * buffers are caller-owned and copied; device memory is owned by the context.
 * One context = one HIP stream on one device.  A context may be used from
 * one thread at a time (thread-compatible, not thread-safe).
 *
 * Layouts (all little-endian, row-major, double = IEEE binary64):
 *   state    psi[l][c][i] complex (re, im interleaved), c = inst * n_orb + orb,
 *            l = 0..n_l-1, i = 0..n_r-1, zero for l < m_orb
 *            (the reference's ksoSetTD.d_x[l][i + k N_r], main.cu:363-391).
 *   field    E[inst][step] at the step midpoint t_n + dt/2.
 *   observables per call: dip/acc/norm[step - step0][c].
 */
#ifndef GGBTD_H
#define GGBTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GGB_ABI_VERSION 4

#define GGB_OK 0
#define GGB_E_ARG (-1)      /* invalid argument / size */
#define GGB_E_HIP (-2)      /* HIP runtime error */
#define GGB_E_STATE (-3)    /* call out of order (e.g. run before upload) */
#define GGB_E_NOMEM (-4)    /* device allocation failed */

typedef struct ggb_ctx ggb_ctx;

/* Sizes and switches of one context.  Replaces the params.cuh macros N_r,
 * N_l, N_kso, N_har, dt, USE_HARTREE (init.py:493-604). */
typedef struct ggb_problem {
    int32_t abi_version;  /* must be GGB_ABI_VERSION */
    int32_t n_r;          /* radial Gauss-Lobatto nodes N (>= 4) */
    int32_t n_l;          /* angular momenta, Lmax + 1 (<= 128) */
    int32_t n_theta;      /* Gauss-Legendre theta nodes (>= n_l, <= 256) */
    int32_t n_har;        /* Hartree multipoles L = 0..n_har-1 (<= 256) */
    int32_t n_orb;        /* orbitals per instance */
    int32_t n_inst;       /* instances sharing the static data */
    int32_t n_m;          /* distinct |m| values; orbital m in [0, n_m) */
    int32_t nk_max;       /* row length of the eigenvector table, multiple of 16 */
    int32_t interacting;  /* 1: Hartree + LDA update each step; 0: dV = 0 */
    int32_t precision;    /* 0: FP64 everywhere; 1: kinetic tables, transforms and
                             v_xc in FP32 on the FP32 matrix cores with FP32
                             accumulation; state, Hartree scans and observable
                             sums stay FP64; 2: as 1, with the compensated
                             split-FP32 kinetic product (GGB_KIN_SPLIT16 only) */
    int32_t device;       /* HIP device ordinal */
    double dt;            /* time step (a.u.) */
} ggb_problem;

/* Static per-grid data.  Replaces the loaders h_initialiseMesh
 * (h_initialiseMesh.cuh:1-81) and h_initialisePropagators
 * (h_initialisePropagators.cuh:13-118). */
typedef struct ggb_static {
    const double *r;           /* [n_r] radial nodes */
    const double *wr;          /* [n_r] Lobatto weight * dr/dx */
    const double *v_gs;        /* [n_r] ground-state KS potential in the kinetic propagator */
    const double *mask;        /* [n_r] absorber exp(-W_cap dt/2) per half step */
    const double *Q;           /* [n_l][n_r][nk_max] eigenvectors of H_l as columns, zero padded */
    const double *lam;         /* [n_l][nk_max] eigenvalues (padding ignored) */
    const int32_t *nkeep;      /* [n_l] retained eigenpairs per l (<= nk_max) */
    const double *B;           /* [n_m][n_theta][n_l] sqrt(omega_j) Theta_lm(chi_j) */
    const double *PL;          /* [n_theta][n_har] Legendre P_L(chi_j) */
    const double *chi;         /* [n_theta] cos(theta_j) */
    const double *omega;       /* [n_theta] Gauss-Legendre weights */
    const int32_t *m;          /* [n_orb] |m| of each orbital */
    const double *occ;         /* [n_orb] occupations */
    const double *rho_frozen;  /* [n_r] frozen-core u-unit density, or NULL */
    /* FP32 mixed form (precision 1) only, else ignored; all three or NULL:
     * the ground-state orbital each propagated orbital starts from, an
     * eigenvector of its block's kinetic propagator.  The kinetic products
     * act on psi minus its component along it, which is advanced in FP64. */
    const double *phi;         /* [n_orb][n_r] radial function, u units, unit norm */
    const int32_t *l0;         /* [n_orb] its l */
    const double *eps;         /* [n_orb] its eigenvalue of H_l (Hartree) */
} ggb_static;

int ggb_abi_version(void);

/* Device bytes a context of these sizes allocates, excluding the field and
 * the per-call observable buffer (8 * n_inst * n_field and 24 * C * nsteps
 * bytes); used for admission by memory footprint.  No counterpart in the
 * reference. */
int ggb_footprint(const ggb_problem *p, uint64_t *bytes);

/* Create a context and allocate its device memory. */
int ggb_create(const ggb_problem *p, ggb_ctx **out);

/* Upload mesh, propagator and angular tables. */
int ggb_upload_static(ggb_ctx *ctx, const ggb_static *s);

/* Upload orbitals psi[l][c][i] (complex interleaved, c over all instances)
 * and build the potential of that state.  Replaces h_initialiseTDKSOSet
 * (h_initialiseTDKSOSet.cuh:19-107) and the t=0 potentials of
 * h_timePropagation.cuh:47-57. */
int ggb_upload_orbitals(ggb_ctx *ctx, const double *psi);

/* Upload the field E[inst][step] for n_steps steps.  Replaces
 * h_initialiseField (h_initialiseField.cuh:1-40). */
int ggb_set_field(ggb_ctx *ctx, const double *E, int32_t n_steps);

/* Switch the field to a zeroed ring buffer of `capacity` steps per
 * instance: step n then reads sample n mod capacity, so a long-running
 * context can admit new instances at any step (continuous batching).
 * No counterpart in the reference (its scans run one after another,
 * main.cu:460-481). */
int ggb_field_ring(ggb_ctx *ctx, int32_t capacity);

/* Write n field samples of one instance starting at step `start` (wrapping
 * in ring mode).  Per-scan counterpart of h_initialiseField. */
int ggb_set_field_slice(ggb_ctx *ctx, int32_t inst, int32_t start, int32_t n, const double *E);

/* Replace orbital columns c0 .. c0+nc-1 (psi[l][nc][i], complex interleaved)
 * without touching the others, e.g. to seat a new instance in a free slot;
 * follow with ggb_rebuild_potential.  Per-scan counterpart of
 * h_initialiseTDKSOSet (main.cu:463). */
int ggb_upload_columns(ggb_ctx *ctx, int32_t c0, int32_t nc, const double *psi);

/* Mark which instances are in use (active[n_inst], nonzero = in use; all are
 * at creation).  The angular, density and potential kernels skip idle
 * instances (their columns keep stale values), and the kinetic products skip
 * every tile of orbital columns that all belong to idle instances (8
 * columns in the narrow kernels, 64 in the tiled ones), so a serving lane with free
 * slots does less work.  A
 * slot being seated must be marked in use before ggb_rebuild_potential.  No
 * counterpart in the reference, which runs its scans one after another. */
int ggb_set_active(ggb_ctx *ctx, const int32_t *active);

/* Rebuild the potential of every instance from the current orbitals.  For
 * instances whose orbitals did not change this reproduces the potential
 * they already hold, bit for bit. */
int ggb_rebuild_potential(ggb_ctx *ctx);

/* Propagate steps step0 .. step0+nsteps-1 and copy the per-step observables
 * <r cos theta>, <cos theta / r^2> and <psi|psi> of every orbital column into
 * dip/acc/norm[nsteps][n_inst*n_orb] (any may be NULL).  Replaces the step
 * loop of h_timePropagation (h_timePropagation.cuh:59-112) with the
 * observables of dipoleMoment.cuh / dipoleAcceleration.cuh. */
int ggb_run(ggb_ctx *ctx, int32_t step0, int32_t nsteps, double *dip, double *acc,
            double *norm);

/* Record, in later ggb_run calls, the per-(orbital column, l) channel
 * dipole of the state after each step,
 *     chan[c][l] = <x_l | r (c(l, m) x_{l+1} + c(l-1, m) x_{l-1})>
 * (complex, conjugate on the left, c(l, m) = <l m|cos theta|l+1 m>; zero for
 * l < m and for idle instances), on = 1, or stop (0, the default).  Replaces
 * h_calculateDipoleMoment (dipoleMoment.cuh:49-75), whose per-channel values
 * h_writeDipoleMoment writes one file per (orbital, l)
 * (h_writeDipoleMoment.cuh:12-22); sum_l Re chan[c][l] is the dip of ggb_run. */
int ggb_set_channels(ggb_ctx *ctx, int32_t on);

/* Copy the channel dipoles of the last ggb_run (which must have run nsteps
 * steps with channels on) into chan[nsteps][n_inst*n_orb][n_l] complex
 * (re, im interleaved). */
int ggb_get_channels(ggb_ctx *ctx, int32_t nsteps, double *chan);

/* Copy the current orbitals out (same layout as ggb_upload_orbitals). */
int ggb_download_orbitals(ggb_ctx *ctx, double *psi);

/* Copy the current potential dV[inst][i][j] = v_H + v_xc - v_gs on the
 * (r, theta) grid (the one the next step applies). */
int ggb_download_potential(ggb_ctx *ctx, double *dv);

/* Run one sub-step in isolation (for parity tests):
 * 0 = half potential then laser even/odd, 1 = kinetic propagator,
 * 2 = laser odd/even, half potential, density, 3 = potential rebuild from the
 * current state.  `step` selects the field sample. */
int ggb_debug_substep(ggb_ctx *ctx, int32_t which, int32_t step);

/* Replay the step loop of ggb_run from a hipGraph of `steps` steps (0 =
 * plain stream launches, the default).  The graph is captured on the next
 * ggb_run and re-captured after any call that may move a buffer; runs with
 * timing enabled use plain launches.  Calls on one context copy on that
 * context's stream only, so other threads may drive other contexts during a
 * capture; ggb_create, ggb_upload_static, ggb_spectrum, ggb_ground_state and
 * ggb_destroy allocate, free or use the legacy stream, and must not overlap
 * a capture in another thread.  No counterpart in the reference. */
int ggb_set_graph(ggb_ctx *ctx, int32_t steps);

/* Groups of contexts replayed together: the chunks of several contexts (the
 * lanes of a batched scan, one context per grid; every context on the same
 * device, each used by no other thread meanwhile) run from one hipGraph
 * launch per `steps` steps, each context on its own stream inside the graph.
 * Counterpart of the reference's sequential scan loop (main.cu:460-481),
 * which runs one scan after another. */
typedef struct ggb_group ggb_group;

/* Create a group of n contexts (1..64, distinct) replayed `steps` steps
 * (1..4096) per graph launch.  Errors: ggb_group_last_error(NULL). */
int ggb_group_create(ggb_ctx *const *ctxs, int32_t n, int32_t steps, ggb_group **out);

/* Propagate every context of the group by nsteps steps from its own step0[k]
 * (field index as in ggb_run) and copy each context's observables into
 * dip[k]/acc[k]/norm[k] ([nsteps][n_inst*n_orb] each; any array or entry may
 * be NULL).  Whole graphs of `steps` steps run from one launch, the rest as
 * plain launches; results equal ggb_run of each context alone, bit for bit.
 * The graph is captured on the first call and again after any call that
 * changes what a context's captured step holds (ggb_set_field,
 * ggb_set_kinetic_variant, ggb_set_channels, ...). */
int ggb_group_run(ggb_group *g, const int32_t *step0, int32_t nsteps, double *const *dip,
                  double *const *acc, double *const *norm);

/* Message of the last failure on g (or of the last failed ggb_group_create). */
const char *ggb_group_last_error(const ggb_group *g);

void ggb_group_destroy(ggb_group *g);

/* Kinetic-pass kernel variants (tile shapes).  DEFAULT: NARROW up to 8
 * orbital columns, else TILED16.  NARROW: one 16x16 tile per work-group with
 * the contraction split over its four waves; TILED16 / TILED32: 64x128
 * work-group tiles, K chunks of 16 / 32 staged in LDS (TILED32 in the FP32
 * mixed form only; FP64 uses 16). */
#define GGB_KIN_DEFAULT 0
#define GGB_KIN_NARROW 1
#define GGB_KIN_TILED16 2
#define GGB_KIN_TILED32 3
#define GGB_KIN_SPLIT16 4   /* precision 2 only: TILED16 tiles, operands split into
                               FP32 hi + lo parts (a_hi b_hi + a_hi b_lo + a_lo b_hi,
                               three FP32 matrix-core products), FP64 chunk sums */
#define GGB_KIN_NVARIANT 5

/* Use kinetic variant `variant` (GGB_KIN_*) from now on.  No counterpart in
 * the reference (one cuBLAS batched GEMM). */
int ggb_set_kinetic_variant(ggb_ctx *ctx, int32_t variant);

/* Autotuner: time both kinetic passes of every variant valid for this
 * context (reps launches each, HIP events, on the context's own tables and
 * orbitals, which are restored afterwards), keep the fastest, and return it
 * in *chosen and the mean ms per variant in ms[GGB_KIN_NVARIANT] (0 for variants not
 * tried; ms may be NULL).  Needs ggb_upload_static and ggb_upload_orbitals.
 * No counterpart in the reference. */
int ggb_autotune(ggb_ctx *ctx, int32_t reps, int32_t *chosen, double *ms);

/* Step phases timed by ggb_set_timing: 0 = half potential + laser,
 * 1 = kinetic pass Q^T, 2 = kinetic pass Q, 3 = laser + half potential +
 * density tiles + observable partials, 4 = density, v_xc and multipole
 * moments (+ observable sums), 5 = Hartree scans, 6 = dV and the phase
 * factors of the next step. */
#define GGB_NPHASE 7

/* Time the step phases of later ggb_run calls with HIP events on the
 * context stream and reset the accumulated times.  Mode 1: the product
 * kernel of each kinetic pass (phases 1 and 2) of every 16th step, with
 * events stamped by the kernel's own dispatch (hipExtLaunchKernel: kernel
 * start to end, cheap enough for a timed run); mode 2: events between every
 * phase of every step; 0: off.  No counterpart in the reference, which has
 * no timing code. */
int ggb_set_timing(ggb_ctx *ctx, int32_t mode);

/* Accumulated milliseconds and launch counts per phase [GGB_NPHASE]. */
int ggb_get_timing(const ggb_ctx *ctx, double *ms_total, int64_t *count);

/* 16x16x4 matrix-core products of fixed inputs A[i][k] = i + 0.25 k,
 * B[k][j] = 0.5 k - j, in FP64 (out[0][16][16]) and FP32 (out[1][16][16])
 * (lane-layout self test of the instructions the kernels use). */
int ggb_selftest_mfma(int32_t device, double *out);

/* Copy `bytes` (multiple of 16) device-to-device with a plain streaming
 * kernel (k_copy) twice and report the second launch's time: the known-size
 * kernel against which the profiler's HBM byte counters are calibrated. */
int ggb_selftest_copy(int32_t device, uint64_t bytes, double *ms);

/* HHG spectrum of one time series on the device: Kaiser window (beta,
 * sym = False) on signal[n], the 8th-order central second difference when
 * `differentiate` (a dipole; else an acceleration), zero padding by
 * n_pad * n on each side (M = (2 n_pad + 1) n), power[M] =
 * |fftshift(fft(fftshift(.)))|^2 at omega_q = 2 pi (q - M/2) / (M dt)
 * (power may be NULL), and intensity[k] = the harmonic-mask sum of
 * orders[k] for fundamental omega0.  Replaces the spectrum part of
 * analysis.py:447-460 (with internal.py:361-379) and the harmonic masks of
 * argonInSitu/arg.py:42-67.  Errors: ggb_last_error(NULL). */
int ggb_spectrum(int32_t device, const double *signal, int32_t n, double dt,
                 int32_t differentiate, double beta, int32_t n_pad, double omega0,
                 const double *orders, int32_t n_orders, double *power, double *intensity);

/* ggb_spectrum without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * results and errors as ggb_spectrum (bit for bit), computed on the legacy
 * stream in a fixed device scratch of 60 MiB that each device holds from its
 * first call until the process ends.  Limits: (2 n_pad + 1) n <= 2^19 (FFT
 * length <= 2^20; e.g. n <= 30840 samples at n_pad = 8) and n_orders <= 256,
 * else GGB_E_ARG with a message.  All calls on one device share the scratch:
 * calls that overlap on one device race (serialise them); the scratch keeps
 * the last call's contents, and every buffer is written before it is read. */
int ggb_spectrum_noalloc(int32_t device, const double *signal, int32_t n, double dt,
                         int32_t differentiate, double beta, int32_t n_pad, double omega0,
                         const double *orders, int32_t n_orders, double *power,
                         double *intensity);

/* Ground state on the device: the orbitals and v_gs that the propagation
 * starts from.  The reference loads them from workingData/
 * (h_initialiseTDKSOSet.cuh:19-107), written by its diagonalisation SCF
 * (src/scripts/phys/groundState.py:182-457); the host counterpart here is the
 * imaginary-time iteration of ggbtddft_amd/groundstate.py, whose fixed point
 * (the self-consistent spherical LDA ground state on the grid) this call
 * reaches by preconditioned Rayleigh-Ritz sweeps on the device, FP64.
 * Inputs: shell_l/occ [n_shell] (atom order), the grid's r, wr (= Lobatto
 * weight * dr/dx) [n_r] and kinetic matrix T [n_r][n_r] (symmetric), and for
 * each distinct l of the shells, ascending, the eigenvectors
 * Q0[g][i][k] (column k = k-th eigenvector) and eigenvalues lam0[g][k] of
 * the bare-nucleus radial Hamiltonian -T/2 + l(l+1)/2r^2 - Z/r (the
 * preconditioner, and the start: shell k of each l starts as column k).
 * Outputs: u[n_shell][n_r] (real, unit norm, u units), eps[n_shell], v_gs and
 * rho_u [n_r] of the final orbitals, the iteration count and whether the
 * host iteration's test (|d eps| < tol, |v_out - v| < 1e3 tol) was met. */
typedef struct {
    int32_t n_r;         /* 4..1536 */
    int32_t n_shell;     /* 1..64, at most 8 of one l */
    int32_t sweeps;      /* preconditioned passes over each l's shells per iteration (8) */
    int32_t max_iter;    /* SCF iterations at most */
    int32_t interacting; /* 0: bare nucleus, v = 0 */
    int32_t device;
    double Z, mix, gamma, tol; /* nuclear charge, potential damping (0.3), preconditioner shift (2), 1e-11 */
} ggb_gs_problem;

int ggb_ground_state(const ggb_gs_problem *p, const int32_t *shell_l, const double *occ,
                     const double *r, const double *wr, const double *T, const double *Q0,
                     const double *lam0, double *u, double *eps, double *v_gs, double *rho_u,
                     int32_t *iterations, int32_t *converged);

/* Kinetic propagator tables on the device.  For l = 0..n_l-1 the radial
 * Hamiltonian H_l = -T/2 + diag(l(l+1)/(2 r^2) + v) (T [n_r][n_r] symmetric,
 * v = -Z/r + v_gs) is formed on the device and all n_l are diagonalised by
 * one batched FP64 divide-and-conquer call (rocSOLVER dsyevd); the
 * eigenpairs with eigenvalue <= ecut (ascending) are returned as
 * Q[n_l][n_r][nk_cap] (column a = the a-th kept eigenvector, zero beyond
 * nkeep[l]; any sign), lam[n_l][nk_cap] and nkeep[n_l].  *nk_needed (may be
 * NULL) gets the largest kept count rounded up to 16; GGB_E_ARG when it
 * exceeds nk_cap.  Errors: ggb_last_error(NULL).  Replaces the host
 * precompute of the propagators (the reference's propagatorsR,
 * internal.py:232-333, which writes Crank-Nicolson tables per l). */
int ggb_propagator_tables(int32_t device, int32_t n_r, int32_t n_l, const double *r,
                          const double *T, const double *v, double ecut, int32_t nk_cap,
                          double *Q, double *lam, int32_t *nkeep, int32_t *nk_needed);

/* Message of the last failure on ctx (or of the last failed ggb_create when
 * ctx is NULL). */
const char *ggb_last_error(const ggb_ctx *ctx);

void ggb_destroy(ggb_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* GGBTD_H */
