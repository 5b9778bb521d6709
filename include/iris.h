/*
 * iris.h -- C ABI of the MI355X FDTD engine for IrisFDTD-Academic's time loop.
 *
 * The reference runs its whole time-stepping loop inside one Fortran
 * subroutine, `driver(e,h,eps_inv,bbox,fft_data,fft_data_o,store_pts,fdata)`
 * (reference code/driver.f90:25-58 interface, :66 definition, loop
 * :338-500, sole caller code/main.f90:179).  Its arguments are Fortran
 * pointer arrays and every scalar travels through module globals
 * (code/parameters.f90:82-257), so it has no C-callable form.  This header
 * replaces that one call with a handle-based API:
 *
 *   iris_create   <- driver setup (code/driver.f90:174-335): UPML tables,
 *                    Fourier boxes, TF/SF ranges, field allocation
 *   iris_step     <- loop body (code/driver.f90:355-450): E/UPML/max wraps,
 *                    H/UPML/min wraps, TF/SF, on-the-fly DFT
 *   iris_flux     <- scattering_fourier_out numbers (code/dataprocess.f90:213-277);
 *                    the host writes the <box>_scatt<j>.dat files itself
 *   iris_get_dft  <- the emg(j)%fdata accumulators (code/dataprocess.f90:143-147)
 *   iris_get_field<- e_cur / h_cur arrays (code/main.f90:166-167)
 *   iris_get_plane<- the field reads of the snapshot writer (code/sdl.f90:210-313,
 *                    called with e_cur, h_prev at code/driver.f90:468-495)
 *   iris_diagnostics <- energy and divergence monitors (code/energy.f90:57-201)
 *   iris_destroy  <- driver teardown (code/driver.f90:503-552)
 *
 * Conventions: every entry point returns 0 on success and a negative
 * IRIS_E* code on failure; nothing exits the process.  Host buffers are
 * borrowed for the duration of the call; device memory belongs to the handle.
 * A handle owns one HIP stream; calls on one handle must come from one host
 * thread at a time; distinct handles are independent.  Complex values are
 * interleaved (re, im) doubles in the Fortran storage order of the
 * reference's arrays.
 */
#ifndef IRIS_H
#define IRIS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IRIS_ABI_VERSION 1

/* error codes */
#define IRIS_OK 0
#define IRIS_EINVAL -1   /* bad argument or inconsistent parameters      */
#define IRIS_ENOMEM -2   /* device allocation failed                     */
#define IRIS_EDEVICE -3  /* HIP runtime error (no device, launch failure) */
#define IRIS_ESTATE -4   /* call out of order (e.g. step past nt)         */
#define IRIS_EUNSUP -5   /* parameter combination not supported           */

/*
 * POD mirror of the reference's module globals after setparam
 * (code/files.f90:39-344), set_emconstants (code/mesh.f90:75-109) and the
 * driver's TF/SF setup (code/driver.f90:310-335).  All indices use the
 * reference's 1-based Fortran convention.  Layout: 34 int32 then doubles, no
 * padding, so Fortran can mirror it with a bind(C) derived type.
 */
typedef struct iris_params {
  int32_t abi_version;    /* IRIS_ABI_VERSION                                   */
  int32_t precision;      /* 8: FP64 fields; 4: FP32 fields, FP64 DFT           */
  int32_t nx, ny, nz;     /* ixmax, iymax, izmax                                */
  int32_t pmlzi, pmlzf;   /* z-UPML: [1,pmlzi] and [pmlzf,nz]                   */
  int32_t hi, hf;         /* structure z-range [hi,hf]                          */
  int32_t i_pol;          /* 1: E//y, 2: E//x                                   */
  int32_t em_source;      /* 1: TF/SF plane-wave source on                      */
  int32_t plane_wave;     /* 1: Gaussian pulse (the only working source)        */
  int32_t cosine;         /* 0: complex analytic source, 1: real source         */
  int32_t iz_min, iz_max; /* TF/SF planes tfsf_zi, tfsf_zf                      */
  int32_t ixo;            /* pulse x-origin, pmlzi-50 (code/tfsf.f90:264)       */
  int32_t nt;             /* texit                                              */
  int32_t to_fourier;     /* first DFT step                                     */
  int32_t sample_fourier; /* DFT sampling period                                */
  int32_t no_downloads;   /* output cadence in steps                            */
  int32_t nof;            /* number of wavelengths                              */
  int32_t nbox;           /* 1 if the two Fourier boxes coincide, else 2        */
  int32_t box[2][6];      /* [box][sxi,sxf,syi,syf,szi,szf]                     */
  /* doubles */
  double q;               /* Q: cell size in atomic units (f_abohr * q_nm)      */
  double dt;              /* time step, atomic units                            */
  double c0, eps0, mu0;   /* code/physconst.f90:31-33 as FP64 values            */
  double pi, f_abohr;     /* float32-rounded pi and nm->a.u. factor             */
  double eps[3];          /* eps(1) cover, eps(2) structure (real part), eps(3) */
  double eps_in[3];       /* 1/eps(i)                                           */
  double epsa_inv[2][3][3]; /* [region I/III][row][col] (code/mesh.f90:90-95)    */
  double eamp;            /* Epump in atomic units                              */
  double wlpump;          /* carrier wavelength in nm                           */
  double gpump;           /* pulse width in mesh units                          */
  double zopump;          /* pulse z-origin in mesh units                       */
  double ratio_angle;     /* incidence parameter, 1 = normal                    */
  double ux, uz;          /* cos/sin(ratio_angle*pi/2) as the host evaluated    */
  double win;             /* source window factor (FTwindow=0 -> 1.0)           */
  double ak[3][2];        /* Bloch phases akx,aky,akz as (re,im)                */
} iris_params;

typedef struct iris_handle iris_handle;

/*
 * Create a simulation.  Arrays follow the reference's shapes:
 *   eps_inv : complex, (nx+2)*(ny+2)*(hf-hi+3), x fastest, z from hi-1 to hf+1
 *             (eps_inv(1,1,0:nx+1,0:ny+1,hi-1:hf+1), code/mesh.f90:98)
 *   fpzmn, fmzmn : 3*pmlzi doubles, (1:3,1:pmlzi)         (code/driver.f90:189)
 *   fpzmx, fmzmx : 3*(nz-pmlzf+1) doubles, (1:3,pmlzf:nz)  (code/driver.f90:190)
 *   fw   : nof doubles, angular frequency times dt        (code/files.f90:217)
 * Fields start at zero (code/main.f90:169-170).
 */
int iris_create(const iris_params* p, const double* eps_inv, const double* fpzmn,
                const double* fmzmn, const double* fpzmx, const double* fmzmx,
                const double* fw, iris_handle** out);

/* Run time steps it_first..it_last inclusive (1-based, it_first must be the
 * next step); steps 2-11 of the reference loop body for each. Asynchronous
 * with respect to the host: it returns once the work is queued. */
int iris_step(iris_handle* h, int32_t it_first, int32_t it_last);

/* Wait until all queued work on the handle's stream is done. */
int iris_sync(iris_handle* h);

/* Face fluxes Jabs(6,nof) of box nobox (1 or 2) at the current step, scaled
 * by eps0*Q/(2*dt); jabs[iw*6 + face], face 0..5 = J1..J6.  Synchronous. */
int iris_flux(iris_handle* h, int32_t nobox, double* jabs);

/* Number of complex points npts of face pair `pair` (1: x, 2: y, 3: z) of box
 * nobox, as the reference allocates emg(pair)%fdata (1 for an absent pair). */
int iris_dft_points(const iris_handle* h, int32_t nobox, int32_t pair, int64_t* npts);

/* Copy emg(pair)%fdata of box nobox: npts*nof complex, points fastest.
 * Synchronous. */
int iris_get_dft(iris_handle* h, int32_t nobox, int32_t pair, double* out);

/* Copy the E (which=0) or H (which=1) field as complex
 * (3,0:nx+1,0:ny+1,0:nz+1), component fastest.  Synchronous. */
int iris_get_field(iris_handle* h, int32_t which, double* out);

/* Overwrite the E or H field (same layout as iris_get_field); used to start
 * from a given state.  Synchronous. */
int iris_set_field(iris_handle* h, int32_t which, const double* in);

/* Field selectors of iris_get_plane: base + component (0: x, 1: y, 2: z). */
#define IRIS_PLANE_E 0      /* E after the step most recently completed       */
#define IRIS_PLANE_H_PREV 3 /* H of the step before it: what the reference's   */
                            /* snapshots show (h_prev, code/driver.f90:471)   */
#define IRIS_PLANE_H 6      /* H after the step most recently completed       */

/* Copy one field component on one plane: axis 1 is the plane x = index with
 * free coordinates (y, z); axis 2 is y = index with (x, z); axis 3 is
 * z = index with (x, y) -- the reference's snapshot planes (code/sdl.f90:
 * 92-110).  range = {a0, a1, b0, b1} are the inclusive ranges of the first
 * and second free coordinate; index and ranges lie in 1..n of their axis.
 * out receives (a1-a0+1)*(b1-b0+1) complex values, first free coordinate
 * fastest.  IRIS_PLANE_H_PREV needs the fused pass (the in-place two-pass
 * engine overwrites H; IRIS_EUNSUP).  Synchronous. */
int iris_get_plane(iris_handle* h, int32_t what, int32_t axis, int32_t index, const int32_t range[4],
                   double* out);

/* iris_get_plane without any device or host allocation.  No counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph
 * capture elsewhere in the process, tight output loops, harnesses that
 * forbid allocation).  Same arguments, layout and errors as iris_get_plane,
 * plus IRIS_EINVAL when the plane holds more than 1048576 (2^20) values:
 * the values are staged in a fixed 16 MiB device array that the library
 * keeps on every GPU it runs on until the process ends.  Every call on one
 * GPU shares that GPU's array, so calls that overlap on one GPU (from
 * several threads, with handles on the same device) race; calls on one
 * handle are ordered by its stream.  Synchronous. */
int iris_get_plane_noalloc(iris_handle* h, int32_t what, int32_t axis, int32_t index, const int32_t range[4],
                           double* out);

/* Field monitors after the step most recently completed, over x in 1..nx,
 * y in 1..ny, z in zi..zf (within pmlzi+1..pmlzf-1, outside the UPML):
 *   out[0] = sum Re(eps_i conj(E_i^n) E_i^(n-1))    (electric part)
 *   out[1] = sum |H_i^(n-1)|^2 / dtcq2               (magnetic part)
 *   out[2] = max |div D|, div D = sum_i eps_i E_i^n(c) - eps_i E_i^n(c - e_i)
 *   out[3] = max |div B|, div B = sum_i H_i^n(c + e_i) - H_i^n(c)
 * eps_i is the inverse of the factor the E update applies (real part for a
 * complex eps_inv); neighbours across the x/y period take the Bloch factor.
 * out[0]+out[1] and both divergences are invariants of the update in a
 * lossless, source-free region.  Replaces the reference's diagnostics
 * calc_energy_density, calc_div_D, calc_div_B (code/energy.f90:57-201, not
 * called by the reference; its energy routine indexes eps_inv(i,i,...)
 * outside the array and its div D keeps only real parts without eps).
 * Needs the fused pass (IRIS_EUNSUP otherwise).  Synchronous. */
int iris_diagnostics(iris_handle* h, int32_t zi, int32_t zf, double out[4]);

/* The step most recently completed (0 before the first step). */
int iris_current_step(const iris_handle* h, int32_t* it);

/* Release device memory and the stream. NULL is accepted. */
int iris_destroy(iris_handle* h);

/* Message for the last failure on this handle (or of the last failed
 * iris_create when h is NULL).  Never NULL. */
const char* iris_last_error(const iris_handle* h);

/* Kernel classes timed by the profiling hook. */
#define IRIS_K_EPASS 0      /* E update + UPML + max-side wraps          */
#define IRIS_K_HPASS 1      /* H update + UPML + min-side wraps          */
#define IRIS_K_TFSF 2       /* TF/SF incident field and corrections      */
#define IRIS_K_DFT_SAMPLE 3 /* copy of the box faces into the sample ring */
#define IRIS_K_DFT_ACCUM 4  /* fold of buffered samples into emg         */
#define IRIS_K_FLUX 5       /* face-flux reduction                       */
#define IRIS_K_YEE 6        /* fused E+H pass (replaces EPASS/HPASS/TFSF) */
#define IRIS_K_YEE2 7       /* two-step pass over the plain z-chunks        */
#define IRIS_K_YEE_REST 8   /* within a two-step pass: one single step of   */
                            /* the other z-chunks                           */
#define IRIS_K_ADJOINT 9    /* one adjoint step (all its kernels)           */
#define IRIS_K_CHECKPOINT 10 /* state snapshots, restores and H records     */
#define IRIS_K_YEE2_PML 11  /* two-step pass over chunks reaching into the   */
                            /* UPML slabs                                    */
#define IRIS_NKERNEL 12

/* Enable (1) or disable (0) per-kernel timing with HIP events recorded
 * around every launch on the handle's stream.  No counterpart in the
 * reference (its only timing is a progress print, code/driver.f90:340-353). */
int iris_profile(iris_handle* h, int32_t enable);

/* Accumulated device time in ms and launch counts per kernel class since the
 * last reset (arrays of IRIS_NKERNEL); waits for pending work.  reset != 0
 * clears the totals after reading. */
int iris_profile_read(iris_handle* h, double* ms, int64_t* launches, int32_t reset);

/* Layout of the step passes: info[0] = planes per z-chunk; info[1] = planes
 * advanced two steps per pass by the plain two-step kernel (0 when the
 * two-step pass is off: a grid narrower than a tile, or one where it wastes
 * more lanes than single steps, or IRIS_TWO_STEP=0); info[2] = its z-chunks;
 * info[3] = 1 when the pass is on; info[4] = planes advanced by its UPML
 * variant; info[5] = UPML planes among those.  No counterpart in the
 * reference; used to account the algorithmic bytes of the passes. */
int iris_pass_info(const iris_handle* h, int64_t info[6]);

/* Tiles of the two-step pass with two x-cells a lane (FP32 fields, even nx)
 * over an nx x ny plane, as a handle would take them; host arithmetic only,
 * no device.  only_ox: 124, 60 or 28 for that main shape alone, 0 for the
 * default choice.  out[0..4] = main tiles: output cells a row, lanes a row,
 * rows a wave, tile columns, tile rows; out[5..9] = the same for the narrow
 * tiles beside them; out[10] = x of the narrow tiles' first column, less 1.
 * IRIS_EINVAL where no layout fits (odd nx, or a grid too small).  No
 * counterpart in the reference. */
int iris_pair_layout(int32_t nx, int32_t ny, int32_t only_ox, int32_t out[11]);

/* Layout tuner (no counterpart in the reference).  Before the first step,
 * times `steps` steps (2..1000) of candidate layouts of the step passes --
 * chunk lengths, the two-step pass and its UPML variant on or off, the
 * tile widths of both, the dealing of tiles to XCDs -- and keeps the
 * fastest.  Only layouts with the
 * same results bit for bit are compared (with a Bloch phase the choice of the
 * two-step pass, of its UPML variant and the tile widths of both stay as they
 * are).  The field state is restored and the handle
 * stays at step 0.  Needs device memory for one copy of the state; call it
 * before iris_checkpoint.  info[0] = two-step pass on, info[1] = UPML planes
 * in it, info[2] / info[3] = planes per single-step / two-step chunk,
 * info[4] / info[5] = tile rows per XCD run of the single-step / two-step
 * pass, info[6] = layouts timed, info[7] = ns per step of the one kept,
 * info[8] = footprint cells per row of a two-step tile (64, 32, or 128 for
 * two cells a lane with FP32 fields; 0 when the two-step pass is off). */
int iris_autotune(iris_handle* h, int32_t steps, int64_t info[9]);

/* Steps as a captured HIP graph.  iris_step runs whole periods of steps --
 * P = ring slots x sample_fourier (64 x with FP32 fields, 32 x with FP64):
 * the DFT sample ring fills and folds once and the field buffers return to
 * their assignment -- as replays of one captured graph of the period's
 * launches, where the period starts with an empty ring at or after
 * to_fourier and a whole period fits in the call.  Results are bitwise those
 * of the launches one by one.  Not while profiling or recording snapshots;
 * IRIS_GRAPH=0 at iris_create turns it off; so does a profiler's tool
 * library in the environment (ROCP_TOOL_LIBRARIES, set by rocprofv3) unless
 * IRIS_GRAPH=1.  iris_batch_step does the same
 * for the pairs of steps of its two-step handles, with one graph per batch.
 * info[0] = 1 when the handle has a graph captured, info[1] = replays so far
 * that ran the handle's steps (its own graph's and a batch's), info[2] = P.  No counterpart in the
 * reference (its loop, code/driver.f90:338-500, calls the step routines
 * one by one). */
int iris_graph_info(const iris_handle* h, int64_t info[3]);

/* ---- Gradient of a face-flux objective.  The reference has no adjoint;
 * the differentiable form stands beside the plain solver on the same handle.
 * Its forward run is the plain one (iris_step), so the flux it returns is
 * the reference's (code/dataprocess.f90:241-277). ---- */

/* Prepare the forward run for iris_adjoint.  The last R steps before nt
 * record H on the structure planes hi-1..hf as they are taken (these steps
 * are never recomputed); before them the whole field state (E, H and the
 * UPML aux fields D, B) is copied every `segment` steps, and iris_adjoint
 * recomputes each segment from its copy.  R is the number of records that
 * fit in record_bytes (0: none; < 0: as many as fit in 90% of the free
 * device memory beside the snapshots and the adjoint's buffers), at most nt.
 * Call before the first step.  Allocates everything up front (IRIS_ENOMEM
 * when it does not fit).  segment < 0 chooses it: the larger of
 * ceil(sqrt(nt)) and sqrt(nt * snapshot bytes / record bytes), where the
 * snapshots and one segment's records take the least memory together.
 * segment = 0 frees it all.  No counterpart in the reference. */
int iris_checkpoint(iris_handle* h, int32_t segment, int64_t record_bytes);

/* Layout chosen by iris_checkpoint: info[0] = segment, info[1] = snapshots,
 * info[2] = recorded steps R, info[3] = first recorded step nt - R,
 * info[4] = bytes of one step's record, info[5] = bytes of one snapshot. */
int iris_checkpoint_info(const iris_handle* h, int64_t info[6]);

/* Gradient of J = sum_{iw, face} weights[iw*6 + face] * Jabs of box nobox
 * (the iris_flux numbers at step nt) with respect to the per-cell factor
 * eps_inv that the E update applies on the structure planes (reference
 * code/fdtd.f90:115-117), on x in 1..nx, y in 1..ny, z in hi..hf.  grad
 * receives nx*ny*(hf-hi+1) complex values, x fastest then y then z:
 * (dJ/d Re eps_inv, dJ/d Im eps_inv).  The engine runs the discrete adjoint
 * of its own time step backwards from nt, over the recorded steps first and
 * then segment by segment, recomputing each segment forward from its
 * snapshot; it restores the state at nt afterwards and may be called again
 * with other weights.  Needs the handle at step nt with iris_checkpoint
 * called before the first step (IRIS_ESTATE otherwise).  Synchronous.  No
 * counterpart in the reference. */
int iris_adjoint(iris_handle* h, int32_t nobox, const double* weights, double* grad);

/* ---- Batches: many independent runs advanced together.  No counterpart
 * in the reference, which runs one structure at a time (its <<LOOP>> flow,
 * code/program.f90:34-57, runs them one after another).  A sweep over
 * incidence angles, Bloch phases or structures puts one handle per run in a
 * batch; each step is then one launch of the fused pass over the tiles of
 * every run (plus one launch for the DFT samples), with z-chunks as long as
 * the whole batch allows.  Runs that the two-step pass covers take it
 * batched as well: each pair of steps is one launch per kernel over all of
 * them (their single-step chunks, plain and UPML two-step chunks, DFT
 * samples).  The batch lays its handles out for itself while it holds them:
 * z-chunks as long as a large grid's (the launches span every handle), and
 * the UPML slabs of two-step runs in the two-step pass (its UPML variant),
 * which a small grid alone leaves to single steps; iris_batch_destroy puts
 * each handle's own layout back.  Results are bitwise those of stepping each
 * handle alone, except on the UPML planes of two-step runs with a Bloch
 * phase, which then round as the two-step pass does (equal to rounding);
 * IRIS_BATCH_UPML=0 keeps each handle's own choice there, and every result
 * bitwise. ---- */
typedef struct iris_batch iris_batch;

/* Group n handles (created with iris_create, same precision and device, the
 * same to_fourier and sample_fourier; no gradient snapshots).  The handles
 * stay usable on their own; they must outlive the batch. */
int iris_batch_create(iris_handle* const* handles, int32_t n, iris_batch** out);

/* Steps it_first..it_last of every handle of the batch (all at step
 * it_first-1, it_last within each nt).  Queued on the batch's stream; later
 * calls on a handle wait for it.  IRIS_ESTATE when a handle's pass layout
 * was rebuilt after iris_batch_create (iris_autotune): the batch's tables
 * hold the old one. */
int iris_batch_step(iris_batch* b, int32_t it_first, int32_t it_last);

/* info[0] = handles, info[1] = planes per z-chunk of the batched pass,
 * info[2] / info[3] = workgroups of its lean / general launch, info[4] =
 * handles that take the (batched) two-step pass. */
int iris_batch_info(const iris_batch* b, int64_t info[5]);

/* Waits for the batch's work, puts each handle's own layout back (not where
 * it was rebuilt since, e.g. by iris_autotune) and frees the batch. */
int iris_batch_destroy(iris_batch* b);

/* Message for the last failure of the batch (or of the last failed
 * iris_batch_create when b is NULL).  Never NULL. */
const char* iris_batch_last_error(const iris_batch* b);

/* Incident field of an oblique TF/SF source (ratio_angle != 1).  The
 * reference evaluates its Gaussian pulse along u = (cos D, 0, sin D),
 * D = ratio_angle*pi/2 (code/tfsf.f90:231-355), and itself only runs
 * ratio_angle = 1 (code/files.f90:267).  Off normal incidence the engine
 * replaces that pulse by the solution of its own update for the Bloch mode
 * exp(i ak_x ix / nx) in the source medium (a 1D auxiliary grid with
 * absorbing ends, driven by the pulse's time signal), so that every
 * frequency satisfies the x-period's Bloch condition and travels at its own
 * angle (sin theta = kx / (k n), kx = ak_x / (nx Q)); ux, uz only scale the
 * amplitude.  Needs plane_wave = 1, the TF/SF plane below the structure and
 * a real Bloch phase in x only (IRIS_EUNSUP otherwise); IRIS_EINVAL at
 * ratio_angle = 1, where the reference's analytic pulse is used.  out
 * receives (nt+1) x 2 x 3 complex values: for step it = 0..nt, plane
 * iz_min-1 then iz_min, the components 1, 3, 5 (i_pol = 2: Ex, Ez, Hy) or
 * 2, 4, 6 (i_pol = 1: Ey, Hx, Hz) at ix = 0 as the TF/SF corrections read
 * them after step it (times exp(i ak_x ix / nx) at ix).  Host computation
 * only (no device work).  No counterpart in the reference. */
int iris_incident_table(const iris_params* p, double* out);

/* Library version string. */
const char* iris_version(void);

#ifdef __cplusplus
}
#endif

#endif /* IRIS_H */
