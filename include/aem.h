/*
 * aem.h -- C ABI of libaem, the MI355X (gfx950) FDTD time-stepping library.
 *
 * libaem replaces the compiled solver executable of the reference
 * (main_fdtd/fdtd_solver.f90, run as "./FDTD inputs.txt" by
 * main_fdtd/master.py:438-441).  The reference has no function-level API:
 * its boundary is a process that reads a positional text deck
 * (fdtd_solver.f90:726-1298) and writes data.dat / *_geometry.bin
 * (fdtd_solver.f90:2355-2357, 8536-8645).  The entry points below split
 * that process into the steps a host program (ctypes, cgo, JNI ...) binds:
 *
 *   aem_plan_from_inputs   parse + setup        (fdtd_solver.f90:719-4369)
 *   aem_sim_create         device allocation and upload (the "!$acc data"
 *                          region, fdtd_solver.f90:4380-4430)
 *   aem_run                leapfrog steps       (fdtd_solver.f90:4440-8299)
 *   aem_finalize           post-loop products   (fdtd_solver.f90:8340-8534)
 *   aem_get_series         port V, incident, far-field, plane sums
 *   aem_write_outputs      data.dat + _geometry.bin (fdtd_solver.f90:8536-8645)
 *
 * Conventions
 *   - Every function returns an aem_status; 0 is success.  The library
 *     never exits the process: deck errors that make the reference STOP
 *     ("ERROR: ...") return AEM_ERR_INPUT with the same message in
 *     aem_last_error().
 *   - Handles are not thread-safe.  Each aem_sim owns one HIP stream and
 *     all of its device buffers; the caller owns every host buffer it
 *     passes in.
 *   - No PyTorch types cross this boundary: plain pointers and sizes only.
 */
#ifndef AEM_H
#define AEM_H

#include <stdint.h>
#include "aem_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum aem_status {
    AEM_OK = 0,
    AEM_ERR_INPUT = 1,        /* malformed deck / reference STOP condition */
    AEM_ERR_IO = 2,           /* file missing or unwritable */
    AEM_ERR_DEVICE = 3,       /* HIP runtime error */
    AEM_ERR_STATE = 4,        /* call out of order, bad handle or argument */
    AEM_ERR_UNSUPPORTED = 5,  /* deck feature outside this library (gridded ports, kmax far field) */
    AEM_ERR_NUMERIC = 6       /* non-finite field detected by the watchdog */
} aem_status;

typedef enum aem_series {
    AEM_SERIES_PORT_V = 0,        /* Voltage_out(port,:)  (f90:8239-8243); idx = port (0-based) */
    AEM_SERIES_PORT_I = 1,        /* Current(port,1) per step: the gap loop integral (f90:4800-4857) */
    AEM_SERIES_INCIDENT = 2,      /* incident(:) (f90:8366-8407) -- valid after aem_finalize */
    AEM_SERIES_E_THETA = 3,       /* E_theta_out(angle,:) (f90:8529-8531) -- after aem_finalize */
    AEM_SERIES_E_PHI = 4,         /* E_phi_out(angle,:) */
    AEM_SERIES_REFLECTED = 5,     /* E_reflected(:) normalised (f90:8342-8361) -- after aem_finalize */
    AEM_SERIES_TRANSMITTED = 6,   /* E_transmitted(:) */
    AEM_SERIES_FF_PC = 7,         /* incident at the far-field phase centre (f90:8636-8642) */
    AEM_SERIES_SPICE_V = 8,       /* Spice_Voltage_out(idx,:) (f90:8250-8254): node voltage minus bias per
                                     SPICE port, then the raw voltage of each additional node */
    AEM_SERIES_SPICE_I = 9,       /* the gap current handed to the circuit per SPICE port (f90:5456-5496) */
    /* kmax build (plan option kmax): AEM_SERIES_REFLECTED / _TRANSMITTED hold the TE sums
     * E_reflected_TE / E_transmitted_TE (f90:7105-7190, normalised f90:8411-8435), these the TM sums */
    AEM_SERIES_REFLECTED_TM = 10,
    AEM_SERIES_TRANSMITTED_TM = 11
} aem_series;

typedef enum aem_field {
    AEM_EX = 0, AEM_EY = 1, AEM_EZ = 2, AEM_HX = 3, AEM_HY = 4, AEM_HZ = 5
} aem_field;

/* A kmax simulation's fields and series are complex (the reference's kmax
 * build declares them complex): a series kind or field component OR-ed with
 * AEM_IMAG selects the imaginary part, without it the real part.  For any
 * other simulation AEM_IMAG is an error (AEM_ERR_STATE). */
#define AEM_IMAG 0x100

typedef struct aem_plan aem_plan;   /* host-side setup result (no GPU needed) */
typedef struct aem_sim  aem_sim;    /* device state of one simulation */

/* Thread-local text of the last error (never NULL). */
const char* aem_last_error(void);

/* Library version string. */
const char* aem_version(void);

/* ---- setup (host only; replaces fdtd_solver.f90:719-4369) ---------------- */

/* Parse an inputs.txt deck (grammar of fdtd_solver.f90:726-1298, documented
 * in main_fdtd/master.py:51-295) and build the plan.  Relative file names in
 * the deck (optional bulk geometry) resolve against the deck's directory.
 * Replaces: the process entry "./FDTD inputs.txt" (fdtd_solver.f90:726-727). */
aem_status aem_plan_from_inputs(const char* inputs_path, aem_plan** out);

/* Same, from deck text held in memory; base_dir resolves relative files. */
aem_status aem_plan_from_text(const char* deck_text, const char* base_dir, aem_plan** out);

/* Fill opt with the reference's setup constants (aem_plan_options in
 * aem_plan.h: CPML thickness, grading, kappa/alpha maxima and reflection
 * target, TF/SF inset, far-field box gap). */
void aem_plan_options_init(aem_plan_options* opt);

/* aem_plan_from_text / aem_plan_from_inputs with setup constants other than
 * the reference's (NULL = the reference's).  No counterpart in the
 * reference, where these are source constants (f90:34-39, 1380-1397,
 * 1638-1643, 1686-1694).  Values outside the accepted range (cpml_cells
 * 2..64, cpml_grading 1..8, kappa_max >= 1, alpha_max >= 0, 0 < reflection
 * < 1, tfsf_inset 3..64, ntff_gap 2..tfsf_inset-1, which keeps the far-field
 * box between the CPML and the TF/SF box) return AEM_ERR_INPUT; so does a grid
 * too small to hold the boxes. */
aem_status aem_plan_from_text_ex(const char* deck_text, const char* base_dir, const aem_plan_options* opt,
                                 aem_plan** out);
aem_status aem_plan_from_inputs_ex(const char* inputs_path, const aem_plan_options* opt, aem_plan** out);

/* Read-only view of the plan (valid until aem_plan_destroy). */
const aem_plan_view* aem_plan_view_get(const aem_plan* plan);

/* Write <name>_geometry.bin exactly as fdtd_solver.f90:2355-2357 does. */
aem_status aem_plan_write_geometry(const aem_plan* plan, const char* path);

void aem_plan_destroy(aem_plan* plan);

/* ---- device lifetime ------------------------------------------------------ */

/* Allocate device state for a plan on the current HIP device and upload
 * coefficients (the reference's "!$acc data copyin" region,
 * fdtd_solver.f90:4380-4430).  The plan must outlive the sim. */
aem_status aem_sim_create(const aem_plan* plan, aem_sim** out);

/* Convenience: aem_plan_from_inputs + aem_sim_create (the sim owns the plan). */
aem_status aem_create_from_inputs(const char* inputs_path, aem_sim** out);

/* Far-field retarded-time binning (f90:7643-7927):
 *   ORDERED    FP32 sums in the serial reference's order -- its output bit
 *              for bit (the default);
 *   ATOMIC     FP64 accumulation through global atomics;
 *   HISTOGRAM  FP64 accumulation through a per-step LDS histogram of the
 *              bin window (fastest with many angles; required for batched
 *              far-field work).
 *   OFF        no time-domain far field (its series are unavailable); for
 *              runs that only need the frequency-domain objective.
 * ATOMIC and HISTOGRAM keep the reference's FP32 index and weight arithmetic
 * and agree with it to FP32 rounding. */
enum {
    AEM_NTFF_ORDERED = 0,
    AEM_NTFF_ATOMIC = 1,
    AEM_NTFF_HISTOGRAM = 2,
    AEM_NTFF_OFF = 3
};

/* Execution options (no counterpart in the reference, whose build flags
 * choose serial/OpenMP/OpenACC).  Zero-initialise, then set fields. */
typedef struct aem_sim_options {
    int32_t ntff_binning;   /* AEM_NTFF_* */
    int32_t planes;         /* z planes marched per thread in the bulk kernels; 0 = default (4) */
    int32_t no_cpml_fusion; /* 1: the separate CPML kernels after ports/TF/SF and one TF/SF launch per face pair */
    int32_t no_single_pass; /* 1: separate H and E bulk launches even where one single-pass launch qualifies */
    int32_t ntff_row_scan;  /* 1: ordered binning by the per-bin scan of staged face rows instead of the
                               contributor lists built at creation (same results; the lists are the default
                               and fall back to the scan when they would exceed 4 GiB) */
} aem_sim_options;

/* aem_sim_create with options (NULL = defaults). */
aem_status aem_sim_create_ex(const aem_plan* plan, const aem_sim_options* options, aem_sim** out);

void aem_destroy(aem_sim* sim);

/* ---- stepping (fdtd_solver.f90:4440-8299) --------------------------------- */

/* Advance n_steps leapfrog steps (H phase, E phase, per-step outputs).  The
 * total may not exceed the deck's time_steps (output arrays are sized by
 * it).  use_graph!=0 replays a captured HIP graph of one step.
 * A deck that exports the full field ("yes" after the video flag,
 * f90:913) appends one record per step to <name>_full_{E,H}{x,y,z}.bin in
 * the current directory (f90:4348-4355, 6992-7003).
 * Every 256 steps and after the last one, every field value is checked:
 * a NaN or infinity (an unstable deck, e.g. a Courant factor below the
 * reference's 0.98, f90:1312) returns AEM_ERR_NUMERIC once the steps are
 * done.  The reference has no such check. */
aem_status aem_run(aem_sim* sim, int64_t n_steps, int32_t use_graph);

/* Steps completed so far. */
int64_t aem_steps_done(const aem_sim* sim);

/* Block until all queued device work of this sim is complete. */
aem_status aem_sync(aem_sim* sim);

/* ---- outputs ---------------------------------------------------------------- */

/* SPICE co-simulation (the reference's use_spice_version build, whose
 * circuit is ngspice driven from the solver: fdtd_solver.f90:5435-5665,
 * 6285-6520, main_fdtd/circuit.F90:140-174).  A deck with SPICE ports is
 * stepped by this call instead of aem_run, one step per call, and the host
 * runs the circuit between calls:
 *   - a call first finishes the step that waits at the exchange: the node
 *     voltages v_in (n_spice + n_spice_extra floats, ports then extra nodes,
 *     as the circuit reports them) set each gap field to (V - bias) / gap
 *     after the lumped ports' E phase, and the rest of the step runs;
 *   - then, when i_out is not NULL, it runs the next step's H phase and
 *     returns the gap current of each SPICE port in i_out (n_spice floats),
 *     and that step waits at the exchange.
 * The first call passes v_in = NULL (no step waits); the last passes
 * i_out = NULL.  Host staging is pinned memory owned by the sim.  The plan
 * view lists the ports (aem_spice_view: location, bias, the grid capacitance
 * the reference assigns to the port's capacitor) and the circuit names. */
aem_status aem_port_exchange(aem_sim* sim, const float* v_in, float* i_out);

/* Post-loop work (fdtd_solver.f90:8340-8534): plane-sum normalisation,
 * incident wave (analytic or the clear 1-D line), NTFF combine and output
 * window.  Call once after the last aem_run. */
aem_status aem_finalize(aem_sim* sim);

/* Copy one time series (length <= time_steps) to host memory. */
aem_status aem_get_series(aem_sim* sim, int32_t kind, int32_t idx, float* out, int64_t n);

/* Copy one field component (plan->ncell floats, common layout) to host. */
aem_status aem_get_field(aem_sim* sim, int32_t comp, float* out, int64_t n);

/* Overwrite one field component from host memory (ncell floats). */
aem_status aem_set_field(aem_sim* sim, int32_t comp, const float* in, int64_t n);

/* Frames [first_step, first_step + n_steps) of the video slice of one
 * component (comp 0..5 = Ex..Hz), n = n_steps * vid_size1 * vid_size2
 * floats, layout [step][size_b][size_a] -- the Ex_video...Hz_video arrays
 * the reference fills every step and writes after the loop
 * (fdtd_solver.f90:6833-6990, 8788-8807).  Needs a deck with video "yes". */
aem_status aem_get_video(aem_sim* sim, int32_t comp, int64_t first_step, int64_t n_steps, float* out, int64_t n);

/* Raw NTFF accumulators U/W (fdtd_solver.f90:7643-7927) as FP64:
 * which = 0..5 -> Ux,Uy,Uz,Wx,Wy,Wz; out has n_angles*len_ff entries
 * ([angle][bin]). */
aem_status aem_get_ntff_raw(aem_sim* sim, int32_t which, double* out, int64_t n);

/* far field phase correction time (f90:8629-8634) per angle, after finalize. */
aem_status aem_get_ff_time(aem_sim* sim, float* out, int64_t n);

/* Discrete Fourier transform of an output series (same kinds and
 * availability as aem_get_series) at n_freq arbitrary frequencies in Hz,
 * computed on the device in FP64 with Kahan-compensated sums:
 *   X(f) = sum_{n=0}^{N-1} x[n] exp(-2 pi i f n dt),
 * the convention of numpy.fft.fft used by the reference's post-processor
 * (utility_scripts/post_processor.py:191-215), where X at f = k/(P dt)
 * equals FFT bin k of the series zero-padded to P samples.  N is the steps
 * run so far for port series and time_steps for the post-loop series.
 * re_im receives n_freq (re, im) pairs.  Replaces the post-processor's FFT
 * step of the S-parameter, gain and RCS calculations (:191-266). */
aem_status aem_get_dft(aem_sim* sim, int32_t kind, int32_t idx, const double* freqs_hz, int32_t n_freq,
                       double* re_im);

/* aem_get_dft without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as aem_get_dft, bit for bit; at most 16384
 * frequencies (more return AEM_ERR_STATE).  The staging lives in a static
 * device array of 384 KiB, one per device, held until the process ends and
 * shared by every call on that device: calls that overlap on one device
 * race. */
aem_status aem_get_dft_noalloc(aem_sim* sim, int32_t kind, int32_t idx, const double* freqs_hz, int32_t n_freq,
                               double* re_im);

/* ---- differentiation --------------------------------------------------------
 * Replaces the reference's adjoint workflow
 * (optimization_scripts/adjoint_optimization/adjoint_optimization.py:55-134,
 * a second forward run with an adjoint source and FFTs of field slices) by
 * the exact discrete adjoint of this solver's update.  Supported decks: CPML
 * or PEC boundaries, simple (dielectric/conductive) materials, plane-wave or
 * no excitation; not lumped ports, thin sheets, Drude media or periodic
 * boundaries (AEM_ERR_UNSUPPORTED). */

/* Far-field objective F = r E_pol(theta, phi, f) e^{jkr}, pol 0 = theta,
 * 1 = phi, accumulated during the run by the frequency-domain
 * near-to-far-field transform of the surface currents of the deck's
 * far-field box:  L = sum M e^{jk r.r'} dS, N = sum J e^{jk r.r'} dS with M
 * sampled at n dt and J at (n - 1/2) dt, E_theta r = -jk/(4 pi)(L_phi +
 * eta N_theta), E_phi r = jk/(4 pi)(L_theta - eta N_phi).  The deck needs a
 * far-field box (at least one angle); call before the first step. */
aem_status aem_objective_set(aem_sim* sim, float theta_deg, float phi_deg, double freq_hz, int32_t pol);
aem_status aem_objective_get(aem_sim* sim, double* re_im);

/* Running-DFT far field (no counterpart in the reference, whose far field is
 * the time-domain transform of f90:7643-7927 followed by the post-processor's
 * FFT, utility_scripts/post_processor.py:245-266): the frequency-domain
 * near-to-far-field transform of the same face currents as the objective,
 * accumulated every step in FP64 with Kahan compensation, at n_ang
 * directions (degrees) and the n_freq frequencies f0 + q df.  Call before the
 * first step; the deck needs a far-field box. */
aem_status aem_ffdft_set(aem_sim* sim, int32_t n_ang, const float* theta_deg, const float* phi_deg, double f0_hz,
                         double df_hz, int32_t n_freq);
/* out[((a * n_freq + q) * 4) + 0..3] = re, im of r E_theta and of r E_phi
 * (e^{jkr} dropped, phase reference at the centre of the phase-centre cell);
 * n = n_ang * n_freq * 4. */
aem_status aem_ffdft_get(aem_sim* sim, double* out, int64_t n);

/* Record the design box of cells [lo, hi] (1-based, inclusive, in the
 * interior outside the CPML slabs) during the run: every step stores E and
 * the update's curl term on the box's edges (time_steps * 6 * edges floats).
 * Call before the first step. */
aem_status aem_adjoint_record(aem_sim* sim, const int32_t* lo, const int32_t* hi);

/* After the full run: gradient of J = |F|^2 with respect to the relative
 * permittivity (all three axes together) of the deck's materials 1..n_mat
 * (declaration order), changed in the cells of the design box only, by the
 * discrete adjoint (transposed update run backwards from the last step) and
 * the setup's four-cell edge averaging (fdtd_solver.f90:2516-2527). */
aem_status aem_adjoint_eps_gradient(aem_sim* sim, int32_t n_mat, double* grad_eps);

/* Design region: the cells [lo, hi] (as for aem_adjoint_record, which this
 * implies) take their relative permittivity per cell from
 * aem_set_design_eps instead of the deck's materials (initially the deck's
 * values).  Conductivity stays the deck's.  The reference reaches per-voxel
 * permittivity only through optional_geom_bulk.bin with one registered
 * material per voxel (fdtd_solver.f90:2051-2095); here the box's edges keep
 * their own Ga/Gb.  Call before the first step. */
aem_status aem_design_box(aem_sim* sim, const int32_t* lo, const int32_t* hi);
/* eps: one isotropic value (>= 1) per cell of the design box, x fastest;
 * n = the box's cell count.  Call before the first step. */
aem_status aem_set_design_eps(aem_sim* sim, const float* eps, int64_t n);
/* After the full run: dJ/d eps_r of every cell of the design box (x fastest,
 * n = the box's cell count), J = |F|^2 of the objective. */
aem_status aem_adjoint_design_gradient(aem_sim* sim, double* grad, int64_t n);

/* Gradient of a general real objective J(F) of the far-field objective F
 * (aem_objective_set), given dJ_dF = {dJ/dRe F, dJ/dIm F} evaluated at the
 * run's F (aem_objective_get): with a design box, dJ/d eps_r of each of its
 * cells (n = the box's cell count, x fastest); otherwise dJ/d eps_r of the
 * deck's materials 1..n (n at most their count) changed in the recorded box.
 * J = |F|^2 is dJ_dF = 2F.  Replaces the gradient step of the reference's
 * adjoint workflow (optimization_scripts/adjoint_optimization/
 * adjoint_optimization.py:76-134), whose objective is a function of a
 * frequency-domain quantity; the caller's J (a goal mismatch, a ratio)
 * enters only through dJ_dF. */
aem_status aem_adjoint(aem_sim* sim, const double* dJ_dF, double* grad, int64_t n);

/* Gradient of a general real objective J of the lumped-port series
 * (AEM_SERIES_PORT_V, AEM_SERIES_PORT_I: one value per step and port),
 * given dJ_dV[q * time_steps + n] = dJ/dV_q(step n+1) and dJ_dI likewise
 * (either may be null), evaluated on the run's series; grad and n as for
 * aem_adjoint.  The far-field objective, if set, does not enter.  The
 * transposed step includes the ports' 1-D lines, the gap fields and the
 * loop currents (fdtd_solver.f90:4800-5013, 6120-6223).  The reference's
 * adjoint workflow (adjoint_optimization.py:76-80) minimises
 * J = |S11_goal - S11|^2 with S11 = V(f0)/Vinc(f0) - 1 (post_processor.py:
 * 255-262), V(f0) = sum_n V(step n+1) exp(-2 pi i f0 n dt) (aem_get_dft):
 * dJ/dV(step n+1) = Re(conj(g) exp(-2 pi i f0 n dt)) with
 * g = 2 (S11 - S11_goal) / conj(Vinc(f0)).  Requires a design record
 * (aem_adjoint_record or aem_design_box) and a completed run; ports must lie
 * a cell clear of the CPML slabs, of the outer boundary and of Drude media
 * (AEM_ERR_UNSUPPORTED at aem_adjoint_record otherwise). */
aem_status aem_adjoint_series(aem_sim* sim, const double* dJ_dV, const double* dJ_dI, double* grad, int64_t n);

/* Checkpoints for the adjoint: the design record keeps `every` steps instead
 * of all of them, the state the step depends on (fields, CPML psi, port
 * lines, Drude currents, the step counter) is saved in device memory before
 * every `every`-th step, and each gradient call recomputes each interval from
 * its snapshot (field update and record only) before transposing it; the
 * state at the end of the run is put back afterwards.  Gradients equal those
 * without checkpoints bit for bit.  Memory: ceil(time_steps/every) + 1
 * snapshots of the state instead of time_steps - every record steps; cost: one
 * more forward pass of the field update per gradient.  No counterpart in the
 * reference (its workflow keeps whole field slices, adjoint_optimization.py:
 * 66-72).  Call once, before aem_adjoint_record / aem_design_box and the
 * first step. */
aem_status aem_adjoint_checkpoint(aem_sim* sim, int32_t every);

/* Low-level adjoint access: apply the transposed step (no source) n_steps
 * times to the adjoint state, and read/write its six field components
 * (layout of aem_get_field).  The adjoint CPML state starts at zero. */
aem_status aem_adjoint_step(aem_sim* sim, int64_t n_steps);
aem_status aem_adjoint_get_field(aem_sim* sim, int32_t comp, float* out, int64_t n);
aem_status aem_adjoint_set_field(aem_sim* sim, int32_t comp, const float* in, int64_t n);

/* ---- batched stepping of independent simulations -------------------------- */

/* Group n created, unfinalized simulations for batched stepping.  Replaces
 * the reference's process fan-out for populations and sweeps
 * (optimization_scripts/GA.py:11-65, main_ga.py:34-72, run_parallel.sh:10):
 * one graph-captured step advances every member, with the bulk updates of
 * all members in one launch per kernel variant and the ports, far-field
 * currents/histogram binning (members created with AEM_NTFF_HISTOGRAM) and
 * step counters in one launch each.  Each member computes exactly what it
 * computes when run alone.  Members keep their own state: read series and
 * fields, finalize and destroy them individually, but do not aem_run a
 * member or destroy it while the batch exists.  Members may differ in grid,
 * materials, sources and time_steps. */
typedef struct aem_batch aem_batch;
aem_status aem_batch_create(aem_sim* const* sims, int32_t n, aem_batch** out);

/* Advance every member n_steps (each member's total may not exceed its
 * time_steps).  use_graph!=0 replays a captured HIP graph of one batched
 * step.  Returns after the device has finished. */
aem_status aem_batch_run(aem_batch* batch, int64_t n_steps, int32_t use_graph);

/* Device time per batched step of the last aem_batch_run (ms). */
float aem_batch_last_run_ms(const aem_batch* batch);

int32_t aem_batch_size(const aem_batch* batch);

void aem_batch_destroy(aem_batch* batch);

/* Write data.dat (positional format read by utility_scripts/post_processor.py)
 * and <name>_geometry.bin, and with video on <name>_video_{E,H}{x,y,z}.bin
 * (one Fortran unformatted record each, f90:8788-8807).  path NULL -> the
 * deck's own file name; <name> is path without its 4-character extension. */
aem_status aem_write_outputs(aem_sim* sim, const char* path);

/* Average device time per step of the last aem_run (ms), from HIP events
 * recorded on the sim's stream around the timed steps. */
float aem_last_run_ms(const aem_sim* sim);

/* Instrumentation with no counterpart in the reference (its only meter is
 * the Mcells/s line, fdtd_solver.f90:8315): advance n_steps eagerly with HIP
 * events on the sim's stream between the five phases of a step and return
 * the average ms per phase in ms[0..4]: the core H bulk launch, the rest of
 * H (slab boxes, sheets, TF/SF, ports, separate CPML, PBC), the core E bulk
 * launch, the rest of E, per-step outputs.  A single-pass step (see
 * aem_single_pass) times its core launch as phase 0 and its slab launches
 * in phase 1. */
aem_status aem_time_phases(aem_sim* sim, int64_t n_steps, float* ms, int32_t n_ms);

/* Instrumentation, no counterpart in the reference: the number of cells of
 * the core bulk launch timed as phases 0 and 2 by aem_time_phases (the box
 * free of CPML slabs when the CPML correction is fused, else the whole
 * updated grid).  -1 for a null handle. */
int64_t aem_core_cells(const aem_sim* sim);

/* Instrumentation, no counterpart in the reference: 1 if a full step of this
 * sim runs its core box as one single-pass launch (H and E of the step,
 * timed as phase 0 by aem_time_phases; phase 2 is then empty), 0 if as
 * separate H and E bulk launches, -1 for a null handle.  The single pass
 * needs a deck with the CPML folded into the bulk and no sheets, TF/SF, PBC
 * or Drude media, and is off for batch members and for material-gradient
 * records (aem_adjoint_record without a design box). */
int32_t aem_single_pass(const aem_sim* sim);

#ifdef __cplusplus
}
#endif

#endif /* AEM_H */
