/*
 * oracle.h -- CPU restatement of the reference FDTD leapfrog loop.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product (libaem) never links, loads or calls it.
 *
 * The oracle consumes the same read-only plan view as the device path
 * (include/aem_plan.h) and advances the reference's time loop
 * (main_fdtd/fdtd_solver.f90:4440-8299) sequentially on the CPU, keeping the
 * reference's FP32 operation order and loop order (the serial build).  It is
 * compiled twice: liboracle.so (FP32, "real" = float) and liboracle64.so
 * (FP64 arithmetic on the same FP32 inputs).
 */
#ifndef AEM_ORACLE_H
#define AEM_ORACLE_H

#include <stdint.h>
#include "../include/aem_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_state oracle_state;

oracle_state* oracle_create(const aem_plan_view* plan);
void          oracle_destroy(oracle_state* s);
/* advance n steps; returns 0 or -1 when the deck's time_steps would be exceeded */
int           oracle_step(oracle_state* s, int n);
int           oracle_steps_done(const oracle_state* s);
/* SPICE decks: one step per call, split at the gap exchange as
 * aem_port_exchange (include/aem.h) splits it; returns -1 on misuse */
int           oracle_port_exchange(oracle_state* s, const float* v_in, float* i_out);
/* post-loop products (f90:8340-8534); call once after the last step */
int           oracle_finalize(oracle_state* s);
/* copy a field component (ncell entries of the common layout) as float */
int           oracle_get_field(const oracle_state* s, int comp, float* out, int64_t n);
/* the same in double (the FP64 build's full precision) */
int           oracle_get_field64(const oracle_state* s, int comp, double* out, int64_t n);
/* overwrite a field component (ncell entries) -- for initial-value tests */
int           oracle_set_field(oracle_state* s, int comp, const float* in, int64_t n);
/* copy a series (kinds 0..7 as aem_series in include/aem.h; 20/21 = SPICE
 * voltages/currents, which are aem_series 8/9; 22/23 = the kmax TM sums,
 * aem_series 10/11) as float.  For a kmax plan (two parts, real and
 * imaginary) a kind or field component OR-ed with 0x100 selects the
 * imaginary part in every accessor. */
int           oracle_get_series(const oracle_state* s, int kind, int idx, float* out, int64_t n);
/* the port and incident series (kinds 0..2) in double: the FP64 build's
 * full precision, for finite differences */
int           oracle_get_series64(const oracle_state* s, int kind, int idx, double* out, int64_t n);
/* raw NTFF accumulators as double: which 0..5 = Ux,Uy,Uz,Wx,Wy,Wz */
int           oracle_get_ntff_raw(const oracle_state* s, int which, double* out, int64_t n);
/* far-field phase correction time per angle (after finalize) */
int           oracle_get_ff_time(const oracle_state* s, float* out, int64_t n);
/* video slice of component comp (0..5 = Ex..Hz) at step n (1-based), vid_size1*vid_size2 floats */
int           oracle_get_video(const oracle_state* s, int comp, int step, float* out, int64_t n);
/* 4 for the FP32 build, 8 for the FP64 build */
int           oracle_real_bytes(void);

#ifdef __cplusplus
}
#endif
#endif
