/*
 * nbody_amd_diag.h -- measurement probes of libnbody_amd.so.  They have no
 * counterpart in the reference and are not part of the drop-in interface
 * (include/nbody_amd.h); the benchmark uses them to state the arithmetic
 * ceiling of this build's force kernels beside their measured FLOP rates.
 */
#ifndef NBODY_AMD_DIAG_H
#define NBODY_AMD_DIAG_H

#include "nbody_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FP32 issue-rate probe on the current device (16 independent v_fma_f32
 * chains per lane, VGPR operands, 8 256-lane blocks per CU).  The library is
 * built without packed FP32, so this, not the published FP32 peak, is the
 * arithmetic ceiling of its force kernels.  Synchronous; *ms (may be NULL)
 * receives the timed launch's duration. */
int nb_probe_fp32_fma(double* tflops, double* ms);
/* The issue cost of one v_rsq_f32 in FMA slots on the current device, from
 * the same chains with one rsqrt before every 12 FMAs (a softened pair term's
 * mix) timed against FMAs alone.  A term of k other instructions and one
 * rsqrt therefore issues at most at k + 1 instructions per (k + *fma_slots)
 * FMA slots.  Synchronous. */
int nb_probe_fp32_rsq_cost(double* fma_slots);

/* Mean duration in ms of an ensemble's batched force launch (one persistent
 * k_ens_force over all systems), from `reps` back-to-back launches between HIP
 * events on the ensemble's stream.  The launches recompute the cached
 * acceleration from the same positions, so the state is unchanged.
 * Synchronous. */
int nb_ensemble_time_force(nb_ensemble* e, int reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_AMD_DIAG_H */
