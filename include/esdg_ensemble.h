/*
 * esdg_ensemble.h -- batched ensembles of independent shallow-water scenarios on one GPU.
 *
 * No counterpart in the reference, which runs one scenario per process (main.cpp:925-1051 drives
 * one deviceclass).  Each scenario is what one deviceclass::DGtimeloop would advance
 * (src/deviceclass.cpp:868-1130), with its own mesh size, degree, constants, boundary variant,
 * CFL and end time; every scenario's result equals that of a single esdg_solver with fused = 2
 * (and hence of the reference kernel sequence) bit for bit.  All scenarios advance together: one
 * launch per RK stage covers every element of every scenario, the dt of each scenario is computed
 * on the device, and pairs of steps are replayed from a captured hipGraph.
 *
 * Host arrays are uploaded once per distinct (pointer, size): scenarios that pass the same mesh
 * and bathymetry arrays share one device copy.
 *
 * Supported per scenario: 2 <= N <= 7 (3 <= ngl <= 8), flux differencing, ES or rotated ES flux,
 * walls / inner-outer / ocean boundaries, positivity limiter on or off, artificial viscosity on or
 * off (the modal sensor, or for N = 7 the learned indicator, esdg_ensemble_set_indicator),
 * low-storage RK of order 3 or 4 (the same order for the whole ensemble).  A scenario with
 * artificial viscosity equals a single esdg_solver with fused = 1 bit for bit (the fused schedule
 * is the one that runs artificial viscosity; its results equal the reference kernel sequence's).
 */
#ifndef ESDG_ENSEMBLE_H
#define ESDG_ENSEMBLE_H

#include "esdg.h"
#include "esdg_mlp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct esdg_ensemble esdg_ensemble;

/* Task ordering and launch shape of the stage kernel (the autotuner's search space). */
typedef struct esdg_ensemble_schedule {
    int order;          /* 0: scenario by scenario; 1: element tiles of scenarios that share a
                           mesh interleaved (tile-major); 2: highest degree first */
    int blocks_per_cu;  /* 0: one block per element tile; k > 0: persistent grid of k blocks per CU */
    int per_degree;     /* 1: one launch per degree group and stage instead of one mixed launch */
    int xcd;            /* 1: each XCD takes a contiguous part of the task list */
    int graph;          /* 1: replay pairs of steps from a captured hipGraph */
    int tile;           /* element-tile size: a block of the stage kernel advances 1, 2 or 4 groups of
                           64/N+1 elements of one scenario, one group after another (0 reads as 1) */
} esdg_ensemble_schedule;

/* params[i], inputs[i], rp[i] describe scenario i (rp[i].fused is ignored).  Only the inputs the
 * stage reads are used: GLw, nx, ny, scal, Yxi, Yeta, Xxi, Xeta, B, Bx, By, D, Jac, q, EleSizes,
 * ElemEdgeMasterSlave, ElemEdgeOrientation, ElemToEdge, EdgeData. */
int esdg_ensemble_create(esdg_ctx* ctx, int nscn, const esdg_params* params, const esdg_inputs* inputs,
                         const esdg_run_params* rp, esdg_ensemble** out);
int esdg_ensemble_set_schedule(esdg_ensemble* ens, const esdg_ensemble_schedule* sc);
int esdg_ensemble_get_schedule(const esdg_ensemble* ens, esdg_ensemble_schedule* sc);
/* Times trial_steps steps (after one warm-up step) of every candidate schedule, keeps the fastest
 * and restores the ensemble's state, time and step counts.  The candidates: order x blocks_per_cu
 * {0, 4, 8} x per_degree x xcd at tile 1, then tiles 2 and 4 with the winner's other settings.
 * best / best_ms_per_step may be null. */
int esdg_ensemble_autotune(esdg_ensemble* ens, int trial_steps, esdg_ensemble_schedule* best,
                           double* best_ms_per_step);
/* Advances every scenario by `steps` steps; a scenario whose time has reached its T stays put
 * and does not count the steps.  Blocks until done. */
int esdg_ensemble_run(esdg_ensemble* ens, int steps);
/* Issues `steps` steps on the context's stream without waiting (for timing loops). */
int esdg_ensemble_run_async(esdg_ensemble* ens, int steps);
int esdg_ensemble_get_q(esdg_ensemble* ens, int scn, double* host_q);   /* blocking D2H copy */
/* Time, steps taken and last dt of scenario scn.  Returns ESDG_ERR_STATE if its dt was ever not
 * positive and finite (the scenario then stopped). */
int esdg_ensemble_status(esdg_ensemble* ens, int scn, double* t, int* steps_taken, double* last_dt);
/* Device bytes held by the ensemble (state, shared geometry, tables). */
int esdg_ensemble_bytes(const esdg_ensemble* ens, size_t* bytes);
/* The learned indicator (esdg_solver_set_indicator, esdg_mlp.h) for every scenario with artificial
 * viscosity and N = 7; the others keep the modal sensor.  m = null restores the sensor everywhere.
 * m must outlive its use by the ensemble. */
int esdg_ensemble_set_indicator(esdg_ensemble* ens, esdg_mlp* m, float lo, float hi);
int esdg_ensemble_indicator_audit(esdg_ensemble* ens, int audit);
/* Counts since the last call, summed over the learned scenarios, as esdg_solver_indicator_counts. */
int esdg_ensemble_indicator_counts(esdg_ensemble* ens, int64_t* viscous, int64_t* false_negatives,
                                   int64_t* sensor_positives, int64_t* not_smooth, int64_t* element_stages);
int esdg_ensemble_destroy(esdg_ensemble* ens);

#ifdef __cplusplus
}
#endif
#endif
