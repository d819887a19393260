// esdg_mlp_internal.h -- the learned indicator's entry points used by the solver (esdg_deviceclass.hip).
#pragma once

#include "esdg_internal.h"
#include "esdg_mlp.h"
#include "esdg_stage.h"

namespace esdg {

// Features + inference over the elements of q; p (may be null) per element and estat (may be
// null): 0 smooth (p <= lo), 1 troubled (p >= hi), 2 unsure.
int mlp_classify(esdg_ctx* ctx, const esdg_mlp* m, int Nelem, const double* q, const double* B, const double* VdmInv,
                 float lo, float hi, float* p, signed char* estat);
// mlp_classify, and in the same pass the modal sensor with zero viscosity where the class is smooth
// (from the modal coefficients the features already formed).  estat as mlp_classify; counters
// (may be null, 5 entries, accumulated): [0] element-stages with viscosity, [1] sensor positives
// declared smooth (audit), [2] sensor positives (audit), [3] element-stages not declared smooth,
// [4] element-stages classified.
// clock (may be null): the fused schedule's device clock {dt, t, ...}; the pass then does nothing
// (and counts nothing) for a step issued after the end time T: with clock_mode 1 (the step's first
// evaluation, before its dt is known) when t >= T, with clock_mode 2 (before a later stage) when
// the step's dt is 0.  N = 7 only.
int mlp_classify_sensor(esdg_ctx* ctx, const esdg_mlp* m, int Nelem, const double* q, const double* B,
                        const double* VdmInv, float lo, float hi, signed char* estat, int audit,
                        unsigned long long* counters, double* visc, double* visc_plot,
                        const double* clock = nullptr, double T = 0.0, int clock_mode = 0,
                        hipStream_t stream = nullptr);   // null: the context's stream
// The same pass over the learned scenarios of a batched ensemble (tasks: 32-element tiles of N = 7
// scenarios, in buffer src of their state): each viscous element marks itself and its face
// neighbours with *mark + stage; in stage 0 the viscosities enter each scenario's maximum.
int mlp_ensemble_indicator(esdg_ctx* ctx, const esdg_mlp* m, const StageScn* scns, const StageTask* tasks, int ntask,
                           int src, const int* mark, int stage, float lo, float hi, int audit,
                           unsigned long long* counters);


}  // namespace esdg
