// esdg_mlp_internal.h -- the learned indicator's entry points used by the solver (esdg_deviceclass.hip).
#pragma once

#include "esdg_internal.h"
#include "esdg_mlp.h"

namespace esdg {

// Features + inference over the elements of q; p (may be null) per element and estat (may be
// null): 0 smooth (p <= lo), 1 troubled (p >= hi), 2 unsure.
int mlp_classify(esdg_ctx* ctx, const esdg_mlp* m, int Nelem, const double* q, const double* B, const double* VdmInv,
                 float lo, float hi, float* p, signed char* estat);
// visc[e] = 0 where estat[e] == 0
int mlp_mask_viscosity(esdg_ctx* ctx, int Nelem, const signed char* estat, double* visc);

}  // namespace esdg
