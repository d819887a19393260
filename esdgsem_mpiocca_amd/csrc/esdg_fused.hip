// esdg_fused.hip -- fused per-stage kernels.
#include "esdg_fused.h"

namespace esdg {

int FusedStage::init(esdg_ctx* ctx, int Ne, int Nf, bool with_av) {
    Nelem = Ne; Nfaces = Nf; av = with_av;
    return set_error(ctx, ESDG_ERR_ARG, "fused stage is not available in this build");
}

int FusedStage::setup(esdg_ctx*, const int*, const int*, const int*, const int*, const double*, const double*,
                      const double*, const double*) {
    return ESDG_OK;
}

int FusedStage::stage(esdg_ctx* ctx, int, double, double, const FusedArgs&) {
    return set_error(ctx, ESDG_ERR_STATE, "fused stage is not available in this build");
}

void FusedStage::release(esdg_ctx* ctx) {
    for (double** p : {&qn, &fstar, &fgrad, &gtr, &fvisc})
        if (*p) { esdg_free(ctx, *p); *p = nullptr; }
}

void* FusedStage::buffer(const std::string&) { return nullptr; }

}  // namespace esdg
