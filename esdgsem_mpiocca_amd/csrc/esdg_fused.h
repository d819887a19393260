// esdg_fused.h -- fused per-stage kernels (a faster schedule of the same per-stage arithmetic).
#pragma once

#include <functional>
#include <string>

#include "esdg_internal.h"

namespace esdg {

// Device pointers the fused stage reads and writes (owned by the deviceclass mirror).
struct FusedArgs {
    int Nelem, Nfaces, artificial_viscosity, positivity;
    double *q, *gRK, *Qt;
    const double *Jac, *Yxi, *Yeta, *Xxi, *Xeta, *B, *Bx, *By, *D, *Dhat, *Dstrong, *VdmInv, *GLw, *EleSizes;
    const double *nx, *ny, *scal, *bL, *bR;
    const int *EdgeData, *EdgeReversed, *ElemToEdge, *EMS, *EOR;
    // element viscosity of the stage: the caller evaluates it (the modal sensor, or the learned
    // indicator with the sensor) on the stage's state before the stage runs
    const double* ViscPara;
    // Optional: the state update (with positivity, after limiting the new state) also evaluates, on
    // the state it writes, the modal sensor into next_visc / next_visc_plot (may be ViscPara itself:
    // each element's value is read before it is replaced) and the element wave speeds into
    // next_lambda.
    double *next_visc = nullptr, *next_visc_plot = nullptr, *next_lambda = nullptr;
    // Optional: dt from device memory (the device clock: clock[0] is the step's dt, 0 for a step
    // issued after T), the update using C_stage * dt; a step with dt = 0 changes nothing.
    const double* dt_dev = nullptr;
    double C_stage = 0.0;
    // Optional (with AV): the pass that evaluates the stage's element viscosity (ViscPara), issued
    // by the stage itself on the stream it is given.  The face fluxes do not read the viscosity, so
    // with the context's second stream (and no per-launch profiling) it runs there beside the
    // face-flux pass, and the face marks follow once both are done.
    std::function<int(hipStream_t)> visc_pass;
};

// The device clock of the fused schedule: clock = {dt, t, steps, last dt, error, bad dt, -, -}.
// One thread: dt = min(T - t, dti / lambda_max [, dvc / eps_max]) from red = {lambda_max, eps_max}
// (src/deviceclass.cpp:899-930, with the host's operation order), t += dt; once t >= T, or after
// a dt that is not positive and finite (flagged), dt = 0.
int fused_clock(esdg_ctx* ctx, const double* red, double* clock, double dti, double dvc, int av, double T);

class FusedStage {
  public:
    int init(esdg_ctx* ctx, int Nelem, int Nfaces, bool av);
    // One RK stage: residual, low-storage accumulate and update (gRK = rkA gRK + Qt; q += cdt gRK).
    // With AV: face -> AV element pass -> viscous face -> volume + surface + viscous + RK [-> limiter].
    int stage(esdg_ctx* ctx, int s, double rkA, double cdt, const FusedArgs& a);
    // side traces of q (the face pass reads them; each stage's update refreshes them): called
    // whenever q was set from outside the stages.  E2E, ED: the mesh's ElemToEdge and EdgeData
    // (the trace slots of the element sides are derived from them once)
    int begin(esdg_ctx* ctx, const double* q, const int* E2E, const int* ED);
    void release(esdg_ctx* ctx);
    void* buffer(const std::string& name);

    bool ready = false;
    int Nelem = 0, Nfaces = 0;
    bool av = false;
    double* fstar = nullptr;    // face fluxes F* [f][3][ngl]
    double* grad = nullptr;     // side traces of the scaled velocity gradients [e][side][4][ngl] (AV)
    double* fvisc = nullptr;    // viscous face fluxes [f][2][ngl] (AV)
    double* qvisc = nullptr;    // viscous volume term [e][2][ngl^2] (hu, hv) (AV)
    double* tr = nullptr;       // side traces of the state in face slots [f][L, R][3][ngl]
    int* slot = nullptr;        // face slot (2 f + 0 / 1) of each element side [e][side]
    bool slots_built = false;
    int* active = nullptr;      // == seq: the element has viscosity in reach this stage (AV)
    int* alist = nullptr;       // the elements marked this stage, each once, in marking order (AV)
    int* acount = nullptr;      // [0], [1]: length of alist for stage mark m at [m & 1]; [2]: the
                                // stage mark m, advanced on the device before each stage (AV)
    int seq = 0;                // stages issued since the marks were last cleared (host count)
    bool capturing = false;     // a step is being captured: no host-side resets inside it
    // the marks restart (every element's mark cleared) when the counter nears its range's end
    int reset_marks(esdg_ctx* ctx);
    // `stages` stages issued by a graph replay: counted, and the marks restarted first if needed
    int advance(esdg_ctx* ctx, int stages);
};

}  // namespace esdg
