// esdg_stage.h -- element-centric single-kernel RK stage over a batch of scenarios.
//
// One launch per RK stage covers every element of every scenario in the batch.  A block (one
// wave) owns EPW = 64/ngl elements of one scenario (a task); inside, the degree of that scenario
// selects the instantiation.  Each element computes the entropy-stable fluxes of its own four
// sides from the stage-start state (the neighbour's trace is read from the same buffer, so the two
// elements of a face produce identical bits), then the flux-differencing volume term, the surface
// lift, the low-storage RK step into the other state buffer and, on request, the positivity
// limiter and the wave speed of the new state (the next step's lambda_max, reduced per scenario
// with an atomic max).  dt is computed on the device; a whole step needs no host synchronisation,
// so steps are captured in a hipGraph.
#pragma once

#include <vector>

#include "esdg_internal.h"

namespace esdg {

struct StageScn {
    int ngl, Nelem, Nfaces, ES, num_flux, positivity, bc;
    double g, geomFace, tol, zero_tol, h_zero;
    double dti;         // (min_ele_size / ngl) * CFL            (dt_i = dti / lambda_max)
    double T;
    const double *Jac, *Yxi, *Yeta, *Xxi, *Xeta, *Bx, *By, *D, *nx, *ny, *scal, *bL, *bR, *EleSizes, *GLw;
    const int *EdgeData, *E2E, *EMS, *EOR;
    double* q[2];       // ping-pong state buffers
    double* gRK;
    double* clock;      // [0] dt of the current step (0 once t >= T), [1] t, [2] steps taken,
                        // [3] dt of the last step taken, [4] 1 if a dt was not positive and finite
    unsigned long long* lmax;   // bit pattern of max lambda (non-negative doubles order as integers)
};

struct StageTask {
    int scn, e0, ngl, pad;
};

// How the tasks of a stage are ordered and launched (the autotuner's search space).
struct StageSchedule {
    int order = 0;          // 0 scenario by scenario; 1 tiles of geometry-sharing scenarios
                            // interleaved (tile-major); 2 degree descending, then scenario by scenario
    int blocks_per_cu = 0;  // 0: one block per task; k > 0: persistent grid of k blocks per CU
    int per_degree = 0;     // 1: one launch per degree group instead of one mixed-degree launch
    int xcd = 0;            // 1: each XCD takes a contiguous range of the task list
    int graph = 1;          // replay pairs of steps from a captured hipGraph
};

// Host side: device copies of the scenario table and the task list, plus the per-step sequence.
class StageBatch {
  public:
    int init(esdg_ctx* ctx, const std::vector<StageScn>& scns);
    int set_schedule(esdg_ctx* ctx, const StageSchedule& sc);
    // initial lambda_max of every scenario from the state in buffer `src`
    int lambda(esdg_ctx* ctx, int src);
    // one time step: dt kernel, then the RK stages; src flips once per stage
    int step(esdg_ctx* ctx, const double* A, const double* C, int stages, int& src);
    // `steps` steps: pairs of steps replayed from a captured graph (schedule.graph), the rest eager
    int run(esdg_ctx* ctx, const double* A, const double* C, int stages, int steps, int& src);
    void release(esdg_ctx* ctx);

    struct Range {
        int beg, n, lo, hi, grid, persistent;
    };
    std::vector<StageScn> host;
    StageSchedule sched;
    std::vector<Range> ranges;
    StageScn* d_scn = nullptr;
    StageTask* d_task = nullptr;
    int ntask = 0;
    int ncu = 0;
    hipGraphExec_t exec[2] = {nullptr, nullptr};   // graph of two steps, one per starting parity

  private:
    void drop_graphs();
};

}  // namespace esdg
