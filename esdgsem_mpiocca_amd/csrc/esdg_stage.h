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

#include "esdg_device.h"
#include "esdg_internal.h"

struct esdg_mlp;

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
    // artificial viscosity (av = 1): the modal sensor, or with learned = 1 (N = 7) the learned
    // indicator with the sensor as fallback, evaluated on each stage's state; BR1 gradients and the
    // viscous terms only where viscosity is in reach (active[e] == the stage's mark)
    int av, learned;
    double dvc;         // DFL / ngl^2 * min_ele_size^2            (dt_v = dvc / eps_max)
    SensorConsts sc;
    const double *B, *VdmInv, *Dhat, *Dstrong;
    double* visc;       // element viscosity of the stage's state [Nelem]
    double* gtr;        // side traces of the scaled velocity gradients [Nelem][4][4][ngl]
    double* Qv;         // viscous volume term [Nelem][2][ngl^2] (hu, hv)
    double* SPV;        // viscous face fluxes [Nfaces][2][ngl]
    int* active;        // == mark: viscosity on the element or a face neighbour this stage
    unsigned long long* emax;   // bit pattern of the step's max viscosity
    // per element side, what its face flux needs from the connectivity (built by StageBatch::init
    // from E2E, EOR, EMS and EdgeData; one record per side, shared by the scenarios of a mesh):
    // x face, y EdgeData[0] (left element), z EdgeData[1] (right element), w bits: [0,2) left
    // side, [2,4) right side, 4 EdgeData[4] (reversed), 5 EOR, 6 EMS > 0, [8,16) EdgeData[5] + 2,
    // [16,24) EdgeData[6] + 2
    const int4* srec;
};

struct StageTask {
    int scn, e0, ngl, pad;
};

// An element with viscosity marks itself and its face neighbours for this stage (the elements whose
// viscous terms can be non-zero), as the fused face pass does for both elements of every face with
// viscosity on either side.  Plain stores of one value: concurrent marks agree.
__device__ __forceinline__ void mark_reach(const StageScn& S, int e, int m) {
    const gptr<int> active = glob(S.active);
    active[e] = m;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const gptr<const int> E = glob(S.EdgeData) + 8 * (int64_t)(glob(S.E2E)[e * 4 + s] - 1);
        if (E[5] == 0) active[E[0]] = m;
        if (E[6] == 0) active[E[1]] = m;
    }
}

// How the tasks of a stage are ordered and launched (the autotuner's search space).
struct StageSchedule {
    int order = 0;          // 0 scenario by scenario; 1 tiles of geometry-sharing scenarios
                            // interleaved (tile-major); 2 degree descending, then scenario by scenario
    int blocks_per_cu = 0;  // 0: one block per task; k > 0: persistent grid of k blocks per CU
    int per_degree = 0;     // 1: one launch per degree group instead of one mixed-degree launch
    int xcd = 0;            // 1: each XCD takes a contiguous range of the task list
    int graph = 1;          // replay pairs of steps from a captured hipGraph
    int tile = 1;           // consecutive tasks (element groups of 64/ngl elements) per block: 1, 2 or 4
};

// Host side: device copies of the scenario table and the task list, plus the per-step sequence.
class StageBatch {
  public:
    int init(esdg_ctx* ctx, const std::vector<StageScn>& scns);
    int set_schedule(esdg_ctx* ctx, const StageSchedule& sc);
    // initial lambda_max of every scenario from the state in buffer `src`
    int lambda(esdg_ctx* ctx, int src);
    // one time step: [marks, viscosity,] dt kernel, then the RK stages [each after its viscosity
    // and viscous terms]; src flips once per stage
    int step(esdg_ctx* ctx, const double* A, const double* C, int stages, int& src);
    // `steps` steps: pairs of steps replayed from a captured graph (schedule.graph), the rest eager
    int run(esdg_ctx* ctx, const double* A, const double* C, int stages, int steps, int& src);
    void release(esdg_ctx* ctx);

    struct Range {
        int beg, n, lo, hi, grid, persistent;
        int learned;   // 1: the tasks of learned-indicator scenarios (see `split`)
    };
    std::vector<StageScn> host;
    StageSchedule sched;
    std::vector<Range> ranges;
    StageScn* d_scn = nullptr;
    StageTask* d_task = nullptr;
    int ntask = 0;
    int ncu = 0;
    hipGraphExec_t exec[2] = {nullptr, nullptr};   // graph of two steps, one per starting parity
    // artificial viscosity: element tiles of the AV scenarios for the sensor (row tiles), the
    // learned indicator (32-element tiles of learned scenarios), the gradient pass (256-element tiles) and
    // the viscous face pass; the step's mark counter lives on the device (graph replays advance it)
    bool has_av = false;
    StageTask *d_sens = nullptr, *d_ind = nullptr, *d_grad = nullptr, *d_vface = nullptr;
    int nsens = 0, nind = 0, ngrad = 0, nvface = 0;
    // With both modal-sensor and learned scenarios the stage runs as two streams of work: the
    // other scenarios (sensor pass, their AV terms, their stage launches) on the second stream and
    // the learned ones (indicator pass, AV terms, stage launches) on the main stream, so the
    // indicator overlaps the other scenarios' stage.  The gradient and viscous-face task lists hold
    // the other scenarios' tasks first (ngrad_o, nvface_o of them).
    bool split = false;
    int ngrad_o = 0, nvface_o = 0;
    int* d_mark = nullptr;
    unsigned long long* d_icount = nullptr;        // learned indicator counts (5, as the solver's)
    const esdg_mlp* mlp = nullptr;
    float ind_lo = 0.5f, ind_hi = 0.5f;
    int ind_audit = 0;
    int64_t marks_used = 0;
    // the learned indicator for the AV scenarios with N = 7 (m = null: the modal sensor everywhere)
    int set_indicator(esdg_ctx* ctx, const esdg_mlp* m, float lo, float hi);

    void drop_graphs();

  private:
    std::vector<int4*> srec_bufs;   // side records, one buffer per distinct mesh
    int build_av_tasks(esdg_ctx* ctx);
    int reset_marks(esdg_ctx* ctx);
    int av_pass(esdg_ctx* ctx, int src, int stage);     // viscosity of the stage's state + marks
    int av_terms(esdg_ctx* ctx, int src, int stage);    // gradients, viscous volume and face terms
    // the same over the learned (1) or the other (0) scenarios only, on stream st
    int av_terms_group(esdg_ctx* ctx, int src, int stage, int learned, hipStream_t st);
};

}  // namespace esdg
