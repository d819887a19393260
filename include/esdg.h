/*
 * esdg.h -- C ABI of the MI355X (gfx950) shallow-water DG hot path.
 *
 * Drop-in for the OCCA kernel calls of the reference's deviceclass (include/deviceclass.h:140-214,
 * src/deviceclass.cpp:868-1130): one entry per kernel call, with the call site's argument order
 * and meaning.  All array pointers are device pointers from esdg_malloc (FP64 data, int32
 * connectivity; layouts documented in esdg_setup.h).  Launches are asynchronous on the context's
 * stream; esdg_memcpy and esdg_finish synchronise, as occa::memory::copyTo and
 * occa::device::finish do.  Every function returns 0 on success or a negative esdg_status; the
 * message is available from esdg_last_error().  No C++ exception crosses this ABI.
 *
 * Compile-time kernel constants of the reference (info.addDefine, src/deviceclass.cpp:164-218)
 * are the fields of esdg_params and are fixed at esdg_create.  ngl = N+1 is a runtime value that
 * selects an ahead-of-time compiled kernel instance (2 <= ngl <= 10).
 *
 * Thread-safety: one context is used from one host thread at a time (the reference's one rank
 * per device).  All of a context's work is ordered on its stream, esdg_stream(ctx).  Some work
 * runs on a second stream that the context owns: with the learned indicator its pass beside the
 * face fluxes, and in a batched ensemble with both learned and modal-sensor scenarios the latter's
 * passes and stage launches.  That work is forked from and joined back into the main stream inside the same call,
 * so synchronising esdg_stream(ctx) (or esdg_finish) covers it.
 */
#ifndef ESDG_H
#define ESDG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum esdg_status {
    ESDG_OK = 0,
    ESDG_ERR_ARG = -1,        /* invalid argument (null pointer, size, unsupported ngl ...) */
    ESDG_ERR_HIP = -2,        /* a HIP runtime call failed */
    ESDG_ERR_NOMEM = -3,      /* device or host allocation failed */
    ESDG_ERR_STATE = -4       /* call out of order (e.g. a solver without inputs) */
} esdg_status;

/* Riemann solver selection, the reference's NumFlux (src/deviceclass.cpp:442-477). */
enum esdg_num_flux {
    ESDG_FLUX_ES = 0,          /* okl/RiemannSolvers/calcESFlux.okl */
    ESDG_FLUX_LF = 1,          /* okl/RiemannSolvers/calcLaxFriedrich.okl */
    ESDG_FLUX_ES_ROTATED = 3,  /* okl/RiemannSolvers/calcESFluxRotated.okl */
    ESDG_FLUX_ROE = 4,         /* okl/RiemannSolvers/calcRoeFlux.okl */
    ESDG_FLUX_OCEAN = 5        /* okl/RiemannSolvers/calcCombinedFluxOcean.okl (Roe on open-ocean
                                  boundary faces, rotated ES elsewhere) */
};

/* Boundary condition of CollectEdgeData (src/deviceclass.cpp:363-415). */
enum esdg_bc {
    ESDG_BC_WALLS = 0,         /* okl/GatherEdgeData/SolidWalls.okl */
    ESDG_BC_INNER_OUTER = 1,   /* okl/GatherEdgeData/InnerOuter.okl (zero momentum outside) */
    ESDG_BC_OCEAN = 2,         /* okl/GatherEdgeData/OceanBoundary.okl (open where b <= 6.7) */
    ESDG_BC_CONVERGENCE = 3,   /* okl/GatherEdgeData/Dirichlet_ConvTest.okl: the exterior state is the
                                  manufactured solution at the stage time; the time loop adds its
                                  source term okl/ManufacturedSolutions/S_ConvTest.okl (Testcase 1) */
    ESDG_BC_THREE_MOUND = 4,   /* okl/GatherEdgeData/3MoundInflow.okl: walls on faces with n_x = 0,
                                  inflow (1/2, 1/2, 0), from t > 300 (1, 1, 0), at x = 0, outflow
                                  elsewhere (Testcase 32) */
    /* The other manufactured solutions h = H0 + cos(kx) sin(ky) cos t - b, hu = h/2, hv = 3h/2
     * (src/deviceclass.cpp:363-387): Dirichlet exterior state and source term as kind 3. */
    ESDG_BC_CONV_TEST3 = 5,    /* Dirichlet_ConvTest3.okl + S_ConvTest3.okl: H0 = 20, k = 1 (Testcase 8) */
    ESDG_BC_CONV_TEST4 = 6,    /* Dirichlet_ConvTest4.okl + S_ConvTest4.okl: H0 = 20, k = 1e-3 (Testcase 88,
                                  coordinates in km) */
    ESDG_BC_CONV_TEST5 = 7     /* Dirichlet_ConvTest5.okl + S_ConvTest5.okl: H0 = 20000, k = 1e-6
                                  (Testcase 89, coordinates in m) */
};

typedef struct esdg_params {
    int ngl;              /* N+1 */
    int ES;               /* entropy-stable dissipation in the numerical flux */
    int num_flux;         /* esdg_num_flux */
    int bc_kind;          /* esdg_bc */
    double g;             /* g_const */
    double geomFace;      /* 1/w_GL[0] (main.cpp:921) */
    double PosPresTOL;
    double ZeroTOL;       /* 1e-5 in the reference (src/deviceclass.cpp:148) */
    double eps0, sigmaMin, sigmaMax;
    double h_zero;        /* still-water level h_0 (ocean boundary) */
} esdg_params;

typedef struct esdg_ctx esdg_ctx;

/* ---- context and memory (occa::device, occa::memory) -------------------------------- */
int  esdg_create(esdg_ctx** ctx, int hip_device, const esdg_params* params);
int  esdg_destroy(esdg_ctx* ctx);
int  esdg_malloc(esdg_ctx* ctx, void** d, size_t bytes);      /* zero-filled, 256-B aligned */
int  esdg_free(esdg_ctx* ctx, void* d);
enum esdg_copy_kind { ESDG_H2D = 0, ESDG_D2H = 1, ESDG_D2D = 2 };
int  esdg_memcpy(esdg_ctx* ctx, void* dst, const void* src, size_t bytes, int kind); /* blocking */
int  esdg_fill(esdg_ctx* ctx, double* d, double value, size_t count);   /* async fill kernel */
int  esdg_finish(esdg_ctx* ctx);                               /* occa::device::finish */
size_t esdg_bytes_allocated(const esdg_ctx* ctx);
const char* esdg_last_error(const esdg_ctx* ctx);              /* ctx may be NULL */
void* esdg_stream(esdg_ctx* ctx);                              /* the hipStream_t of ctx */

/* Per-kernel timing with HIP events on the context's stream (for measurement tools). */
int  esdg_profile_enable(esdg_ctx* ctx, int on);
int  esdg_profile_reset(esdg_ctx* ctx);
/* Synchronises, then returns launches and total milliseconds of kernel 'name'. */
int  esdg_profile_query(esdg_ctx* ctx, const char* name, int64_t* launches, double* total_ms);

/* ---- kernels: one per deviceclass kernel call ----------------------------------------- */

/* CollectEdgeData_Bottom (okl/GatherEdgeData/CollectEdgeData_Bottom.okl; deviceclass.cpp:875) */
int esdg_collect_edge_data_bottom(esdg_ctx* ctx, int Nfaces, const int* EdgeData, const double* B,
                                  double* bL, double* bR);
/* CollectEdgeData (okl/GatherEdgeData/{SolidWalls,InnerOuter,OceanBoundary}.okl;
 * deviceclass.cpp:957).  x, y, bL, bR and t are read only by boundary kinds that need them. */
int esdg_collect_edge_data(esdg_ctx* ctx, int Nfaces, const int* EdgeData, const double* q,
                           const double* x, const double* y, const double* nx, const double* ny,
                           const double* bL, const double* bR, double* qL, double* qR, double t);
/* VolumeKernel, flux-differencing form (okl/DG/VolumeKernelFluxDiffV99.okl; deviceclass.cpp:968) */
int esdg_volume_kernel(esdg_ctx* ctx, int Nelem, const double* Jac, const double* Yxi,
                       const double* Yeta, const double* Xxi, const double* Xeta, const double* q,
                       const double* D, const double* Bx, const double* By, double* Qt);
/* VolumeKernel, standard strong form (okl/DG/VolumeKernelV99.okl, FluxDifferencing = 0) */
int esdg_volume_kernel_std(esdg_ctx* ctx, int Nelem, const double* Jac, const double* Yxi,
                           const double* Yeta, const double* Xxi, const double* Xeta,
                           const double* q, const double* D, const double* Bx, const double* By,
                           double* Qt);
/* calcNumFluxes (the okl/RiemannSolvers kernel selected by esdg_params.num_flux; deviceclass.cpp:1004) */
int esdg_calc_num_fluxes(esdg_ctx* ctx, int Nfaces, const int* EdgeData, const int* EdgeReversed,
                         const double* nx, const double* ny, const double* scal, const double* qL,
                         const double* qR, const double* bL, const double* bR,
                         double* SurfaceParts);
/* SurfaceKernel (okl/DG/SurfaceKernel.okl; deviceclass.cpp:1006) */
int esdg_surface_kernel(esdg_ctx* ctx, int Nelem, const double* Jac, const int* ElemEdgeMasterSlave,
                        const int* ElemEdgeOrientation, const int* ElemToEdge, const double* FStar,
                        const double* DBSurf1, const double* DBSurf2, double* Qt);
/* FindLambdaMax (okl/DG/FindLambdaMax.okl; deviceclass.cpp:899) */
int esdg_find_lambda_max(esdg_ctx* ctx, int Nelem, const double* q, double* LambdaMax);
/* calcRK (okl/RungeKutta/rkLS.okl; deviceclass.cpp:1111): gRK = rkA*gRK + Qt */
int esdg_calc_rk(esdg_ctx* ctx, int Nelem, const double* Qt, double rkA, double rkB, double* gRK);
/* UpdateKernel, low storage (okl/RungeKutta/UpdateKernelLS.okl; deviceclass.cpp:1112) */
int esdg_update_kernel_ls(esdg_ctx* ctx, int Nelem, double rkC, double dt, const double* gRK,
                          double* q);
/* UpdateKernel, SSP (okl/RungeKutta/UpdateKernelSSP_V1.okl; deviceclass.cpp:1107) */
int esdg_update_kernel_ssp(esdg_ctx* ctx, int Nelem, double rkA, double rkB, double rkC, double dt,
                           const double* Qt, const double* qTmp, double* q);
/* ShockCapturing (okl/ShockCapturing/ShockCapturing.okl; deviceclass.cpp:918, 1045) */
int esdg_shock_capturing(esdg_ctx* ctx, int Nelem, const double* q, const double* B,
                         const double* VdmInv, const double* EleSizes, double* ViscPara,
                         double* ViscForPlot);
/* preservePosivitity (okl/Positivity/PosPres.okl; deviceclass.cpp:1120) */
int esdg_pos_pres(esdg_ctx* ctx, int Nelem, const double* EleSizes, const double* GLw,
                  const double* Jac, double* q);
/* BR1 gradient chain (okl/BR1_Gradient kernels; deviceclass.cpp:1048-1057) */
int esdg_calc_num_fluxes_gradient(esdg_ctx* ctx, int Nfaces, const int* EdgeData, const double* nx,
                                  const double* ny, const double* scal, const double* qL,
                                  const double* qR, double* SurfGradientX, double* SurfGradientY);
int esdg_calc_gradient(esdg_ctx* ctx, int Nelem, const double* Jac, const double* Yxi,
                       const double* Yeta, const double* Xxi, const double* Xeta, const double* q,
                       const double* B, const double* Dhat, double* qGradientX, double* qGradientY);
int esdg_surface_kernel_gradient(esdg_ctx* ctx, int Nelem, const double* Jac,
                                 const int* ElemEdgeMasterSlave, const int* ElemEdgeOrientation,
                                 const int* ElemToEdge, const double* SurfGradientX,
                                 const double* SurfGradientY, double* qGradientX,
                                 double* qGradientY);
int esdg_scale_gradient(esdg_ctx* ctx, int Nelem, const double* q, double* qGradientX,
                        double* qGradientY);
int esdg_collect_edge_data_gradient(esdg_ctx* ctx, int Nfaces, const int* EdgeData,
                                    const double* qGradientX, const double* qGradientY,
                                    const double* ViscPara, double* ViscParaL, double* ViscParaR,
                                    double* qGradientXL, double* qGradientXR, double* qGradientYL,
                                    double* qGradientYR);
/* viscous terms (okl/ViscoseParts kernels; deviceclass.cpp:1073-1094) */
int esdg_volume_kernel_viscose(esdg_ctx* ctx, int Nelem, const double* Jac, const double* Yxi,
                               const double* Yeta, const double* Xxi, const double* Xeta,
                               const double* qGradientX, const double* qGradientY,
                               const double* Dstrong, const double* ViscPara, double* Qt);
int esdg_calc_num_fluxes_viscose(esdg_ctx* ctx, int Nfaces, const int* EdgeData, const double* nx,
                                 const double* ny, const double* scal, const double* ViscParaL,
                                 const double* ViscParaR, const double* qGradientXL,
                                 const double* qGradientXR, const double* qGradientYL,
                                 const double* qGradientYR, double* SurfacePartsVisc);
int esdg_surface_kernel_visc(esdg_ctx* ctx, int Nelem, const double* Jac,
                             const int* ElemEdgeMasterSlave, const int* ElemEdgeOrientation,
                             const int* ElemToEdge, const double* SurfacePartsVisc, double* Qt);

/* addS (src/deviceclass.cpp:1100-1102, okl/ManufacturedSolutions/S_ConvTest{,3,4,5}.okl): Qt += source
 * of the manufactured solution h = H0 + cos(kx) sin(ky) cos t - b, u = 1/2, v = 3/2 at time t, with
 * H0, k of the context's bc_kind (ESDG_BC_CONV_TEST3..5; H0 = 8, k = 1 for every other kind). */
int esdg_add_source(esdg_ctx* ctx, int Nelem, const double* Bx, const double* By, const double* B,
                    const double* x, const double* y, double t, double* Qt);

/* ---- analysis and solution integrals --------------------------------------------------- */
/* ArrivalTimes (okl/Analysis/ArrivalTimes.okl:16-62; called after every step when
 * calcArrivalTimes, src/deviceclass.cpp:1132-1135): where ArrivalTimings == 1000.0 (its initial
 * value, :324-333) and |h + b - h_zero| > 5e-5, set it to t.  Layout [Nelem][ngl][ngl]. */
int esdg_arrival_times(esdg_ctx* ctx, int Nelem, double t, const double* q, const double* B,
                       double* ArrivalTimings);
/* MaximumElevation (okl/Analysis/MaximumElevation.okl:16-55; :1138-1141): MaximumElevation =
 * max(MaximumElevation, h + b - h_zero) (fmax; initial value 0, :337-342).  t is unused. */
int esdg_maximum_elevation(esdg_ctx* ctx, int Nelem, double t, const double* q, const double* B,
                           double* MaximumElevation);
/* The gauge samples of the time series DGtimeloop records at its plot times (createTimeSeries,
 * src/deviceclass.cpp:841-858, 1196-1222): node_ids[k] is the global node index of gauge k
 * (element node_ids[k] / ngl^2, local node node_ids[k] mod ngl^2); out[k] = h + b - h_zero there.
 * node_ids and out are device arrays; esdg_plot_time_series (esdg_output.h) writes the records. */
int esdg_gauge_sample(esdg_ctx* ctx, int ngauges, const int* node_ids, const double* q, const double* B,
                      double* out);
/* basis::calcTotalMass (src/basis.cpp:1088-1111) on device arrays: sum of h J w_i w_j (Jac = 1/J);
 * basis::calcTotalEntropy (:1004-1055): sum of (h|u|^2/2 + g h^2/2 + g h b) J w_i w_j, u = v = 0
 * where h <= 1e-12.  Device reductions in a fixed order; the result (host) differs from the
 * reference's serial sum by rounding only.  Blocking. */
int esdg_total_mass_dev(esdg_ctx* ctx, int Nelem, const double* q, const double* Jac, const double* GLw,
                        double* host_mass);
int esdg_total_entropy_dev(esdg_ctx* ctx, int Nelem, const double* q, const double* B, const double* Jac,
                           const double* GLw, double* host_entropy);

/* Device-wide max of x[0..n) into *d_out (device pointer).  Replaces the host max loop of
 * GetGlobalLambdaMax / GetGlobalViscParaMax (src/MPI_Communication.cpp:490-541); max is exact. */
int esdg_reduce_max(esdg_ctx* ctx, int n, const double* x, double* d_out);

/* The max of esdg_reduce_max without a context and without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Launches on `stream` (a hipStream_t; null is
 * the default stream) of the current device; the block partials live in a static device array of
 * ESDG_NOALLOC_MAX_N / 256 doubles (128 KiB) that each device holds until the process ends, so n
 * must be <= ESDG_NOALLOC_MAX_N (larger n is rejected with ESDG_ERR_ARG).  The array's address is
 * looked up once per device (by esdg_create for its device, else by the first call there).  Every
 * call on a device shares that array: calls that overlap on one device (two streams, two host
 * threads) race.  Errors are reported through esdg_last_error(NULL). */
#define ESDG_NOALLOC_MAX_N 4194304
int esdg_reduce_max_noalloc(void* stream, int n, const double* x, double* d_out);

/* ---- the time loop (deviceclass::DGtimeloop, src/deviceclass.cpp:630-1130) ------------- */

typedef struct esdg_run_params {
    double CFL, DFL, T;
    double min_ele_size;          /* globalMinEleSize: sqrt of the smallest element area */
    int artificial_viscosity;     /* ArtificialViscosity */
    int positivity;               /* PositivityPreserving */
    int rk_order, rk_ssp;         /* rkorder (3 or 4), rkSSP */
    int flux_differencing;        /* FluxDifferencing */
    int fused;                    /* 0: the reference's kernel sequence; 1: fused stage kernels;
                                     2: single-kernel stage (2 <= N <= 7; with artificial viscosity
                                     the sensor, gradient and viscous face passes run before it) */
    int arrival_times;            /* calcArrivalTimes: buffer "ArrivalTimings" (fused 0 or 1) */
    int maximum_elevation;        /* calcMaximumElevation: buffer "MaximumElevation" (fused 0 or 1) */
} esdg_run_params;

/* Host arrays handed over once, as deviceclass::copyDeviceVariables (deviceclass.cpp:577-626). */
typedef struct esdg_inputs {
    int Nelem, Nfaces;
    const double *GLw, *nx, *ny, *scal, *Yxi, *Yeta, *Xxi, *Xeta, *B, *Bx, *By;
    const double *D, *Dstrong, *Dhat, *Jac, *x, *y, *q, *EleSizes, *VdmInv;
    const int *ElemEdgeMasterSlave, *ElemEdgeOrientation, *ElemToEdge, *EdgeData, *EdgeReversed;
} esdg_inputs;

typedef struct esdg_solver esdg_solver;

/* initDeviceVariables + copyDeviceVariables + buildDeviceKernels in one call. */
int esdg_solver_create(esdg_ctx* ctx, const esdg_inputs* in, const esdg_run_params* rp,
                       esdg_solver** out);
/* DGtimeloop with a step cap: at most max_steps steps or until t >= T.  Returns steps taken
 * and the final time.  dt is computed on the device; the fused and single-kernel schedules keep
 * the clock there and read it back every 64 steps (the fused schedule reads dt back every step
 * when analysis buffers are on). */
int esdg_solver_run(esdg_solver* s, int max_steps, int* steps_taken, double* t_out);
int esdg_solver_get_q(esdg_solver* s, double* host_q);        /* blocking D2H copy of q */
int esdg_solver_set_q(esdg_solver* s, const double* host_q);  /* blocking H2D copy, t = 0 */
int esdg_solver_time(const esdg_solver* s, double* t);
/* Binary checkpoint of a solver between steps, and restart from one (no counterpart in the
 * reference, which only writes plot files).  The file holds t, the last dt, q, the low-storage
 * RK register and, when on, the ArrivalTimings / MaximumElevation buffers, so that a restarted
 * run continues bit for bit.  Restart checks that the file matches the solver (ngl, Nelem,
 * analysis buffers).  Layout: "ESDGCKP1", int32 {1, ngl, Nelem, arrival, maximum, 0, 0, 0},
 * double {t, last dt}, then the arrays as doubles in the state layouts, host byte order. */
int esdg_solver_checkpoint(esdg_solver* s, const char* path);
int esdg_solver_restart(esdg_solver* s, const char* path);
int esdg_solver_last_dt(const esdg_solver* s, double* dt);
/* Device pointer of a named solver buffer ("q","Qt","gRK","ViscPara","LambdaMax", ...). */
int esdg_solver_buffer(esdg_solver* s, const char* name, void** dptr);
int esdg_solver_destroy(esdg_solver* s);

#ifdef __cplusplus
}
#endif
#endif
