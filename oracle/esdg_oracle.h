/*
 * esdg_oracle.h -- TEST ORACLE (test infrastructure only; never linked into the product).
 *
 * A plain-C, single-threaded restatement of the reference's per-stage shallow-water kernels
 * (the okl kernels) and of the deviceclass::DGtimeloop driver order (src/deviceclass.cpp:630-1130),
 * with OCCA Serial-mode semantics: every inner-loop block runs to completion, in index
 * order, before the next block of the same outer iteration.  Only tests/, the smoke check and
 * bench.py's CPU-baseline leg may load it.  Layouts are those of include/esdg_setup.h.
 */
#ifndef ESDG_ORACLE_H
#define ESDG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_params {
    int ngl;            /* N+1 */
    int ES;             /* entropy-stable dissipation on (1) or off (0) */
    double g;           /* g_const */
    double geomFace;    /* 1/w_GL[0] */
    double PosPresTOL;
    double ZeroTOL;     /* hard-coded 1e-5 in the reference (src/deviceclass.cpp:148) */
    double eps0, sigmaMin, sigmaMax;
    int bc_kind;        /* exterior state of boundary faces: 0 walls, 1 inner/outer, 2 ocean,
                           3 manufactured (ConvTest), 4 three-mound inflow, 5-7 ConvTest3/4/5 */
    double h_zero;      /* still-water level of the ocean boundary (define h_zero) */
} oracle_params;

/* kernels (one per okl kernel; argument order as the deviceclass call sites) */
void oracle_collect_edge_data_bottom(const oracle_params* p, int Nfaces, const int* EdgeData,
                                     const double* B, double* bL, double* bR);
void oracle_collect_edge_data_walls(const oracle_params* p, int Nfaces, const int* EdgeData,
                                    const double* q, const double* nx, const double* ny,
                                    double* qL, double* qR);
/* CollectEdgeData with the boundary variant p->bc_kind (deviceclass.cpp:363-415) */
void oracle_collect_edge_data(const oracle_params* p, int Nfaces, const int* EdgeData, const double* q,
                              const double* nx, const double* ny, const double* bL, const double* bR,
                              double* qL, double* qR);
/* time-dependent boundaries: 3, 5-7 Dirichlet_ConvTest{,3,4,5}.okl (manufactured solution at time t),
 * 4 3MoundInflow.okl (walls / inflow at x = 0 / outflow) */
void oracle_collect_edge_data_xt(const oracle_params* p, int Nfaces, const int* EdgeData, const double* q,
                                 const double* x, const double* y, const double* nx, const double* ny,
                                 const double* bL, const double* bR, double t, double* qL, double* qR);
/* S_ConvTest{,3,4,5}.okl addS: Qt += manufactured source at time t (variant by p->bc_kind) */
void oracle_add_source(const oracle_params* p, int Nelem, const double* Bx, const double* By, const double* B,
                       const double* x, const double* y, double t, double* Qt);
void oracle_volume_fluxdiff(const oracle_params* p, int Nelem, const double* Jac, const double* Yxi,
                            const double* Yeta, const double* Xxi, const double* Xeta, const double* q,
                            const double* D, const double* Bx, const double* By, double* Qt);
void oracle_num_flux_es(const oracle_params* p, int Nfaces, const int* EdgeReversed,
                        const double* nx, const double* ny, const double* scal, const double* qL,
                        const double* qR, const double* bL, const double* bR, double* SurfaceParts);
void oracle_num_flux_es_rotated(const oracle_params* p, int Nfaces, const int* EdgeReversed,
                                const double* nx, const double* ny, const double* scal,
                                const double* qL, const double* qR, const double* bL,
                                const double* bR, double* SurfaceParts);
void oracle_num_flux_roe(const oracle_params* p, int Nfaces, const int* EdgeReversed,
                         const double* nx, const double* ny, const double* scal, const double* qL,
                         const double* qR, const double* bL, const double* bR, double* SP);
void oracle_num_flux_ocean(const oracle_params* p, int Nfaces, const int* EdgeData, const int* EdgeReversed,
                           const double* nx, const double* ny, const double* scal, const double* qL,
                           const double* qR, const double* bL, const double* bR, double* SP);
void oracle_num_flux_lf(const oracle_params* p, int Nfaces, const int* EdgeReversed,
                        const double* nx, const double* ny, const double* scal, const double* qL,
                        const double* qR, const double* bL, const double* bR, double* SurfaceParts);
void oracle_volume_std(const oracle_params* p, int Nelem, const double* Jac, const double* Yxi,
                       const double* Yeta, const double* Xxi, const double* Xeta, const double* q,
                       const double* D, const double* Bx, const double* By, double* Qt);
void oracle_surface(const oracle_params* p, int Nelem, const double* Jac, const int* EMS,
                    const int* EOR, const int* E2E, const double* FStar, const double* DBSurf1,
                    const double* DBSurf2, double* Qt);
void oracle_find_lambda_max(const oracle_params* p, int Nelem, const double* q, double* LambdaMax);
void oracle_rk_ls_calc(const oracle_params* p, int Nelem, const double* Qt, double rkA, double* gRK);
void oracle_rk_ls_update(const oracle_params* p, int Nelem, double rkC, double dt, const double* gRK,
                         double* q);
void oracle_rk_ssp_update(const oracle_params* p, int Nelem, double a, double b, double c, double dt,
                          const double* Qt, const double* qTmp, double* q);
void oracle_shock_capturing(const oracle_params* p, int Nelem, const double* q, const double* b,
                            const double* VdmInv, double* ViscPara, double* ViscForPlot);
void oracle_pos_pres(const oracle_params* p, int Nelem, const double* EleSizes, const double* GLw,
                     const double* Jac, double* q);
void oracle_num_fluxes_gradient(const oracle_params* p, int Nfaces, const int* EdgeData,
                                const double* nx, const double* ny, const double* scal,
                                const double* qL, const double* qR, double* SGX, double* SGY);
void oracle_calc_gradient(const oracle_params* p, int Nelem, const double* Jac, const double* Yxi,
                          const double* Yeta, const double* Xxi, const double* Xeta, const double* q,
                          const double* Dhat, double* gX, double* gY);
void oracle_surface_gradient(const oracle_params* p, int Nelem, const double* Jac, const int* EMS,
                             const int* EOR, const int* E2E, const double* SGX, const double* SGY,
                             double* gX, double* gY);
void oracle_scale_gradient(const oracle_params* p, int Nelem, const double* q, double* gX, double* gY);
void oracle_collect_edge_data_gradient(const oracle_params* p, int Nfaces, const int* EdgeData,
                                       const double* gX, const double* gY, const double* ViscPara,
                                       double* VL, double* VR, double* gXL, double* gXR,
                                       double* gYL, double* gYR);
void oracle_volume_viscose(const oracle_params* p, int Nelem, const double* Jac, const double* Yxi,
                           const double* Yeta, const double* Xxi, const double* Xeta,
                           const double* gX, const double* gY, const double* Dstrong,
                           const double* ViscPara, double* Qt);
void oracle_num_fluxes_viscose(const oracle_params* p, int Nfaces, const int* EdgeData,
                               const double* nx, const double* ny, const double* scal,
                               const double* VL, const double* VR, const double* gXL,
                               const double* gXR, const double* gYL, const double* gYR,
                               double* SPV);
void oracle_surface_visc(const oracle_params* p, int Nelem, const double* Jac, const int* EMS,
                         const int* EOR, const int* E2E, const double* SPV, double* Qt);

/* DGtimeloop (src/deviceclass.cpp:868-1130) with a step cap: advances q in place for at most
 * max_steps steps or until t >= T.  num_flux: 0 ES, 3 rotated ES.  Returns the number of steps
 * taken; *t_out gets the final time, dt_hist (nullable, max_steps entries) the step sizes. */
/* okl/Analysis/ArrivalTimes.okl:16-62 and MaximumElevation.okl:16-55 (element-wise) */
void oracle_arrival_times(const oracle_params* p, int Nelem, double t, const double* q, const double* b,
                          double* arrival);
void oracle_maximum_elevation(const oracle_params* p, int Nelem, const double* q, const double* b,
                              double* max_elevation);
/* basis::calcTotalMass (src/basis.cpp:1088-1111) and basis::calcTotalEntropy with
 * calcEntropyPointwise (:1004-1055): one serial sum in the reference's order; out[0] */
void oracle_total_mass(const oracle_params* p, int Nelem, const double* q, const double* Jac,
                       const double* GLw, double* out);
void oracle_total_entropy(const oracle_params* p, int Nelem, const double* q, const double* b,
                          const double* Jac, const double* GLw, double* out);

typedef struct oracle_run {
    int Nelem, Nfaces;
    double CFL, DFL, T, min_ele_size;
    int num_flux, artificial_viscosity, positivity, rk_order, rk_ssp;
    int max_steps;
    /* constant inputs */
    const double *Jac, *Yxi, *Yeta, *Xxi, *Xeta, *Bx, *By, *B, *D, *Dhat, *Dstrong, *VdmInv;
    const double *nx, *ny, *scal, *EleSizes, *GLw;
    const int *EdgeData, *EdgeReversed, *ElemToEdge, *EMS, *EOR;
    const double *x, *y;   /* node coordinates (the convergence-test boundary and source) */
} oracle_run;

/* threads used by the element and face loops (OpenMP) */
int oracle_threads(void);
int oracle_timeloop(const oracle_params* p, const oracle_run* r, double* q, double* t_out,
                    double* dt_hist);

#ifdef __cplusplus
}
#endif
#endif
