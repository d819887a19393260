// esdg_deviceclass.hip -- host C++ mirror of the reference's deviceclass (include/deviceclass.h,
// src/deviceclass.cpp) on top of the C ABI: device buffers, kernel selection and the time loop.
//
// Kept from the reference: buffer set and names, kernel call order and arguments, dt rule,
// stage-time bookkeeping.  Changed: the per-stage debug copies of Qt to the host
// (src/deviceclass.cpp:971-972, 1010-1011) are gone; the dt reductions run on the device and
// only two doubles come back per step; DBSurf1/2, never written when DiscBottom = 0 yet read by
// SurfaceKernel (okl/DG/SurfaceKernel.okl:71-72), are zero-filled.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <tuple>
#include <vector>

#include "esdg_device_setup.h"
#include "esdg_fused.h"
#include "esdg_internal.h"
#include "esdg_mlp_internal.h"
#include "esdg_stage.h"

namespace esdg {

class deviceclass {
  public:
    explicit deviceclass(esdg_ctx* ctx) : ctx_(ctx) {}
    ~deviceclass() { freeOccaVars(); }

    // initDeviceVariables (src/deviceclass.cpp:119-342): sizes and allocations.
    int initDeviceVariables(const esdg_inputs& in, const esdg_run_params& rp) {
        Nelem = in.Nelem;
        Nfaces = in.Nfaces;
        rp_ = rp;
        ngl = ctx_->p.ngl;
        ngl2 = ngl * ngl;
        const size_t nodes = (size_t)Nelem * ngl2, fn = (size_t)Nfaces * ngl;
        int rc = ESDG_OK;
        auto A = [&](double** p, size_t n) { if (rc == ESDG_OK) rc = esdg_malloc(ctx_, (void**)p, n * sizeof(double)); };
        auto I = [&](int** p, size_t n) { if (rc == ESDG_OK) rc = esdg_malloc(ctx_, (void**)p, n * sizeof(int)); };
        A(&o_q, nodes * 3); A(&o_Qt, nodes * 3); A(&o_gRK, nodes * 3);
        if (rp.rk_ssp) A(&o_Qtmp, nodes * 3);
        A(&o_D, ngl2); A(&o_Dstrong, ngl2); A(&o_Dhat, ngl2); A(&o_VdmInv, ngl2); A(&o_GLw, ngl);
        A(&o_Jac, nodes); A(&o_Yxi, nodes); A(&o_Yeta, nodes); A(&o_Xxi, nodes); A(&o_Xeta, nodes);
        A(&o_B, nodes); A(&o_Bx, nodes); A(&o_By, nodes); A(&o_x, nodes); A(&o_y, nodes);
        A(&o_nx, fn); A(&o_ny, fn); A(&o_scal, fn);
        A(&o_qL, fn * 3); A(&o_qR, fn * 3); A(&o_bL, fn); A(&o_bR, fn); A(&o_SurfaceParts, fn * 3);
        A(&o_DBSurf1, fn); A(&o_DBSurf2, fn);
        A(&o_EleSizes, Nelem); A(&o_LambdaMax, Nelem); A(&o_ViscPara, Nelem); A(&o_ViscForPlot, Nelem);
        A(&o_red, 4);
        I(&o_EdgeData, (size_t)Nfaces * 8); I(&o_EdgeReversed, Nfaces);
        I(&o_ElemToEdge, (size_t)Nelem * 4); I(&o_ElemEdgeMasterSlave, (size_t)Nelem * 4);
        I(&o_ElemEdgeOrientation, (size_t)Nelem * 4);
        if (rp.artificial_viscosity) {
            A(&o_qGradientX, nodes * 2); A(&o_qGradientY, nodes * 2);
            A(&o_SurfGradientX, fn * 2); A(&o_SurfGradientY, fn * 2);
            A(&o_qGradientXL, fn * 2); A(&o_qGradientXR, fn * 2);
            A(&o_qGradientYL, fn * 2); A(&o_qGradientYR, fn * 2);
            A(&o_SurfacePartsVisc, fn * 2);
            A(&o_ViscParaL, Nfaces); A(&o_ViscParaR, Nfaces);
        }
        // analysis buffers (src/deviceclass.cpp:324-342): arrival times start at 1000.0, maximum
        // elevations at 0
        if (rp.arrival_times) {
            A(&o_ArrivalTimings, nodes);
            if (rc == ESDG_OK) rc = esdg_fill(ctx_, o_ArrivalTimings, 1000.0, nodes);
        }
        if (rp.maximum_elevation) A(&o_MaximumElevation, nodes);
        if (rc == ESDG_OK && rp.fused == 1) rc = fused_.init(ctx_, Nelem, Nfaces, rp.artificial_viscosity != 0);
        if (rp.fused == 2) {
            // single-kernel stage: second state buffer (ping-pong), device clock, lambda_max cell
            A(&o_q2, nodes * 3); A(&o_clock, 8); A(reinterpret_cast<double**>(&o_lmax), 1);
            if (rp.artificial_viscosity) {
                // the stage's gradient side traces, marks and viscosity maximum (the viscous volume
                // term goes to qGradientX, the viscous face fluxes to SurfacePartsVisc)
                A(&o_gtr, (size_t)Nelem * 16 * ngl);
                A(reinterpret_cast<double**>(&o_active), ((size_t)Nelem + 1) / 2);
                A(reinterpret_cast<double**>(&o_emax), 1);
                if (rc == ESDG_OK) rc = esdg_fill(ctx_, reinterpret_cast<double*>(o_active), 0.0, ((size_t)Nelem + 1) / 2);
                if (rc == ESDG_OK) rc = esdg_fill(ctx_, reinterpret_cast<double*>(o_emax), 0.0, 1);
                if (rc == ESDG_OK) rc = esdg_fill(ctx_, o_ViscPara, 0.0, Nelem);
            }
        }
        if (rp.fused == 1) A(&o_clock, 8);   // the fused schedule's device clock
        return rc;
    }

    // copyDeviceVariables (src/deviceclass.cpp:577-626); kind ESDG_H2D for host inputs, ESDG_D2D
    // for inputs already on the device
    int copyDeviceVariables(const esdg_inputs& in, int kind = ESDG_H2D) {
        const size_t nodes = (size_t)Nelem * ngl2, fn = (size_t)Nfaces * ngl;
        int rc = ESDG_OK;
        auto C = [&](void* d, const void* h, size_t bytes) {
            if (rc == ESDG_OK) {
                if (!h) rc = set_error(ctx_, ESDG_ERR_ARG, "esdg_solver_create: missing input array");
                else rc = esdg_memcpy(ctx_, d, h, bytes, kind);
            }
        };
        const size_t dd = sizeof(double);
        C(o_nx, in.nx, fn * dd); C(o_ny, in.ny, fn * dd); C(o_scal, in.scal, fn * dd);
        C(o_Yxi, in.Yxi, nodes * dd); C(o_Yeta, in.Yeta, nodes * dd); C(o_Xxi, in.Xxi, nodes * dd);
        C(o_Xeta, in.Xeta, nodes * dd); C(o_B, in.B, nodes * dd); C(o_Bx, in.Bx, nodes * dd); C(o_By, in.By, nodes * dd);
        C(o_D, in.D, ngl2 * dd); C(o_Dstrong, in.Dstrong, ngl2 * dd); C(o_Dhat, in.Dhat, ngl2 * dd);
        C(o_Jac, in.Jac, nodes * dd); C(o_x, in.x, nodes * dd); C(o_y, in.y, nodes * dd);
        C(o_q, in.q, nodes * 3 * dd); C(o_EleSizes, in.EleSizes, Nelem * dd); C(o_VdmInv, in.VdmInv, ngl2 * dd);
        C(o_GLw, in.GLw, ngl * dd);
        C(o_ElemEdgeMasterSlave, in.ElemEdgeMasterSlave, Nelem * 4 * sizeof(int));
        C(o_ElemEdgeOrientation, in.ElemEdgeOrientation, Nelem * 4 * sizeof(int));
        C(o_ElemToEdge, in.ElemToEdge, Nelem * 4 * sizeof(int));
        C(o_EdgeData, in.EdgeData, (size_t)Nfaces * 8 * sizeof(int));
        C(o_EdgeReversed, in.EdgeReversed, Nfaces * sizeof(int));
        return rc;
    }

    // buildDeviceKernels (src/deviceclass.cpp:349-544): kernels are compiled ahead of time; here the
    // run flags are checked against what is implemented.
    int buildDeviceKernels() {
        if (rp_.rk_order != 3 && rp_.rk_order != 4) return set_error(ctx_, ESDG_ERR_ARG, "rk_order must be 3 or 4");
        if (rp_.rk_order == 4 && rp_.rk_ssp) return set_error(ctx_, ESDG_ERR_ARG, "rk_order 4 has no SSP variant (rkSSP must be 0)");
        if (rp_.artificial_viscosity && ngl < 3) return set_error(ctx_, ESDG_ERR_ARG, "artificial viscosity needs N >= 2");
        if (rp_.fused < 0 || rp_.fused > 2) return set_error(ctx_, ESDG_ERR_ARG, "fused must be 0, 1 or 2");
        if (rp_.fused && (is_manufactured(ctx_->p.bc_kind) || ctx_->p.bc_kind == ESDG_BC_THREE_MOUND))
            return set_error(ctx_, ESDG_ERR_ARG, "time-dependent boundaries run on the reference kernel sequence");
        if (rp_.fused && (rp_.rk_ssp || !rp_.flux_differencing))
            return set_error(ctx_, ESDG_ERR_ARG, "the fused stage supports flux differencing with low-storage RK only");
        if (rp_.fused == 2 && (rp_.arrival_times || rp_.maximum_elevation))
            return set_error(ctx_, ESDG_ERR_ARG, "the analysis buffers are updated by the fused 0 and 1 schedules");
        if (rp_.fused == 2 && (ngl < 3 || ngl > 8))
            return set_error(ctx_, ESDG_ERR_ARG, "the single-kernel stage supports 2 <= N <= 7");
        const int nf = ctx_->p.num_flux;
        if (nf != ESDG_FLUX_ES && nf != ESDG_FLUX_ES_ROTATED && nf != ESDG_FLUX_LF && nf != ESDG_FLUX_ROE &&
            nf != ESDG_FLUX_OCEAN)
            return set_error(ctx_, ESDG_ERR_ARG, "num_flux must be 0, 1, 3, 4 or 5");
        // RK coefficients (src/RungeKutta.cpp:18-90)
        if (rp_.rk_order == 4) {
            stages = 5;
            A[0] = 0.0;
            A[1] = -(567301805773.0 / 1357537059087.0);
            A[2] = -(2404267990393.0 / 2016746695238.0);
            A[3] = -(3550918686646.0 / 2091501179385.0);
            A[4] = -(1275806237668.0 / 842570457699.0);
            Bc[0] = 0.0;
            Bc[1] = (1432997174477.0 / 9575080441755.0);
            Bc[2] = (2526269341429.0 / 6820363962896.0);
            Bc[3] = (2006345519317.0 / 3224310063776.0);
            Bc[4] = (2802321613138.0 / 2924317926251.0);
            Cc[0] = (1432997174477.0 / 9575080441755.0);
            Cc[1] = (5161836677717.0 / 13612068292357.0);
            Cc[2] = (1720146321549.0 / 2090206949498.0);
            Cc[3] = (3134564353537.0 / 4481467310338.0);
            Cc[4] = (2277821191437.0 / 14882151754819.0);
        } else if (rp_.rk_ssp) {
            stages = 3;
            A[0] = 1.0; A[1] = 3.0 / 4.0; A[2] = 1.0 / 3.0;
            Bc[0] = 0.0; Bc[1] = 1.0 / 4.0; Bc[2] = 2.0 / 3.0;
            Cc[0] = 1.0; Cc[1] = 1.0 / 4.0; Cc[2] = 2.0 / 3.0;
        } else {
            stages = 3;
            A[0] = 0.0; A[1] = -5.0 / 9.0; A[2] = -153.0 / 128.0;
            Bc[0] = 0.0; Bc[1] = 1.0 / 3.0; Bc[2] = 3.0 / 4.0;
            Cc[0] = 1.0 / 3.0; Cc[1] = 15.0 / 16.0; Cc[2] = 8.0 / 15.0;
        }
        return ESDG_OK;
    }

    // The element viscosity of the current state (ShockCapturing, src/deviceclass.cpp:918 and
    // :1045): the modal sensor, or with the learned indicator the network's classes and the sensor
    // where the network does not settle the element as smooth.
    // clock_mode (the device clock of the fused schedule; 0 without it): the learned indicator skips
    // a step issued after T (1: the step's first evaluation, 2: before a later stage)
    // stream (learned indicator only; null: the context's stream)
    int viscosity(int clock_mode = 0, hipStream_t stream = nullptr) {
        if (!ind_) return esdg_shock_capturing(ctx_, Nelem, o_q, o_B, o_VdmInv, o_EleSizes, o_ViscPara, o_ViscForPlot);
        return mlp_classify_sensor(ctx_, ind_, Nelem, o_q, o_B, o_VdmInv, ind_lo, ind_hi, o_estat, ind_audit, o_icount,
                                   o_ViscPara, o_ViscForPlot, clock_mode ? o_clock : nullptr, rp_.T, clock_mode, stream);
    }

    // One time step's dt (src/deviceclass.cpp:899-928), reductions on the device.
    // The fused stage may have evaluated FindLambdaMax and ShockCapturing of the current state
    // already (lambda_fresh, visc_fresh): then the time step reduces those values.
    int timestep(double* dt_out) {
        int rc = lambda_fresh ? ESDG_OK : esdg_find_lambda_max(ctx_, Nelem, o_q, o_LambdaMax);
        if (rc == ESDG_OK) rc = esdg_reduce_max(ctx_, Nelem, o_LambdaMax, o_red);
        if (rc == ESDG_OK && rp_.artificial_viscosity) {
            if (!visc_fresh) rc = viscosity();
            if (rc == ESDG_OK) rc = esdg_reduce_max(ctx_, Nelem, o_ViscPara, o_red + 1);
        }
        double red[2] = {0.0, 0.0};
        if (rc == ESDG_OK) rc = esdg_memcpy(ctx_, red, o_red, sizeof red, ESDG_D2H);
        if (rc != ESDG_OK) return rc;
        const double dt_i = rp_.min_ele_size / (ngl) * rp_.CFL / red[0];
        double dt;
        if (rp_.artificial_viscosity) {
            const double dt_v = rp_.DFL / (std::pow(ngl, 2)) * std::pow(rp_.min_ele_size, 2) / red[1];
            dt = std::fmin(rp_.T - t, std::fmin(dt_i, dt_v));
        } else {
            dt = std::fmin(rp_.T - t, dt_i);
        }
        *dt_out = dt;
        return ESDG_OK;
    }

    // The reference's kernel sequence for one RK stage (src/deviceclass.cpp:938-1126).
    int stage(int s, double dt) {
        int rc = esdg_collect_edge_data(ctx_, Nfaces, o_EdgeData, o_q, o_x, o_y, o_nx, o_ny, o_bL, o_bR, o_qL, o_qR,
                                        t + (rp_.rk_ssp ? rkD[s] : Bc[s]) * dt);
        if (rc) return rc;
        if (rp_.flux_differencing)
            rc = esdg_volume_kernel(ctx_, Nelem, o_Jac, o_Yxi, o_Yeta, o_Xxi, o_Xeta, o_q, o_D, o_Bx, o_By, o_Qt);
        else
            rc = esdg_volume_kernel_std(ctx_, Nelem, o_Jac, o_Yxi, o_Yeta, o_Xxi, o_Xeta, o_q, o_D, o_Bx, o_By, o_Qt);
        if (rc) return rc;
        rc = esdg_calc_num_fluxes(ctx_, Nfaces, o_EdgeData, o_EdgeReversed, o_nx, o_ny, o_scal, o_qL, o_qR, o_bL, o_bR,
                                  o_SurfaceParts);
        if (rc) return rc;
        rc = esdg_surface_kernel(ctx_, Nelem, o_Jac, o_ElemEdgeMasterSlave, o_ElemEdgeOrientation, o_ElemToEdge,
                                 o_SurfaceParts, o_DBSurf1, o_DBSurf2, o_Qt);
        if (rc) return rc;
        if (rp_.artificial_viscosity) {
            if (s > 0) rc = esdg_shock_capturing(ctx_, Nelem, o_q, o_B, o_VdmInv, o_EleSizes, o_ViscPara, o_ViscForPlot);
            if (!rc) rc = esdg_calc_num_fluxes_gradient(ctx_, Nfaces, o_EdgeData, o_nx, o_ny, o_scal, o_qL, o_qR,
                                                        o_SurfGradientX, o_SurfGradientY);
            if (!rc) rc = esdg_calc_gradient(ctx_, Nelem, o_Jac, o_Yxi, o_Yeta, o_Xxi, o_Xeta, o_q, o_B, o_Dhat, o_qGradientX,
                                             o_qGradientY);
            if (!rc) rc = esdg_surface_kernel_gradient(ctx_, Nelem, o_Jac, o_ElemEdgeMasterSlave, o_ElemEdgeOrientation,
                                                       o_ElemToEdge, o_SurfGradientX, o_SurfGradientY, o_qGradientX,
                                                       o_qGradientY);
            if (!rc) rc = esdg_scale_gradient(ctx_, Nelem, o_q, o_qGradientX, o_qGradientY);
            if (!rc) rc = esdg_collect_edge_data_gradient(ctx_, Nfaces, o_EdgeData, o_qGradientX, o_qGradientY, o_ViscPara,
                                                          o_ViscParaL, o_ViscParaR, o_qGradientXL, o_qGradientXR,
                                                          o_qGradientYL, o_qGradientYR);
            if (!rc) rc = esdg_volume_kernel_viscose(ctx_, Nelem, o_Jac, o_Yxi, o_Yeta, o_Xxi, o_Xeta, o_qGradientX,
                                                     o_qGradientY, o_Dstrong, o_ViscPara, o_Qt);
            if (!rc) rc = esdg_calc_num_fluxes_viscose(ctx_, Nfaces, o_EdgeData, o_nx, o_ny, o_scal, o_ViscParaL,
                                                       o_ViscParaR, o_qGradientXL, o_qGradientXR, o_qGradientYL,
                                                       o_qGradientYR, o_SurfacePartsVisc);
            if (!rc) rc = esdg_surface_kernel_visc(ctx_, Nelem, o_Jac, o_ElemEdgeMasterSlave, o_ElemEdgeOrientation,
                                                   o_ElemToEdge, o_SurfacePartsVisc, o_Qt);
            if (rc) return rc;
        }
        // manufactured source of the convergence tests (src/deviceclass.cpp:1098-1102)
        if (is_manufactured(ctx_->p.bc_kind)) {
            rc = esdg_add_source(ctx_, Nelem, o_Bx, o_By, o_B, o_x, o_y, t + (rp_.rk_ssp ? rkD[s] : Bc[s]) * dt, o_Qt);
            if (rc) return rc;
        }
        if (rp_.rk_ssp) {
            rc = esdg_update_kernel_ssp(ctx_, Nelem, A[s], Bc[s], Cc[s], dt, o_Qt, o_Qtmp, o_q);
        } else {
            rc = esdg_calc_rk(ctx_, Nelem, o_Qt, A[s], Bc[s], o_gRK);
            if (!rc) rc = esdg_update_kernel_ls(ctx_, Nelem, Cc[s], dt, o_gRK, o_q);
        }
        if (!rc && rp_.positivity) rc = esdg_pos_pres(ctx_, Nelem, o_EleSizes, o_GLw, o_Jac, o_q);
        return rc;
    }

    // DGtimeloop (src/deviceclass.cpp:868-1130) with a step cap.
    int DGtimeloop(int max_steps, int* steps_taken) {
        int rc = ESDG_OK;
        if (!bottom_done) {
            rc = esdg_collect_edge_data_bottom(ctx_, Nfaces, o_EdgeData, o_B, o_bL, o_bR);
            if (rc) return rc;
            bottom_done = true;
        }
        if (rp_.fused == 2) return stage_timeloop(max_steps, steps_taken);
        if (rp_.fused == 1) {
            rc = fused_.begin(ctx_, o_q, o_ElemToEdge, o_EdgeData);
            if (rc) return rc;
            // without per-step host work (the analysis buffers take t from the host) dt and t stay
            // on the device and the host runs ahead
            if (!rp_.arrival_times && !rp_.maximum_elevation) return clock_timeloop(max_steps, steps_taken);
        }
        // q may have been set since the last call: its wave speeds and viscosity are evaluated anew
        lambda_fresh = visc_fresh = false;
        // The fused state update (or the limiter after it) evaluates the sensor (every stage) and
        // the wave speeds (last stage) of the state it writes, in place of the separate passes.
        const bool fuse_next = rp_.fused == 1;
        const bool fuse_sensor = fuse_next && rp_.artificial_viscosity && !ind_;
        int n = 0;
        while (t < rp_.T && n < max_steps) {
            double dt = 0.0;
            rc = timestep(&dt);
            if (rc) return rc;
            if (!(dt > 0.0) || !std::isfinite(dt))
                return set_error(ctx_, ESDG_ERR_STATE, "time step is not positive and finite (dt = " + std::to_string(dt) + ")");
            if (rp_.rk_ssp) {
                rc = hip_check(ctx_, hipMemcpyAsync(o_Qtmp, o_q, (size_t)Nelem * ngl2 * 3 * sizeof(double),
                                                    hipMemcpyDeviceToDevice, ctx_->stream), "Qtmp copy");
                if (rc) return rc;
            }
            for (int s = 0; s < stages; ++s) {
                // the fused stage takes the viscosity of its state from here (stage 0: the time step's)
                if (rp_.fused && rp_.artificial_viscosity && s > 0 && !visc_fresh) {
                    rc = viscosity();
                    if (rc) return rc;
                }
                if (rp_.fused) {
                    FusedArgs fa = fused_args();
                    if (fuse_sensor) { fa.next_visc = o_ViscPara; fa.next_visc_plot = o_ViscForPlot; }
                    if (fuse_next && s == stages - 1) fa.next_lambda = o_LambdaMax;
                    rc = fused_.stage(ctx_, s, A[s], Cc[s] * dt, fa);
                    visc_fresh = fuse_sensor;
                    lambda_fresh = fuse_next && s == stages - 1;
                } else {
                    rc = stage(s, dt);
                }
                if (rc) return rc;
            }
            t += dt;
            last_dt = dt;
            ++n;
            // src/deviceclass.cpp:1132-1141
            if (rp_.arrival_times) rc = esdg_arrival_times(ctx_, Nelem, t, o_q, o_B, o_ArrivalTimings);
            if (!rc && rp_.maximum_elevation) rc = esdg_maximum_elevation(ctx_, Nelem, t, o_q, o_B, o_MaximumElevation);
            if (rc) return rc;
        }
        if (steps_taken) *steps_taken = n;
        return esdg_finish(ctx_);
    }

    // The fused schedule with dt and t on the device (fused_clock): steps are issued in chunks of
    // up to 64 without a host synchronisation and the clock is read back after each chunk.  Steps
    // issued after t reaches T change nothing and are not counted.  Same arithmetic as DGtimeloop.
    int clock_timeloop(int max_steps, int* steps_taken) {
        lambda_fresh = visc_fresh = false;
        // with the learned indicator the network's classes come before the sensor, so the
        // indicator pass runs before every stage instead of the sensor fused into the update
        const bool fuse_sensor = rp_.artificial_viscosity && !ind_;
        const double dti = rp_.min_ele_size / (ngl) * rp_.CFL;
        const double dvc = rp_.DFL / (std::pow(ngl, 2)) * std::pow(rp_.min_ele_size, 2);
        double ck[8] = {0.0, t, 0.0, last_dt, 0.0, 0.0, 0.0, 0.0};
        int rc = esdg_memcpy(ctx_, o_clock, ck, sizeof ck, ESDG_H2D);
        // one step: the reductions and dt on the device, then the stages
        auto issue_step = [&]() {
            int r = lambda_fresh ? ESDG_OK : esdg_find_lambda_max(ctx_, Nelem, o_q, o_LambdaMax);
            if (!r) r = esdg_reduce_max(ctx_, Nelem, o_LambdaMax, o_red);
            if (!r && rp_.artificial_viscosity) {
                if (!visc_fresh) r = viscosity(1);
                if (!r) r = esdg_reduce_max(ctx_, Nelem, o_ViscPara, o_red + 1);
            }
            if (!r) r = fused_clock(ctx_, o_red, o_clock, dti, dvc, rp_.artificial_viscosity, rp_.T);
            for (int s = 0; s < stages && !r; ++s) {
                FusedArgs fa = fused_args();
                if (rp_.artificial_viscosity && s > 0 && !visc_fresh) {
                    // the learned indicator of this stage's state: issued by the stage beside its
                    // face fluxes (which do not read the viscosity)
                    if (ind_) fa.visc_pass = [this](hipStream_t st) { return viscosity(2, st); };
                    else r = viscosity(2);
                }
                if (r) break;
                if (fuse_sensor) { fa.next_visc = o_ViscPara; fa.next_visc_plot = o_ViscForPlot; }
                if (s == stages - 1) fa.next_lambda = o_LambdaMax;
                fa.dt_dev = o_clock;
                fa.C_stage = Cc[s];
                r = fused_.stage(ctx_, s, A[s], 0.0, fa);
                visc_fresh = fuse_sensor;
                lambda_fresh = s == stages - 1;
            }
            return r;
        };
        int issued = 0;
        while (!rc && issued < max_steps && ck[1] < rp_.T) {
            const int chunk = std::min(max_steps - issued, 64);
            for (int k = 0; k < chunk && !rc; ++k) {
                // From the second step of a run the step's launches and arguments repeat (dt, t
                // and the stage marks live on the device): that step is captured once as a graph
                // and replayed.  Profiling runs, and a capture the runtime refuses, stay eager.
                if (lambda_fresh && !ctx_->profile && graph_ok) {
                    if (graph_stale()) drop_step_graph();
                    if (!step_exec) {
                        rc = capture_step(issue_step);
                        if (rc) break;
                    }
                    if (step_exec) {
                        rc = fused_.advance(ctx_, stages);
                        if (!rc) rc = hip_check(ctx_, hipGraphLaunch(step_exec, ctx_->stream), "step graph launch");
                        continue;
                    }
                }
                rc = issue_step();
            }
            issued += chunk;
            if (!rc) rc = esdg_memcpy(ctx_, ck, o_clock, sizeof ck, ESDG_D2H);
            if (!rc && ck[4] != 0.0)
                rc = set_error(ctx_, ESDG_ERR_STATE, "time step is not positive and finite (dt = " + std::to_string(ck[5]) + ")");
        }
        if (rc) return rc;
        t = ck[1];
        last_dt = ck[3];
        if (steps_taken) *steps_taken = (int)ck[2];
        return esdg_finish(ctx_);
    }

    // Captures one step of the fused schedule into step_exec.  A capture that fails leaves the
    // solver eager (graph_ok = false) without an error; the step is then issued eagerly.
    template <class F>
    int capture_step(F&& issue_step) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(ctx_->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            graph_ok = false;
            return ESDG_OK;
        }
        fused_.capturing = true;
        const int r = issue_step();
        fused_.capturing = false;
        const hipError_t e = hipStreamEndCapture(ctx_->stream, &g);
        if (r) {
            if (g) (void)hipGraphDestroy(g);
            return r;
        }
        if (e != hipSuccess || !g || hipGraphInstantiate(&step_exec, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            step_exec = nullptr;
            graph_ok = false;
        }
        if (g) (void)hipGraphDestroy(g);
        // the capture issued nothing: the step is launched from the graph by the caller
        graph_key = current_graph_key();
        return ESDG_OK;
    }
    // what a captured step depends on beyond its buffers: the learned indicator's settings
    std::tuple<const void*, float, float, int> current_graph_key() const {
        return std::make_tuple((const void*)ind_, ind_lo, ind_hi, ind_audit);
    }
    bool graph_stale() const { return step_exec && graph_key != current_graph_key(); }
    void drop_step_graph() {
        if (step_exec) (void)hipGraphExecDestroy(step_exec);
        step_exec = nullptr;
    }

    // The same time loop with the single-kernel stage: dt and t live on the device, steps run in
    // chunks (pairs replayed from a hipGraph) and the clock is read back after each chunk.  Steps
    // issued after t reaches T leave the state unchanged and are not counted.
    int stage_timeloop(int max_steps, int* steps_taken) {
        int rc = ESDG_OK;
        if (!stage_ready) {
            StageScn S{};
            S.ngl = ngl; S.Nelem = Nelem; S.Nfaces = Nfaces; S.ES = ctx_->p.ES; S.num_flux = ctx_->p.num_flux;
            S.positivity = rp_.positivity; S.bc = ctx_->p.bc_kind; S.h_zero = ctx_->p.h_zero;
            S.g = ctx_->p.g; S.geomFace = ctx_->p.geomFace; S.tol = ctx_->p.PosPresTOL; S.zero_tol = ctx_->p.ZeroTOL;
            S.dti = rp_.min_ele_size / (ngl) * rp_.CFL;
            S.T = rp_.T;
            S.Jac = o_Jac; S.Yxi = o_Yxi; S.Yeta = o_Yeta; S.Xxi = o_Xxi; S.Xeta = o_Xeta; S.Bx = o_Bx; S.By = o_By;
            S.D = o_D; S.nx = o_nx; S.ny = o_ny; S.scal = o_scal; S.bL = o_bL; S.bR = o_bR;
            S.EleSizes = o_EleSizes; S.GLw = o_GLw;
            S.EdgeData = o_EdgeData; S.E2E = o_ElemToEdge; S.EMS = o_ElemEdgeMasterSlave; S.EOR = o_ElemEdgeOrientation;
            S.q[0] = o_q; S.q[1] = o_q2; S.gRK = o_gRK; S.clock = o_clock; S.lmax = o_lmax;
            S.emax = o_emax;
            if (rp_.artificial_viscosity) {
                S.av = 1;
                S.dvc = rp_.DFL / (std::pow(ngl, 2)) * std::pow(rp_.min_ele_size, 2);
                S.sc = SensorConsts{ctx_->p.ZeroTOL, ctx_->p.eps0, ctx_->p.sigmaMin, ctx_->p.sigmaMax};
                S.B = o_B; S.VdmInv = o_VdmInv; S.Dhat = o_Dhat; S.Dstrong = o_Dstrong;
                S.visc = o_ViscPara; S.gtr = o_gtr; S.Qv = o_qGradientX; S.SPV = o_SurfacePartsVisc;
                S.active = o_active;
            }
            rc = stage_.init(ctx_, std::vector<StageScn>{S});
            if (rc) return rc;
            stage_ready = true;
            stage_ind = nullptr;
        }
        // the learned indicator as last set on the solver
        if (rp_.artificial_viscosity && (stage_ind != ind_ || stage_lo != ind_lo || stage_hi != ind_hi)) {
            rc = stage_.set_indicator(ctx_, ind_, ind_lo, ind_hi);
            if (rc) return rc;
            stage_ind = ind_; stage_lo = ind_lo; stage_hi = ind_hi;
        }
        if (stage_.ind_audit != ind_audit) {
            stage_.ind_audit = ind_audit;
            stage_.drop_graphs();
        }
        // the state is in o_q at every entry; the clock restarts from the host's t
        double ck[8] = {0.0, t, 0.0, last_dt, 0.0, 0.0, 0.0, 0.0};
        rc = esdg_memcpy(ctx_, o_clock, ck, sizeof ck, ESDG_H2D);
        if (!rc) rc = esdg_fill(ctx_, reinterpret_cast<double*>(o_lmax), 0.0, 1);
        int src = 0;
        if (!rc) rc = stage_.lambda(ctx_, src);
        int issued = 0;
        while (!rc && issued < max_steps && ck[1] < rp_.T && ck[4] == 0.0) {
            const int chunk = std::min(max_steps - issued, 64);
            rc = stage_.run(ctx_, A, Cc, stages, chunk, src);
            if (!rc) rc = esdg_memcpy(ctx_, ck, o_clock, sizeof ck, ESDG_D2H);
            issued += chunk;
        }
        if (!rc && src == 1)
            rc = hip_check(ctx_, hipMemcpyAsync(o_q, o_q2, (size_t)Nelem * ngl2 * 3 * sizeof(double), hipMemcpyDeviceToDevice,
                                                ctx_->stream), "state copy");
        // ViscPara / ViscForPlot of the final state, as the fused schedule leaves them
        if (!rc && rp_.artificial_viscosity && !ind_)
            rc = esdg_shock_capturing(ctx_, Nelem, o_q, o_B, o_VdmInv, o_EleSizes, o_ViscPara, o_ViscForPlot);
        if (!rc) rc = esdg_finish(ctx_);
        if (rc) return rc;
        t = ck[1];
        last_dt = ck[3];
        if (steps_taken) *steps_taken = (int)ck[2];
        if (ck[4] != 0.0) return set_error(ctx_, ESDG_ERR_STATE, "time step is not positive and finite");
        return ESDG_OK;
    }

    int freeOccaVars() {
        drop_step_graph();
        void* ptrs[] = {o_q, o_Qt, o_gRK, o_Qtmp, o_D, o_Dstrong, o_Dhat, o_VdmInv, o_GLw, o_Jac, o_Yxi, o_Yeta, o_Xxi, o_Xeta,
                        o_B, o_Bx, o_By, o_x, o_y, o_nx, o_ny, o_scal, o_qL, o_qR, o_bL, o_bR, o_SurfaceParts, o_DBSurf1,
                        o_DBSurf2, o_EleSizes, o_LambdaMax, o_ViscPara, o_ViscForPlot, o_red, o_EdgeData, o_EdgeReversed,
                        o_ElemToEdge, o_ElemEdgeMasterSlave, o_ElemEdgeOrientation, o_qGradientX, o_qGradientY,
                        o_SurfGradientX, o_SurfGradientY, o_qGradientXL, o_qGradientXR, o_qGradientYL, o_qGradientYR,
                        o_SurfacePartsVisc, o_ViscParaL, o_ViscParaR, o_ArrivalTimings, o_MaximumElevation};
        for (void* p : ptrs)
            if (p) esdg_free(ctx_, p);
        fused_.release(ctx_);
        stage_.release(ctx_);
        for (void* p : {(void*)o_q2, (void*)o_clock, (void*)o_lmax, (void*)o_estat, (void*)o_icount, (void*)o_gtr,
                        (void*)o_active, (void*)o_emax})
            if (p) esdg_free(ctx_, p);
        o_q2 = o_clock = o_gtr = nullptr;
        o_lmax = o_icount = o_emax = nullptr;
        o_active = nullptr;
        o_estat = nullptr;
        o_q = o_Qt = o_gRK = o_Qtmp = nullptr;
        return ESDG_OK;
    }

    FusedArgs fused_args() {
        FusedArgs a;
        a.Nelem = Nelem; a.Nfaces = Nfaces;
        a.artificial_viscosity = rp_.artificial_viscosity; a.positivity = rp_.positivity;
        a.q = o_q; a.gRK = o_gRK; a.Qt = o_Qt;
        a.Jac = o_Jac; a.Yxi = o_Yxi; a.Yeta = o_Yeta; a.Xxi = o_Xxi; a.Xeta = o_Xeta;
        a.B = o_B; a.Bx = o_Bx; a.By = o_By; a.D = o_D; a.Dhat = o_Dhat; a.Dstrong = o_Dstrong;
        a.VdmInv = o_VdmInv; a.GLw = o_GLw; a.EleSizes = o_EleSizes;
        a.nx = o_nx; a.ny = o_ny; a.scal = o_scal; a.bL = o_bL; a.bR = o_bR;
        a.EdgeData = o_EdgeData; a.EdgeReversed = o_EdgeReversed; a.ElemToEdge = o_ElemToEdge;
        a.EMS = o_ElemEdgeMasterSlave; a.EOR = o_ElemEdgeOrientation;
        a.ViscPara = o_ViscPara;
        return a;
    }

    void* buffer(const std::string& n) {
        if (n == "q") return o_q;
        if (n == "Qt") return o_Qt;
        if (n == "gRK") return o_gRK;
        if (n == "qL") return o_qL;
        if (n == "qR") return o_qR;
        if (n == "bL") return o_bL;
        if (n == "bR") return o_bR;
        if (n == "SurfaceParts") return o_SurfaceParts;
        if (n == "LambdaMax") return o_LambdaMax;
        if (n == "ViscPara") return o_ViscPara;
        if (n == "ViscForPlot") return o_ViscForPlot;
        if (n == "ArrivalTimings") return o_ArrivalTimings;
        if (n == "MaximumElevation") return o_MaximumElevation;
        if (n == "qGradientX") return o_qGradientX;
        if (n == "qGradientY") return o_qGradientY;
        return fused_.buffer(n);
    }

    esdg_ctx* ctx_;
    esdg_run_params rp_{};
    int Nelem = 0, Nfaces = 0, ngl = 0, ngl2 = 0, stages = 0;
    double A[5] = {0}, Bc[5] = {0}, Cc[5] = {0};
    const double rkD[3] = {0.0, 1.0, 0.5};
    double t = 0.0, last_dt = 0.0;
    bool bottom_done = false;
    bool lambda_fresh = false, visc_fresh = false;   // o_LambdaMax / o_ViscPara hold those of o_q
    FusedStage fused_;
    StageBatch stage_;
    bool stage_ready = false;
    // learned troubled-cell indicator (esdg_mlp.h)
    esdg_mlp* ind_ = nullptr;
    float ind_lo = 0.0f, ind_hi = 1.0f;
    int ind_audit = 0;
    signed char* o_estat = nullptr;
    unsigned long long* o_icount = nullptr;
    double *o_q2 = nullptr, *o_clock = nullptr;
    unsigned long long* o_lmax = nullptr;
    // single-kernel stage with artificial viscosity
    double* o_gtr = nullptr;
    int* o_active = nullptr;
    unsigned long long* o_emax = nullptr;
    const esdg_mlp* stage_ind = nullptr;     // the indicator the stage batch was last given
    float stage_lo = 0.0f, stage_hi = 0.0f;
    // a captured step of the fused schedule (clock_timeloop) and the settings it was captured with
    hipGraphExec_t step_exec = nullptr;
    bool graph_ok = true;
    std::tuple<const void*, float, float, int> graph_key{nullptr, 0.0f, 0.0f, 0};
    double *o_q = nullptr, *o_Qt = nullptr, *o_gRK = nullptr, *o_Qtmp = nullptr;
    double *o_D = nullptr, *o_Dstrong = nullptr, *o_Dhat = nullptr, *o_VdmInv = nullptr, *o_GLw = nullptr;
    double *o_Jac = nullptr, *o_Yxi = nullptr, *o_Yeta = nullptr, *o_Xxi = nullptr, *o_Xeta = nullptr;
    double *o_B = nullptr, *o_Bx = nullptr, *o_By = nullptr, *o_x = nullptr, *o_y = nullptr;
    double *o_nx = nullptr, *o_ny = nullptr, *o_scal = nullptr;
    double *o_qL = nullptr, *o_qR = nullptr, *o_bL = nullptr, *o_bR = nullptr, *o_SurfaceParts = nullptr;
    double *o_DBSurf1 = nullptr, *o_DBSurf2 = nullptr;
    double *o_EleSizes = nullptr, *o_LambdaMax = nullptr, *o_ViscPara = nullptr, *o_ViscForPlot = nullptr, *o_red = nullptr;
    double *o_ArrivalTimings = nullptr, *o_MaximumElevation = nullptr;
    int *o_EdgeData = nullptr, *o_EdgeReversed = nullptr, *o_ElemToEdge = nullptr, *o_ElemEdgeMasterSlave = nullptr,
        *o_ElemEdgeOrientation = nullptr;
    double *o_qGradientX = nullptr, *o_qGradientY = nullptr, *o_SurfGradientX = nullptr, *o_SurfGradientY = nullptr;
    double *o_qGradientXL = nullptr, *o_qGradientXR = nullptr, *o_qGradientYL = nullptr, *o_qGradientYR = nullptr;
    double *o_SurfacePartsVisc = nullptr, *o_ViscParaL = nullptr, *o_ViscParaR = nullptr;
};

}  // namespace esdg

struct esdg_solver {
    esdg::deviceclass dc;
    explicit esdg_solver(esdg_ctx* c) : dc(c) {}
};

using namespace esdg;

extern "C" {

static int solver_create(esdg_ctx* ctx, const esdg_inputs* in, const esdg_run_params* rp, esdg_solver** out,
                         int kind) {
    ESDG_REQUIRE(ctx, ctx && in && rp && out, "esdg_solver_create: null argument");
    *out = nullptr;
    ESDG_REQUIRE(ctx, in->Nelem > 0 && in->Nfaces > 0, "esdg_solver_create: empty mesh");
    ESDG_REQUIRE(ctx, (int64_t)in->Nelem * ctx->p.ngl * ctx->p.ngl * 3 < (int64_t)INT32_MAX * 8,
                 "esdg_solver_create: mesh too large");
    esdg_solver* s = new (std::nothrow) esdg_solver(ctx);
    if (!s) return set_error(ctx, ESDG_ERR_NOMEM, "esdg_solver_create: out of host memory");
    int rc = s->dc.initDeviceVariables(*in, *rp);
    if (!rc) rc = s->dc.buildDeviceKernels();
    if (!rc) rc = s->dc.copyDeviceVariables(*in, kind);
    if (rc) { delete s; return rc; }
    *out = s;
    return ESDG_OK;
}

int esdg_solver_create(esdg_ctx* ctx, const esdg_inputs* in, const esdg_run_params* rp, esdg_solver** out) {
    return solver_create(ctx, in, rp, out, ESDG_H2D);
}

int esdg_solver_create_dev(esdg_ctx* ctx, const esdg_inputs* in, const esdg_run_params* rp, esdg_solver** out) {
    return solver_create(ctx, in, rp, out, ESDG_D2D);
}

int esdg_solver_run(esdg_solver* s, int max_steps, int* steps_taken, double* t_out) {
    if (!s) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_run: null solver");
    int rc = s->dc.DGtimeloop(max_steps, steps_taken);
    if (t_out) *t_out = s->dc.t;
    return rc;
}

int esdg_solver_get_q(esdg_solver* s, double* host_q) {
    if (!s || !host_q) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_get_q: null argument");
    return esdg_memcpy(s->dc.ctx_, host_q, s->dc.o_q, (size_t)s->dc.Nelem * s->dc.ngl2 * 3 * sizeof(double), ESDG_D2H);
}

int esdg_solver_set_q(esdg_solver* s, const double* host_q) {
    if (!s || !host_q) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_set_q: null argument");
    s->dc.t = 0.0;
    return esdg_memcpy(s->dc.ctx_, s->dc.o_q, host_q, (size_t)s->dc.Nelem * s->dc.ngl2 * 3 * sizeof(double), ESDG_H2D);
}

namespace {
struct CheckpointHeader {
    char magic[8];
    int32_t version, ngl, Nelem, arrival, maximum, pad[3];
    double t, last_dt;
};
}  // namespace

int esdg_solver_checkpoint(esdg_solver* s, const char* path) {
    if (!s || !path) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_checkpoint: null argument");
    auto& d = s->dc;
    const size_t nq = (size_t)d.Nelem * d.ngl2 * 3, nn = (size_t)d.Nelem * d.ngl2;
    std::vector<double> q(nq), g(nq), a(d.o_ArrivalTimings ? nn : 0), m(d.o_MaximumElevation ? nn : 0);
    int rc = esdg_memcpy(d.ctx_, q.data(), d.o_q, nq * sizeof(double), ESDG_D2H);
    if (!rc) rc = esdg_memcpy(d.ctx_, g.data(), d.o_gRK, nq * sizeof(double), ESDG_D2H);
    if (!rc && !a.empty()) rc = esdg_memcpy(d.ctx_, a.data(), d.o_ArrivalTimings, nn * sizeof(double), ESDG_D2H);
    if (!rc && !m.empty()) rc = esdg_memcpy(d.ctx_, m.data(), d.o_MaximumElevation, nn * sizeof(double), ESDG_D2H);
    if (rc) return rc;
    CheckpointHeader h{};
    std::memcpy(h.magic, "ESDGCKP1", 8);
    h.version = 1; h.ngl = d.ngl; h.Nelem = d.Nelem; h.arrival = !a.empty(); h.maximum = !m.empty();
    h.t = d.t; h.last_dt = d.last_dt;
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(d.ctx_, ESDG_ERR_ARG, std::string("esdg_solver_checkpoint: cannot open ") + path);
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
    for (const std::vector<double>* v : {&q, &g, &a, &m})
        if (ok && !v->empty()) ok = std::fwrite(v->data(), sizeof(double), v->size(), f) == v->size();
    ok = (std::fclose(f) == 0) && ok;
    return ok ? ESDG_OK : set_error(d.ctx_, ESDG_ERR_ARG, std::string("esdg_solver_checkpoint: write failed: ") + path);
}

int esdg_solver_restart(esdg_solver* s, const char* path) {
    if (!s || !path) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_restart: null argument");
    auto& d = s->dc;
    const size_t nq = (size_t)d.Nelem * d.ngl2 * 3, nn = (size_t)d.Nelem * d.ngl2;
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(d.ctx_, ESDG_ERR_ARG, std::string("esdg_solver_restart: cannot open ") + path);
    CheckpointHeader h{};
    bool ok = std::fread(&h, sizeof h, 1, f) == 1;
    std::string why;
    if (!ok || std::memcmp(h.magic, "ESDGCKP1", 8) != 0 || h.version != 1) why = "not a checkpoint of this format";
    else if (h.ngl != d.ngl || h.Nelem != d.Nelem) why = "degree or element count differ from the solver's";
    else if ((h.arrival != 0) != (d.o_ArrivalTimings != nullptr) || (h.maximum != 0) != (d.o_MaximumElevation != nullptr))
        why = "analysis buffers differ from the solver's";
    std::vector<double> q(nq), g(nq), a(h.arrival ? nn : 0), m(h.maximum ? nn : 0);
    if (why.empty())
        for (std::vector<double>* v : {&q, &g, &a, &m})
            if (!v->empty() && std::fread(v->data(), sizeof(double), v->size(), f) != v->size()) why = "file too short";
    std::fclose(f);
    if (!why.empty()) return set_error(d.ctx_, ESDG_ERR_ARG, "esdg_solver_restart: " + why + ": " + path);
    int rc = esdg_memcpy(d.ctx_, d.o_q, q.data(), nq * sizeof(double), ESDG_H2D);
    if (!rc) rc = esdg_memcpy(d.ctx_, d.o_gRK, g.data(), nq * sizeof(double), ESDG_H2D);
    if (!rc && !a.empty()) rc = esdg_memcpy(d.ctx_, d.o_ArrivalTimings, a.data(), nn * sizeof(double), ESDG_H2D);
    if (!rc && !m.empty()) rc = esdg_memcpy(d.ctx_, d.o_MaximumElevation, m.data(), nn * sizeof(double), ESDG_H2D);
    if (rc) return rc;
    d.t = h.t;
    d.last_dt = h.last_dt;
    return ESDG_OK;
}

int esdg_solver_time(const esdg_solver* s, double* t) {
    if (!s || !t) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_time: null argument");
    *t = s->dc.t;
    return ESDG_OK;
}

int esdg_solver_last_dt(const esdg_solver* s, double* dt) {
    if (!s || !dt) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_last_dt: null argument");
    *dt = s->dc.last_dt;
    return ESDG_OK;
}

int esdg_solver_buffer(esdg_solver* s, const char* name, void** dptr) {
    if (!s || !name || !dptr) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_buffer: null argument");
    *dptr = s->dc.buffer(name);
    if (!*dptr) return set_error(s->dc.ctx_, ESDG_ERR_ARG, std::string("esdg_solver_buffer: no buffer ") + name);
    return ESDG_OK;
}

int esdg_solver_set_indicator(esdg_solver* s, esdg_mlp* m, float lo, float hi) {
    if (!s) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_set_indicator: null solver");
    deviceclass& d = s->dc;
    if (!m) {
        d.ind_ = nullptr;
        return ESDG_OK;
    }
    if (d.rp_.fused == 0 || !d.rp_.artificial_viscosity)
        return set_error(d.ctx_, ESDG_ERR_ARG, "the learned indicator needs a fused schedule (1 or 2) with artificial viscosity");
    if (d.ngl != 8) return set_error(d.ctx_, ESDG_ERR_ARG, "the learned indicator is defined for N = 7");
    if (!(lo <= hi)) return set_error(d.ctx_, ESDG_ERR_ARG, "esdg_solver_set_indicator: lo > hi");
    int rc = ESDG_OK;
    if (!d.o_estat) rc = esdg_malloc(d.ctx_, (void**)&d.o_estat, (size_t)d.Nelem);
    if (!rc && !d.o_icount) rc = esdg_malloc(d.ctx_, (void**)&d.o_icount, 5 * sizeof(unsigned long long));
    if (rc) return rc;
    d.ind_ = m;
    d.ind_lo = lo;
    d.ind_hi = hi;
    return ESDG_OK;
}

int esdg_solver_indicator_audit(esdg_solver* s, int audit) {
    if (!s) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_indicator_audit: null solver");
    s->dc.ind_audit = audit != 0;
    return ESDG_OK;
}

int esdg_solver_indicator_counts(esdg_solver* s, int64_t* viscous, int64_t* false_negatives, int64_t* sensor_positives,
                                 int64_t* not_smooth, int64_t* element_stages) {
    if (!s) return set_error(nullptr, ESDG_ERR_ARG, "esdg_solver_indicator_counts: null solver");
    deviceclass& d = s->dc;
    unsigned long long c[5] = {0, 0, 0, 0, 0};
    unsigned long long* ic = d.rp_.fused == 2 ? d.stage_.d_icount : d.o_icount;
    if (ic) {
        int rc = esdg_memcpy(d.ctx_, c, ic, sizeof c, ESDG_D2H);
        if (!rc) rc = esdg_fill(d.ctx_, reinterpret_cast<double*>(ic), 0.0, 5);
        if (!rc) rc = esdg_finish(d.ctx_);
        if (rc) return rc;
    }
    if (viscous) *viscous = (int64_t)c[0];
    if (false_negatives) *false_negatives = (int64_t)c[1];
    if (sensor_positives) *sensor_positives = (int64_t)c[2];
    if (not_smooth) *not_smooth = (int64_t)c[3];
    if (element_stages) *element_stages = (int64_t)c[4];
    return ESDG_OK;
}

int esdg_solver_destroy(esdg_solver* s) {
    delete s;
    return ESDG_OK;
}

}  // extern "C"
