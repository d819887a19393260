// esdg_ensemble.hip -- batched ensembles of scenarios on the single-kernel stage (esdg_ensemble.h).
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "esdg_ensemble.h"
#include "esdg_mlp.h"
#include "esdg_internal.h"
#include "esdg_setup.h"
#include "esdg_stage.h"

using namespace esdg;

struct esdg_ensemble {
    esdg_ctx* ctx = nullptr;
    std::vector<StageScn> scn;
    std::vector<size_t> qcount;                          // doubles of one state vector per scenario
    std::map<std::pair<const void*, size_t>, void*> up;  // host (pointer, bytes) -> device copy
    std::map<std::tuple<const void*, const void*, int, int>, std::pair<double*, double*>> btr;
    std::vector<void*> owned;
    double* clocks = nullptr;             // 8 doubles per scenario
    unsigned long long* lmax = nullptr;   // one per scenario
    unsigned long long* emax = nullptr;   // one per scenario (artificial viscosity)
    StageBatch batch;
    int src = 0;
    bool need_lambda = true;
    int stages = 0;
    double A[5] = {0}, Bc[5] = {0}, C[5] = {0};

    ~esdg_ensemble() {
        batch.release(ctx);
        for (void* p : owned) esdg_free(ctx, p);
    }

    void* alloc(size_t bytes, int* rc) {
        void* p = nullptr;
        *rc = esdg_malloc(ctx, &p, bytes ? bytes : 8);
        if (*rc == ESDG_OK) owned.push_back(p);
        return p;
    }

    // device copy of a host array, shared between scenarios that pass the same array
    const void* upload(const void* h, size_t bytes, int* rc) {
        if (!h) {
            *rc = set_error(ctx, ESDG_ERR_ARG, "esdg_ensemble_create: missing input array");
            return nullptr;
        }
        auto it = up.find({h, bytes});
        if (it != up.end()) return it->second;
        void* d = alloc(bytes, rc);
        if (*rc == ESDG_OK) *rc = esdg_memcpy(ctx, d, h, bytes, ESDG_H2D);
        if (*rc == ESDG_OK) up[{h, bytes}] = d;
        return d;
    }

    int start() {
        if (!need_lambda) return ESDG_OK;
        int rc = esdg_fill(ctx, reinterpret_cast<double*>(lmax), 0.0, scn.size());
        if (!rc) rc = batch.lambda(ctx, src);
        if (!rc) need_lambda = false;
        return rc;
    }
};

namespace {

int check_scenario(esdg_ctx* ctx, int i, const esdg_params& p, const esdg_inputs& in, const esdg_run_params& r) {
    const std::string at = "esdg_ensemble_create: scenario " + std::to_string(i) + ": ";
    if (p.ngl < 3 || p.ngl > 8) return set_error(ctx, ESDG_ERR_ARG, at + "needs 2 <= N <= 7");
    if (in.Nelem <= 0 || in.Nfaces <= 0) return set_error(ctx, ESDG_ERR_ARG, at + "empty mesh");
    if ((int64_t)in.Nelem * p.ngl * p.ngl * 3 >= (int64_t)INT32_MAX * 8) return set_error(ctx, ESDG_ERR_ARG, at + "mesh too large");
    if (p.num_flux < ESDG_FLUX_ES || p.num_flux > ESDG_FLUX_OCEAN || p.num_flux == 2)
        return set_error(ctx, ESDG_ERR_ARG, at + "num_flux must be 0, 1, 3, 4 or 5");
    if (p.bc_kind < ESDG_BC_WALLS || p.bc_kind > ESDG_BC_OCEAN)
        return set_error(ctx, ESDG_ERR_ARG, at + "bc_kind must be walls, inner/outer or ocean");
    if (r.artificial_viscosity && !(r.DFL > 0.0)) return set_error(ctx, ESDG_ERR_ARG, at + "DFL must be > 0");
    if (r.arrival_times || r.maximum_elevation)
        return set_error(ctx, ESDG_ERR_ARG, at + "the analysis buffers are not kept for ensembles");
    if (!r.flux_differencing) return set_error(ctx, ESDG_ERR_ARG, at + "needs flux differencing");
    if (r.rk_ssp || (r.rk_order != 3 && r.rk_order != 4)) return set_error(ctx, ESDG_ERR_ARG, at + "needs low-storage RK 3 or 4");
    if (!(r.CFL > 0.0) || !(r.min_ele_size > 0.0)) return set_error(ctx, ESDG_ERR_ARG, at + "CFL and min_ele_size must be > 0");
    return ESDG_OK;
}

// CollectEdgeData_Bottom for one scenario's degree (the kernel takes ngl from the context)
int bottom_traces(esdg_ctx* ctx, int ngl, int Nfaces, const int* ED, const double* B, double* bL, double* bR) {
    const int keep = ctx->p.ngl;
    ctx->p.ngl = ngl;
    const int rc = esdg_collect_edge_data_bottom(ctx, Nfaces, ED, B, bL, bR);
    ctx->p.ngl = keep;
    return rc;
}

}  // namespace

extern "C" {

int esdg_ensemble_create(esdg_ctx* ctx, int nscn, const esdg_params* params, const esdg_inputs* inputs,
                         const esdg_run_params* rp, esdg_ensemble** out) {
    ESDG_REQUIRE(ctx, ctx && params && inputs && rp && out, "esdg_ensemble_create: null argument");
    ESDG_REQUIRE(ctx, nscn > 0, "esdg_ensemble_create: no scenarios");
    *out = nullptr;
    for (int i = 0; i < nscn; ++i) {
        int rc = check_scenario(ctx, i, params[i], inputs[i], rp[i]);
        if (rc) return rc;
        if (rp[i].rk_order != rp[0].rk_order)
            return set_error(ctx, ESDG_ERR_ARG, "esdg_ensemble_create: every scenario needs the same rk_order");
    }
    esdg_ensemble* E = new (std::nothrow) esdg_ensemble();
    if (!E) return set_error(ctx, ESDG_ERR_NOMEM, "esdg_ensemble_create: out of host memory");
    E->ctx = ctx;
    int rc = ESDG_OK;
    double Bdummy[5];
    if (esdg_rk_coefficients(rp[0].rk_order, 0, &E->stages, E->A, Bdummy, E->C) != 0)
        rc = set_error(ctx, ESDG_ERR_ARG, "esdg_ensemble_create: RK coefficients");
    if (!rc) E->clocks = (double*)E->alloc((size_t)nscn * 8 * sizeof(double), &rc);
    if (!rc) E->lmax = (unsigned long long*)E->alloc((size_t)nscn * sizeof(unsigned long long), &rc);
    if (!rc) E->emax = (unsigned long long*)E->alloc((size_t)nscn * sizeof(unsigned long long), &rc);
    if (!rc) rc = esdg_fill(ctx, reinterpret_cast<double*>(E->emax), 0.0, nscn);
    const size_t dd = sizeof(double), di = sizeof(int);
    for (int i = 0; i < nscn && !rc; ++i) {
        const esdg_params& p = params[i];
        const esdg_inputs& in = inputs[i];
        const esdg_run_params& r = rp[i];
        const int ngl = p.ngl, ngl2 = ngl * ngl;
        const size_t nodes = (size_t)in.Nelem * ngl2, fn = (size_t)in.Nfaces * ngl;
        StageScn S{};
        S.ngl = ngl; S.Nelem = in.Nelem; S.Nfaces = in.Nfaces; S.ES = p.ES; S.num_flux = p.num_flux;
        S.positivity = r.positivity; S.bc = p.bc_kind;
        S.g = p.g; S.geomFace = p.geomFace; S.tol = p.PosPresTOL; S.zero_tol = p.ZeroTOL; S.h_zero = p.h_zero;
        S.dti = r.min_ele_size / (ngl) * r.CFL;
        S.T = r.T;
        S.Jac = (const double*)E->upload(in.Jac, nodes * dd, &rc);
        if (!rc) S.Yxi = (const double*)E->upload(in.Yxi, nodes * dd, &rc);
        if (!rc) S.Yeta = (const double*)E->upload(in.Yeta, nodes * dd, &rc);
        if (!rc) S.Xxi = (const double*)E->upload(in.Xxi, nodes * dd, &rc);
        if (!rc) S.Xeta = (const double*)E->upload(in.Xeta, nodes * dd, &rc);
        if (!rc) S.Bx = (const double*)E->upload(in.Bx, nodes * dd, &rc);
        if (!rc) S.By = (const double*)E->upload(in.By, nodes * dd, &rc);
        if (!rc) S.D = (const double*)E->upload(in.D, ngl2 * dd, &rc);
        if (!rc) S.nx = (const double*)E->upload(in.nx, fn * dd, &rc);
        if (!rc) S.ny = (const double*)E->upload(in.ny, fn * dd, &rc);
        if (!rc) S.scal = (const double*)E->upload(in.scal, fn * dd, &rc);
        if (!rc) S.EleSizes = (const double*)E->upload(in.EleSizes, in.Nelem * dd, &rc);
        if (!rc) S.GLw = (const double*)E->upload(in.GLw, ngl * dd, &rc);
        if (!rc) S.EdgeData = (const int*)E->upload(in.EdgeData, (size_t)in.Nfaces * 8 * di, &rc);
        if (!rc) S.E2E = (const int*)E->upload(in.ElemToEdge, (size_t)in.Nelem * 4 * di, &rc);
        if (!rc) S.EMS = (const int*)E->upload(in.ElemEdgeMasterSlave, (size_t)in.Nelem * 4 * di, &rc);
        if (!rc) S.EOR = (const int*)E->upload(in.ElemEdgeOrientation, (size_t)in.Nelem * 4 * di, &rc);
        const double* dB = nullptr;
        if (!rc) dB = (const double*)E->upload(in.B, nodes * dd, &rc);
        if (rc) break;
        // face traces of b, once per distinct (B, EdgeData, Nfaces, ngl)
        const auto key = std::make_tuple((const void*)in.B, (const void*)in.EdgeData, in.Nfaces, ngl);
        auto it = E->btr.find(key);
        if (it == E->btr.end()) {
            double* bL = (double*)E->alloc(fn * dd, &rc);
            double* bR = rc ? nullptr : (double*)E->alloc(fn * dd, &rc);
            if (!rc) rc = bottom_traces(ctx, ngl, in.Nfaces, S.EdgeData, dB, bL, bR);
            if (rc) break;
            it = E->btr.emplace(key, std::make_pair(bL, bR)).first;
        }
        S.bL = it->second.first;
        S.bR = it->second.second;
        // state: two ping-pong buffers and the RK register, private to the scenario
        S.q[0] = (double*)E->alloc(nodes * 3 * dd, &rc);
        if (!rc) S.q[1] = (double*)E->alloc(nodes * 3 * dd, &rc);
        if (!rc) S.gRK = (double*)E->alloc(nodes * 3 * dd, &rc);
        if (!rc) {
            if (!in.q) rc = set_error(ctx, ESDG_ERR_ARG, "esdg_ensemble_create: missing q");
            else rc = esdg_memcpy(ctx, S.q[0], in.q, nodes * 3 * dd, ESDG_H2D);
        }
        S.clock = E->clocks + 8 * (size_t)i;
        S.lmax = E->lmax + i;
        S.emax = E->emax + i;
        if (r.artificial_viscosity) {
            // the element viscosity and the BR1 / viscous work buffers, private to the scenario;
            // the dt limit of src/deviceclass.cpp:922 with the solver's constant
            S.av = 1;
            S.dvc = r.DFL / (std::pow(ngl, 2)) * std::pow(r.min_ele_size, 2);
            S.sc = SensorConsts{p.ZeroTOL, p.eps0, p.sigmaMin, p.sigmaMax};
            S.B = dB;
            S.VdmInv = (const double*)E->upload(in.VdmInv, ngl2 * dd, &rc);
            if (!rc) S.Dhat = (const double*)E->upload(in.Dhat, ngl2 * dd, &rc);
            if (!rc) S.Dstrong = (const double*)E->upload(in.Dstrong, ngl2 * dd, &rc);
            if (!rc) S.visc = (double*)E->alloc(in.Nelem * dd, &rc);
            if (!rc) S.gtr = (double*)E->alloc((size_t)in.Nelem * 16 * ngl * dd, &rc);
            if (!rc) S.Qv = (double*)E->alloc(nodes * 2 * dd, &rc);
            if (!rc) S.SPV = (double*)E->alloc(fn * 2 * dd, &rc);
            if (!rc) S.active = (int*)E->alloc(((size_t)in.Nelem + 1) / 2 * dd, &rc);
            if (!rc) rc = esdg_fill(ctx, S.visc, 0.0, in.Nelem);
            if (!rc) rc = esdg_fill(ctx, reinterpret_cast<double*>(S.active), 0.0, ((size_t)in.Nelem + 1) / 2);
            if (rc) break;
        }
        E->scn.push_back(S);
        E->qcount.push_back(nodes * 3);
    }
    if (!rc) rc = E->batch.init(ctx, E->scn);
    if (rc) {
        delete E;
        return rc;
    }
    *out = E;
    return ESDG_OK;
}

int esdg_ensemble_set_schedule(esdg_ensemble* E, const esdg_ensemble_schedule* sc) {
    if (!E || !sc) return set_error(E ? E->ctx : nullptr, ESDG_ERR_ARG, "esdg_ensemble_set_schedule: null argument");
    StageSchedule s;
    s.order = sc->order; s.blocks_per_cu = sc->blocks_per_cu; s.per_degree = sc->per_degree != 0;
    s.xcd = sc->xcd != 0; s.graph = sc->graph != 0; s.tile = sc->tile ? sc->tile : 1;
    return E->batch.set_schedule(E->ctx, s);
}

int esdg_ensemble_get_schedule(const esdg_ensemble* E, esdg_ensemble_schedule* sc) {
    if (!E || !sc) return set_error(E ? E->ctx : nullptr, ESDG_ERR_ARG, "esdg_ensemble_get_schedule: null argument");
    const StageSchedule& s = E->batch.sched;
    sc->order = s.order; sc->blocks_per_cu = s.blocks_per_cu; sc->per_degree = s.per_degree;
    sc->xcd = s.xcd; sc->graph = s.graph; sc->tile = s.tile;
    return ESDG_OK;
}

int esdg_ensemble_run_async(esdg_ensemble* E, int steps) {
    if (!E) return set_error(nullptr, ESDG_ERR_ARG, "esdg_ensemble_run: null ensemble");
    ESDG_REQUIRE(E->ctx, steps >= 0, "esdg_ensemble_run: steps < 0");
    int rc = E->start();
    if (!rc) rc = E->batch.run(E->ctx, E->A, E->C, E->stages, steps, E->src);
    return rc;
}

int esdg_ensemble_run(esdg_ensemble* E, int steps) {
    int rc = esdg_ensemble_run_async(E, steps);
    if (!rc) rc = esdg_finish(E->ctx);
    return rc;
}

int esdg_ensemble_get_q(esdg_ensemble* E, int i, double* host_q) {
    if (!E || !host_q) return set_error(E ? E->ctx : nullptr, ESDG_ERR_ARG, "esdg_ensemble_get_q: null argument");
    ESDG_REQUIRE(E->ctx, i >= 0 && i < (int)E->scn.size(), "esdg_ensemble_get_q: scenario out of range");
    return esdg_memcpy(E->ctx, host_q, E->scn[i].q[E->src], E->qcount[i] * sizeof(double), ESDG_D2H);
}

int esdg_ensemble_status(esdg_ensemble* E, int i, double* t, int* steps_taken, double* last_dt) {
    if (!E) return set_error(nullptr, ESDG_ERR_ARG, "esdg_ensemble_status: null ensemble");
    ESDG_REQUIRE(E->ctx, i >= 0 && i < (int)E->scn.size(), "esdg_ensemble_status: scenario out of range");
    double ck[8];
    int rc = esdg_memcpy(E->ctx, ck, E->clocks + 8 * (size_t)i, sizeof ck, ESDG_D2H);
    if (rc) return rc;
    if (t) *t = ck[1];
    if (steps_taken) *steps_taken = (int)ck[2];
    if (last_dt) *last_dt = ck[3];
    if (ck[4] != 0.0)
        return set_error(E->ctx, ESDG_ERR_STATE, "scenario " + std::to_string(i) + ": time step not positive and finite");
    return ESDG_OK;
}

int esdg_ensemble_bytes(const esdg_ensemble* E, size_t* bytes) {
    if (!E || !bytes) return set_error(E ? E->ctx : nullptr, ESDG_ERR_ARG, "esdg_ensemble_bytes: null argument");
    size_t b = 0;
    for (void* p : E->owned) {
        auto it = E->ctx->allocs.find(p);
        if (it != E->ctx->allocs.end()) b += it->second;
    }
    *bytes = b;
    return ESDG_OK;
}

int esdg_ensemble_autotune(esdg_ensemble* E, int trial_steps, esdg_ensemble_schedule* best, double* best_ms) {
    if (!E) return set_error(nullptr, ESDG_ERR_ARG, "esdg_ensemble_autotune: null ensemble");
    esdg_ctx* ctx = E->ctx;
    ESDG_REQUIRE(ctx, trial_steps >= 1, "esdg_ensemble_autotune: trial_steps < 1");
    int rc = E->start();
    if (rc) return rc;
    // snapshot: current state buffer of every scenario, the clocks and lambda_max cells
    const int ns = (int)E->scn.size();
    size_t total = 0;
    for (size_t c : E->qcount) total += c;
    double* snap = nullptr;
    rc = esdg_malloc(ctx, (void**)&snap, (total + 9 * (size_t)ns) * sizeof(double));
    if (rc) return rc;
    auto copy_all = [&](bool save) {
        size_t off = 0;
        int r = ESDG_OK;
        for (int i = 0; i < ns && !r; ++i) {
            double* q = E->scn[i].q[E->src];
            r = hip_check(ctx, hipMemcpyAsync(save ? snap + off : q, save ? q : snap + off, E->qcount[i] * sizeof(double),
                                              hipMemcpyDeviceToDevice, ctx->stream), "autotune snapshot");
            off += E->qcount[i];
        }
        if (!r) r = hip_check(ctx, hipMemcpyAsync(save ? snap + off : E->clocks, save ? E->clocks : snap + off,
                                                  (size_t)ns * 8 * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream),
                              "autotune snapshot");
        off += (size_t)ns * 8;
        if (!r) r = hip_check(ctx, hipMemcpyAsync(save ? (void*)(snap + off) : (void*)E->lmax,
                                                  save ? (const void*)E->lmax : (const void*)(snap + off),
                                                  (size_t)ns * sizeof(unsigned long long), hipMemcpyDeviceToDevice,
                                                  ctx->stream),
                              "autotune snapshot");
        if (!r) r = esdg_finish(ctx);
        return r;
    };
    rc = copy_all(true);
    // the learned indicator's counts are restored after the trials as well
    unsigned long long icount[5] = {0, 0, 0, 0, 0};
    if (!rc && E->batch.d_icount) rc = esdg_memcpy(ctx, icount, E->batch.d_icount, sizeof icount, ESDG_D2H);
    const int src0 = E->src;
    const StageSchedule keep = E->batch.sched;
    StageSchedule best_s = keep;
    double best_t = 1e300;
    std::vector<StageSchedule> cand;
    for (int order = 0; order < 3; ++order)
        for (int bpc : {0, 4, 8})
            for (int pd = 0; pd < 2; ++pd)
                for (int xcd = 0; xcd < 2; ++xcd) {
                    StageSchedule s;
                    s.order = order; s.blocks_per_cu = bpc; s.per_degree = pd; s.xcd = xcd; s.graph = keep.graph;
                    cand.push_back(s);
                }
    auto trial = [&](const StageSchedule& s) {
        rc = E->batch.set_schedule(ctx, s);
        if (!rc) rc = E->batch.run(ctx, E->A, E->C, E->stages, 2, E->src);   // warm-up, captures the graph
        if (!rc) rc = esdg_finish(ctx);
        if (rc) return;
        const auto t0 = std::chrono::steady_clock::now();
        rc = E->batch.run(ctx, E->A, E->C, E->stages, trial_steps, E->src);
        if (!rc) rc = esdg_finish(ctx);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
                          trial_steps;
        if (!rc && ms < best_t) {
            best_t = ms;
            best_s = s;
        }
        // back to the snapshot (the state sits in buffer src0 again)
        E->src = src0;
        if (!rc) rc = copy_all(false);
    };
    for (const StageSchedule& s : cand) {
        if (rc) break;
        trial(s);
    }
    // the element-tile size, on the winner's launch shape
    const StageSchedule shape = best_s;
    for (int tile : {2, 4}) {
        if (rc) break;
        StageSchedule s = shape;
        s.tile = tile;
        trial(s);
    }
    esdg_free(ctx, snap);
    if (!rc && E->batch.d_icount) rc = esdg_memcpy(ctx, E->batch.d_icount, icount, sizeof icount, ESDG_H2D);
    if (!rc) rc = E->batch.set_schedule(ctx, best_s);
    if (rc) return rc;
    if (best) esdg_ensemble_get_schedule(E, best);
    if (best_ms) *best_ms = best_t;
    return ESDG_OK;
}

int esdg_ensemble_set_indicator(esdg_ensemble* E, esdg_mlp* m, float lo, float hi) {
    if (!E) return set_error(nullptr, ESDG_ERR_ARG, "esdg_ensemble_set_indicator: null ensemble");
    return E->batch.set_indicator(E->ctx, m, lo, hi);
}

int esdg_ensemble_indicator_audit(esdg_ensemble* E, int audit) {
    if (!E) return set_error(nullptr, ESDG_ERR_ARG, "esdg_ensemble_indicator_audit: null ensemble");
    E->batch.ind_audit = audit != 0;
    E->batch.drop_graphs();   // the audit flag is a kernel argument of the captured steps
    return ESDG_OK;
}

int esdg_ensemble_indicator_counts(esdg_ensemble* E, int64_t* viscous, int64_t* false_negatives,
                                   int64_t* sensor_positives, int64_t* not_smooth, int64_t* element_stages) {
    if (!E) return set_error(nullptr, ESDG_ERR_ARG, "esdg_ensemble_indicator_counts: null ensemble");
    unsigned long long c[5] = {0, 0, 0, 0, 0};
    if (E->batch.d_icount) {
        int rc = esdg_memcpy(E->ctx, c, E->batch.d_icount, sizeof c, ESDG_D2H);
        if (!rc) rc = esdg_fill(E->ctx, reinterpret_cast<double*>(E->batch.d_icount), 0.0, 5);
        if (!rc) rc = esdg_finish(E->ctx);
        if (rc) return rc;
    }
    if (viscous) *viscous = (int64_t)c[0];
    if (false_negatives) *false_negatives = (int64_t)c[1];
    if (sensor_positives) *sensor_positives = (int64_t)c[2];
    if (not_smooth) *not_smooth = (int64_t)c[3];
    if (element_stages) *element_stages = (int64_t)c[4];
    return ESDG_OK;
}

int esdg_ensemble_destroy(esdg_ensemble* E) {
    if (!E) return ESDG_OK;
    (void)esdg_finish(E->ctx);
    delete E;
    return ESDG_OK;
}

}  // extern "C"
