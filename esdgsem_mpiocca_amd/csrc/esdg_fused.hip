// esdg_fused.hip -- the fused stage: the reference's per-stage arithmetic in fewer HBM passes.
//
// Reference sequence per RK stage (src/deviceclass.cpp:957-1122) and what the fused schedule
// does with it:
//   CollectEdgeData + calcNumFluxes [+ calcNumFluxesGradient]      -> fused_face (one pass over faces,
//                                                                     traces read from the side-trace
//                                                                     buffer the previous update wrote)
//   [ShockCapturing]                                                -> the sensor pass, run by the caller
//                                                                     before the stage (the time step's
//                                                                     sensor serves stage 0, as there)
//   [calcGradient + SurfaceKernelGradient + scaleGradient + VolumeKernelViscose]
//                                                                   -> fused_av (one element pass; the
//                                                                     viscous volume term goes to Qv;
//                                                                     blocks with no viscosity in reach
//                                                                     leave at once)
//   [CollectEdgeDataGradient + calcNumFluxesViscose]                -> fused_viscous_face
//   VolumeKernel + SurfaceKernel [+ Qv + SurfaceKernelVisc] + calcRK + UpdateKernel
//                                                                   -> fused_volume (esdg_volume.hip),
//                                                                     RK in place
//   [PosPres]                                                       -> pos_pres
// The viscous terms depend only on the stage-start state and the face fluxes, so the AV pass runs
// before the volume pass.  Every value is computed with the reference kernel's expressions in the
// reference's order, and Qt receives its contributions in the reference's order (volume, surface,
// viscous volume, viscous surface), so the fused stage reproduces the kernel-by-kernel results.
#include <climits>
#include <type_traits>

#include "esdg_av_device.h"
#include "esdg_device.h"
#include "esdg_fused.h"

namespace esdg {

int fused_volume(esdg_ctx* ctx, const FusedArgs& a, bool rk, double rkA, double cdt, const double* fstar, double* Qt,
                 const double* Qv, const double* SPV, double* tr, const int* slot, const int* active = nullptr,
                 const int* list = nullptr, const int* acount = nullptr);
int write_traces(esdg_ctx* ctx, int Nelem, const double* q, double* tr, const int* slot);
int trace_slots(esdg_ctx* ctx, int Nelem, const int* E2E, const int* ED, int* slot);

// ---------------------------------------------------------------------------------------------
// fused_face: CollectEdgeData (SolidWalls.okl:16-207) + calcNumFluxes (calcESFlux.okl /
// calcESFluxRotated.okl), one thread per face node; with AV it also marks the elements with
// viscosity in reach (the AV pass evaluates the BR1 face fluxes of those elements' sides itself).
// ---------------------------------------------------------------------------------------------
// KIND: the numerical flux, a template argument so that each instance holds only its own flux
// (fewer registers, more waves in flight).
template <bool AV, int KIND>
__global__ void __launch_bounds__(256) k_fused_face(int Nfaces, int ngl, FluxConsts c, GatherConsts gc,
                                                    const int* __restrict__ ED, const double* __restrict__ tr,
                                                    const double* __restrict__ nxa, const double* __restrict__ nya,
                                                    const double* __restrict__ scal, const double* __restrict__ BL,
                                                    const double* __restrict__ BR, double* __restrict__ SP,
                                                    const double* __restrict__ visc, int* __restrict__ active,
                                                    int Nelem, int* __restrict__ alist, int* __restrict__ acount) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Nfaces * ngl) return;
    const int f = (int)(t / ngl), i = (int)(t % ngl);
    const int* E = ED + 8 * (int64_t)f;
    const int rv = E[4];
    const int ir = (1 - rv) * (ngl - 1 - i) + rv * i;
    const int64_t xidR = (int64_t)f * ngl + ir;
    const double nx = nxa[t], ny = nya[t], sc = scal[t];
    const double bl = BL[t], br = BR[t];
    // qL at node i and qR at node ir, exactly as the gather kernel stores them (the element side
    // traces come from the side-trace buffer: contiguous per side, unlike the columns of q)
    const FaceState L = gathered_tr(tr, ngl, f, E, true, i, nx, ny, gc, bl, br);
    FaceState R;
    double brR;
    if (64 % ngl == 0) {
        // a face's ngl nodes are consecutive lanes of one wave: each lane gathers qR at its own
        // node, and node ir's values come from its lane, so no load waits for EdgeData's
        // orientation
        const FaceState Ro = gathered_tr(tr, ngl, f, E, false, i, nx, ny, gc, bl, br);
        const int src = (int)(threadIdx.x & 63) - i + ir;
        R = FaceState{__shfl(Ro.h, src), __shfl(Ro.hu, src), __shfl(Ro.hv, src)};
        brR = __shfl(br, src);
    } else {
        R = gathered_tr(tr, ngl, f, E, false, ir, nxa[xidR], nya[xidR], gc, BL[xidR], BR[xidR]);
        brR = BR[xidR];
    }
    double F0, F1, F2;
    numerical_flux(KIND, c, E[6], nx, ny, L.h, L.hu, L.hv, R.h, R.hu, R.hv, bl, brR, F0, F1, F2);
    const int64_t eid = (int64_t)f * ngl * ESDG_NEQ + i;
    SP[eid] = sc * c.geomFace * F0;
    SP[eid + ngl] = sc * c.geomFace * F1;
    SP[eid + 2 * ngl] = sc * c.geomFace * F2;
    if (AV) {
        // the stage's mark lives on the device (acount[2], advanced before each stage); the next
        // stage's list length starts at zero (the previous stage, which used that slot, has
        // finished)
        const int seq = acount[2];
        if (t == 0) acount[(seq + 1) & 1] = 0;
        // elements with viscosity in reach (on themselves or a face neighbour) need the viscous
        // work: a face with viscosity on either side marks both its elements for this stage
        const int cpuL = E[5], cpuR = E[6];
        if (i == 0) {
            const bool hasL = cpuL == 0, hasR = cpuR == 0;
            if ((hasL && visc[E[0]] != 0.0) || (hasR && visc[E[1]] != 0.0)) {
                // the first face to mark an element this stage also appends it to the stage's list;
                // the list length of stage seq is acount[seq & 1]
                auto mark = [&](int e) {
                    if (atomicExch(&active[e], seq) != seq) {
                        const int k = atomicAdd(acount + (seq & 1), 1);
                        if (k < Nelem) alist[k] = e;
                    }
                };
                if (hasL) mark(E[0]);
                if (hasR) mark(E[1]);
            }
        }
    }
}

// The face marks of k_fused_face<true, *> on their own (one thread per face), for a stage whose
// face fluxes were computed while the viscosity was still being evaluated.
__global__ void __launch_bounds__(256) k_fused_face_marks(int Nfaces, const int* __restrict__ ED,
                                                          const double* __restrict__ visc, int* __restrict__ active,
                                                          int Nelem, int* __restrict__ alist, int* __restrict__ acount) {
    const int f = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (f >= Nfaces) return;
    const int seq = acount[2];
    if (f == 0) acount[(seq + 1) & 1] = 0;
    const int* E = ED + 8 * (int64_t)f;
    const int cpuL = E[5], cpuR = E[6];
    const bool hasL = cpuL == 0, hasR = cpuR == 0;
    if ((hasL && visc[E[0]] != 0.0) || (hasR && visc[E[1]] != 0.0)) {
        auto mark = [&](int e) {
            if (atomicExch(&active[e], seq) != seq) {
                const int k = atomicAdd(acount + (seq & 1), 1);
                if (k < Nelem) alist[k] = e;
            }
        };
        if (hasL) mark(E[0]);
        if (hasR) mark(E[1]);
    }
}

// ---------------------------------------------------------------------------------------------
// fused_av: calcGradient.okl:18-156 + SurfaceKernelGradient.okl:16-148 (with the BR1 face fluxes
// of calcNumFluxesGradient.okl evaluated per element side from the state's side traces) +
// scaleGradient.okl:17-60 + VolumeKernelViscose.okl:17-150 for one block of elements, one thread
// per node, all intermediate fields in LDS.  Writes the scaled gradients (read by the viscous face
// kernel) as element-side traces and the viscous volume term Qv(hu, hv).
// ---------------------------------------------------------------------------------------------
//
// The element viscosity comes in (evaluated on the stage's state before the stage).  A block
// whose elements and their face neighbours all have zero viscosity has nothing to contribute: its
// gradient traces feed only faces of zero viscosity (whose fluxes are zero without reading them)
// and its viscous volume term is skipped where the viscosity is zero.  The face pass marks the
// others (active[e] == seq) and lists each once (alist[0 .. *acount)); the pass runs the listed
// elements, EPB per workgroup iteration, in list order (each element's results depend on nothing
// but its own inputs, so the order does not matter).
template <int NGL>
__global__ void __launch_bounds__(256) k_fused_av(int Nelem, AvMesh M, const double* __restrict__ tr,
                                                  const double* __restrict__ Dh, const double* __restrict__ Ds,
                                                  const double* __restrict__ visc, const int* __restrict__ alist,
                                                  const int* __restrict__ acount) {
    using EB = ElemBlock<NGL>;
    constexpr int NN = NGL * NGL;
    const int seq = acount[2];
    __shared__ double sDh[NN], sDs[NN];
    __shared__ double sA[EB::THREADS], sB[EB::THREADS], sC[EB::THREADS], sE[EB::THREADS];
    __shared__ double sEps[EB::EPB];
    if (threadIdx.x < NN) {
        sDh[threadIdx.x] = Dh[threadIdx.x];
        sDs[threadIdx.x] = Ds[threadIdx.x];
    }
    const int el = threadIdx.x / NN;
    const int cnt = min(acount[seq & 1], Nelem);
    __syncthreads();
    for (int grp = blockIdx.x; grp * EB::EPB < cnt; grp += gridDim.x) {
        const int idx = grp * EB::EPB + el;
        const bool act = idx < cnt;
        const int e = act ? alist[idx] : 0;
        av_element_block<NGL>(M, TraceFromSides{tr}, e, act, act ? visc[e] : 0.0, sDh, sDs, sA, sB, sC, sE, sEps);
    }
}

// ---------------------------------------------------------------------------------------------
// fused_viscous_face: CollectEdgeDataGradient.okl:17-194 + calcNumFluxesViscose.okl:17-107
// ---------------------------------------------------------------------------------------------
// Only elements marked this stage (active[e] == seq: viscosity on themselves or a face neighbour)
// read these fluxes; every face of an unmarked element has zero viscosity on both sides, so its
// flux is +-0 and the volume pass skips it.  The pass runs over the side nodes of the listed
// (marked) elements; each face node is evaluated once, by the face's left element if that is
// marked, else by its right element.  Faces of no marked element are not written.
__global__ void __launch_bounds__(256) k_fused_viscous_face(int Nelem, int ngl, double geomFace,
                                                            const int* __restrict__ ED,
                                                            const int* __restrict__ E2E,
                                                            const double* __restrict__ gtr,
                                                            const double* __restrict__ visc,
                                                            const double* __restrict__ nxa,
                                                            const double* __restrict__ nya,
                                                            const double* __restrict__ scal,
                                                            const int* __restrict__ active,
                                                            const int* __restrict__ alist,
                                                            const int* __restrict__ acount, double* __restrict__ SPV) {
    const int seq = acount[2];
    const int64_t items = (int64_t)min(acount[seq & 1], Nelem) * 4 * ngl;
    for (int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; it < items;
         it += (int64_t)gridDim.x * blockDim.x) {
        const int li = (int)(it / (4 * ngl)), rem = (int)(it % (4 * ngl));
        const int side = rem / ngl, i = rem % ngl;
        const int e = alist[li];
        const int f = E2E[e * 4 + side] - 1;
        const int* E = ED + 8 * (int64_t)f;
        const int iel = E[0], ier = E[1], isl = E[2], isr = E[3], rv = E[4], cpuL = E[5], cpuR = E[6];
        const int owner = (cpuL == 0 && active[iel] == seq) ? iel : ier;
        if (e != owner) continue;
        viscous_face_node(f, i, ngl, geomFace, iel, ier, isl, isr, rv, cpuL, cpuR, gtr, visc, nxa, nya, scal, SPV);
    }
}


template <int NGL>
void launch_fused_av(esdg_ctx* ctx, const FusedArgs& A, const FusedStage& F) {
    using EB = ElemBlock<NGL>;
    // the list length is known on the device only: enough workgroups for a few percent of the
    // elements in one pass, the rest loop
    const int64_t nblk = (A.Nelem + EB::EPB - 1) / EB::EPB;
    const unsigned grid = (unsigned)std::max<int64_t>(std::min<int64_t>(nblk, 4096), 1);
    const AvMesh M{A.q, A.Jac, A.Yxi, A.Yeta, A.Xxi, A.Xeta, A.ElemToEdge, A.EMS, A.EOR, A.EdgeData,
                   A.nx, A.ny, A.scal, A.bL, A.bR, GatherConsts{ctx->p.bc_kind, ctx->p.h_zero},
                   ctx->p.geomFace, ctx->p.PosPresTOL, F.grad, F.qvisc};
    k_fused_av<NGL><<<grid, EB::THREADS, 0, ctx->stream>>>(A.Nelem, M, F.tr, A.Dhat, A.Dstrong, A.ViscPara, F.alist,
                                                          F.acount);
}


// The stage mark (acount[2]): one more per stage.
__global__ void k_fused_mark(int* acount) { acount[2] += 1; }

// Marks restart from 0: every element's mark and the counter cleared (outside any captured step).
int FusedStage::reset_marks(esdg_ctx* ctx) {
    int rc = esdg_fill(ctx, reinterpret_cast<double*>(active), 0.0, ((size_t)Nelem + 1) / 2);
    if (!rc) rc = esdg_fill(ctx, reinterpret_cast<double*>(acount), 0.0, 2);
    if (!rc) seq = 0;
    return rc;
}

int FusedStage::advance(esdg_ctx* ctx, int stages) {
    if (!av) return ESDG_OK;
    if (seq >= INT_MAX - 64 - stages) {
        int rc = reset_marks(ctx);
        if (rc) return rc;
    }
    seq += stages;
    return ESDG_OK;
}

__global__ void k_fused_clock(const double* __restrict__ red, double* __restrict__ ck, double dti, double dvc, int av,
                              double T) {
    const double t = ck[1];
    double dt = 0.0;
    if (t < T && ck[4] == 0.0) {
        const double dt_i = dti / red[0];
        dt = av ? fmin(T - t, fmin(dt_i, dvc / red[1])) : fmin(T - t, dt_i);
        if (!(dt > 0.0) || isinf(dt)) {
            ck[4] = 1.0;
            ck[5] = dt;
            dt = 0.0;
        } else {
            ck[1] = t + dt;
            ck[2] += 1.0;
            ck[3] = dt;
        }
    }
    ck[0] = dt;
}

int fused_clock(esdg_ctx* ctx, const double* red, double* clock, double dti, double dvc, int av, double T) {
    KernelScope ks(ctx, "fused_clock");
    k_fused_clock<<<1, 1, 0, ctx->stream>>>(red, clock, dti, dvc, av, T);
    return launch_check(ctx, "fused_clock");
}

int FusedStage::init(esdg_ctx* ctx, int Ne, int Nf, bool with_av) {
    Nelem = Ne; Nfaces = Nf; av = with_av;
    const int ngl = ctx->p.ngl;
    const size_t fn = (size_t)Nf * ngl;
    int rc = esdg_malloc(ctx, (void**)&fstar, fn * 3 * sizeof(double));
    if (!rc) rc = esdg_malloc(ctx, (void**)&tr, (size_t)Nf * 6 * ngl * sizeof(double));
    if (!rc) rc = esdg_malloc(ctx, (void**)&slot, (size_t)Ne * 4 * sizeof(int));
    slots_built = false;
    if (!rc && av) rc = esdg_malloc(ctx, (void**)&active, ((size_t)Ne + 1) / 2 * sizeof(double));
    if (!rc && av) rc = esdg_malloc(ctx, (void**)&alist, ((size_t)Ne + 1) / 2 * sizeof(double));
    if (!rc && av) rc = esdg_malloc(ctx, (void**)&acount, 256);
    if (!rc && av) rc = esdg_fill(ctx, reinterpret_cast<double*>(acount), 0.0, 32);
    if (!rc && av) rc = esdg_malloc(ctx, (void**)&grad, (size_t)Ne * 16 * ngl * sizeof(double));
    if (!rc && av) rc = esdg_malloc(ctx, (void**)&fvisc, fn * 2 * sizeof(double));
    if (!rc && av) rc = esdg_malloc(ctx, (void**)&qvisc, (size_t)Ne * 2 * ngl * ngl * sizeof(double));
    ready = (rc == ESDG_OK);
    return rc;
}

int FusedStage::begin(esdg_ctx* ctx, const double* q, const int* E2E, const int* ED) {
    if (!ready) return set_error(ctx, ESDG_ERR_STATE, "fused stage used before init");
    // an interrupted stage may have left a list behind (both list lengths: 2 ints)
    if (acount) {
        int rc = esdg_fill(ctx, reinterpret_cast<double*>(acount), 0.0, 1);
        if (rc) return rc;
    }
    if (!slots_built) {
        int rc = trace_slots(ctx, Nelem, E2E, ED, slot);
        if (rc) return rc;
        slots_built = true;
    }
    return write_traces(ctx, Nelem, q, tr, slot);
}

int FusedStage::stage(esdg_ctx* ctx, int, double rkA, double cdt, const FusedArgs& A) {
    if (!ready) return set_error(ctx, ESDG_ERR_STATE, "fused stage used before init");
    const int ngl = ctx->p.ngl;
    const int kind = ctx->p.num_flux;
    FluxConsts fc{ctx->p.g, ctx->p.geomFace, ctx->p.PosPresTOL, ctx->p.ES};
    GatherConsts gc{ctx->p.bc_kind, ctx->p.h_zero};
    const unsigned fb = blocks_for((int64_t)A.Nfaces * ngl, 256);
    if (av) {
        // stage mark of the elements with viscosity in reach, advanced on the device (replays of a
        // captured step take fresh marks); the marks of earlier stages are other values, so
        // nothing needs clearing (except before the counter would wrap)
        int rc = ESDG_OK;
        if (!capturing && seq >= INT_MAX - 64) rc = reset_marks(ctx);
        ++seq;
        if (!rc) {
            KernelScope ks(ctx, "fused_mark");
            k_fused_mark<<<1, 1, 0, ctx->stream>>>(acount);
            rc = launch_check(ctx, "fused_mark");
        }
        if (rc) return rc;
    }
    // the stage's viscosity pass (when the stage issues it): beside the face fluxes on the second
    // stream, or first, in order
    const bool overlap = av && A.visc_pass && ctx->side && !ctx->profile;
    if (av && A.visc_pass && !overlap) {
        const int rc = A.visc_pass(ctx->stream);
        if (rc) return rc;
    }
    if (overlap) {
        (void)hipEventRecord(ctx->fork, ctx->stream);
        (void)hipStreamWaitEvent(ctx->side, ctx->fork, 0);
        const int rc = A.visc_pass(ctx->side);
        (void)hipEventRecord(ctx->join, ctx->side);
        if (rc) return rc;
    }
    {
        KernelScope ks(ctx, "fused_face");
        auto launch = [&](auto av_tag, auto kind_tag) {
            constexpr bool AVK = decltype(av_tag)::value;
            constexpr int KK = decltype(kind_tag)::value;
            k_fused_face<AVK, KK><<<fb, 256, 0, ctx->stream>>>(A.Nfaces, ngl, fc, gc, A.EdgeData, tr, A.nx, A.ny,
                                                             A.scal, A.bL, A.bR, fstar, AVK ? A.ViscPara : nullptr,
                                                             AVK ? active : nullptr, A.Nelem,
                                                             AVK ? alist : nullptr, AVK ? acount : nullptr);
        };
        auto by_kind = [&](auto av_tag) {
            switch (kind) {
                case ESDG_FLUX_ES: launch(av_tag, std::integral_constant<int, ESDG_FLUX_ES>{}); break;
                case ESDG_FLUX_LF: launch(av_tag, std::integral_constant<int, ESDG_FLUX_LF>{}); break;
                case ESDG_FLUX_ES_ROTATED: launch(av_tag, std::integral_constant<int, ESDG_FLUX_ES_ROTATED>{}); break;
                case ESDG_FLUX_ROE: launch(av_tag, std::integral_constant<int, ESDG_FLUX_ROE>{}); break;
                default: launch(av_tag, std::integral_constant<int, ESDG_FLUX_OCEAN>{}); break;
            }
        };
        if (av && !overlap) by_kind(std::true_type{});
        else by_kind(std::false_type{});
        int rc = launch_check(ctx, "fused_face");
        if (rc) return rc;
    }
    if (overlap) {
        (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
        KernelScope ks(ctx, "fused_face_marks");
        k_fused_face_marks<<<blocks_for(A.Nfaces, 256), 256, 0, ctx->stream>>>(A.Nfaces, A.EdgeData, A.ViscPara,
                                                                                active, A.Nelem, alist, acount);
        const int rc = launch_check(ctx, "fused_face_marks");
        if (rc) return rc;
    }
    if (!av) {
        // volume + surface + RK (+ the limiter on the new state) in one element pass, which writes
        // the side traces of the state it stores
        return fused_volume(ctx, A, true, rkA, cdt, fstar, nullptr, nullptr, nullptr, tr, slot);
    }
    int rc = ESDG_OK;
    {
        KernelScope ks(ctx, "fused_av");
        ESDG_NGL_DISPATCH(ngl, launch_fused_av, ctx, A, *this);
        rc = launch_check(ctx, "fused_av");
        if (rc) return rc;
    }
    {
        KernelScope ks(ctx, "fused_viscous_face");
        const unsigned vb = (unsigned)std::max<int64_t>(std::min<int64_t>(((int64_t)A.Nelem * 4 * ngl + 255) / 256, 2048), 1);
        k_fused_viscous_face<<<vb, 256, 0, ctx->stream>>>(A.Nelem, ngl, ctx->p.geomFace, A.EdgeData, A.ElemToEdge,
                                                          grad, A.ViscPara, A.nx, A.ny, A.scal, active, alist,
                                                          acount, fvisc);
        rc = launch_check(ctx, "fused_viscous_face");
        if (rc) return rc;
    }
    // volume + surface + viscous terms + RK (+ the limiter on the new state) in one element pass
    return fused_volume(ctx, A, true, rkA, cdt, fstar, nullptr, qvisc, fvisc, tr, slot, active, alist, acount);
}

void FusedStage::release(esdg_ctx* ctx) {
    for (double** p : {&fstar, &grad, &fvisc, &qvisc, &tr})
        if (*p) { esdg_free(ctx, *p); *p = nullptr; }
    for (int** p : {&active, &alist, &acount, &slot})
        if (*p) { esdg_free(ctx, *p); *p = nullptr; }
    slots_built = false;
    seq = 0;
    ready = false;
}

void* FusedStage::buffer(const std::string& name) {
    if (name == "FStar") return fstar;
    if (name == "Grad") return grad;
    if (name == "FVisc") return fvisc;
    if (name == "QVisc") return qvisc;
    if (name == "Traces") return tr;
    return nullptr;
}

}  // namespace esdg
