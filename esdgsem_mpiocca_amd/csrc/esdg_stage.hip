// esdg_stage.hip -- element-centric single-kernel RK stage over a batch of scenarios (see
// esdg_stage.h).  Per element and stage it performs, in the reference's arithmetic and order:
//   CollectEdgeData + calcNumFluxes for its 4 sides    (SolidWalls/InnerOuter/OceanBoundary.okl,
//                                                       calcESFlux.okl / calcESFluxRotated.okl)
//   VolumeKernel (flux differencing)                   (VolumeKernelFluxDiffV99.okl)
//   SurfaceKernel                                      (SurfaceKernel.okl, DBSurf = 0)
//   calcRK + UpdateKernel                              (rkLS.okl, UpdateKernelLS.okl)
//   [PosPres]                                          (PosPres.okl)
//   [FindLambdaMax of the new state, last stage]      (FindLambdaMax.okl)
// and the dt of DGtimeloop (src/deviceclass.cpp:899-930) runs on the device.
#include "esdg_av_device.h"
#include "esdg_device.h"
#include "esdg_mlp_internal.h"
#include "esdg_stage.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>

namespace esdg {

// Flux-differencing D of each degree (ngl 3..8) in constant memory: read through the scalar cache
// into SGPRs, as a kernel argument would be, whatever scenario of the batch a block serves.  It
// depends on the degree only; StageBatch::init fills it and refuses a batch whose D differs from
// the one already loaded for that degree.
__constant__ double c_stage_D[9][64];

// Per element: the four planes of the volume sweep (VolCfg::ESZ), reused for the face fluxes
// [side][v][k] once the sweeps have read them, and 16 doubles for the limiter's element means
// (keeping consecutive elements 8 doubles apart modulo 16, as ESZ does).  At N = 7 a block of 8
// elements then fits 8 blocks per CU.
template <int NGL>
struct StageCfg {
    using V = VolCfg<NGL>;
    static constexpr int EPW = V::EPW;
    static constexpr int FSZ = 12 * NGL;              // face fluxes [side][v][k] of one element
    static_assert(FSZ <= 4 * V::PLANE, "the face fluxes fit in the element's planes");
    static constexpr int ELDS = V::ESZ + 16;          // per element
    static constexpr int LDS = EPW * ELDS;            // doubles of LDS per block
};

constexpr int cmax(int a, int b) { return a > b ? a : b; }
template <int LO, int HI>
struct MaxLds {
    static constexpr int value = cmax(StageCfg<LO>::LDS, MaxLds<LO + 1, HI>::value);
};
template <int HI>
struct MaxLds<HI, HI> {
    static constexpr int value = StageCfg<HI>::LDS;
};

// Volume term, surface lift, RK update, limiter and wave speed of one block of elements; the
// face fluxes of the elements' sides come from `faces`, into LDS (FS) after the sweeps.  Pointer
// parameters are __restrict__ so that the geometry loads schedule as in a kernel taking them as
// arguments, and global (gptr) so that they compile to global loads and stores.
template <int NGL, bool AV, class Faces, class Early>
__device__ __forceinline__ void stage_core(int e, int r, bool act, double* P, const Faces& faces, const Early& early,
                                           double g, double tol,
                                           double zero_tol, int positivity, double rkA, double rkC, double dt, bool last,
                                           gptr<const double> __restrict__ qi, gptr<double> __restrict__ qo,
                                           gptr<double> __restrict__ gRK, gptr<const double> __restrict__ Yeta,
                                           gptr<const double> __restrict__ Xeta, gptr<const double> __restrict__ Yxi,
                                           gptr<const double> __restrict__ Xxi, gptr<const double> __restrict__ Jac,
                                           gptr<const double> __restrict__ Bx, gptr<const double> __restrict__ By,
                                           gptr<const double> __restrict__ EleSizes,
                                           gptr<const double> __restrict__ GLw,
                                           gptr<unsigned long long> __restrict__ lmax, const StageScn* SV, int mark,
                                           const double* D) {
    using C = VolCfg<NGL>;
    double* FS = P;                  // face fluxes, after the sweeps (the planes are free then)
    double* MS = P + C::ESZ;         // the limiter's element means
    const int lane = threadIdx.x;
    double h[NGL], hu[NGL], hv[NGL], u[NGL], v[NGL], Y[NGL], X[NGL];
    double a1[NGL], a2[NGL], a3[NGL];
#pragma unroll
    for (int c = 0; c < NGL; ++c) { a1[c] = 0.0; a2[c] = 0.0; a3[c] = 0.0; }

    // ---- row phase: eta-row r ---------------------------------------------------------------
    if (act) {
        const int64_t qb = (int64_t)e * C::NN * 3 + r * NGL;
        const int64_t xb = (int64_t)e * C::NN + r * NGL;
#pragma unroll
        for (int c = 0; c < NGL; ++c) {
            h[c] = qi[qb + c];
            hu[c] = qi[qb + C::NN + c];
            hv[c] = qi[qb + 2 * C::NN + c];
            Y[c] = Yeta[xb + c];
            X[c] = Xeta[xb + c];
        }
        early();   // LDS stores of values loaded beside the row data
#pragma unroll
        for (int c = 0; c < NGL; ++c) {
            const double inv = (h[c] > 0) ? 1.0 / h[c] : 0.0;
            u[c] = hu[c] * inv;
            v[c] = hv[c] * inv;
            P[0 * C::PLANE + r * C::RS + c] = h[c];
            P[1 * C::PLANE + r * C::RS + c] = hu[c];
            P[2 * C::PLANE + r * C::RS + c] = hv[c];
            P[3 * C::PLANE + r * C::RS + c] = inv;
        }
        line_sweep<NGL, false>(D, g, h, hu, hv, u, v, Y, X, a1, a2, a3);
    }
    __syncthreads();

    // ---- column phase: xi-column r ----------------------------------------------------------
    const int64_t xb = (int64_t)e * C::NN + r;
    if (act) {
#pragma unroll
        for (int k = 0; k < NGL; ++k) {
            h[k] = P[0 * C::PLANE + k * C::RS + r];
            hu[k] = P[1 * C::PLANE + k * C::RS + r];
            hv[k] = P[2 * C::PLANE + k * C::RS + r];
            const double inv = P[3 * C::PLANE + k * C::RS + r];
            u[k] = hu[k] * inv;
            v[k] = hv[k] * inv;
            Y[k] = Yxi[xb + k * NGL];
            X[k] = Xxi[xb + k * NGL];
        }
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int c = 0; c < NGL; ++c) {
            P[0 * C::PLANE + r * C::RS + c] = a1[c];
            P[1 * C::PLANE + r * C::RS + c] = a2[c];
            P[2 * C::PLANE + r * C::RS + c] = a3[c];
        }
    }
    __syncthreads();
    double b1[NGL], b2[NGL], b3[NGL], rJ[NGL];
    if (act) {
#pragma unroll
        for (int k = 0; k < NGL; ++k) {
            b1[k] = P[0 * C::PLANE + k * C::RS + r];
            b2[k] = P[1 * C::PLANE + k * C::RS + r];
            b3[k] = P[2 * C::PLANE + k * C::RS + r];
        }
        line_sweep<NGL, true>(D, g, h, hu, hv, u, v, Y, X, b1, b2, b3);
#pragma unroll
        for (int k = 0; k < NGL; ++k) {
            unscale_sums(b1[k], b2[k], b3[k]);
            rJ[k] = Jac[xb + k * NGL];
            b1[k] = -rJ[k] * b1[k];
            b2[k] = -rJ[k] * b2[k] - g * h[k] * Bx[xb + k * NGL];
            b3[k] = -rJ[k] * b3[k] - g * h[k] * By[xb + k * NGL];
        }
    }
    // the face fluxes of the element's sides (from the stage-start state) into the planes, which
    // every lane has read by now
    __syncthreads();
    if (act) faces(FS);
    __syncthreads();
    if (act) {
        // surface lift: bottom/top sides first, then right/left (SurfaceKernel.okl).  FS holds
        // -ms F* (ms = +-1, so the sign is exact and -ms F* rJ equals the okl's -ms rJ F* bitwise).
        {
            const double* F0 = FS;
            const double* F2 = FS + 2 * 3 * NGL;
            b1[0] += F0[r] * rJ[0];
            b2[0] += F0[NGL + r] * rJ[0];
            b3[0] += F0[2 * NGL + r] * rJ[0];
            b1[NGL - 1] += F2[r] * rJ[NGL - 1];
            b2[NGL - 1] += F2[NGL + r] * rJ[NGL - 1];
            b3[NGL - 1] += F2[2 * NGL + r] * rJ[NGL - 1];
        }
        if (r == 0 || r == NGL - 1) {
            const double* F = FS + ((r == 0) ? 3 : 1) * 3 * NGL;
#pragma unroll
            for (int k = 0; k < NGL; ++k) {
                b1[k] += F[k] * rJ[k];
                b2[k] += F[NGL + k] * rJ[k];
                b3[k] += F[2 * NGL + k] * rJ[k];
            }
        }
        if constexpr (AV) {
            // the viscous terms of an element with viscosity in reach, in the reference's order
            // after the surface lift: the viscous volume term (where the element's viscosity is not
            // zero), then the lift of the viscous face fluxes (VolumeKernelViscose.okl,
            // SurfaceKernelVisc.okl), as the fused volume pass adds them
            const gptr<const int> E2E = glob(SV->E2E), EOR = glob(SV->EOR), EMS = glob(SV->EMS);
            const gptr<const double> SPV = glob((const double*)SV->SPV);
            if (SV->av && glob(SV->active)[e] == mark) {
                if (glob(SV->visc)[e] != 0.0) {
                    const gptr<const double> qv = glob((const double*)SV->Qv) + (int64_t)e * C::NN * 2 + r;
#pragma unroll
                    for (int k = 0; k < NGL; ++k) {
                        b2[k] += qv[k * NGL];
                        b3[k] += qv[C::NN + k * NGL];
                    }
                }
#pragma unroll
                for (int ph = 0; ph < 2; ++ph) {
#pragma unroll
                    for (int k = 0; k < NGL; ++k) {
                        int side = -1, t = 0;
                        if (ph == 0) {
                            if (k == 0) { side = 0; t = r; }
                            else if (k == NGL - 1) { side = 2; t = r; }
                        } else {
                            if (r == NGL - 1) { side = 1; t = k; }
                            else if (r == 0) { side = 3; t = k; }
                        }
                        if (side < 0) continue;
                        const int ifa = E2E[e * 4 + side] - 1;
                        const int rev = EOR[e * 4 + side];
                        const int ms = EMS[e * 4 + side];
                        const int inc = rev * t + (1 - rev) * (NGL - 1 - t);
                        const int64_t fl = (int64_t)ifa * NGL * 2 + inc;
                        b2[k] += -ms * rJ[k] * SPV[fl];
                        b3[k] += -ms * rJ[k] * SPV[fl + NGL];
                    }
                }
            }
        }
        // low-storage RK: gRK = A gRK + Qt; q_new = q + (C dt) gRK, q at node (k, r) being this
        // lane's column input (h, hu, hv): the same bits, without a second read of q
        const double cdt = rkC * dt;
        const int64_t ob = (int64_t)e * C::NN * 3 + r;
#pragma unroll
        for (int k = 0; k < NGL; ++k) {
            const int64_t id = ob + k * NGL;
            const double g1 = rkA * gRK[id] + b1[k];
            const double g2 = rkA * gRK[id + C::NN] + b2[k];
            const double g3 = rkA * gRK[id + 2 * C::NN] + b3[k];
            gRK[id] = g1;
            gRK[id + C::NN] = g2;
            gRK[id + 2 * C::NN] = g3;
            b1[k] = h[k] + cdt * g1;
            b2[k] = hu[k] + cdt * g2;
            b3[k] = hv[k] + cdt * g3;
        }
    }
    if (positivity) {
        // PosPres.okl on the new state: element means in the okl's serial order (xi outer, eta inner)
        __syncthreads();
        if (act) {
#pragma unroll
            for (int k = 0; k < NGL; ++k) {
                P[0 * C::PLANE + k * C::RS + r] = b1[k];
                P[1 * C::PLANE + k * C::RS + r] = b2[k];
                P[2 * C::PLANE + k * C::RS + r] = b3[k];
                P[3 * C::PLANE + k * C::RS + r] = Jac[xb + k * NGL];
            }
        }
        __syncthreads();
        if (act && r == 0) {
            double hmin = 0.0, ha = 0.0, hua = 0.0, hva = 0.0;
            for (int l = 0; l < NGL; ++l)
                for (int k = 0; k < NGL; ++k) {
                    const int o = k * C::RS + l;
                    const double w = GLw[l] * GLw[k] / P[3 * C::PLANE + o];
                    ha += P[o] * w;
                    hua += P[C::PLANE + o] * w;
                    hva += P[2 * C::PLANE + o] * w;
                    hmin = fmin(hmin, P[o]);
                }
            const double size = EleSizes[e];
            ha = ha / size;
            if (fabs(ha) < zero_tol) ha = 0.0;
            MS[0] = ha;
            MS[1] = hua / size;
            MS[2] = hva / size;
            MS[3] = (ha == hmin) ? 1.0 : fmin(1.0, ha / (ha - hmin));
            MS[4] = hmin;
        }
        __syncthreads();
        if (act) {
            const double ha = MS[0], hua = MS[1], hva = MS[2], th = MS[3], hmin = MS[4];
#pragma unroll
            for (int k = 0; k < NGL; ++k) {
                if (hmin < 0.0) {
                    double hn = 0.0, hun = 0.0, hvn = 0.0;
                    if (ha > 0.0) {
                        hn = fabs(th * (b1[k] - ha) + ha);
                        if (hn > tol) {
                            hun = th * (b2[k] - hua) + hua;
                            hvn = th * (b3[k] - hva) + hva;
                        }
                    }
                    if (ha < 0.0) { hn = -INFINITY; hun = 0.0; hvn = 0.0; }
                    b1[k] = hn; b2[k] = hun; b3[k] = hvn;
                } else if (b1[k] < tol) {
                    b2[k] = 0.0; b3[k] = 0.0;
                }
            }
        }
    }
    double lam = 0.0;
    if (act) {
        const int64_t ob = (int64_t)e * C::NN * 3 + r;
#pragma unroll
        for (int k = 0; k < NGL; ++k) {
            const int64_t id = ob + k * NGL;
            qo[id] = b1[k];
            qo[id + C::NN] = b2[k];
            qo[id + 2 * C::NN] = b3[k];
        }
        if (last) {
            // FindLambdaMax.okl on the new state (the next step's dt)
#pragma unroll
            for (int k = 0; k < NGL; ++k) {
                const double hh = b1[k];
                const double Qinv = (hh > tol) ? 1.0 / hh : 0.0;
                const double c = sqrt(g * hh);
                lam = fmax(lam, c + fmax(fabs(b2[k] * Qinv), fabs(b3[k] * Qinv)));
            }
        }
    }
    if (last) {
        for (int off = 32; off > 0; off >>= 1) lam = fmax(lam, __shfl_xor(lam, off));
        if (lane == 0) atomicMax((unsigned long long*)lmax, (unsigned long long)__double_as_longlong(lam));
    }
}


template <int NGL, bool AV>
__device__ __forceinline__ void stage_body(const StageScn& S, int e0, double* lds, double* sD, int src, double rkA,
                                           double rkC, bool last, int mark) {
    using C = VolCfg<NGL>;
    using SC = StageCfg<NGL>;
    const int lane = threadIdx.x;
    // D of this degree in LDS (wave-uniform reads, broadcast): keeps it out of the scalar registers
    for (int t = lane; t < NGL * NGL; t += 64) sD[t] = c_stage_D[NGL][t];
    const int el = lane / NGL, r = lane % NGL;
    const int e = e0 + el;
    const bool act = (el < C::EPW) && (e < S.Nelem);
    double* P = lds + (act ? el : 0) * SC::ELDS;     // planes: h, hu, hv, 1/h (later: the row sums)
    const gptr<const double> qi = glob((const double*)S.q[src]);
    const gptr<double> qo = glob(S.q[src ^ 1]);
    const double g = S.g;

    // The element's four side records (StageScn::srec) are loaded beside the row data and stored
    // to the limiter's LDS area (free until the epilogue), so the face-flux reads after the sweeps
    // need no dependent round trips through ElemToEdge and EdgeData.
    constexpr int FPL = (4 + NGL - 1) / NGL;   // sides per lane: lane r loads sides r, r + NGL, ...
    int4 rec[FPL];
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
        const int sd = r + j * NGL;
        rec[j] = make_int4(0, 0, 0, 0);
        if (act && sd < 4) {
            // one 16-byte load (a vector type: int4's assignment would take a generic reference)
            using i4v = int __attribute__((ext_vector_type(4)));
            const i4v v = ((gptr<const i4v>)glob(S.srec))[(int64_t)e * 4 + sd];
            rec[j] = make_int4(v.x, v.y, v.z, v.w);
        }
    }
    int4* RS = reinterpret_cast<int4*>(P + C::ESZ);
    auto early = [&]() {
#pragma unroll
        for (int j = 0; j < FPL; ++j)
            if (r + j * NGL < 4) RS[r + j * NGL] = rec[j];
    };

    // ---- fluxes of the element's four sides: lane r takes side node r of every side (called by
    // stage_core once the sweeps are done, with the face-flux area FS) ----------------------------
    auto faces = [&](double* FS) {
        const FluxConsts fc{g, S.geomFace, S.tol, S.ES};
        const GatherConsts gc{S.bc, S.h_zero};
        const gptr<const double> nxp = glob(S.nx), nyp = glob(S.ny), scp = glob(S.scal), bLp = glob(S.bL),
                                 bRp = glob(S.bR);
        // one side's decoded record and loads: side s+1's loads are issued before side s's flux
        // is evaluated, so a wave waits for memory once per element rather than once per side
        struct Side {
            int E[8];
            int ir, rev;     // the reversed face node (EdgeData[4]) of the lane's node; EOR
            unsigned w;
            FaceState Li, Ri;
            double nx, ny, sc, bl, br;
        };
        auto load = [&](int s, Side& d) {
            const int4 R4 = RS[s];
            const int f = R4.x;
            d.w = (unsigned)R4.w;
            d.rev = (int)((d.w >> 5) & 1u);
            const int i = d.rev * r + (1 - d.rev) * (NGL - 1 - r);      // face node of side node r
            // EdgeData's entries as gathered() reads them
            d.E[0] = R4.y; d.E[1] = R4.z; d.E[2] = (int)(d.w & 3u); d.E[3] = (int)((d.w >> 2) & 3u);
            d.E[4] = (int)((d.w >> 4) & 1u); d.E[5] = (int)((d.w >> 8) & 255u) - 2;
            d.E[6] = (int)((d.w >> 16) & 255u) - 2; d.E[7] = 0;
            const int rv = d.E[4];
            d.ir = (1 - rv) * (NGL - 1 - i) + rv * i;
            const int64_t t = (int64_t)f * NGL + i;
            face_side_traces(qi, NGL, d.E, i, d.Li, d.Ri);
            d.nx = nxp[t]; d.ny = nyp[t]; d.sc = scp[t]; d.bl = bLp[t]; d.br = bRp[t];
        };
        // The element's NGL lanes (one wave) hold the face's NGL nodes: each lane also evaluates
        // the right trace at its own node, and lane r takes node ir's from the lane holding it
        auto flux = [&](int s, const Side& d) {
            const FaceState L = gathered_from(d.Li, d.Ri, d.E, true, d.nx, d.ny, gc, d.bl, d.br);
            const FaceState Ro = gathered_from(d.Li, d.Ri, d.E, false, d.nx, d.ny, gc, d.bl, d.br);
            const int src = lane - r + (d.rev * d.ir + (1 - d.rev) * (NGL - 1 - d.ir));
            const FaceState R{__shfl(Ro.h, src), __shfl(Ro.hu, src), __shfl(Ro.hv, src)};
            const double brr = __shfl(d.br, src);
            double F0, F1, F2;
            numerical_flux(S.num_flux, fc, d.E[6], d.nx, d.ny, L.h, L.hu, L.hv, R.h, R.hu, R.hv, d.bl, brr, F0, F1,
                           F2);
            const double nms = ((d.w >> 6) & 1u) ? -1.0 : 1.0;   // -EMS
            FS[(s * 3 + 0) * NGL + r] = nms * (d.sc * S.geomFace * F0);
            FS[(s * 3 + 1) * NGL + r] = nms * (d.sc * S.geomFace * F1);
            FS[(s * 3 + 2) * NGL + r] = nms * (d.sc * S.geomFace * F2);
        };
        Side a, b;
        load(0, a);
        load(1, b);
        flux(0, a);
        load(2, a);
        flux(1, b);
        load(3, b);
        flux(2, a);
        flux(3, b);
    };

    __syncthreads();   // sD written
    stage_core<NGL, AV>(e, r, act, P, faces, early, g, S.tol, S.zero_tol, S.positivity, rkA, rkC, glob(S.clock)[0], last,
                        qi, qo, glob(S.gRK), glob(S.Yeta), glob(S.Xeta), glob(S.Yxi), glob(S.Xxi), glob(S.Jac),
                        glob(S.Bx), glob(S.By), glob(S.EleSizes), glob(S.GLw), glob(S.lmax), &S, mark, sD);
}

template <int LO, int HI, bool AV>
__device__ __forceinline__ void stage_task(const StageScn* __restrict__ scns, const StageTask& tk, double* lds,
                                           double* sD, int src, double rkA, double rkC, bool last, int mark) {
    const StageScn& S = scns[tk.scn];
    switch (tk.ngl) {
#define ESDG_CASE(n)                                                                \
    case n:                                                                         \
        if constexpr (n >= LO && n <= HI) stage_body<n, AV>(S, tk.e0, lds, sD, src, rkA, rkC, last, mark);    \
        break;
        ESDG_CASE(3) ESDG_CASE(4) ESDG_CASE(5) ESDG_CASE(6) ESDG_CASE(7) ESDG_CASE(8)
#undef ESDG_CASE
        default: break;
    }
}

// One RK stage over tasks[0, n).  Non-persistent: block b takes one task.  Persistent: the tasks
// are cut into chunks of `tile` consecutive tasks and a grid of gridDim.x blocks loops over the
// chunks, one block taking a chunk's tasks one after another (with tile > 1 and a grid of one
// block per chunk this is the element-tile size: a block advances tile element groups).  With
// xcd = 1 the task list is cut into 8 contiguous parts and the blocks that the dispatcher places
// on XCD x (b mod 8 = x) take part x, so tasks that are adjacent in the list (one element tile of
// several scenarios that share geometry) meet in the same L2.  Every block leaves after a bounded
// number of tasks.
template <int LO, int HI, bool PERSIST, bool AV>
__global__ void __launch_bounds__(64, 2) k_stage(const StageScn* __restrict__ scns, const StageTask* __restrict__ tasks,
                                                 int n, int xcd, int tile, int src, double rkA, double rkC,
                                                 int last, const int* __restrict__ markp, int stage) {
    __shared__ __attribute__((aligned(16))) double lds[MaxLds<LO, HI>::value];
    __shared__ double sD[HI * HI];
    const int mark = AV ? *markp + stage : 0;
    const int b = blockIdx.x;
    int beg = 0, cnt = n, i0 = b, stride = gridDim.x;
    if (xcd) {
        const int q = n >> 3, rr = n & 7, x = b & 7;
        cnt = q + (x < rr ? 1 : 0);
        beg = x * q + (x < rr ? x : rr);
        i0 = b >> 3;
        stride = gridDim.x >> 3;
    }
    if constexpr (!PERSIST) {
        if (i0 < cnt) stage_task<LO, HI, AV>(scns, tasks[beg + i0], lds, sD, src, rkA, rkC, last != 0, mark);
    } else {
        const int nch = (cnt + tile - 1) / tile;
        for (int c = i0; c < nch; c += stride) {
            const int i1 = min(cnt, (c + 1) * tile);
            for (int i = c * tile; i < i1; ++i) {
                if (i != i0 * tile) __syncthreads();   // LDS of the previous task is free
                stage_task<LO, HI, AV>(scns, tasks[beg + i], lds, sD, src, rkA, rkC, last != 0, mark);
            }
        }
    }
}

// lambda_max of the state in buffer src, per scenario (FindLambdaMax.okl + the host max loop)
__global__ void __launch_bounds__(256) k_stage_lambda(const StageScn* __restrict__ scns,
                                                      const StageTask* __restrict__ tasks, int src) {
    const StageTask tk = tasks[blockIdx.x];
    const StageScn& S = scns[tk.scn];
    const int ngl2 = tk.ngl * tk.ngl;
    const int epw = 64 / tk.ngl;
    double lam = 0.0;
    for (int t = threadIdx.x; t < epw * ngl2; t += blockDim.x) {
        const int e = tk.e0 + t / ngl2;
        if (e >= S.Nelem) continue;
        const int64_t id = (int64_t)e * ngl2 * 3 + t % ngl2;
        const double* q = S.q[src];
        const double hh = q[id];
        const double Qinv = (hh > S.tol) ? 1.0 / hh : 0.0;
        const double c = sqrt(S.g * hh);
        lam = fmax(lam, c + fmax(fabs(q[id + ngl2] * Qinv), fabs(q[id + 2 * ngl2] * Qinv)));
    }
    for (int off = 32; off > 0; off >>= 1) lam = fmax(lam, __shfl_xor(lam, off));
    if ((threadIdx.x & 63) == 0) atomicMax(S.lmax, (unsigned long long)__double_as_longlong(lam));
}

// DGtimeloop's dt (src/deviceclass.cpp:899-930) per scenario: while t < T, dt = min(T - t,
// dti / lambda_max[, dvc / eps_max]) and t += dt; once t >= T (or after a dt that is not positive
// and finite, which is flagged) dt = 0 and the stages leave the state unchanged.  lambda_max and
// eps_max are reset for the next step's reductions.
__global__ void k_stage_dt(const StageScn* __restrict__ scns, int nscn) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nscn) return;
    const StageScn& S = scns[s];
    const double lm = __longlong_as_double((long long)*S.lmax);
    *S.lmax = 0ull;
    double em = 0.0;
    if (S.av) {
        em = __longlong_as_double((long long)*S.emax);
        *S.emax = 0ull;
    }
    double* ck = S.clock;
    const double t = ck[1];
    double dt = 0.0;
    if (t < S.T && ck[4] == 0.0) {
        const double dt_i = S.dti / lm;
        // with artificial viscosity also the viscous limit dvc / max eps (infinite where no element
        // has viscosity), in the order of src/deviceclass.cpp:920-926
        dt = S.av ? fmin(S.T - t, fmin(dt_i, S.dvc / em)) : fmin(S.T - t, dt_i);
        if (!(dt > 0.0) || isinf(dt)) {
            ck[4] = 1.0;
            dt = 0.0;
        } else {
            ck[1] = t + dt;
            ck[2] += 1.0;
            ck[3] = dt;
        }
    }
    ck[0] = dt;
}

// ---------------------------------------------------------------------------------------------
// Artificial viscosity of the scenarios with av = 1.  Before each stage: the element viscosity of
// the stage's state (ShockCapturing.okl, or the learned indicator's pass in esdg_mlp.hip), each
// viscous element marking itself and its face neighbours with the stage's mark; the gradient pass
// (calcGradient + SurfaceKernelGradient + scaleGradient + VolumeKernelViscose) over the marked
// elements; the viscous face fluxes (CollectEdgeDataGradient + calcNumFluxesViscose) of the faces
// of marked elements.  The stage kernel then adds the viscous terms of the marked elements.  The
// same work as the fused schedule's AV passes (esdg_fused.hip), element by element, with the face
// traces read from the state instead of a side-trace buffer: the same bits.
// ---------------------------------------------------------------------------------------------

// Elements per task of the gradient pass (one flag per thread decides whether a tile has work)
constexpr int GRAD_TILE = 256;

// The marks of a step: stage s of the step marks with *mark + s.  Advanced by 8 at each step on
// the device, so that replays of a captured step take fresh marks.
__global__ void k_stage_mark(int* mark) { *mark += 8; }

template <int LO, int HI>
struct MaxRowLds {
    static constexpr int value = cmax(RowBlock<LO>::EPB * RowBlock<LO>::ES, MaxRowLds<LO + 1, HI>::value);
};
template <int HI>
struct MaxRowLds<HI, HI> {
    static constexpr int value = RowBlock<HI>::EPB * RowBlock<HI>::ES;
};

template <int NGL>
__device__ __forceinline__ void ens_sensor_body(const StageScn& S, int e0, int src, int m, bool stage0, double* sV,
                                                double* sT, double (*sSum)[4], unsigned long long* sMax) {
    if (threadIdx.x < NGL * NGL) sV[threadIdx.x] = glob(S.VdmInv)[threadIdx.x];
    const double v = sensor_rows<NGL>(S.Nelem, e0, S.sc, glob((const double*)S.q[src]), glob(S.B), sV, sT, sSum);
    const int el = threadIdx.x / NGL, r = threadIdx.x % NGL;
    const int e = e0 + el;
    if (el < RowBlock<NGL>::EPB && e < S.Nelem && r == 0) {
        glob(S.visc)[e] = v;
        if (v != 0.0) mark_reach(S, e, m);
        // non-negative doubles order as their bit patterns; a NaN stays out, as fmax leaves it
        if (stage0 && v > 0.0) atomicMax(sMax, (unsigned long long)__double_as_longlong(v));
    }
}

// ShockCapturing of the stage's state over RowBlock tiles (tasks: {scenario, first element, ngl}).
// A scenario whose time has reached T (or whose step has dt = 0) is skipped; in stage 0 the
// block's largest viscosity enters the scenario's eps_max.
template <int LO, int HI>
__global__ void __launch_bounds__(256) k_ens_sensor(const StageScn* __restrict__ scns,
                                                    const StageTask* __restrict__ tasks, int src,
                                                    const int* __restrict__ markp, int stage) {
    __shared__ double sV[HI * HI];
    __shared__ double sT[MaxRowLds<LO, HI>::value];
    __shared__ double sSum[RowBlock<LO>::EPB][4];
    __shared__ unsigned long long sMax;
    const StageTask tk = tasks[blockIdx.x];
    const StageScn& S = scns[tk.scn];
    if (stage > 0 && S.clock[0] == 0.0) return;
    if (stage == 0 && S.clock[1] >= S.T) return;
    const int m = *markp + stage;
    if (threadIdx.x == 0) sMax = 0ull;   // published by the sensor's first barrier
    switch (tk.ngl) {
#define ESDG_CASE(n)                                                                          \
    case n:                                                                                   \
        if constexpr (n >= LO && n <= HI) ens_sensor_body<n>(S, tk.e0, src, m, stage == 0, sV, sT, sSum, &sMax); \
        break;
        ESDG_CASE(3) ESDG_CASE(4) ESDG_CASE(5) ESDG_CASE(6) ESDG_CASE(7) ESDG_CASE(8)
#undef ESDG_CASE
        default: break;
    }
    if (stage == 0) {
        __syncthreads();
        if (threadIdx.x == 0 && sMax) atomicMax(S.emax, sMax);
    }
}

template <int NGL>
__device__ __forceinline__ void ens_grad_body(const StageScn& S, int e0, int src, int m, double* sDh, double* sDs,
                                              double* sA, double* sB, double* sC, double* sE, double* sEps) {
    using EB = ElemBlock<NGL>;
    constexpr int NN = NGL * NGL;
    const int e1 = min(e0 + GRAD_TILE, S.Nelem);
    // most tiles hold no marked element: one flag per thread, one barrier, and the block is done
    const int et = e0 + (int)threadIdx.x;
    if (!__syncthreads_or(et < e1 && S.active[et] == m)) return;
    if (threadIdx.x < NN) {
        sDh[threadIdx.x] = S.Dhat[threadIdx.x];
        sDs[threadIdx.x] = S.Dstrong[threadIdx.x];
    }
    const AvMesh M{S.q[src], S.Jac, S.Yxi, S.Yeta, S.Xxi, S.Xeta, S.E2E, S.EMS, S.EOR, S.EdgeData,
                   S.nx, S.ny, S.scal, S.bL, S.bR, GatherConsts{S.bc, S.h_zero}, S.geomFace, S.tol, S.gtr, S.Qv};
    const int el = threadIdx.x / NN;
    for (int g = e0; g < e1; g += EB::EPB) {
        const int e = g + el;
        const bool act = el < EB::EPB && e < e1 && S.active[e] == m;
        // the first barrier also publishes sDh, sDs; a group without a marked element is skipped
        if (!__syncthreads_or(act)) continue;
        av_element_block<NGL>(M, TraceFromState{S.q[src]}, e, act, act ? S.visc[e] : 0.0, sDh, sDs, sA, sB, sC, sE,
                              sEps);
    }
}

// The gradient pass over tiles of GRAD_TILE elements (tasks: {scenario, first element, ngl}): the
// marked elements of each tile, in groups of ElemBlock<ngl>::EPB elements, one thread per node.
template <int LO, int HI>
__global__ void __launch_bounds__(256) k_ens_grad(const StageScn* __restrict__ scns,
                                                  const StageTask* __restrict__ tasks, int src,
                                                  const int* __restrict__ markp, int stage) {
    __shared__ double sDh[HI * HI], sDs[HI * HI];
    __shared__ double sA[256], sB[256], sC[256], sE[256];
    __shared__ double sEps[ElemBlock<LO>::EPB];
    const StageTask tk = tasks[blockIdx.x];
    const StageScn& S = scns[tk.scn];
    const int m = *markp + stage;
    switch (tk.ngl) {
#define ESDG_CASE(n)                                                                          \
    case n:                                                                                   \
        if constexpr (n >= LO && n <= HI) ens_grad_body<n>(S, tk.e0, src, m, sDh, sDs, sA, sB, sC, sE, sEps); \
        break;
        ESDG_CASE(3) ESDG_CASE(4) ESDG_CASE(5) ESDG_CASE(6) ESDG_CASE(7) ESDG_CASE(8)
#undef ESDG_CASE
        default: break;
    }
}

// The viscous face fluxes over tiles of 256 faces (tasks: {scenario, first face, ngl}), one thread
// per face: a face is evaluated (all its nodes) when either of its elements is marked; the faces a
// marked element reads are exactly those.
__global__ void __launch_bounds__(256) k_ens_vface(const StageScn* __restrict__ scns,
                                                   const StageTask* __restrict__ tasks,
                                                   const int* __restrict__ markp, int stage) {
    const StageTask tk = tasks[blockIdx.x];
    const StageScn& S = scns[tk.scn];
    const int m = *markp + stage;
    const int ngl = tk.ngl;
    const int f = tk.e0 + (int)threadIdx.x;
    if (f >= S.Nfaces) return;
    const int* E = S.EdgeData + 8 * (int64_t)f;
    const int iel = E[0], ier = E[1], cpuL = E[5], cpuR = E[6];
    if (!((cpuL == 0 && S.active[iel] == m) || (cpuR == 0 && S.active[ier] == m))) return;
    const int isl = E[2], isr = E[3], rv = E[4];
    for (int i = 0; i < ngl; ++i)
        viscous_face_node(f, i, ngl, S.geomFace, iel, ier, isl, isr, rv, cpuL, cpuR, S.gtr, S.visc, S.nx, S.ny, S.scal,
                          S.SPV);
}

// Loads each degree's D into c_stage_D once per device; a later batch must bring the same D.
static int load_degree_operators(esdg_ctx* ctx, const std::vector<StageScn>& scns) {
    static std::mutex mtx;
    static std::map<int, std::map<int, std::vector<double>>> loaded;   // device -> ngl -> D
    std::lock_guard<std::mutex> lock(mtx);
    for (const StageScn& S : scns) {
        const int n2 = S.ngl * S.ngl;
        std::vector<double> D(n2);
        int rc = esdg_memcpy(ctx, D.data(), S.D, n2 * sizeof(double), ESDG_D2H);
        if (rc) return rc;
        auto& slot = loaded[ctx->device][S.ngl];
        if (slot.empty()) {
            rc = hip_check(ctx, hipSetDevice(ctx->device), "set device");
            if (!rc)
                rc = hip_check(ctx, hipMemcpyToSymbol(HIP_SYMBOL(c_stage_D), D.data(), n2 * sizeof(double),
                                                      (size_t)S.ngl * 64 * sizeof(double), hipMemcpyHostToDevice),
                               "D to constant memory");
            if (rc) return rc;
            slot = D;
        } else if (std::memcmp(slot.data(), D.data(), n2 * sizeof(double)) != 0) {
            return set_error(ctx, ESDG_ERR_ARG, "the single-kernel stage: D differs from the one loaded for this degree");
        }
    }
    return ESDG_OK;
}

// One record per element side (StageScn::srec) from the connectivity arrays.
__global__ void __launch_bounds__(256) k_side_records(int Nelem, const int* __restrict__ E2E, const int* __restrict__ EOR,
                                                      const int* __restrict__ EMS, const int* __restrict__ ED,
                                                      int4* __restrict__ rec) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Nelem * 4) return;
    const int f = E2E[t] - 1;
    const int* E = ED + 8 * (int64_t)f;
    const unsigned w = (unsigned)(E[2] & 3) | ((unsigned)(E[3] & 3) << 2) | ((unsigned)(E[4] & 1) << 4) |
                       ((unsigned)(EOR[t] & 1) << 5) | ((EMS[t] > 0 ? 1u : 0u) << 6) |
                       (((unsigned)(E[5] + 2) & 255u) << 8) | (((unsigned)(E[6] + 2) & 255u) << 16);
    rec[t] = make_int4(f, E[0], E[1], (int)w);
}

int StageBatch::init(esdg_ctx* ctx, const std::vector<StageScn>& scns) {
    host = scns;
    if (scns.empty()) return set_error(ctx, ESDG_ERR_ARG, "the single-kernel stage needs at least one scenario");
    for (const StageScn& S : scns) {
        if (S.ngl < 3 || S.ngl > 8) return set_error(ctx, ESDG_ERR_ARG, "the single-kernel stage supports 2 <= N <= 7");
        if (S.Nelem <= 0) return set_error(ctx, ESDG_ERR_ARG, "the single-kernel stage needs at least one element");
    }
    int rc = load_degree_operators(ctx, scns);
    if (rc) return rc;
    // side records, once per distinct connectivity
    for (size_t a = 0; a < host.size() && !rc; ++a) {
        StageScn& S = host[a];
        S.srec = nullptr;
        for (size_t b = 0; b < a; ++b)
            if (host[b].E2E == S.E2E && host[b].EdgeData == S.EdgeData && host[b].Nelem == S.Nelem) {
                S.srec = host[b].srec;
                break;
            }
        if (S.srec) continue;
        int4* d = nullptr;
        rc = esdg_malloc(ctx, (void**)&d, (size_t)S.Nelem * 4 * sizeof(int4));
        if (rc) break;
        srec_bufs.push_back(d);
        k_side_records<<<(unsigned)(((int64_t)S.Nelem * 4 + 255) / 256), 256, 0, ctx->stream>>>(S.Nelem, S.E2E, S.EOR,
                                                                                              S.EMS, S.EdgeData, d);
        rc = launch_check(ctx, "side records");
        S.srec = d;
    }
    if (rc) return rc;
    rc = hip_check(ctx, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device), "CU count");
    if (rc) return rc;
    has_av = false;
    for (const StageScn& S : scns) has_av = has_av || S.av;
    rc = esdg_malloc(ctx, (void**)&d_scn, host.size() * sizeof(StageScn));
    if (!rc) rc = esdg_memcpy(ctx, d_scn, host.data(), host.size() * sizeof(StageScn), ESDG_H2D);
    if (!rc && has_av) {
        rc = esdg_malloc(ctx, (void**)&d_mark, 64);
        if (!rc) rc = esdg_malloc(ctx, (void**)&d_icount, 5 * sizeof(unsigned long long));
        if (!rc) rc = esdg_fill(ctx, reinterpret_cast<double*>(d_mark), 0.0, 8);
        if (!rc) rc = esdg_fill(ctx, reinterpret_cast<double*>(d_icount), 0.0, 5);
        if (!rc) rc = build_av_tasks(ctx);
    }
    if (!rc) rc = set_schedule(ctx, StageSchedule{});
    return rc;
}

// Element and face tiles of the AV scenarios for the sensor (RowBlock tiles), the learned
// indicator (32-element tiles), the gradient pass (GRAD_TILE elements) and the viscous face pass
// (256 faces).
int StageBatch::build_av_tasks(esdg_ctx* ctx) {
    std::vector<StageTask> sens, ind, grad, vface;
    // gradient and viscous-face tasks: the non-learned scenarios' first, then the learned ones'
    for (int pass = 0; pass < 2; ++pass) {
        for (int s = 0; s < (int)host.size(); ++s) {
            const StageScn& S = host[s];
            if (!S.av || S.learned != pass) continue;
            const int ngl = S.ngl;
            if (S.learned)
                for (int e0 = 0; e0 < S.Nelem; e0 += 32) ind.push_back(StageTask{s, e0, ngl, 0});
            else
                for (int e0 = 0; e0 < S.Nelem; e0 += 256 / ngl) sens.push_back(StageTask{s, e0, ngl, 0});
            for (int e0 = 0; e0 < S.Nelem; e0 += GRAD_TILE) grad.push_back(StageTask{s, e0, ngl, 0});
            for (int f0 = 0; f0 < S.Nfaces; f0 += 256) vface.push_back(StageTask{s, f0, ngl, 0});
        }
        if (pass == 0) {
            ngrad_o = (int)grad.size();
            nvface_o = (int)vface.size();
        }
    }
    auto put = [&](const std::vector<StageTask>& v, StageTask*& d, int& n) -> int {
        if (d) esdg_free(ctx, d);
        d = nullptr;
        n = (int)v.size();
        if (v.empty()) return ESDG_OK;
        int rc = esdg_malloc(ctx, (void**)&d, v.size() * sizeof(StageTask));
        if (!rc) rc = esdg_memcpy(ctx, d, v.data(), v.size() * sizeof(StageTask), ESDG_H2D);
        return rc;
    };
    int rc = put(sens, d_sens, nsens);
    if (!rc) rc = put(ind, d_ind, nind);
    if (!rc) rc = put(grad, d_grad, ngrad);
    if (!rc) rc = put(vface, d_vface, nvface);
    return rc;
}

int StageBatch::set_indicator(esdg_ctx* ctx, const esdg_mlp* m, float lo, float hi) {
    if (m && !(lo <= hi)) return set_error(ctx, ESDG_ERR_ARG, "set_indicator: lo > hi");
    int learned = 0;
    for (StageScn& S : host) {
        S.learned = (m && S.av && S.ngl == 8) ? 1 : 0;
        learned += S.learned;
    }
    if (m && !learned)
        return set_error(ctx, ESDG_ERR_ARG, "the learned indicator needs a scenario with N = 7 and artificial viscosity");
    mlp = m;
    ind_lo = lo;
    ind_hi = hi;
    drop_graphs();
    int rc = esdg_memcpy(ctx, d_scn, host.data(), host.size() * sizeof(StageScn), ESDG_H2D);
    if (!rc && has_av) rc = build_av_tasks(ctx);
    if (!rc) rc = set_schedule(ctx, sched);   // the stage launches split by group (or not)
    return rc;
}

// A fresh start of the marks (the step counter would wrap): every mark array back to 0.
int StageBatch::reset_marks(esdg_ctx* ctx) {
    int rc = esdg_fill(ctx, reinterpret_cast<double*>(d_mark), 0.0, 8);
    for (const StageScn& S : host)
        if (!rc && S.av) rc = esdg_fill(ctx, reinterpret_cast<double*>(S.active), 0.0, ((size_t)S.Nelem + 1) / 2);
    marks_used = 0;
    return rc;
}

int StageBatch::av_pass(esdg_ctx* ctx, int src, int stage) {
    if (nind && !mlp) return set_error(ctx, ESDG_ERR_STATE, "learned scenarios without a network");
    // The modal-sensor and the learned scenarios are disjoint (each pass writes its own scenarios'
    // viscosities, marks and maxima): with both present the sensor pass runs on the second stream
    // beside the indicator pass (not under per-launch profiling, whose events are on the main
    // stream).
    const bool overlap = nsens && nind && ctx->side && !ctx->profile;
    if (overlap) {
        (void)hipEventRecord(ctx->fork, ctx->stream);
        (void)hipStreamWaitEvent(ctx->side, ctx->fork, 0);
    }
    if (nsens) {
        KernelScope ks(ctx, "ens_sensor");
        k_ens_sensor<3, 8><<<nsens, 256, 0, overlap ? ctx->side : ctx->stream>>>(d_scn, d_sens, src, d_mark, stage);
        int rc = launch_check(ctx, "ens_sensor");
        if (rc) return rc;
    }
    if (overlap) (void)hipEventRecord(ctx->join, ctx->side);
    if (nind) {
        int rc = mlp_ensemble_indicator(ctx, mlp, d_scn, d_ind, nind, src, d_mark, stage, ind_lo, ind_hi, ind_audit,
                                        d_icount);
        if (rc) return rc;
    }
    if (overlap) (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
    return ESDG_OK;
}

int StageBatch::av_terms_group(esdg_ctx* ctx, int src, int stage, int learned, hipStream_t st) {
    const int gb = learned ? ngrad_o : 0, gn = learned ? ngrad - ngrad_o : ngrad_o;
    const int vb = learned ? nvface_o : 0, vn = learned ? nvface - nvface_o : nvface_o;
    if (gn) {
        k_ens_grad<3, 8><<<gn, 256, 0, st>>>(d_scn, d_grad + gb, src, d_mark, stage);
        int rc = launch_check(ctx, "ens_grad");
        if (rc) return rc;
    }
    if (vn) {
        k_ens_vface<<<vn, 256, 0, st>>>(d_scn, d_vface + vb, d_mark, stage);
        return launch_check(ctx, "ens_viscous_face");
    }
    return ESDG_OK;
}

int StageBatch::av_terms(esdg_ctx* ctx, int src, int stage) {
    {
        KernelScope ks(ctx, "ens_grad");
        k_ens_grad<3, 8><<<ngrad, 256, 0, ctx->stream>>>(d_scn, d_grad, src, d_mark, stage);
        int rc = launch_check(ctx, "ens_grad");
        if (rc) return rc;
    }
    KernelScope ks(ctx, "ens_viscous_face");
    k_ens_vface<<<nvface, 256, 0, ctx->stream>>>(d_scn, d_vface, d_mark, stage);
    return launch_check(ctx, "ens_viscous_face");
}

void StageBatch::drop_graphs() {
    for (auto& x : exec)
        if (x) { (void)hipGraphExecDestroy(x); x = nullptr; }
}

int StageBatch::set_schedule(esdg_ctx* ctx, const StageSchedule& sc) {
    if (sc.order < 0 || sc.order > 2 || sc.blocks_per_cu < 0 || sc.blocks_per_cu > 64 ||
        (sc.tile != 1 && sc.tile != 2 && sc.tile != 4))
        return set_error(ctx, ESDG_ERR_ARG, "stage schedule: order must be 0..2, blocks_per_cu 0..64, tile 1, 2 or 4");
    const int ns = (int)host.size();
    // scenario visiting order
    std::vector<int> scn(ns);
    for (int s = 0; s < ns; ++s) scn[s] = s;
    if (sc.order == 2 || sc.per_degree)
        std::stable_sort(scn.begin(), scn.end(), [&](int a, int b) { return host[a].ngl > host[b].ngl; });
    std::vector<StageTask> tasks;
    auto tiles = [&](int s) { const int epw = 64 / host[s].ngl; return (host[s].Nelem + epw - 1) / epw; };
    auto task = [&](int s, int t) { return StageTask{s, t * (64 / host[s].ngl), host[s].ngl, 0}; };
    if (sc.order == 1) {
        // groups of scenarios on the same geometry, in order of first appearance; tile-major inside
        std::vector<int> done(ns, 0);
        for (int a = 0; a < ns; ++a) {
            const int s0 = scn[a];
            if (done[s0]) continue;
            std::vector<int> grp;
            for (int b = a; b < ns; ++b) {
                const int s1 = scn[b];
                if (!done[s1] && host[s1].Jac == host[s0].Jac && host[s1].ngl == host[s0].ngl &&
                    host[s1].Nelem == host[s0].Nelem) {
                    grp.push_back(s1);
                    done[s1] = 1;
                }
            }
            for (int t = 0; t < tiles(s0); ++t)
                for (int s1 : grp) tasks.push_back(task(s1, t));
        }
    } else {
        for (int s : scn)
            for (int t = 0; t < tiles(s); ++t) tasks.push_back(task(s, t));
    }
    // with both modal-sensor and learned scenarios: the other scenarios' tasks first, then the
    // learned ones', each group in the schedule's order and launched on its own stream
    split = has_av && nsens > 0 && nind > 0;
    int nother = (int)tasks.size();
    if (split) {
        std::stable_partition(tasks.begin(), tasks.end(), [&](const StageTask& t) { return !host[t.scn].learned; });
        nother = 0;
        while (nother < (int)tasks.size() && !host[tasks[nother].scn].learned) ++nother;
    }
    // launch ranges
    std::vector<Range> rg;
    int cur_learned = 0;
    auto add_range = [&](int beg, int n, int lo, int hi) {
        if (n <= 0) return;
        Range r{beg, n, lo, hi, n, 0, cur_learned};
        // chunks of sc.tile tasks; with xcd every XCD's part is chunked on its own
        const int nch = sc.xcd ? 8 * (((n + 7) / 8 + sc.tile - 1) / sc.tile) : (n + sc.tile - 1) / sc.tile;
        const int pgrid = sc.blocks_per_cu * ncu;
        if (sc.blocks_per_cu > 0 && pgrid < nch && (!sc.xcd || pgrid >= 8)) {
            r.grid = sc.xcd ? (pgrid / 8) * 8 : pgrid;
            r.persistent = 1;
        } else if (sc.tile > 1) {
            r.grid = nch;
            r.persistent = 1;
        }
        rg.push_back(r);
    };
    auto add_group = [&](int g0, int g1) {
        if (sc.per_degree) {
            int beg = g0;
            while (beg < g1) {
                int end = beg;
                while (end < g1 && tasks[end].ngl == tasks[beg].ngl) ++end;
                add_range(beg, end - beg, tasks[beg].ngl, tasks[beg].ngl);
                beg = end;
            }
        } else {
            int lo = 99, hi = 0;
            for (int i = g0; i < g1; ++i) { lo = std::min(lo, tasks[i].ngl); hi = std::max(hi, tasks[i].ngl); }
            add_range(g0, g1 - g0, lo, hi);
        }
    };
    add_group(0, nother);
    cur_learned = 1;
    add_group(nother, (int)tasks.size());
    if (d_task && (int)tasks.size() != ntask) {
        esdg_free(ctx, d_task);
        d_task = nullptr;
    }
    int rc = ESDG_OK;
    if (!d_task) rc = esdg_malloc(ctx, (void**)&d_task, tasks.size() * sizeof(StageTask));
    if (!rc) rc = esdg_memcpy(ctx, d_task, tasks.data(), tasks.size() * sizeof(StageTask), ESDG_H2D);
    if (rc) return rc;
    ntask = (int)tasks.size();
    ranges = rg;
    sched = sc;
    drop_graphs();
    return ESDG_OK;
}

int StageBatch::lambda(esdg_ctx* ctx, int src) {
    KernelScope ks(ctx, "stage_lambda");
    k_stage_lambda<<<ntask, 64, 0, ctx->stream>>>(d_scn, d_task, src);
    return launch_check(ctx, "stage_lambda");
}

template <int LO, int HI, bool AV>
static void launch_stage(hipStream_t st, const StageBatch& b, const StageBatch::Range& r, int src, double rkA,
                         double rkC, int last, int stage) {
    if (r.persistent)
        k_stage<LO, HI, true, AV><<<r.grid, 64, 0, st>>>(b.d_scn, b.d_task + r.beg, r.n, b.sched.xcd,
                                                                   b.sched.tile, src, rkA, rkC, last, b.d_mark, stage);
    else
        k_stage<LO, HI, false, AV><<<r.grid, 64, 0, st>>>(b.d_scn, b.d_task + r.beg, r.n, b.sched.xcd, 1,
                                                                    src, rkA, rkC, last, b.d_mark, stage);
}

template <bool AV>
static void launch_ranges(hipStream_t st, const StageBatch& b, const StageBatch::Range& r, int src, double rkA,
                          double rkC, int last, int stage) {
    if (r.lo != r.hi) launch_stage<3, 8, AV>(st, b, r, src, rkA, rkC, last, stage);
    else switch (r.lo) {
        case 3: launch_stage<3, 3, AV>(st, b, r, src, rkA, rkC, last, stage); break;
        case 4: launch_stage<4, 4, AV>(st, b, r, src, rkA, rkC, last, stage); break;
        case 5: launch_stage<5, 5, AV>(st, b, r, src, rkA, rkC, last, stage); break;
        case 6: launch_stage<6, 6, AV>(st, b, r, src, rkA, rkC, last, stage); break;
        case 7: launch_stage<7, 7, AV>(st, b, r, src, rkA, rkC, last, stage); break;
        default: launch_stage<8, 8, AV>(st, b, r, src, rkA, rkC, last, stage); break;
    }
}

int StageBatch::step(esdg_ctx* ctx, const double* A, const double* C, int stages, int& src) {
    const int ns = (int)host.size();
    if (has_av) {
        // the step's marks, then the viscosity of the step's state (its maximum enters the dt)
        KernelScope ks(ctx, "stage_mark");
        k_stage_mark<<<1, 1, 0, ctx->stream>>>(d_mark);
        int rc = launch_check(ctx, "stage_mark");
        if (!rc) rc = av_pass(ctx, src, 0);
        if (rc) return rc;
    }
    {
        KernelScope ks(ctx, "stage_dt");
        k_stage_dt<<<(ns + 63) / 64, 64, 0, ctx->stream>>>(d_scn, ns);
        int rc = launch_check(ctx, "stage_dt");
        if (rc) return rc;
    }
    for (int s = 0; s < stages; ++s) {
        const int last = (s == stages - 1);
        if (split && ctx->side && !ctx->profile) {
            // the other scenarios on the second stream, the learned ones on the main stream
            (void)hipEventRecord(ctx->fork, ctx->stream);
            (void)hipStreamWaitEvent(ctx->side, ctx->fork, 0);
            int rc = ESDG_OK;
            if (s > 0 && nsens) {
                k_ens_sensor<3, 8><<<nsens, 256, 0, ctx->side>>>(d_scn, d_sens, src, d_mark, s);
                rc = launch_check(ctx, "ens_sensor");
            }
            if (!rc) rc = av_terms_group(ctx, src, s, 0, ctx->side);
            for (const Range& r : ranges)
                if (!rc && !r.learned) {
                    launch_ranges<true>(ctx->side, *this, r, src, A[s], C[s], last, s);
                    rc = launch_check(ctx, "stage");
                }
            (void)hipEventRecord(ctx->join, ctx->side);
            if (!rc && s > 0)
                rc = mlp_ensemble_indicator(ctx, mlp, d_scn, d_ind, nind, src, d_mark, s, ind_lo, ind_hi, ind_audit,
                                            d_icount);
            if (!rc) rc = av_terms_group(ctx, src, s, 1, ctx->stream);
            for (const Range& r : ranges)
                if (!rc && r.learned) {
                    launch_ranges<true>(ctx->stream, *this, r, src, A[s], C[s], last, s);
                    rc = launch_check(ctx, "stage");
                }
            (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
            if (rc) return rc;
            src ^= 1;
            continue;
        }
        if (has_av) {
            int rc = (s > 0) ? av_pass(ctx, src, s) : ESDG_OK;
            if (!rc) rc = av_terms(ctx, src, s);
            if (rc) return rc;
        }
        for (const Range& r : ranges) {
            KernelScope ks(ctx, "stage");
            if (has_av) launch_ranges<true>(ctx->stream, *this, r, src, A[s], C[s], last, s);
            else launch_ranges<false>(ctx->stream, *this, r, src, A[s], C[s], last, s);
            int rc = launch_check(ctx, "stage");
            if (rc) return rc;
        }
        src ^= 1;
    }
    return ESDG_OK;
}

int StageBatch::run(esdg_ctx* ctx, const double* A, const double* C, int stages, int steps, int& src) {
    int n = 0;
    if (has_av) {
        // marks advance by 8 per step: start them afresh long before the int would wrap
        if (marks_used + (int64_t)steps + 2 > (int64_t)(1 << 27)) {
            int rc = reset_marks(ctx);
            if (rc) return rc;
        }
        marks_used += steps;
    }
    if (sched.graph && !ctx->profile && steps >= 2) {
        // two steps flip the buffers an even number of times, so the graph that starts from buffer
        // `src` ends there and can be replayed
        const int start = src;
        if (!exec[start]) {
            hipGraph_t gph = nullptr;
            int rc = hip_check(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "graph capture");
            if (rc) return rc;
            int s2 = start;
            for (int k = 0; k < 2 && rc == ESDG_OK; ++k) rc = step(ctx, A, C, stages, s2);
            const hipError_t e = hipStreamEndCapture(ctx->stream, &gph);
            if (rc) {
                if (gph) (void)hipGraphDestroy(gph);
                return rc;
            }
            rc = hip_check(ctx, e, "graph capture end");
            if (rc) return rc;
            rc = hip_check(ctx, hipGraphInstantiate(&exec[start], gph, nullptr, nullptr, 0), "graph instantiate");
            (void)hipGraphDestroy(gph);
            if (rc) return rc;
        }
        for (; n + 2 <= steps; n += 2) {
            int rc = hip_check(ctx, hipGraphLaunch(exec[start], ctx->stream), "graph launch");
            if (rc) return rc;
        }
    }
    for (; n < steps; ++n) {
        int rc = step(ctx, A, C, stages, src);
        if (rc) return rc;
    }
    return ESDG_OK;
}

void StageBatch::release(esdg_ctx* ctx) {
    drop_graphs();
    for (StageTask** p : {&d_task, &d_sens, &d_ind, &d_grad, &d_vface})
        if (*p) { esdg_free(ctx, *p); *p = nullptr; }
    if (d_scn) esdg_free(ctx, d_scn);
    for (int4* p : srec_bufs) esdg_free(ctx, p);
    srec_bufs.clear();
    if (d_mark) esdg_free(ctx, d_mark);
    if (d_icount) esdg_free(ctx, d_icount);
    d_scn = nullptr;
    d_mark = nullptr;
    d_icount = nullptr;
}

}  // namespace esdg
