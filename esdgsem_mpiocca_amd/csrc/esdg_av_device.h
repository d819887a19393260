// esdg_av_device.h -- the artificial-viscosity element and face work, shared by the fused stage
// (esdg_fused.hip: one mesh, face traces from the side-trace buffer) and the batched ensembles
// (esdg_stage.hip: many meshes, face traces gathered from the state).  Both give the same bits.
//
//   br1_face_flux      calcNumFluxesGradient.okl:16-135 at one face node
//   av_element_block   calcGradient.okl:18-156 + SurfaceKernelGradient.okl:16-148 +
//                      scaleGradient.okl:17-60 + VolumeKernelViscose.okl:17-150 for a block of
//                      elements, one thread per node
//   viscous_face_node  CollectEdgeDataGradient.okl:17-194 + calcNumFluxesViscose.okl:17-107 at one
//                      face node
#pragma once

#include "esdg_device.h"
#include "esdg_internal.h"

namespace esdg {

// Face states of face f as CollectEdgeData stores them into qL / qR, read from the face slots of
// the side-trace buffer or straight from the state.
struct TraceFromSides {
    const double* tr;
    __device__ __forceinline__ FaceState operator()(int ngl, int f, const int* E, bool left, int k, double nx,
                                                    double ny, const GatherConsts& gc, double bl, double br) const {
        return gathered_tr(tr, ngl, f, E, left, k, nx, ny, gc, bl, br);
    }
};
struct TraceFromState {
    const double* q;
    __device__ __forceinline__ FaceState operator()(int ngl, int, const int* E, bool left, int k, double nx, double ny,
                                                    const GatherConsts& gc, double bl, double br) const {
        return gathered(q, ngl, E, left, k, nx, ny, gc, bl, br);
    }
};

// calcNumFluxesGradient.okl:16-135 at face node i: the BR1 face fluxes of u and v (x and y parts)
// from qL and qR as the gather stores them, both read unreversed at node i, the exterior state
// taking the place of a missing side.
struct Br1Flux {
    double x0, x1, y0, y1;
};

template <class Src>
__device__ __forceinline__ Br1Flux br1_face_flux(const Src& src, int ngl, int f, const int* E, int i, double nx,
                                                 double ny, double sc, double bl, double br, const GatherConsts& gc,
                                                 double tol, double geomFace) {
    const FaceState L = src(ngl, f, E, true, i, nx, ny, gc, bl, br);
    const FaceState Ri = src(ngl, f, E, false, i, nx, ny, gc, bl, br);
    const FaceState LL = (E[5] == -1) ? Ri : L;
    const FaceState RR = (E[6] == -1) ? L : Ri;
    const double iL = (LL.h > tol) ? 1.0 / LL.h : 0.0;
    const double iR = (RR.h > tol) ? 1.0 / RR.h : 0.0;
    const double uL = LL.hu * iL, uR = RR.hu * iR, vL = LL.hv * iL, vR = RR.hv * iR;
    return Br1Flux{nx * sc * geomFace * 0.5 * (uL + uR), nx * sc * geomFace * 0.5 * (vL + vR),
                   ny * sc * geomFace * 0.5 * (uL + uR), ny * sc * geomFace * 0.5 * (vL + vR)};
}

// The inputs of the AV element work for one mesh.
struct AvMesh {
    const double *q, *Jac, *Yxi, *Yeta, *Xxi, *Xeta;
    const int *E2E, *EMS, *EOR, *ED;
    const double *nx, *ny, *scal, *bL, *bR;
    GatherConsts gc;
    double geomFace, tol;
    double* gtr;    // out: side traces of the scaled gradients [e][side][gX0, gX1, gY0, gY1][k]
    double* Qv;     // out: viscous volume term [e][2][ngl^2] (hu, hv)
};

// One block of EPB elements, NN threads each (node n = (xi a, eta c) = (n % NGL, n / NGL)): the
// BR1 gradient of u and v (weak-form volume part with Dhat, then the surface lift of the BR1 face
// fluxes of the element's sides), scaled by h, its side traces, and the viscous volume term
// Qv = J^-1 Dstrong (eps J grad) in the contravariant form.  Called by every thread of the block;
// act = whether this thread's element takes part, eps its viscosity.  LDS: sDh, sDs (ngl^2), and
// sA, sB, sC, sE (EPB ngl^2 each), sEps (EPB); the call ends with the LDS free.
template <int NGL, class Src>
__device__ __forceinline__ void av_element_block(const AvMesh& M, const Src& src, int e, bool act, double eps,
                                                 const double* sDh, const double* sDs, double* sA, double* sB,
                                                 double* sC, double* sE, double* sEps) {
    constexpr int NN = NGL * NGL;
    const int el = threadIdx.x / NN, n = threadIdx.x % NN;
    if (act && n == 0) sEps[el] = eps;
    const int a = n % NGL, c = n / NGL;
    const int64_t xid = (int64_t)e * NN + n;
    double h = 0.0, hu = 0.0, hv = 0.0;
    if (act) {
        const int64_t id = (int64_t)e * NN * ESDG_NEQ + n;
        h = M.q[id]; hu = M.q[id + NN]; hv = M.q[id + 2 * NN];
        const double Qinv = (h > M.tol) ? 1.0 / h : 0.0;
        sB[threadIdx.x] = hu * Qinv;                // u at node n
        sC[threadIdx.x] = hv * Qinv;                // v at node n
    }
    __syncthreads();
    // ---- BR1 gradient (weak form volume part) ----
    const double* u = sB + el * NN;
    const double* v = sC + el * NN;
    double gx0 = 0.0, gx1 = 0.0, gy0 = 0.0, gy1 = 0.0;
    double ye = 0.0, yx = 0.0, xe = 0.0, xx = 0.0, rJ = 0.0;
    if (act) {
        double u_xi = 0.0, u_eta = 0.0, v_xi = 0.0, v_eta = 0.0;
#pragma unroll
        for (int l = 0; l < NGL; ++l) {
            u_xi += sDh[a * NGL + l] * u[c * NGL + l];
            u_eta += sDh[c * NGL + l] * u[l * NGL + a];
            v_xi += sDh[a * NGL + l] * v[c * NGL + l];
            v_eta += sDh[c * NGL + l] * v[l * NGL + a];
        }
        ye = M.Yeta[xid]; yx = M.Yxi[xid]; xe = M.Xeta[xid]; xx = M.Xxi[xid]; rJ = M.Jac[xid];
        const double u_x = ye * u_xi - yx * u_eta;
        const double v_x = ye * v_xi - yx * v_eta;
        const double u_y = -xe * u_xi + xx * u_eta;
        const double v_y = -xe * v_xi + xx * v_eta;
        gx0 = rJ * u_x; gx1 = rJ * v_x; gy0 = rJ * u_y; gy1 = rJ * v_y;
        // surface lift of the BR1 face fluxes: sides 0/2 first, then 1/3 (SurfaceKernelGradient.okl)
#pragma unroll
        for (int ph = 0; ph < 2; ++ph) {
            int side = -1, t = 0;
            if (ph == 0) {
                if (c == 0) { side = 0; t = a; }
                else if (c == NGL - 1) { side = 2; t = a; }
            } else {
                if (a == NGL - 1) { side = 1; t = c; }
                else if (a == 0) { side = 3; t = c; }
            }
            if (side < 0) continue;
            const int ifa = M.E2E[e * 4 + side] - 1;
            const int rev = M.EOR[e * 4 + side];
            const int ms = M.EMS[e * 4 + side];
            const int inc = rev * t + (1 - rev) * (NGL - 1 - t);
            const int64_t fn = (int64_t)ifa * NGL + inc;
            const Br1Flux F = br1_face_flux(src, NGL, ifa, M.ED + 8 * (int64_t)ifa, inc, M.nx[fn], M.ny[fn], M.scal[fn],
                                            M.bL[fn], M.bR[fn], M.gc, M.tol, M.geomFace);
            if (ph == 0) gx0 += ms * F.x0 * rJ;
            else gx0 += ms * rJ * F.x0;
            gx1 += ms * rJ * F.x1;
            gy0 += ms * rJ * F.y0;
            gy1 += ms * rJ * F.y1;
        }
        // scaleGradient: multiply by h
        gx0 = h * gx0; gx1 = h * gx1; gy0 = h * gy0; gy1 = h * gy1;
        // side traces of the scaled gradients, [e][side][gX0, gX1, gY0, gY1][k] (k in the side's order)
        double* T = M.gtr + (int64_t)e * 16 * NGL;
        auto put = [&](int side, int k) {
            double* t = T + side * 4 * NGL + k;
            t[0] = gx0; t[NGL] = gx1; t[2 * NGL] = gy0; t[3 * NGL] = gy1;
        };
        if (c == 0) put(0, a);
        if (c == NGL - 1) put(2, a);
        if (a == NGL - 1) put(1, c);
        if (a == 0) put(3, c);
    }
    __syncthreads();   // sEps ready; u, v no longer needed
    // ---- viscous volume term: contravariant fluxes, then Dstrong ----
    if (act) {
        const double ep = sEps[el];
        const double f2 = ep * gx0, f3 = ep * gx1, g2 = ep * gy0, g3 = ep * gy1;
        sA[threadIdx.x] = ye * f2 - xe * g2;      // F2
        sB[threadIdx.x] = -yx * f2 + xx * g2;     // G2
        sC[threadIdx.x] = ye * f3 - xe * g3;      // F3
        sE[threadIdx.x] = -yx * f3 + xx * g3;     // G3
    }
    __syncthreads();
    if (act) {
        const double* F2 = sA + el * NN;
        const double* G2 = sB + el * NN;
        const double* F3 = sC + el * NN;
        const double* G3 = sE + el * NN;
        double d2 = 0.0, d3 = 0.0;
#pragma unroll
        for (int l = 0; l < NGL; ++l) {
            d2 += sDs[a * NGL + l] * F2[c * NGL + l];
            d2 += sDs[c * NGL + l] * G2[l * NGL + a];
            d3 += sDs[a * NGL + l] * F3[c * NGL + l];
            d3 += sDs[c * NGL + l] * G3[l * NGL + a];
        }
        double* o = M.Qv + (int64_t)e * NN * 2 + n;
        o[0] = rJ * d2;
        o[NN] = rJ * d3;
    }
    __syncthreads();   // LDS free for the next element block
}

// calcNumFluxesViscose.okl at face node i of face f from the side traces of the scaled gradients
__device__ __forceinline__ void viscous_face_node(int f, int i, int ngl, double geomFace, int iel, int ier, int isl,
                                                  int isr, int rv, int cpuL, int cpuR, const double* __restrict__ gtr,
                                                  const double* __restrict__ visc, const double* __restrict__ nxa,
                                                  const double* __restrict__ nya, const double* __restrict__ scal,
                                                  double* __restrict__ SPV) {
    const int64_t t = (int64_t)f * ngl + i;
    const int64_t eid = (int64_t)f * ngl * 2 + i;
    double N0 = 0.0, N1 = 0.0;
    if (!(cpuL == -1 || cpuR == -1)) {
        // both sides present (procL = procR = 0): the boundary-face branches of the gather never apply;
        // the right element's trace is read at the reversed index, as CollectEdgeDataGradient does.
        // Zero viscosity on both sides makes the flux zero: the traces are then not read.
        const double m = fmax(visc[iel], visc[ier]);
        if (m != 0.0) {
            const int inc = (rv == 0) ? (ngl - 1) - i : i;
            const double* TL = gtr + ((int64_t)iel * 4 + isl) * 4 * ngl + i;
            const double* TR = gtr + ((int64_t)ier * 4 + isr) * 4 * ngl + inc;
            const double nx = nxa[t], ny = nya[t];
            const double f2L = nx * TL[0] + ny * TL[2 * ngl];
            const double f2R = nx * TR[0] + ny * TR[2 * ngl];
            const double f3L = nx * TL[ngl] + ny * TL[3 * ngl];
            const double f3R = nx * TR[ngl] + ny * TR[3 * ngl];
            N0 = 0.5 * m * (f2L + f2R);
            N1 = 0.5 * m * (f3L + f3R);
        }
    }
    const double sc = scal[t];
    SPV[eid] = -sc * geomFace * N0;
    SPV[eid + ngl] = -sc * geomFace * N1;
}

}  // namespace esdg
