// esdg_device.h -- device functions shared by the kernels that follow the reference kernel by
// kernel (esdg_kernels.hip) and the fused stage kernels (esdg_fused.hip).  Each states the okl
// source it restates; the operation order inside every expression is that source's.
#pragma once

#include <hip/hip_runtime.h>

#include "esdg.h"

namespace esdg {

// A pointer into global memory with its address space in the type.  Pointers that a kernel reads
// from a record in memory (rather than receives as arguments) are otherwise generic, and loads
// through them compile to flat instructions, which count against the LDS counter too: every wait
// for one of them then also waits for all LDS traffic, and no wait can be partial.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> glob(T* p) {
    return (gptr<T>)p;
}

// node index of trace point k of element side s (okl/GatherEdgeData/SolidWalls.okl:74-89)
__device__ __forceinline__ int side_node(int ngl, int side, int k) {
    switch (side) {
        case 0: return k;
        case 2: return ngl * ngl - ngl + k;
        case 3: return k * ngl;
        default: return (k + 1) * ngl - 1;
    }
}

template <int NGL>
struct ElemBlock {
    static constexpr int NODES = NGL * NGL;
    static constexpr int EPB = (256 / NODES) > 0 ? (256 / NODES) : 1;
    static constexpr int THREADS = EPB * NODES;
};

// One thread per element row (the sensor and limiter passes): 256 / NGL elements per block.
template <int NGL>
struct RowBlock {
    static constexpr int EPB = 256 / NGL;            // elements per block
    static constexpr int THREADS = EPB * NGL;
    static constexpr int RS = NGL + 1;               // padded LDS row stride of an element's modes
    static constexpr int ES = NGL * RS + 1;          // padded LDS stride between elements
};

__device__ __forceinline__ void wall_mirror(double nx, double ny, double vx, double vy, double& ox, double& oy) {
    const double v_n = nx * vx + ny * vy;
    const double v_t = -ny * vx + nx * vy;
    ox = -nx * v_n - ny * v_t;
    oy = -ny * v_n + nx * v_t;
}

// Exterior-state variants of CollectEdgeData (src/deviceclass.cpp:363-415).
struct GatherConsts {
    int bc;          // esdg_bc
    double h_zero;   // still-water level of the ocean boundary
};

struct FaceState {
    double h, hu, hv;
};

// The two element side traces of face node k (E: the face's EdgeData), read whether or not the
// side is an element: an absent side reads the other side's element instead (or element 0 if
// neither is one), so every address is valid; its values are not used.  The loads need no branch,
// so they issue together with whatever the caller loads beside them.  (QP: a pointer to const
// double, generic or gptr; side_node stays inside the element for any side value.)
template <class QP>
__device__ __forceinline__ void face_side_traces(QP __restrict__ q, int ngl, const int* E, int k, FaceState& Lq,
                                                 FaceState& Rq) {
    const int ngl2 = ngl * ngl;
    const int iel = E[0], ier = E[1], isl = E[2], isr = E[3];
    const bool hl = E[5] == 0, hr = E[6] == 0;
    const int64_t el = hl ? iel : (hr ? ier : 0), er = hr ? ier : (hl ? iel : 0);
    const int64_t idl = el * ngl2 * 3 + side_node(ngl, hl ? isl : isr, k);
    const int64_t idr = er * ngl2 * 3 + side_node(ngl, hr ? isr : isl, k);
    Lq = FaceState{q[idl], q[idl + ngl2], q[idl + 2 * ngl2]};
    Rq = FaceState{q[idr], q[idr + ngl2], q[idr + 2 * ngl2]};
}

// Trace of face node k as CollectEdgeData stores it into qL (left = true) or qR, from the element
// side traces Lq, Rq at that node (SolidWalls.okl:74-132), with the exterior state of a boundary
// face from the wall mirror (SolidWalls.okl:133-175), zero momentum (InnerOuter.okl) or the ocean
// rule (OceanBoundary.okl: still water h_zero - b where b <= 6.7, wall otherwise).  bl, br: the
// face's bL, bR at node k.
__device__ __forceinline__ FaceState gathered_from(const FaceState& Lq, const FaceState& Rq, const int* E, bool left,
                                                   double nx, double ny, const GatherConsts& gc, double bl,
                                                   double br) {
    const int cpuL = E[5], cpuR = E[6];
    FaceState L{0.0, 0.0, 0.0}, R{0.0, 0.0, 0.0};
    if (cpuL == 0) L = Lq;
    if (cpuR == 0) R = Rq;
    if (gc.bc == ESDG_BC_INNER_OUTER) {
        if (cpuL == -1) { L.h = R.h; L.hu = 0.0; L.hv = 0.0; }
        if (cpuR == -1) { R.h = L.h; R.hu = 0.0; R.hv = 0.0; }
    } else if (gc.bc == ESDG_BC_OCEAN) {
        if (cpuL == -1) {
            if (br > 6.7) { L.h = R.h; wall_mirror(nx, ny, R.hu, R.hv, L.hu, L.hv); }
            else { L.h = gc.h_zero - br; L.hu = 0.0; L.hv = 0.0; }
        }
        if (cpuR == -1) {
            if (bl > 6.7) { R.h = L.h; wall_mirror(nx, ny, L.hu, L.hv, R.hu, R.hv); }
            else { R.h = gc.h_zero - bl; R.hu = 0.0; R.hv = 0.0; }
        }
    } else {
        if (cpuL == -1) { L.h = R.h; wall_mirror(nx, ny, R.hu, R.hv, L.hu, L.hv); }
        if (cpuR == -1) { R.h = L.h; wall_mirror(nx, ny, L.hu, L.hv, R.hu, R.hv); }
    }
    return left ? L : R;
}

// Trace of face node k as CollectEdgeData stores it into qL (left = true) or qR, read from the
// state q.
template <class QP>
__device__ __forceinline__ FaceState gathered(QP __restrict__ q, int ngl, const int* E, bool left, int k,
                                              double nx, double ny, const GatherConsts& gc, double bl, double br) {
    FaceState Lq, Rq;
    face_side_traces(q, ngl, E, k, Lq, Rq);
    return gathered_from(Lq, Rq, E, left, nx, ny, gc, bl, br);
}

// Side traces of the state in face slots: the traces of the element side that is face f's left
// (right) side fill slot 2 f (2 f + 1), tr[(slot * 3 + v) * ngl + k] with v = h, hu, hv and k the
// side node in the element's side order (side_node), so that the face pass reads each face's two
// traces as one contiguous run.  slot[e * 4 + side] names the slot of each element side
// (trace_slots).  tr_put stores node n = j*ngl + i of element e into every side it lies on.
__device__ __forceinline__ void tr_put(double* __restrict__ tr, const int* __restrict__ slot, int ngl, int64_t e,
                                       int i, int j, double h, double hu, double hv) {
    auto put = [&](int side, int k) {
        double* t = tr + (int64_t)slot[e * 4 + side] * 3 * ngl + k;
        t[0] = h; t[ngl] = hu; t[2 * ngl] = hv;
    };
    if (j == 0) put(0, i);
    if (j == ngl - 1) put(2, i);
    if (i == ngl - 1) put(1, j);
    if (i == 0) put(3, j);
}

// gathered() with the interior traces of face f read from its slots of the side-trace buffer
// instead of the state
__device__ __forceinline__ FaceState gathered_tr(const double* __restrict__ tr, int ngl, int f, const int* E, bool left,
                                                 int k, double nx, double ny, const GatherConsts& gc, double bl,
                                                 double br) {
    const int cpuL = E[5], cpuR = E[6];
    // both slots of the face are read whether or not their side is present (the buffer holds both
    // slots of every face; an absent side's slot is never written and its values are not used), so
    // the reads need not wait for EdgeData
    const double* tl = tr + (int64_t)f * 6 * ngl + k;
    const double* trr = tr + ((int64_t)f * 6 + 3) * ngl + k;
    const FaceState Lt{tl[0], tl[ngl], tl[2 * ngl]}, Rt{trr[0], trr[ngl], trr[2 * ngl]};
    FaceState L{0.0, 0.0, 0.0}, R{0.0, 0.0, 0.0};
    if (cpuL == 0) L = Lt;
    if (cpuR == 0) R = Rt;
    if (gc.bc == ESDG_BC_INNER_OUTER) {
        if (cpuL == -1) { L.h = R.h; L.hu = 0.0; L.hv = 0.0; }
        if (cpuR == -1) { R.h = L.h; R.hu = 0.0; R.hv = 0.0; }
    } else if (gc.bc == ESDG_BC_OCEAN) {
        if (cpuL == -1) {
            if (br > 6.7) { L.h = R.h; wall_mirror(nx, ny, R.hu, R.hv, L.hu, L.hv); }
            else { L.h = gc.h_zero - br; L.hu = 0.0; L.hv = 0.0; }
        }
        if (cpuR == -1) {
            if (bl > 6.7) { R.h = L.h; wall_mirror(nx, ny, L.hu, L.hv, R.hu, R.hv); }
            else { R.h = gc.h_zero - bl; R.hu = 0.0; R.hv = 0.0; }
        }
    } else {
        if (cpuL == -1) { L.h = R.h; wall_mirror(nx, ny, R.hu, R.hv, L.hu, L.hv); }
        if (cpuR == -1) { R.h = L.h; wall_mirror(nx, ny, L.hu, L.hv, R.hu, R.hv); }
    }
    return left ? L : R;
}

struct FluxConsts {
    double g, geomFace, tol;
    int ES;
};

__device__ __forceinline__ void es_dissipation(double R21, double R22, double R23, double R31, double R32, double R33,
                                               double SL11, double SL22, double SL33, double j1, double j2, double j3,
                                               double& NF0, double& NF1, double& NF2) {
    const double R11 = 1.0, R12 = 0.0, R13 = 1.0;
    const double S11 = R11 * SL11, S12 = 0.0, S13 = R13 * SL33;
    const double S21 = R21 * SL11, S22 = R22 * SL22, S23 = R23 * SL33;
    const double S31 = R31 * SL11, S32 = R32 * SL22, S33 = R33 * SL33;
    const double M11 = S11 * R11 + S12 * R12 + S13 * R13;
    const double M12 = S11 * R21 + S12 * R22 + S13 * R23;
    const double M13 = S11 * R31 + S12 * R32 + S13 * R33;
    const double M21 = S21 * R11 + S22 * R12 + S23 * R13;
    const double M22 = S21 * R21 + S22 * R22 + S23 * R23;
    const double M23 = S21 * R31 + S22 * R32 + S23 * R33;
    const double M31 = S31 * R11 + S32 * R12 + S33 * R13;
    const double M32 = S31 * R21 + S32 * R22 + S33 * R23;
    const double M33 = S31 * R31 + S32 * R32 + S33 * R33;
    NF0 += -0.5 * (M11 * j1 + M12 * j2 + M13 * j3);
    NF1 += -0.5 * (M21 * j1 + M22 * j2 + M23 * j3);
    NF2 += -0.5 * (M31 * j1 + M32 * j2 + M33 * j3);
}

// One face node of the ES flux in Cartesian momentum; returns F* (unscaled).
__device__ __forceinline__ void flux_es(const FluxConsts& c, double nx, double ny, double hL, double huL, double hvL,
                                        double hR, double huR, double hvR, double bL, double bR, double& NF0,
                                        double& NF1, double& NF2) {
    const double g = c.g;
    double uL, vL, uR, vR;
    if (hL >= c.tol) { uL = huL / hL; vL = hvL / hL; } else { hL = 0.0; uL = 0.0; vL = 0.0; }
    if (hR >= c.tol) { uR = huR / hR; vR = hvR / hR; } else { hR = 0.0; uR = 0.0; vR = 0.0; }
    const double avgH = 0.5 * (hL + hR);
    const double avgH2 = 0.5 * (hL * hL + hR * hR);
    const double avgU = 0.5 * (uL + uR);
    const double avgV = 0.5 * (vL + vR);
    const double normV = avgU * nx + avgV * ny;
    NF0 = avgH * normV;
    NF1 = avgH * avgU * normV + (0.5 * g * avgH2) * nx;
    NF2 = avgH * avgV * normV + (0.5 * g * avgH2) * ny;
    if (c.ES) {
        const double cL = g * hL, cR = g * hR;
        const double avgC = sqrt(fabs(0.5 * (cL + cR)));
        const double q1L = g * (hL + bL) - 0.5 * uL * uL - 0.5 * vL * vL;
        const double q1R = g * (hR + bR) - 0.5 * uR * uR - 0.5 * vR * vR;
        const double SL11 = 0.5 / g * fabs(nx * avgU + ny * avgV + avgC);
        const double SL22 = fabs(avgH * (nx * avgU + ny * avgV));
        const double SL33 = 0.5 / g * fabs(nx * avgU + ny * avgV - avgC);
        es_dissipation(avgU + nx * avgC, ny, avgU - nx * avgC, avgV + ny * avgC, nx, avgV - ny * avgC, SL11, SL22, SL33,
                       q1R - q1L, uR - uL, vR - vL, NF0, NF1, NF2);
    }
}

// One face node of the rotated ES flux; returns F* with momentum rotated back.
__device__ __forceinline__ void flux_es_rotated(const FluxConsts& c, double nx, double ny, double hL, double huL,
                                                double hvL, double hR, double huR, double hvR, double bL, double bR,
                                                double& NF0, double& NF1, double& NF2) {
    const double g = c.g;
    double uL, vL, uR, vR;
    if (hL > c.tol) { uL = huL / hL; vL = hvL / hL; } else { uL = 0.0; vL = 0.0; }
    if (hR > c.tol) { uR = huR / hR; vR = hvR / hR; } else { uR = 0.0; vR = 0.0; }
    const double vLn = nx * uL + ny * vL, vLt = -ny * uL + nx * vL;
    const double vRn = nx * uR + ny * vR, vRt = -ny * uR + nx * vR;
    const double avgH = 0.5 * (hL + hR);
    const double avgH2 = 0.5 * (hL * hL + hR * hR);
    const double avgU = 0.5 * (vLn + vRn);
    const double avgV = 0.5 * (vLt + vRt);
    double F0 = avgH * avgU;
    double F1 = avgH * avgU * avgU + (0.5 * g * avgH2);
    double F2 = avgH * avgV * avgU;
    if (c.ES) {
        const double cL = g * hL, cR = g * hR;
        const double avgC = sqrt(fabs(0.5 * (cL + cR)));
        const double q1L = g * (hL + bL) - 0.5 * vLn * vLn - 0.5 * vLt * vLt;
        const double q1R = g * (hR + bR) - 0.5 * vRn * vRn - 0.5 * vRt * vRt;
        const double SL11 = 0.5 / g * fabs(avgU + avgC);
        const double SL22 = fabs(avgH * avgU);
        const double SL33 = 0.5 / g * fabs(avgU - avgC);
        es_dissipation(avgU + avgC, 0.0, avgU - avgC, avgV, 1.0, avgV, SL11, SL22, SL33, q1R - q1L, vRn - vLn,
                       vRt - vLt, F0, F1, F2);
    }
    NF0 = F0;
    NF1 = nx * F1 - ny * F2;
    NF2 = ny * F1 + nx * F2;
}

// Local Lax-Friedrichs: okl/RiemannSolvers/calcLaxFriedrich.okl:17-170.  Dry sides keep h but
// zero momentum and wave speed; the dissipation uses max(c+|u|, c+|v|) over both sides.
__device__ __forceinline__ void flux_lf(const FluxConsts& c, double nx, double ny, double hL, double huL, double hvL,
                                        double hR, double huR, double hvR, double& NF0, double& NF1, double& NF2) {
    const double g = c.g;
    double uL = 0.0, vL = 0.0, cL = 0.0, uR = 0.0, vR = 0.0, cR = 0.0;
    if (hL > c.tol) { uL = huL / hL; vL = hvL / hL; cL = sqrt(g * hL); } else { huL = 0.0; hvL = 0.0; }
    if (hR > c.tol) { uR = huR / hR; vR = hvR / hR; cR = sqrt(g * hR); } else { huR = 0.0; hvR = 0.0; }
    const double F2L = huL * uL + 0.5 * g * hL * hL, F2R = huR * uR + 0.5 * g * hR * hR;
    const double G2L = huL * vL, G2R = huR * vR;
    const double F3L = huL * vL, F3R = huR * vR;
    const double G3L = hvL * vL + 0.5 * g * hL * hL, G3R = hvR * vR + 0.5 * g * hR * hR;
    const double maxL1 = fmax(cL + fabs(uL), cR + fabs(uR));
    const double maxL2 = fmax(cL + fabs(vL), cR + fabs(vR));
    const double lam = fmax(maxL1, maxL2);
    const double FF0 = 0.5 * (huL + huR), FF1 = 0.5 * (F2L + F2R), FF2 = 0.5 * (F3L + F3R);
    const double GG0 = 0.5 * (hvL + hvR), GG1 = 0.5 * (G2L + G2R), GG2 = 0.5 * (G3L + G3R);
    NF0 = (nx * FF0 + ny * GG0) - 0.5 * lam * (hR - hL);
    NF1 = (nx * FF1 + ny * GG1) - 0.5 * lam * (huR - huL);
    NF2 = (nx * FF2 + ny * GG2) - 0.5 * lam * (hvR - hvL);
}

// Roe flux: okl/RiemannSolvers/calcRoeFlux.okl:17-230 (and F_Roe of calcCombinedFluxOcean.okl).
// Dry sides keep h with zero velocity and momentum; |.| of a double is fabs.
__device__ __forceinline__ void flux_roe(const FluxConsts& c, double nx, double ny, double hL, double q2L, double q3L,
                                         double hR, double q2R, double q3R, double& NF0, double& NF1, double& NF2) {
    const double g = c.g, half = 0.5;
    double uL = 0.0, vL = 0.0, huL = 0.0, hvL = 0.0, uR = 0.0, vR = 0.0, huR = 0.0, hvR = 0.0;
    if (hL > c.tol) { uL = q2L / hL; vL = q3L / hL; huL = q2L; hvL = q3L; }
    if (hR > c.tol) { uR = q2R / hR; vR = q3R / hR; huR = q2R; hvR = q3R; }
    const double F1L = huL, F1R = huR, G1L = hvL, G1R = hvR;
    const double F2L = huL * uL + half * g * hL * hL;
    const double F2R = (huR * uR + half * g * hR * hR);
    const double G2L = huL * vL, G2R = huR * vR;
    const double F3L = huL * vL, F3R = huR * vR;
    const double G3L = (hvL * vL + half * g * hL * hL);
    const double G3R = (hvR * vR + half * g * hR * hR);
    const double rh = half * (hL + hR);
    const double ru = (uL * sqrt(hL) + uR * sqrt(hR)) / (sqrt(hL) + sqrt(hR));
    const double rv = (vL * sqrt(hL) + vR * sqrt(hR)) / (sqrt(hL) + sqrt(hR));
    const double rc = sqrt(g * rh);
    const double rn = nx * ru + ny * rv;
    const double l1 = fabs(rn - rc), l2 = fabs(rn), l3 = fabs(rn + rc);
    const double R11 = 1.0, R12 = 0.0, R13 = 1.0;
    const double R21 = ru - rc * nx, R22 = -ny, R23 = ru + rc * nx;
    const double R31 = rv - rc * ny, R32 = nx, R33 = rv + rc * ny;
    const double o2c = half / rc;
    const double L11 = l1 * (rn + rc) * o2c, L12 = l1 * (-nx * o2c), L13 = l1 * (-ny * o2c);
    const double L21 = l2 * (-ny * ru + nx * rv), L22 = l2 * (-ny), L23 = l2 * (nx);
    const double L31 = l3 * (-(rn - rc) * o2c), L32 = l3 * (nx * o2c), L33 = l3 * (ny * o2c);
    const double A11 = R11 * L11 + R12 * L21 + R13 * L31;
    const double A12 = R11 * L12 + R12 * L22 + R13 * L32;
    const double A13 = R11 * L13 + R12 * L23 + R13 * L33;
    const double A21 = R21 * L11 + R22 * L21 + R23 * L31;
    const double A22 = R21 * L12 + R22 * L22 + R23 * L32;
    const double A23 = R21 * L13 + R22 * L23 + R23 * L33;
    const double A31 = R31 * L11 + R32 * L21 + R33 * L31;
    const double A32 = R31 * L12 + R32 * L22 + R33 * L32;
    const double A33 = R31 * L13 + R32 * L23 + R33 * L33;
    const double dH = hR - hL, dHu = huR - huL, dHv = hvR - hvL;
    const double d1 = A11 * dH + A12 * dHu + A13 * dHv;
    const double d2 = A21 * dH + A22 * dHu + A23 * dHv;
    const double d3 = A31 * dH + A32 * dHu + A33 * dHv;
    NF0 = (nx * (half * (F1L + F1R)) + ny * (half * (G1L + G1R))) - half * d1;
    NF1 = (nx * (half * (F2L + F2R)) + ny * (half * (G2L + G2R))) - half * d2;
    NF2 = (nx * (half * (F3L + F3R)) + ny * (half * (G3L + G3R))) - half * d3;
}

// The numerical flux of one face node for NumFlux = kind (src/deviceclass.cpp:442-477).  cpuR is
// EdgeData's right rank (-1 on a boundary face), bL the face's left bottom trace, bR the right one
// at the node read by the flux.  Kind 5 (calcCombinedFluxOcean.okl:330-420) takes the Roe flux on
// boundary faces over the open ocean (bL < 6.7), the rotated ES flux elsewhere.
__device__ __forceinline__ void numerical_flux(int kind, const FluxConsts& c, int cpuR, double nx, double ny,
                                               double hL, double huL, double hvL, double hR, double huR, double hvR,
                                               double bL, double bR, double& F0, double& F1, double& F2) {
    switch (kind) {
        case ESDG_FLUX_ES_ROTATED: flux_es_rotated(c, nx, ny, hL, huL, hvL, hR, huR, hvR, bL, bR, F0, F1, F2); break;
        case ESDG_FLUX_LF: flux_lf(c, nx, ny, hL, huL, hvL, hR, huR, hvR, F0, F1, F2); break;
        case ESDG_FLUX_ROE: flux_roe(c, nx, ny, hL, huL, hvL, hR, huR, hvR, F0, F1, F2); break;
        case ESDG_FLUX_OCEAN:
            if (cpuR == -1 && bL < 6.7) flux_roe(c, nx, ny, hL, huL, hvL, hR, huR, hvR, F0, F1, F2);
            else flux_es_rotated(c, nx, ny, hL, huL, hvL, hR, huR, hvR, bL, bR, F0, F1, F2);
            break;
        default: flux_es(c, nx, ny, hL, huL, hvL, hR, huR, hvR, bL, bR, F0, F1, F2); break;
    }
}

__device__ __forceinline__ int surf_node(int ngl, int side, int k) {
    switch (side) {
        case 0: return k;
        case 2: return ngl * ngl - ngl + k;
        case 1: return k * ngl + ngl - 1;
        default: return k * ngl;
    }
}

struct SensorConsts {
    double ZeroTOL, eps0, sigmaMin, sigmaMax;
};

// One of the four modal energy sums of ShockCapturing.okl:145-196, in the okl's serial order:
// which = 0: zaehler1 (top shell), 1: zaehler2 (second shell), 2: nenner1 (all), 3: nenner2 (below N).
// RS: row stride of the mode array (m[i][j] at m[i * RS + j]).  SQUARED: m holds the squares of
// the modes already (each square is the same rounded product, so the sums are bit-identical).
template <int NGL, int RS = NGL, bool SQUARED = false>
__device__ __forceinline__ double sensor_sum(const double* m, int which) {
    constexpr int N = NGL - 1;
    auto sq = [&](int k) { return SQUARED ? m[k] : m[k] * m[k]; };
    double s = 0.0;
    if (which == 0) {
        for (int i = 0; i < N; ++i) s += sq(i * RS + N) + sq(N * RS + i);
        s += sq(N * RS + N);
    } else if (which == 1) {
        for (int i = 0; i < N - 1; ++i) s += sq(i * RS + N - 1) + sq((N - 1) * RS + i);
        s += sq((N - 1) * RS + N - 1);
    } else if (which == 2) {
        for (int i = 0; i < NGL; ++i)
            for (int j = 0; j < NGL; ++j) s += sq(i * RS + j);
    } else {
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) s += sq(i * RS + j);
    }
    return s;
}

// sigma = log10(max(z1/n1, z2/n2)) and the sine ramp to eps0 (ShockCapturing.okl:198-217)
__device__ __forceinline__ double sensor_ramp(double z1, double z2, double n1, double n2, const SensorConsts& c) {
    const double t1 = (n1 < c.ZeroTOL) ? 0.0 : z1 / n1;
    const double t2 = (n1 < c.ZeroTOL) ? 0.0 : z2 / n2;   // the reference guards term 2 with nenner1
    const double sigma = log10(fmax(t1, t2));
    if (c.sigmaMin >= sigma) return 0.0;
    if (c.sigmaMax >= sigma) {
        const double PI = 3.141592653589793;
        const double delta = 1.0 + sin(PI * (sigma - 0.5 * (c.sigmaMax + c.sigmaMin)) / (c.sigmaMax - c.sigmaMin));
        return c.eps0 * 0.5 * delta;
    }
    return c.eps0;
}

template <int NGL>
__device__ __forceinline__ double sensor_from_modes(const double* m, const SensorConsts& c) {
    return sensor_ramp(sensor_sum<NGL>(m, 0), sensor_sum<NGL>(m, 1), sensor_sum<NGL>(m, 2), sensor_sum<NGL>(m, 3), c);
}

// ShockCapturing.okl:63-217 over the RowBlock<NGL> elements e0 .. e0 + EPB - 1 of a state, one
// thread per element row (eta = r), called by every thread of the block.  The modal coefficients
// m[a][c] = sum_l T[a][l] V[c][l] with T[a][c] = sum_l V[a][l] u(xi = l, eta = c), u = h + b,
// each sum over l in ascending order as in the okl: thread r forms column r of T from its own row
// of u, then row r of m from row r of T; four threads form the energy sums in the okl's order.
// sV holds V (written by the caller before the call; the call's first barrier publishes it), sT
// EPB * ES doubles, sSum EPB x 4.  Returns the element's viscosity on its row-0 thread.
// (QP: a pointer to const double, generic or gptr.)
template <int NGL, class QP>
__device__ __forceinline__ double sensor_rows(int Nelem, int e0, const SensorConsts& sc, QP __restrict__ q,
                                              QP __restrict__ b, const double* sV, double* sT, double (*sSum)[4]) {
    using RB = RowBlock<NGL>;
    constexpr int NN = NGL * NGL;
    const int el = threadIdx.x / NGL, r = threadIdx.x % NGL;
    const int e = e0 + el;
    const bool act = (el < RB::EPB) && (e < Nelem);
    double u[NGL];
    if (act) {
        const int64_t xb = (int64_t)e * NN + r * NGL;
        const int64_t qb = (int64_t)e * NN * 3 + r * NGL;
#pragma unroll
        for (int i = 0; i < NGL; ++i) u[i] = b[xb + i] + q[qb + i];
    }
    __syncthreads();
    double* T = sT + (act ? el : 0) * RB::ES;
    if (act) {
        // column c = r of T
#pragma unroll
        for (int a = 0; a < NGL; ++a) {
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < NGL; ++l) s += sV[a * NGL + l] * u[l];
            T[a * RB::RS + r] = s;
        }
    }
    __syncthreads();
    if (act) {
        // row a = r of m, in place of row r of T (only this thread reads and writes that row); only
        // the squares of the modes enter the sums, so the row's squares are stored (every lane
        // squares its own row, and the four sums below only add)
        double t[NGL];
#pragma unroll
        for (int l = 0; l < NGL; ++l) t[l] = T[r * RB::RS + l];
#pragma unroll
        for (int c = 0; c < NGL; ++c) {
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < NGL; ++l) s += t[l] * sV[c * NGL + l];
            T[r * RB::RS + c] = s * s;
        }
    }
    __syncthreads();
    if (act && r < 4) sSum[el][r] = sensor_sum<NGL, RB::RS, true>(T, r);
    __syncthreads();
    double v = 0.0;
    if (act && r == 0) {
        const double* S4 = sSum[el];
        v = sensor_ramp(S4[0], S4[1], S4[2], S4[3], sc);
    }
    return v;
}

// ---------------------------------------------------------------------------------------------
// Flux-differencing volume term building blocks (okl/DG/VolumeKernelFluxDiffV99.okl:17-57)
// ---------------------------------------------------------------------------------------------

// Symmetric two-point EC flux along one line (Fvol of okl/DG/VolumeKernelFluxDiffV99.okl:17-57 for
// metrics (Y, X) = (Yeta, Xeta); Gvol = -Fvol evaluated with (Yxi, Xxi)).
//
// The okl halves every pair sum (hu_A = (huL + huR)/2, Y_A = (YL + YR)/2, ...).  A product by 1/2,
// 1/4, 1/8, 4 or 8 is exact (no intermediate here comes near the subnormal or overflow range), and
// scaling commutes with rounding, so the halvings can be moved out of the pair:
//   f1 = (1/4) (SY Shu - SX Shv)
//   f2 = (1/8) (SY (Shu Su + P4) - (SX Shv) Su)
//   f3 = (1/8) (SY (Shu Sv) - SX (Shv Sv + P4))
// with S* the plain sums of the two nodes' values and P4 = 4 (h_A - g/2 h2_A) = (g Sh) Sh - g (hL^2 + hR^2).
// Each rounded operation is the okl's rounded operation times a power of two, so f1, f2, f3 are the
// okl's bits.  The sweep keeps its sums scaled by 4 (component 1) and 8 (components 2, 3): every
// term D f is then the okl's term times the same power of two, and so is every partial sum.  The
// caller rescales the finished sums (unscale_sums) -- 40 instead of 48 FP64 operations per pair.
struct PairOut {
    double f1, f2, f3;     // 4 f1, 8 f2, 8 f3 of the okl
};

__device__ __forceinline__ PairOut ec_pair_scaled(double g, double hL, double huL, double hvL, double uL, double vL,
                                                  double YL, double XL, double hR, double huR, double hvR, double uR,
                                                  double vR, double YR, double XR) {
    const double Sh = hL + hR;
    const double P4 = (g * Sh) * Sh - g * (hL * hL + hR * hR);
    const double Shu = huL + huR, Shv = hvL + hvR, Su = uL + uR, Sv = vL + vR, SY = YL + YR, SX = XL + XR;
    const double XV = SX * Shv;
    PairOut o;
    o.f1 = SY * Shu - XV;
    o.f2 = SY * (Shu * Su + P4) - XV * Su;
    o.f3 = SY * (Shu * Sv) - SX * (Shv * Sv + P4);
    return o;
}

template <int NGL>
struct VolCfg {
    static constexpr int EPW = 64 / NGL;        // elements per wave
    static constexpr int NN = NGL * NGL;
    static constexpr int RS = NGL + 1;          // LDS row stride (padding against bank conflicts)
    static constexpr int PLANE = NGL * RS;      // one field of one element in LDS
    // h, hu, hv, 1/h (the row sums reuse three planes), padded so that consecutive elements start
    // 8 doubles (16 banks) apart modulo 16 doubles: the row-wise and the column-wise accesses of
    // the wave's elements then fall into disjoint LDS banks
    static constexpr int ESZ = 4 * PLANE + ((8 - (4 * PLANE) % 16) + 16) % 16;
};

// All pairs of one line of NGL nodes; acc[k] += sgn * D(k,l) F(k,l), l ascending for every k, with
// the sums scaled as ec_pair_scaled's outputs are (4, 8, 8).
template <int NGL, bool NEG>
__device__ __forceinline__ void line_sweep(const double* __restrict__ D, double g,
                                           const double (&h)[NGL], const double (&hu)[NGL], const double (&hv)[NGL],
                                           const double (&u)[NGL], const double (&v)[NGL],
                                           const double (&Y)[NGL], const double (&X)[NGL], double (&a1)[NGL],
                                           double (&a2)[NGL], double (&a3)[NGL]) {
#pragma unroll
    for (int c = 0; c < NGL; ++c) {
#pragma unroll
        for (int l = c; l < NGL; ++l) {
            const PairOut F = ec_pair_scaled(g, h[c], hu[c], hv[c], u[c], v[c], Y[c], X[c], h[l], hu[l], hv[l], u[l],
                                             v[l], Y[l], X[l]);
            const double dcl = NEG ? -D[c * NGL + l] : D[c * NGL + l];
            a1[c] += dcl * F.f1;
            a2[c] += dcl * F.f2;
            a3[c] += dcl * F.f3;
            if (l != c) {
                const double dlc = NEG ? -D[l * NGL + c] : D[l * NGL + c];
                a1[l] += dlc * F.f1;
                a2[l] += dlc * F.f2;
                a3[l] += dlc * F.f3;
            }
        }
    }
}

// The okl's flux-difference sums from line_sweep's scaled ones (exact).
__device__ __forceinline__ void unscale_sums(double& a1, double& a2, double& a3) {
    a1 = 0.25 * a1;
    a2 = 0.125 * a2;
    a3 = 0.125 * a3;
}

}  // namespace esdg
