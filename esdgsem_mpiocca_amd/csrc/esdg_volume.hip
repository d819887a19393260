// esdg_volume.hip -- volume integrals on gfx950.
//
// Flux-differencing volume term (okl/DG/VolumeKernelFluxDiffV99.okl:17-218):
//   Qt(i,j) = -J^-1 [ sum_l D(j,l) F#(u_ij, u_il) + sum_l D(i,l) G#(u_ij, u_lj) ] - g h (0, Bx, By)
// with i the eta row and j the xi column.  The two-point fluxes are symmetric in their two
// states, so each unordered pair of nodes on a line is evaluated once and added to both nodes'
// sums.  The pairs are visited so that every node still accumulates its terms in increasing l,
// as in the reference's serial loop.
//
// Work mapping (wave64, one wave per block): lane = (element el, line r), EPW = 64 / ngl elements
// per wave.  Row phase: the lane owns eta-row r of its element, holds the row's 7 node fields in
// registers (loaded straight from HBM: 8 contiguous doubles per array) and performs all xi-pairs.
// It parks h, hu, hv and 1/h in LDS (u = hu/h and v = hv/h are re-derived bit-identically).  Column phase: the same lane takes xi-column r, reads the column
// from LDS, loads the xi-metrics of that column from HBM (one 64-byte segment per element and row),
// continues the sums of its column nodes with the eta-pairs, and stores Qt.  The row sums cross
// over through LDS.  Per element the kernel reads q, the four metrics, 1/J, Bx, By once and writes
// Qt once: 13 doubles per node, the compulsory traffic.
#include "esdg_device.h"
#include "esdg_fused.h"
#include "esdg_internal.h"

namespace esdg {

// What the flux-differencing kernel does after the volume term of its nodes:
//   VOL_ONLY      store Qt (the reference's VolumeKernel)
//   VOL_SURF      add the surface lift of the face fluxes (okl/DG/SurfaceKernel.okl, DBSurf = 0), store Qt
//   VOL_SURF_RK   as VOL_SURF, then the low-storage RK step gRK = A gRK + Qt, q += cdt gRK in place
//                 (okl/RungeKutta/rkLS.okl + UpdateKernelLS.okl); Qt never reaches HBM
//   VOL_SURF_VISC_RK  as VOL_SURF_RK with the viscous terms of the artificial viscosity added
//                 after the surface lift, in the reference's order: the viscous volume term
//                 (VolumeKernelViscose.okl, precomputed per node; skipped where the element's
//                 viscosity is zero, where it is +-0) and the viscous surface lift
//                 (SurfaceKernelVisc.okl)
enum VolMode { VOL_ONLY = 0, VOL_SURF = 1, VOL_SURF_RK = 2, VOL_SURF_VISC_RK = 3 };

struct SurfIn {
    const int *E2E, *EMS, *EOR;
    const double* FStar;
    double rkA, cdt;
    double* gRK;
    const double* Qv;      // viscous volume term [e][2][ngl^2] (hu, hv)
    const double* visc;    // element viscosity
    const double* SPV;     // viscous face fluxes [f][2][ngl]
    double* tr;            // side traces of the updated state, in face slots (RK modes; may be null)
    const int* slot;       // face slot of each element side
    // Of the updated state, when tr is given (RK modes; each may be null):
    double* visc_next;     // ShockCapturing's element viscosity (and visc_plot, its copy)
    double* visc_plot;
    const double *B, *VdmInv;
    SensorConsts sc;
    double* lam;           // FindLambdaMax's element maximum wave speed
    double tol;            // PosPresTOL (velocities are 0 where h <= tol)
    // VOL_SURF_VISC_RK: elements with active[e] != the stage mark have zero viscosity on
    // themselves and on every face neighbour; their viscous terms are +-0 and are not read
    const int* active;
    // the marked elements (list[0 .. acount[m & 1]), m = acount[2] the stage mark): with them
    // given, VOL_SURF_RK leaves the marked elements out and VOL_SURF_VISC_RK runs over the list
    // only, so that the viscous terms run in wavefronts of marked elements alone
    const int* list;
    const int* acount;
    // the device clock's dt (may be null): the update uses Cs * dt, and a block with dt = 0 (a step
    // issued after T) does nothing
    const double* dtp;
    double Cs;
    // RK modes: the positivity limiter (PosPres) on the updated state before it is stored, with
    // the element sizes, the GLL weights and ZeroTOL; the side traces, sensor and wave speed are
    // then those of the limited state
    int limit;
    const double* EleSizes;
    const double* GLw;
    double zero_tol;
};

// The LDS of one block of the update pass (one wave): the element planes, D (later the sensor's
// VdmInv), and the sensor's and wave speed's reductions.
template <int NGL>
struct VolLds {
    __attribute__((aligned(16))) double s[VolCfg<NGL>::EPW * VolCfg<NGL>::ESZ];
    double sM[VolCfg<NGL>::NN];
    double sSum[VolCfg<NGL>::EPW][4];
    double sLam[64];
};

// The update pass over block blk of its grid (a kernel's whole body: a block that returns is done).
template <int NGL, int MODE>
__device__ __forceinline__ void volume_body(int blk, VolLds<NGL>& L, int Nelem, double g, const double* __restrict__ Jac,
                                                        const double* __restrict__ Yxi, const double* __restrict__ Yeta,
                                                        const double* __restrict__ Xxi, const double* __restrict__ Xeta,
                                                        double* q, const double* __restrict__ Dg,
                                                        const double* __restrict__ Bx, const double* __restrict__ By,
                                                        double* __restrict__ Qt, SurfIn si) {
    using C = VolCfg<NGL>;
    double cdt = si.cdt;
    if constexpr (MODE == VOL_SURF_RK || MODE == VOL_SURF_VISC_RK) {
        if (si.dtp) {
            const double d = *si.dtp;
            if (d == 0.0) return;
            cdt = si.Cs * d;
        }
    }
    double* s = L.s;
    // D in LDS: its entries are read with wave-uniform addresses (broadcast), which keeps the
    // 64-entry matrix out of the scalar registers (no SGPR spills into vector registers).  After
    // the column sweep the same slots take the sensor's VdmInv (the block is one wave, whose LDS
    // accesses execute in program order).
    double* sM = L.sM;
    const int lane = threadIdx.x;
    for (int t = lane; t < C::NN; t += 64) sM[t] = Dg[t];
    const double* D = sM;
    const int el = lane / NGL, r = lane % NGL;
    int e = blk * C::EPW + el;
    bool act = (el < C::EPW) && (e < Nelem);
    int seq = 0;
    if constexpr (MODE == VOL_SURF_RK || MODE == VOL_SURF_VISC_RK) {
        if (si.acount) seq = si.acount[2];
    }
    if constexpr (MODE == VOL_SURF_VISC_RK) {
        if (si.list) {
            // blocks past the end of the list do nothing (the count is the same for every lane)
            const int cnt = min(si.acount[seq & 1], Nelem);
            if (blk * C::EPW >= cnt) return;
            act = (el < C::EPW) && (e < cnt);
            e = act ? si.list[e] : 0;
        }
    } else if constexpr (MODE == VOL_SURF_RK) {
        if (si.list && act && si.active[e] == seq) act = false;
    }
    double* S = s + (act ? el : 0) * C::ESZ;
    // The element's face indices (ElemToEdge, EdgeOrientation, EdgeMasterSlave of its 4 sides) go
    // to padding slots of its LDS planes (column NGL of every row, which no sweep touches; two ints
    // per slot, planes 3, 2, ... in turn), loaded beside the row data so that the surface lift
    // needs no dependent global round trip before its face-flux reads.  The epilogue reuses the
    // slots only after the lift.
    auto fidx = [&](int i) -> int& {
        const int d = i >> 1;
        return reinterpret_cast<int*>(S + (3 - d / NGL) * C::PLANE + (d % NGL) * C::RS + NGL)[i & 1];
    };
    static_assert(4 * NGL * 2 >= 12, "face indices fit the padding slots");
    constexpr int FPL = (4 + NGL - 1) / NGL;   // sides per lane: lane r loads sides r, r + NGL, ...
    int fE2E[FPL], fEOR[FPL], fEMS[FPL];
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
        const int sd = r + j * NGL;
        fE2E[j] = 0; fEOR[j] = 0; fEMS[j] = 0;
        if (MODE != VOL_ONLY && act && sd < 4) {
            fE2E[j] = si.E2E[e * 4 + sd];
            fEOR[j] = si.EOR[e * 4 + sd];
            fEMS[j] = si.EMS[e * 4 + sd];
        }
    }
    double lam = 0.0;   // this column's maximum wave speed of the updated state (si.lam)
    double bsen[NGL];   // b of row r (si.visc_next)

    double h[NGL], hu[NGL], hv[NGL], u[NGL], v[NGL], Y[NGL], X[NGL];
    double a1[NGL], a2[NGL], a3[NGL];
#pragma unroll
    for (int c = 0; c < NGL; ++c) { a1[c] = 0.0; a2[c] = 0.0; a3[c] = 0.0; }

    // ---- row phase: eta-row r, nodes (r, c) -------------------------------------------------
    if (act) {
        const int64_t qb = (int64_t)e * C::NN * ESDG_NEQ + r * NGL;
        const int64_t xb = (int64_t)e * C::NN + r * NGL;
#pragma unroll
        for (int c = 0; c < NGL; ++c) {
            h[c] = q[qb + c];
            hu[c] = q[qb + C::NN + c];
            hv[c] = q[qb + 2 * C::NN + c];
            Y[c] = Yeta[xb + c];
            X[c] = Xeta[xb + c];
        }
#pragma unroll
        for (int c = 0; c < NGL; ++c) {
            const double inv = (h[c] > 0) ? 1.0 / h[c] : 0.0;
            u[c] = hu[c] * inv;
            v[c] = hv[c] * inv;
            S[0 * C::PLANE + r * C::RS + c] = h[c];
            S[1 * C::PLANE + r * C::RS + c] = hu[c];
            S[2 * C::PLANE + r * C::RS + c] = hv[c];
            S[3 * C::PLANE + r * C::RS + c] = inv;
        }
#pragma unroll
        for (int j = 0; j < FPL; ++j) {
            const int sd = r + j * NGL;
            if (MODE != VOL_ONLY && sd < 4) {
                fidx(3 * sd) = fE2E[j];
                fidx(3 * sd + 1) = fEOR[j];
                fidx(3 * sd + 2) = fEMS[j];
            }
        }
    }
    __syncthreads();   // sD written
    if (act) line_sweep<NGL, false>(D, g, h, hu, hv, u, v, Y, X, a1, a2, a3);
    __syncthreads();

    // ---- column phase: xi-column r, nodes (k, r) -------------------------------------------
    double b1[NGL], b2[NGL], b3[NGL];
    double rJ[NGL];
    if (act) {
        const int64_t xb = (int64_t)e * C::NN + r;
#pragma unroll
        for (int k = 0; k < NGL; ++k) {
            h[k] = S[0 * C::PLANE + k * C::RS + r];
            hu[k] = S[1 * C::PLANE + k * C::RS + r];
            hv[k] = S[2 * C::PLANE + k * C::RS + r];
            const double inv = S[3 * C::PLANE + k * C::RS + r];
            u[k] = hu[k] * inv;
            v[k] = hv[k] * inv;
            Y[k] = Yxi[xb + k * NGL];
            X[k] = Xxi[xb + k * NGL];
        }
    }
    __syncthreads();
    if (act) {
        // hand the row sums over: node (r, c) of the row phase -> LDS plane slot [r][c]
#pragma unroll
        for (int c = 0; c < NGL; ++c) {
            S[0 * C::PLANE + r * C::RS + c] = a1[c];
            S[1 * C::PLANE + r * C::RS + c] = a2[c];
            S[2 * C::PLANE + r * C::RS + c] = a3[c];
        }
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int k = 0; k < NGL; ++k) {
            b1[k] = S[0 * C::PLANE + k * C::RS + r];
            b2[k] = S[1 * C::PLANE + k * C::RS + r];
            b3[k] = S[2 * C::PLANE + k * C::RS + r];
        }
    }
    // The RK register gRK of the wave's elements is copied into planes 0-2 (free once the row sums
    // are read) by direct global-to-LDS loads, issued before the column sweep so that the update
    // after it finds gRK in LDS instead of waiting for HBM.  An element's gRK is 3 NN contiguous
    // doubles in HBM and lands contiguously at the start of its LDS area ([v][k NGL + r]); 16-byte
    // loads need that block 16-byte aligned, and the face indices must stay clear of it (in the
    // padding slots of plane 3): NGL 8 and 10 (at NGL 6 the copies' addressing costs the instance
    // 50 VGPRs and a wave per SIMD).
    constexpr int GBYTES = 3 * C::NN * 8;
    constexpr bool kDma = (MODE == VOL_SURF_RK || MODE == VOL_SURF_VISC_RK) && NGL >= 8 && GBYTES % 16 == 0;
    if constexpr (kDma) {
        // the row sums' LDS reads complete before the copies overwrite those slots
        __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
        const unsigned long long am = __ballot(act);
#pragma unroll
        for (int sl = 0; sl < C::EPW; ++sl) {
            if (!((am >> (sl * NGL)) & 1ull)) continue;   // the same for every lane
            const int es = __builtin_amdgcn_readlane(e, sl * NGL);
            const char* src = reinterpret_cast<const char*>(si.gRK + (int64_t)es * 3 * C::NN);
            char* dst = reinterpret_cast<char*>(s + sl * C::ESZ);
#pragma unroll
            for (int c = 0; c * 1024 < GBYTES; ++c)
                if (c * 1024 + lane * 16 < GBYTES)
                    __builtin_amdgcn_global_load_lds(src + c * 1024 + lane * 16,
                                                     (__attribute__((address_space(3))) void*)(dst + c * 1024), 16, 0, 0);
        }
    }
    if (act) {
        line_sweep<NGL, true>(D, g, h, hu, hv, u, v, Y, X, b1, b2, b3);
        const int64_t xb = (int64_t)e * C::NN + r;
        const int64_t ob = (int64_t)e * C::NN * ESDG_NEQ + r;
#pragma unroll
        for (int k = 0; k < NGL; ++k) {
            unscale_sums(b1[k], b2[k], b3[k]);
            rJ[k] = Jac[xb + k * NGL];
            b1[k] = -rJ[k] * b1[k];
            b2[k] = -rJ[k] * b2[k] - g * h[k] * Bx[xb + k * NGL];
            b3[k] = -rJ[k] * b3[k] - g * h[k] * By[xb + k * NGL];
        }
        if (MODE != VOL_ONLY) {
            // surface lift: bottom/top sides first (nodes k = 0, N of this column, trace index r),
            // then right/left (the whole column r = N or 0, trace index k).  A side's face,
            // orientation and sign are the same for all its nodes: read once, so the face-flux
            // loads of a side do not wait for one another.
            const int fa0 = fidx(0) - 1, rv0 = fidx(1), ms0 = fidx(2);
            const int fa2 = fidx(6) - 1, rv2 = fidx(7), ms2 = fidx(8);
            const int sc = (r == NGL - 1) ? 1 : 3;   // the column side of lanes r = N and r = 0
            const int fac = fidx(3 * sc) - 1, rvc = fidx(3 * sc + 1), msc = fidx(3 * sc + 2);
            const int64_t l0 = (int64_t)fa0 * NGL * ESDG_NEQ + rv0 * r + (1 - rv0) * (NGL - 1 - r);
            const int64_t l2 = (int64_t)fa2 * NGL * ESDG_NEQ + rv2 * r + (1 - rv2) * (NGL - 1 - r);
            b1[0] += -ms0 * si.FStar[l0] * rJ[0];
            b2[0] += -ms0 * rJ[0] * si.FStar[l0 + NGL];
            b3[0] += -ms0 * rJ[0] * si.FStar[l0 + 2 * NGL];
            b1[NGL - 1] += -ms2 * si.FStar[l2] * rJ[NGL - 1];
            b2[NGL - 1] += -ms2 * rJ[NGL - 1] * si.FStar[l2 + NGL];
            b3[NGL - 1] += -ms2 * rJ[NGL - 1] * si.FStar[l2 + 2 * NGL];
            if (r == 0 || r == NGL - 1) {
                const int64_t fc = (int64_t)fac * NGL * ESDG_NEQ;
#pragma unroll
                for (int k = 0; k < NGL; ++k) {
                    const int64_t fl = fc + rvc * k + (1 - rvc) * (NGL - 1 - k);
                    b1[k] += -msc * si.FStar[fl] * rJ[k];
                    b2[k] += -msc * rJ[k] * si.FStar[fl + NGL];
                    b3[k] += -msc * rJ[k] * si.FStar[fl + 2 * NGL];
                }
            }
        }
        if (MODE == VOL_SURF_VISC_RK && si.active[e] == seq) {
            if (si.visc[e] != 0.0) {
                const double* qv = si.Qv + (int64_t)e * C::NN * 2 + r;
#pragma unroll
                for (int k = 0; k < NGL; ++k) {
                    b2[k] += qv[k * NGL];
                    b3[k] += qv[C::NN + k * NGL];
                }
            }
            // the viscous face fluxes, each side's face, orientation and sign read once (as in the
            // surface lift above)
            {
                const int fa0 = fidx(0) - 1, rv0 = fidx(1), ms0 = fidx(2);
                const int fa2 = fidx(6) - 1, rv2 = fidx(7), ms2 = fidx(8);
                const int sc = (r == NGL - 1) ? 1 : 3;
                const int fac = fidx(3 * sc) - 1, rvc = fidx(3 * sc + 1), msc = fidx(3 * sc + 2);
                const int64_t l0 = (int64_t)fa0 * NGL * 2 + rv0 * r + (1 - rv0) * (NGL - 1 - r);
                const int64_t l2 = (int64_t)fa2 * NGL * 2 + rv2 * r + (1 - rv2) * (NGL - 1 - r);
                b2[0] += -ms0 * rJ[0] * si.SPV[l0];
                b3[0] += -ms0 * rJ[0] * si.SPV[l0 + NGL];
                b2[NGL - 1] += -ms2 * rJ[NGL - 1] * si.SPV[l2];
                b3[NGL - 1] += -ms2 * rJ[NGL - 1] * si.SPV[l2 + NGL];
                if (r == 0 || r == NGL - 1) {
                    const int64_t fc = (int64_t)fac * NGL * 2;
#pragma unroll
                    for (int k = 0; k < NGL; ++k) {
                        const int64_t fl = fc + rvc * k + (1 - rvc) * (NGL - 1 - k);
                        b2[k] += -msc * rJ[k] * si.SPV[fl];
                        b3[k] += -msc * rJ[k] * si.SPV[fl + NGL];
                    }
                }
            }
        }
        if (MODE == VOL_SURF_RK || MODE == VOL_SURF_VISC_RK) {
            // gRK of this column, from LDS when copied there above (all of it read before the
            // planes take the new state)
            double gk[3][NGL];
            if constexpr (kDma) {
                __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the copies have landed
#pragma unroll
                for (int k = 0; k < NGL; ++k) {
                    gk[0][k] = S[0 * C::NN + k * NGL + r];
                    gk[1][k] = S[1 * C::NN + k * NGL + r];
                    gk[2][k] = S[2 * C::NN + k * NGL + r];
                }
            }
            // b of row r for the sensor below, loaded here so that its latency overlaps the update
            if (si.visc_next) {
                const int64_t bb = (int64_t)e * C::NN + r * NGL;
#pragma unroll
                for (int i = 0; i < NGL; ++i) bsen[i] = si.B[bb + i];
            }
            // the limiter's GLL weights, all loaded before the loop's stores (which the compiler
            // cannot tell apart from them)
            double gw[NGL];
            double gx = 0.0;
            if (si.limit) {
#pragma unroll
                for (int k = 0; k < NGL; ++k) gw[k] = si.GLw[k];
                gx = gw[r];
            }
            double mn = 0.0;
#pragma unroll
            for (int k = 0; k < NGL; ++k) {
                const int64_t id = ob + k * NGL;
                double g1, g2, g3;
                if constexpr (kDma) {
                    g1 = gk[0][k]; g2 = gk[1][k]; g3 = gk[2][k];
                } else {
                    g1 = si.gRK[id]; g2 = si.gRK[id + C::NN]; g3 = si.gRK[id + 2 * C::NN];
                }
                g1 = si.rkA * g1 + b1[k];
                g2 = si.rkA * g2 + b2[k];
                g3 = si.rkA * g3 + b3[k];
                si.gRK[id] = g1;
                si.gRK[id + C::NN] = g2;
                si.gRK[id + 2 * C::NN] = g3;
                // the stage's state at node (k, r) is this lane's column input (h, hu, hv): the same
                // bits as q there, without a second read of q from HBM
                const double n1 = h[k] + cdt * g1;
                const double n2 = hu[k] + cdt * g2;
                const double n3 = hv[k] + cdt * g3;
                if (si.limit) {
                    // to the limiter below: the new state to planes 0-2, its weights
                    // w = GLw[xi] GLw[eta] / Jac to plane 3, the column minimum of h
                    S[0 * C::PLANE + k * C::RS + r] = n1;
                    S[1 * C::PLANE + k * C::RS + r] = n2;
                    S[2 * C::PLANE + k * C::RS + r] = n3;
                    S[3 * C::PLANE + k * C::RS + r] = gx * gw[k] / rJ[k];
                    mn = fmin(mn, n1);
                    continue;
                }
                q[id] = n1;
                q[id + C::NN] = n2;
                q[id + 2 * C::NN] = n3;
                // the new state also goes to this column of the element's LDS planes (each lane
                // writes only the slots it read the row sums from), for the side traces below
                if (si.tr) {
                    S[0 * C::PLANE + k * C::RS + r] = n1;
                    S[1 * C::PLANE + k * C::RS + r] = n2;
                    S[2 * C::PLANE + k * C::RS + r] = n3;
                }
                // FindLambdaMax.okl:41-60 at this node (the element maximum is order-free)
                if (si.lam) {
                    const double Qinv = (n1 > si.tol) ? 1.0 / n1 : 0.0;
                    const double c = sqrt(g * n1);
                    lam = fmax(lam, c + fmax(fabs(n2 * Qinv), fabs(n3 * Qinv)));
                }
            }
            // the column minimum in the padding slot of row r of plane 3
            if (si.limit) S[3 * C::PLANE + r * C::RS + NGL] = mn;
        } else {
#pragma unroll
            for (int k = 0; k < NGL; ++k) {
                Qt[ob + k * NGL] = b1[k];
                Qt[ob + C::NN + k * NGL] = b2[k];
                Qt[ob + 2 * C::NN + k * NGL] = b3[k];
            }
        }
    }
    if constexpr (MODE == VOL_SURF_RK || MODE == VOL_SURF_VISC_RK) {
        if (si.limit) {
            // PosPres (okl/Positivity/PosPres.okl:17-223) on the new state before it is stored,
            // with the arithmetic of k_pos_pres: lane v < 3 of an element sums the products
            // h w (hu w, hv w) of plane v and plane 3 in the okl's serial order (xi outer, eta
            // inner) into the padding slot of row 0 of plane v; then every lane limits its column
            __syncthreads();
            if (act && r < 3) {
                const double* P = S + r * C::PLANE;
                const double* W = S + 3 * C::PLANE;
                double a = 0.0;
                for (int l = 0; l < NGL; ++l)
                    for (int k = 0; k < NGL; ++k) a += P[k * C::RS + l] * W[k * C::RS + l];
                S[r * C::PLANE + NGL] = a;
            }
            __syncthreads();
            if (act) {
                double hmin = 0.0;
#pragma unroll
                for (int k = 0; k < NGL; ++k) hmin = fmin(hmin, S[3 * C::PLANE + k * C::RS + NGL]);
                const double size = si.EleSizes[e];
                double ha = S[0 * C::PLANE + NGL] / size;
                if (fabs(ha) < si.zero_tol) ha = 0.0;
                const double hua = S[1 * C::PLANE + NGL] / size, hva = S[2 * C::PLANE + NGL] / size;
                const double th = (ha == hmin) ? 1.0 : fmin(1.0, ha / (ha - hmin));
                const int64_t ob = (int64_t)e * C::NN * ESDG_NEQ + r;
#pragma unroll
                for (int k = 0; k < NGL; ++k) {
                    double n1 = S[0 * C::PLANE + k * C::RS + r];
                    double n2 = S[1 * C::PLANE + k * C::RS + r];
                    double n3 = S[2 * C::PLANE + k * C::RS + r];
                    if (hmin < 0.0) {
                        double hn = 0.0, hun = 0.0, hvn = 0.0;
                        if (ha > 0.0) {
                            hn = fabs(th * (n1 - ha) + ha);
                            if (hn > si.tol) {
                                hun = th * (n2 - hua) + hua;
                                hvn = th * (n3 - hva) + hva;
                            }
                        }
                        if (ha < 0.0) { hn = -INFINITY; hun = 0.0; hvn = 0.0; }
                        n1 = hn; n2 = hun; n3 = hvn;
                    } else if (n1 < si.tol) {
                        n2 = 0.0; n3 = 0.0;
                    }
                    const int64_t id = ob + k * NGL;
                    q[id] = n1;
                    q[id + C::NN] = n2;
                    q[id + 2 * C::NN] = n3;
                    // the limited state in this lane's column of the planes, for the side traces
                    S[0 * C::PLANE + k * C::RS + r] = n1;
                    S[1 * C::PLANE + k * C::RS + r] = n2;
                    S[2 * C::PLANE + k * C::RS + r] = n3;
                    if (si.lam) {
                        const double Qinv = (n1 > si.tol) ? 1.0 / n1 : 0.0;
                        const double c = sqrt(g * n1);
                        lam = fmax(lam, c + fmax(fabs(n2 * Qinv), fabs(n3 * Qinv)));
                    }
                }
            }
        }
        double* sV = sM;
        if (si.visc_next)
            for (int t = lane; t < C::NN; t += 64) sV[t] = si.VdmInv[t];
        // side traces of the new state, in the face slots of the element sides (read by the next
        // stage's face pass)
        if (si.tr) {
            __syncthreads();
            // lane (el, r) writes side node r of every (side, v) plane of its element into the
            // side's face slot: one unrolled iteration per plane, so the LDS source of each store
            // is fixed at compile time
            if (act) {
#pragma unroll
                for (int side = 0; side < 4; ++side) {
                    double* T = si.tr + (int64_t)si.slot[e * 4 + side] * 3 * NGL + r;
#pragma unroll
                    for (int v = 0; v < 3; ++v) {
                        const int row = (side == 0) ? 0 : (side == 2) ? NGL - 1 : r;
                        const int col = (side == 1) ? NGL - 1 : (side == 3) ? 0 : r;
                        T[v * NGL] = S[v * C::PLANE + row * C::RS + col];
                    }
                }
            }
            // ShockCapturing.okl:63-217 on the updated state, the next stage's viscosity, with the
            // arithmetic of k_shock_capturing: lane r forms column r of T = V (h + b) from row r of
            // h + b, then row r of the squared modes (T V^T)^2 in place; four lanes form the
            // energy sums.  Plane 3 (1/h) is free after the column phase and holds T and then the
            // squares.  The side traces above are the last reader of planes 0-2.
            if (si.visc_next) {
                double (*sSum)[4] = L.sSum;
                double* Tm = S + 3 * C::PLANE;
                const double* V = sV;
                if (act) {
                    double w[NGL];
#pragma unroll
                    for (int i = 0; i < NGL; ++i) w[i] = bsen[i] + S[0 * C::PLANE + r * C::RS + i];
#pragma unroll
                    for (int a = 0; a < NGL; ++a) {
                        double acc = 0.0;
#pragma unroll
                        for (int l = 0; l < NGL; ++l) acc += V[a * NGL + l] * w[l];
                        Tm[a * C::RS + r] = acc;
                    }
                }
                __syncthreads();
                if (act) {
                    double t[NGL];
#pragma unroll
                    for (int l = 0; l < NGL; ++l) t[l] = Tm[r * C::RS + l];
                    // the squares of the modes, one row per lane (only they enter the sums); then
                    // the four sums add them in the okl's order
#pragma unroll
                    for (int c = 0; c < NGL; ++c) {
                        double acc = 0.0;
#pragma unroll
                        for (int l = 0; l < NGL; ++l) acc += t[l] * V[c * NGL + l];
                        Tm[r * C::RS + c] = acc * acc;
                    }
                }
                __syncthreads();
                if (act && r < 4) sSum[el][r] = sensor_sum<NGL, C::RS, true>(Tm, r);
                __syncthreads();
                if (act && r == 0) {
                    const double v = sensor_ramp(sSum[el][0], sSum[el][1], sSum[el][2], sSum[el][3], si.sc);
                    si.visc_next[e] = v;
                    if (si.visc_plot) si.visc_plot[e] = v;
                }
            }
            if (si.lam) {
                double* sLam = L.sLam;
                sLam[lane] = lam;
                __syncthreads();
                if (act && r == 0) {
                    double m = 0.0;
#pragma unroll
                    for (int k = 0; k < NGL; ++k) m = fmax(m, sLam[el * NGL + k]);
                    si.lam[e] = m;
                }
            }
        }
    }
}

template <int NGL, int MODE>
__global__ void __launch_bounds__(64, 2) k_volume_fluxdiff(int Nelem, double g, const double* __restrict__ Jac,
                                                        const double* __restrict__ Yxi, const double* __restrict__ Yeta,
                                                        const double* __restrict__ Xxi, const double* __restrict__ Xeta,
                                                        double* q, const double* __restrict__ Dg,
                                                        const double* __restrict__ Bx, const double* __restrict__ By,
                                                        double* __restrict__ Qt, SurfIn si) {
    __shared__ VolLds<NGL> L;
    volume_body<NGL, MODE>(blockIdx.x, L, Nelem, g, Jac, Yxi, Yeta, Xxi, Xeta, q, Dg, Bx, By, Qt, si);
}

// Both update launches of a stage with artificial viscosity as one: blocks [0, nl) take the listed
// marked elements (VOL_SURF_VISC_RK), the others the unmarked elements (VOL_SURF_RK).  The list's
// blocks come first, so its few working blocks start with the launch; its blocks past the list's
// end exit at once.  One launch instead of two on two streams: no fork and join per stage.
template <int NGL>
__global__ void __launch_bounds__(64, 2) k_volume_split(int nl, int Nelem, double g, const double* __restrict__ Jac,
                                                        const double* __restrict__ Yxi, const double* __restrict__ Yeta,
                                                        const double* __restrict__ Xxi, const double* __restrict__ Xeta,
                                                        double* q, const double* __restrict__ Dg,
                                                        const double* __restrict__ Bx, const double* __restrict__ By,
                                                        double* __restrict__ Qt, SurfIn si) {
    __shared__ VolLds<NGL> L;
    if ((int)blockIdx.x < nl)
        volume_body<NGL, VOL_SURF_VISC_RK>(blockIdx.x, L, Nelem, g, Jac, Yxi, Yeta, Xxi, Xeta, q, Dg, Bx, By, Qt, si);
    else
        volume_body<NGL, VOL_SURF_RK>((int)blockIdx.x - nl, L, Nelem, g, Jac, Yxi, Yeta, Xxi, Xeta, q, Dg, Bx, By, Qt, si);
}

// Standard strong-form volume term (okl/DG/VolumeKernelV99.okl:17-167), FluxDifferencing = 0.
// Thread slot n = i*NGL + j is node (eta row i, xi column j) as in the okl.
template <int NGL>
__global__ void __launch_bounds__(256) k_volume_std(int Nelem, double g, double tol, const double* __restrict__ Jac,
                                                    const double* __restrict__ Yxi, const double* __restrict__ Yeta,
                                                    const double* __restrict__ Xxi, const double* __restrict__ Xeta,
                                                    const double* __restrict__ q, const double* __restrict__ D,
                                                    const double* __restrict__ Bx, const double* __restrict__ By,
                                                    double* __restrict__ Qt) {
    constexpr int NN = NGL * NGL;
    constexpr int EPB = (256 / NN) > 0 ? 256 / NN : 1;
    __shared__ double sD[NN];
    __shared__ double sF[3][EPB * NN], sG[3][EPB * NN];
    const int el = threadIdx.x / NN, n = threadIdx.x % NN;
    const int e = blockIdx.x * EPB + el;
    if (threadIdx.x < NN) sD[threadIdx.x] = D[threadIdx.x];
    const int64_t xid = (int64_t)e * NN + n;
    const int64_t fid = (int64_t)e * NN * ESDG_NEQ + n;
    double h0 = 0.0;
    if (e < Nelem) {
        h0 = q[fid];
        double h = h0, hu = 0.0, hv = 0.0, inv = 0.0;
        if (h > tol) { hu = q[fid + NN]; hv = q[fid + 2 * NN]; inv = 1.0 / h; }
        else h = 0.0;
        const double half_g = 0.5 * g;
        const double F1 = hu, F2 = hu * hu * inv + half_g * h * h, F3 = hv * hu * inv;
        const double G1 = hv, G2 = hu * hv * inv, G3 = hv * hv * inv + half_g * h * h;
        const double ye = Yeta[xid], yx = Yxi[xid], xe = Xeta[xid], xx = Xxi[xid];
        sF[0][threadIdx.x] = ye * F1 - xe * G1;
        sG[0][threadIdx.x] = -yx * F1 + xx * G1;
        sF[1][threadIdx.x] = ye * F2 - xe * G2;
        sG[1][threadIdx.x] = -yx * F2 + xx * G2;
        sF[2][threadIdx.x] = ye * F3 - xe * G3;
        sG[2][threadIdx.x] = -yx * F3 + xx * G3;
    }
    __syncthreads();
    if (e >= Nelem) return;
    const int i = n / NGL, j = n % NGL;
    const int base = el * NN;
    double d1 = 0.0, d2 = 0.0, d3 = 0.0;
#pragma unroll
    for (int l = 0; l < NGL; ++l) {
        const double dil = sD[i * NGL + l], djl = sD[j * NGL + l];
        d1 += djl * sF[0][base + i * NGL + l] + dil * sG[0][base + l * NGL + j];
        d2 += djl * sF[1][base + i * NGL + l] + dil * sG[1][base + l * NGL + j];
        d3 += djl * sF[2][base + i * NGL + l] + dil * sG[2][base + l * NGL + j];
    }
    const double rJ = Jac[xid];
    Qt[fid] = -rJ * d1;
    Qt[fid + NN] = -rJ * d2 - g * h0 * Bx[xid];
    Qt[fid + 2 * NN] = -rJ * d3 - g * h0 * By[xid];
}

template <int NGL>
void launch_volume_fluxdiff(esdg_ctx* ctx, int Nelem, const double* Jac, const double* Yxi, const double* Yeta,
                            const double* Xxi, const double* Xeta, const double* q, const double* D, const double* Bx,
                            const double* By, double* Qt) {
    using C = VolCfg<NGL>;
    SurfIn si{};
    k_volume_fluxdiff<NGL, VOL_ONLY><<<blocks_for(Nelem, C::EPW), 64, 0, ctx->stream>>>(
        Nelem, ctx->p.g, Jac, Yxi, Yeta, Xxi, Xeta, const_cast<double*>(q), D, Bx, By, Qt, si);
}

template <int NGL>
void launch_volume_fused(esdg_ctx* ctx, const FusedArgs& a, bool rk, double rkA, double cdt, const double* fstar,
                         double* Qt, const double* Qv, const double* SPV, double* tr, const int* slot,
                         const int* active, const int* list, const int* acount) {
    using C = VolCfg<NGL>;
    SurfIn si{a.ElemToEdge, a.EMS, a.EOR, fstar, rkA, cdt, a.gRK, Qv, a.ViscPara, SPV, rk ? tr : nullptr, slot};
    si.active = active;
    si.list = list;
    si.acount = acount;
    si.dtp = a.dt_dev;
    si.Cs = a.C_stage;
    if (rk) {
        // PosPres on the new state before it is stored (okl/Positivity/PosPres.okl)
        si.tol = ctx->p.PosPresTOL;
        si.limit = a.positivity;
        si.EleSizes = a.EleSizes;
        si.GLw = a.GLw;
        si.zero_tol = ctx->p.ZeroTOL;
    }
    if (rk && tr) {
        // the updated state's viscosity and wave speed, for the next stage or time step
        si.visc_next = a.next_visc;
        si.visc_plot = a.next_visc ? a.next_visc_plot : nullptr;
        si.B = a.B;
        si.VdmInv = a.VdmInv;
        si.sc = SensorConsts{ctx->p.ZeroTOL, ctx->p.eps0, ctx->p.sigmaMin, ctx->p.sigmaMax};
        si.lam = a.next_lambda;
    }
    if (rk && Qv) {
        // the listed marked elements and the unmarked ones (disjoint sets, each updated once;
        // nothing else is read or written by both) in one launch
        // (at ngl = 10 the two bodies in one kernel spill far more than each alone: two launches
        // there, in order)
        const unsigned nb = blocks_for(a.Nelem, C::EPW);
        if constexpr (NGL <= 9) {
            k_volume_split<NGL><<<2 * nb, 64, 0, ctx->stream>>>(
                (int)nb, a.Nelem, ctx->p.g, a.Jac, a.Yxi, a.Yeta, a.Xxi, a.Xeta, a.q, a.D, a.Bx, a.By, nullptr, si);
        } else {
            k_volume_fluxdiff<NGL, VOL_SURF_VISC_RK><<<nb, 64, 0, ctx->stream>>>(
                a.Nelem, ctx->p.g, a.Jac, a.Yxi, a.Yeta, a.Xxi, a.Xeta, a.q, a.D, a.Bx, a.By, nullptr, si);
            k_volume_fluxdiff<NGL, VOL_SURF_RK><<<nb, 64, 0, ctx->stream>>>(
                a.Nelem, ctx->p.g, a.Jac, a.Yxi, a.Yeta, a.Xxi, a.Xeta, a.q, a.D, a.Bx, a.By, nullptr, si);
        }
    } else if (rk)
        k_volume_fluxdiff<NGL, VOL_SURF_RK><<<blocks_for(a.Nelem, C::EPW), 64, 0, ctx->stream>>>(
            a.Nelem, ctx->p.g, a.Jac, a.Yxi, a.Yeta, a.Xxi, a.Xeta, a.q, a.D, a.Bx, a.By, nullptr, si);
    else
        k_volume_fluxdiff<NGL, VOL_SURF><<<blocks_for(a.Nelem, C::EPW), 64, 0, ctx->stream>>>(
            a.Nelem, ctx->p.g, a.Jac, a.Yxi, a.Yeta, a.Xxi, a.Xeta, a.q, a.D, a.Bx, a.By, Qt, si);
}

// Fused volume + surface lift (+ viscous terms when Qv is given) (+ RK update in place when rk,
// with the side traces of the new state written to tr when given): called by the fused stage.
int fused_volume(esdg_ctx* ctx, const FusedArgs& a, bool rk, double rkA, double cdt, const double* fstar, double* Qt,
                 const double* Qv, const double* SPV, double* tr, const int* slot, const int* active,
                 const int* list, const int* acount) {
    if (Qv && !(active && list && acount))
        return set_error(ctx, ESDG_ERR_STATE, "fused_volume: viscous terms need the stage marks and element list");
    KernelScope ks(ctx, "fused_volume");
    if (tr && !slot) return set_error(ctx, ESDG_ERR_STATE, "fused_volume: side traces need the face slots");
    ESDG_NGL_DISPATCH(ctx->p.ngl, launch_volume_fused, ctx, a, rk, rkA, cdt, fstar, Qt, Qv, SPV, tr, slot, active,
                      list, acount);
    return launch_check(ctx, "fused_volume");
}

template <int NGL>
void launch_volume_std(esdg_ctx* ctx, int Nelem, const double* Jac, const double* Yxi, const double* Yeta,
                       const double* Xxi, const double* Xeta, const double* q, const double* D, const double* Bx,
                       const double* By, double* Qt) {
    constexpr int NN = NGL * NGL;
    constexpr int EPB = (256 / NN) > 0 ? 256 / NN : 1;
    k_volume_std<NGL><<<blocks_for(Nelem, EPB), EPB * NN, 0, ctx->stream>>>(Nelem, ctx->p.g, ctx->p.PosPresTOL, Jac, Yxi,
                                                                            Yeta, Xxi, Xeta, q, D, Bx, By, Qt);
}

}  // namespace esdg

using namespace esdg;

extern "C" {

int esdg_volume_kernel(esdg_ctx* ctx, int Nelem, const double* Jac, const double* Yxi, const double* Yeta,
                       const double* Xxi, const double* Xeta, const double* q, const double* D, const double* Bx,
                       const double* By, double* Qt) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem >= 0 && Jac && Yxi && Yeta && Xxi && Xeta && q && D && Bx && By && Qt,
                 "esdg_volume_kernel: bad argument");
    if (Nelem == 0) return ESDG_OK;
    KernelScope ks(ctx, "volume_kernel");
    ESDG_NGL_DISPATCH(ctx->p.ngl, launch_volume_fluxdiff, ctx, Nelem, Jac, Yxi, Yeta, Xxi, Xeta, q, D, Bx, By, Qt);
    return launch_check(ctx, "volume_kernel");
}

int esdg_volume_kernel_std(esdg_ctx* ctx, int Nelem, const double* Jac, const double* Yxi, const double* Yeta,
                           const double* Xxi, const double* Xeta, const double* q, const double* D, const double* Bx,
                           const double* By, double* Qt) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem >= 0 && Jac && Yxi && Yeta && Xxi && Xeta && q && D && Bx && By && Qt,
                 "esdg_volume_kernel_std: bad argument");
    if (Nelem == 0) return ESDG_OK;
    KernelScope ks(ctx, "volume_kernel_std");
    ESDG_NGL_DISPATCH(ctx->p.ngl, launch_volume_std, ctx, Nelem, Jac, Yxi, Yeta, Xxi, Xeta, q, D, Bx, By, Qt);
    return launch_check(ctx, "volume_kernel_std");
}

}  // extern "C"
