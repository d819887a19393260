// esdg_mlp.hip -- learned troubled-cell indicator (include/esdg_mlp.h): element features, BF16
// MFMA inference with FP32 accumulation, and the training kernels (forward, backward, Adam).
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "esdg_device.h"
#include "esdg_internal.h"
#include "esdg_mlp.h"
#include "esdg_mlp_internal.h"
#include "esdg_stage.h"

namespace esdg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int NIN = ESDG_MLP_NIN;        // 72
constexpr int NH = ESDG_MLP_NHID;        // 32
constexpr int KP = 96;                   // layer-1 K padded to three 32-wide MFMA steps
constexpr int OFF_W1 = 0, OFF_B1 = NIN * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH;
static_assert(OFF_B3 + 1 == ESDG_MLP_NPARAM, "parameter layout");

// round to nearest even BF16 (finite inputs)
__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ inline float rbf(float f) { return bf2f(f2bf(f)); }

// Element features with one thread per element row (eta = r) and FEPB = 32 elements per
// 256-thread block, so that many elements are in flight per CU.  Thread (el, r) returns the
// normalised depths x[8r + i] (i < 8) in xn[] and the modal-shell feature of shell r in *modal:
//   x[n] = (h_n - mean h) / max |h - mean h| (0 for a constant element),
//   modal[k] = max(log10(E_k / E), -16) / 8 (-2 when E = 0), E_k the energy of the modal
//   coefficients m = V u V^T of u = h + b with max(index) = k, E their total.
constexpr int FEPB = 32;
constexpr int FRS = 9;            // padded row stride of an element's 8x8 LDS tile
constexpr int FES = 8 * FRS + 1;  // padded stride between elements
struct FeatLds {
    double V[64];
    double T[FEPB * FES];
    double red[FEPB][8];
};

__device__ void row_features(gptr<const double> __restrict__ q, gptr<const double> __restrict__ B, FeatLds& L, int e, bool act,
                             float* xn, float& modal) {
    const int el = threadIdx.x >> 3, r = threadIdx.x & 7;
    double h[8], u[8];
    double rs = 0.0;
    if (act) {
        const int64_t qb = (int64_t)e * 64 * 3 + r * 8;
        const int64_t xb = (int64_t)e * 64 + r * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            h[i] = q[qb + i];
            u[i] = B[xb + i] + h[i];
            rs += h[i];
        }
    }
    L.red[el][r] = rs;
    __syncthreads();
    double mean = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) mean += L.red[el][k];
    mean *= (1.0 / 64.0);
    double dev[8], mx = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        dev[i] = act ? h[i] - mean : 0.0;
        mx = fmax(mx, fabs(dev[i]));
    }
    __syncthreads();
    L.red[el][r] = mx;
    double* T = L.T + el * FES;
    // column r of T = V u (sum over xi)
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        double t = 0.0;
#pragma unroll
        for (int l = 0; l < 8; ++l) t += L.V[a * 8 + l] * u[l];
        T[a * FRS + r] = act ? t : 0.0;
    }
    __syncthreads();
    double sdev = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) sdev = fmax(sdev, L.red[el][k]);
#pragma unroll
    for (int i = 0; i < 8; ++i) xn[i] = (sdev > 0.0) ? (float)(dev[i] / sdev) : 0.0f;
    // row r of m = T V^T (sum over eta), in place of row r of T (only this thread touches it):
    // only the squares of the modes are used from here on, so the row's squares are stored
    {
        double t[8];
#pragma unroll
        for (int l = 0; l < 8; ++l) t[l] = T[r * FRS + l];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            double m = 0.0;
#pragma unroll
            for (int l = 0; l < 8; ++l) m += L.V[c * 8 + l] * t[l];
            T[r * FRS + c] = m * m;
        }
    }
    __syncthreads();
    // shell r: row r up to the diagonal and column r above it
    double Ek = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (c <= r) Ek += T[r * FRS + c];
#pragma unroll
    for (int a = 0; a < 8; ++a)
        if (a < r) Ek += T[a * FRS + r];
    L.red[el][r] = Ek;
    __syncthreads();
    double Et = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) Et += L.red[el][k];
    modal = (Et > 0.0) ? (float)(fmax(log10(Ek / Et), -16.0) / 8.0) : -2.0f;
}

// features to global FP32 rows [e][72]
__global__ void __launch_bounds__(256) k_mlp_features(int Nelem, const double* __restrict__ q,
                                                      const double* __restrict__ B, const double* __restrict__ V,
                                                      float* __restrict__ feat) {
    __shared__ FeatLds L;
    if (threadIdx.x < 64) L.V[threadIdx.x] = V[threadIdx.x];
    __syncthreads();
    const int el = threadIdx.x >> 3, r = threadIdx.x & 7;
    const int e = blockIdx.x * FEPB + el;
    const bool act = e < Nelem;
    float xn[8], modal;
    row_features(glob(q), glob(B), L, e, act, xn, modal);
    if (act) {
        float* row = feat + (int64_t)e * NIN;
#pragma unroll
        for (int i = 0; i < 8; ++i) row[r * 8 + i] = xn[i];
        row[64 + r] = modal;
    }
}

// One wave: 16 feature rows (bf16, stride KP, in LDS) -> 16 outputs.  Weights in LDS as
// W1T[32][KP], W2T[32][32] (bf16, output-major so a lane's 8 K values are contiguous).
__device__ inline void mlp_tile(const uint16_t* X, uint16_t* H, const uint16_t* W1T, const uint16_t* W2T,
                                const float* w3, const float* b1, const float* b2, float b3, float* zout) {
    const int l = threadIdx.x & 63;
    const int row = l & 15, kq = (l >> 4) * 8;
    floatx4 acc[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < KP / 32; ++kc) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(X + row * KP + kc * 32 + kq);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(W1T + (t * 16 + row) * KP + kc * 32 + kq);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
    // D layout: lane l holds D[4*(l/16) + r][l % 16]
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = (l >> 4) * 4 + r, j = t * 16 + (l & 15);
            H[i * NH + j] = f2bf(fmaxf(acc[t][r] + b1[j], 0.0f));
        }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(H + row * NH + kq);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(W2T + (t * 16 + row) * NH + kq);
        const floatx4 z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        const int j = t * 16 + (l & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) part[r] += rbf(fmaxf(z[r] + b2[j], 0.0f)) * w3[j];
    }
    // sum over the 16 lanes that hold the same rows
#pragma unroll
    for (int r = 0; r < 4; ++r)
        for (int o = 8; o > 0; o >>= 1) part[r] += __shfl_xor(part[r], o);
    if ((l & 15) == 0)
#pragma unroll
        for (int r = 0; r < 4; ++r) zout[(l >> 4) * 4 + r] = part[r] + b3;
    __builtin_amdgcn_wave_barrier();
}

struct MlpDev {
    const uint16_t* W1T;   // [32][KP]
    const uint16_t* W2T;   // [32][32]
    const float* w3;       // [32] bf16-rounded values
    const float* bias;     // b1[32], b2[32], b3
};

__device__ inline void load_weights(const MlpDev& W, uint16_t* sW1, uint16_t* sW2, float* sw3, float* sb) {
    for (int t = threadIdx.x; t < NH * KP; t += blockDim.x) sW1[t] = W.W1T[t];
    for (int t = threadIdx.x; t < NH * NH; t += blockDim.x) sW2[t] = W.W2T[t];
    for (int t = threadIdx.x; t < NH; t += blockDim.x) sw3[t] = W.w3[t];
    for (int t = threadIdx.x; t < 2 * NH + 1; t += blockDim.x) sb[t] = W.bias[t];
}

__device__ inline float sigmoidf_(float z) { return 1.0f / (1.0f + expf(-z)); }

// inference on feature rows: each wave takes 16 rows at a time
__global__ void __launch_bounds__(256) k_mlp_infer(int n, const float* __restrict__ feat, MlpDev W, float* __restrict__ p) {
    __shared__ __attribute__((aligned(16))) uint16_t sW1[NH * KP], sW2[NH * NH];
    __shared__ float sw3[NH], sb[2 * NH + 1];
    __shared__ __attribute__((aligned(16))) uint16_t sX[4][16 * KP], sH[4][16 * NH];
    __shared__ float sZ[4][16];
    load_weights(W, sW1, sW2, sw3, sb);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    for (int t = l; t < 16 * KP; t += 64) sX[w][t] = 0;
    __syncthreads();
    for (int r0 = (blockIdx.x * 4 + w) * 16; r0 < n; r0 += gridDim.x * 64) {
        for (int t = l; t < 16 * NIN; t += 64) {
            const int r = t / NIN, k = t % NIN;
            sX[w][r * KP + k] = (r0 + r < n) ? f2bf(feat[(int64_t)(r0 + r) * NIN + k]) : 0;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        mlp_tile(sX[w], sH[w], sW1, sW2, sw3, sb, sb + NH, sb[2 * NH], sZ[w]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        if (l < 16 && r0 + l < n) p[r0 + l] = sigmoidf_(sZ[w][l]);
        __builtin_amdgcn_wave_barrier();
    }
}

// Features + inference fused over the elements of a state; p per element and, when estat is
// given, the class: 0 smooth (p <= lo), 1 troubled (p >= hi), 2 unsure.  A block takes 32 elements:
// features by row threads into two 16-row BF16 tiles, then waves 0 and 1 run one tile each.
// With MS.visc (and estat), the block then finishes ShockCapturing masked by the classes: the
// feature pass left the squares of the modal coefficients m = V (h + b) V^T in LDS, formed with
// the sensor's arithmetic (ShockCapturing.okl:63-129), so four threads per element form the energy
// sums from them in the okl's order; the viscosity is the sensor's value, or 0 where the class is
// smooth.
// counters (may be null): [0] element-stages with viscosity, [1] sensor positives declared smooth
// and [2] sensor positives (both with audit only), [3] element-stages not declared smooth.
struct MaskedSensor {
    double* visc = nullptr;
    double* visc_plot = nullptr;
    SensorConsts sc{};
    int audit = 0;
    unsigned long long* counters = nullptr;
    // the device clock (see mlp_classify_sensor): skip a step issued after T
    const double* clock = nullptr;
    double T = 0.0;
    int clock_mode = 0;
    // a scenario of a batched ensemble: an element given viscosity marks itself and its face
    // neighbours with `mark`; with stage0 its viscosity also enters the scenario's maximum
    const StageScn* scn = nullptr;
    int mark = 0;
    int stage0 = 0;
};

// The indicator over elements e0 .. e0 + FEPB - 1 of one state (all 256 threads of the block).
// LDS holds the feature pass's tiles (V, the modal coefficients, the row reductions); once the
// sensor's energy sums are taken from the coefficients, the two 16-row BF16 feature tiles and the
// hidden activations take their place.  The network's weights are read from global memory (8 KB,
// cached), not staged per block: about 22 KB of LDS per block, 7 blocks per CU.
__device__ void mlp_indicator_block(int Nelem, int e0, gptr<const double> __restrict__ q,
                                    gptr<const double> __restrict__ B, gptr<const double> __restrict__ V,
                                    const MlpDev& W, float lo, float hi,
                                    float* __restrict__ p, signed char* __restrict__ estat, const MaskedSensor& MS) {
    if (MS.counters && threadIdx.x == 0)
        atomicAdd(MS.counters + 4, (unsigned long long)min(FEPB, Nelem - e0));
    __shared__ __attribute__((aligned(16))) FeatLds L;
    __shared__ float sZ[2][16];
    __shared__ signed char sSt[FEPB];
    __shared__ double sSum[FEPB][4];
    static_assert(sizeof(L.T) >= (2 * 16 * KP + 2 * 16 * NH) * sizeof(uint16_t), "feature tiles fit the T area");
    // V is published by the first barrier inside row_features
    if (threadIdx.x < 64) L.V[threadIdx.x] = V[threadIdx.x];
    const int el = threadIdx.x >> 3, r = threadIdx.x & 7;
    const int e = e0 + el;
    const bool act = e < Nelem;
    float xn[8], modal;
    row_features(q, B, L, e, act, xn, modal);
    // row_features leaves the squares of the modal coefficients m = V (h + b) V^T in L.T: the
    // sensor's sums
    if (MS.visc && act && r < 4) sSum[el][r] = sensor_sum<8, FRS, true>(L.T + el * FES, r);
    __syncthreads();   // L.T is free
    uint16_t* sX = reinterpret_cast<uint16_t*>(L.T);   // [2][16][KP]
    uint16_t* sH = sX + 2 * 16 * KP;                   // [2][16][NH]
    {
        uint16_t* row = sX + (el >> 4) * 16 * KP + (el & 15) * KP;
#pragma unroll
        for (int i = 0; i < 8; ++i) row[r * 8 + i] = act ? f2bf(xn[i]) : (uint16_t)0;
        row[64 + r] = act ? f2bf(modal) : (uint16_t)0;
        for (int k = NIN + r; k < KP; k += 8) row[k] = 0;   // the K padding of layer 1
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (w < 2) {
        mlp_tile(sX + w * 16 * KP, sH + w * 16 * NH, W.W1T, W.W2T, W.w3, W.bias, W.bias + NH, W.bias[2 * NH], sZ[w]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        const int ee = e0 + w * 16 + l;
        if (l < 16 && ee < Nelem) {
            const float pv = sigmoidf_(sZ[w][l]);
            const signed char st = (pv <= lo) ? 0 : (pv >= hi ? 1 : 2);
            if (p) p[ee] = pv;
            if (estat) estat[ee] = st;
            sSt[w * 16 + l] = st;
        }
    }
    if (MS.visc) {
        // the block's counts and largest viscosity gather in LDS; one global atomic each at the end
        __shared__ unsigned long long sCnt[4], sMax;
        if (threadIdx.x < 4) sCnt[threadIdx.x] = 0ull;
        if (threadIdx.x == 4) sMax = 0ull;
        __syncthreads();   // the classes are in sSt
        if (act && r == 0) {
            double v = sensor_ramp(sSum[el][0], sSum[el][1], sSum[el][2], sSum[el][3], MS.sc);
            const int st = sSt[el];
            const double plain = v;
            if (st == 0) v = 0.0;
            if (MS.counters) {
                if (v > 0.0) atomicAdd(&sCnt[0], 1ull);
                if (MS.audit && plain > 0.0 && st == 0) atomicAdd(&sCnt[1], 1ull);
                if (MS.audit && plain > 0.0) atomicAdd(&sCnt[2], 1ull);
                if (st != 0) atomicAdd(&sCnt[3], 1ull);
            }
            glob(MS.visc)[e] = v;
            if (MS.visc_plot) glob(MS.visc_plot)[e] = v;
            if (MS.scn) {
                if (v != 0.0) mark_reach(*MS.scn, e, MS.mark);
                // non-negative doubles order as their bit patterns; a NaN stays out, as fmax leaves it
                if (MS.stage0 && v > 0.0) atomicMax(&sMax, (unsigned long long)__double_as_longlong(v));
            }
        }
        const bool to_max = MS.scn && MS.stage0;
        if (MS.counters || to_max) {
            __syncthreads();
            if (threadIdx.x == 0) {
                if (MS.counters)
                    for (int k = 0; k < 4; ++k)
                        if (sCnt[k]) atomicAdd(MS.counters + k, sCnt[k]);
                if (to_max && sMax) atomicMax(MS.scn->emax, sMax);
            }
        }
    }
}

__global__ void __launch_bounds__(256) k_mlp_indicator(int Nelem, const double* __restrict__ q,
                                                       const double* __restrict__ B, const double* __restrict__ V,
                                                       MlpDev W, float lo, float hi, float* __restrict__ p,
                                                       signed char* __restrict__ estat, MaskedSensor MS) {
    if (MS.clock) {
        // a step issued after T changes nothing: no evaluation, no counts (uniform per block)
        if ((MS.clock_mode == 1 && MS.clock[1] >= MS.T) || (MS.clock_mode == 2 && MS.clock[0] == 0.0)) return;
    }
    mlp_indicator_block(Nelem, (int)blockIdx.x * FEPB, glob(q), glob(B), glob(V), W, lo, hi, p, estat, MS);
}

// The indicator over the learned scenarios of a batched ensemble: block b takes tile tasks[b] (32
// elements of one N = 7 scenario) of the stage state in buffer src.  A scenario whose time has
// reached its end (dt = 0) is skipped.
__global__ void __launch_bounds__(256) k_ens_indicator(const StageScn* __restrict__ scns,
                                                       const StageTask* __restrict__ tasks, int src,
                                                       const int* __restrict__ mark, int stage, MlpDev W, float lo,
                                                       float hi, int audit, unsigned long long* counters) {
    const StageTask tk = tasks[blockIdx.x];
    const StageScn& S = scns[tk.scn];
    if (stage > 0 && S.clock[0] == 0.0) return;
    if (stage == 0 && S.clock[1] >= S.T) return;
    MaskedSensor MS;
    MS.visc = S.visc;
    MS.sc = S.sc;
    MS.audit = audit;
    MS.counters = counters;
    MS.scn = &S;
    MS.mark = *mark + stage;
    MS.stage0 = stage == 0;
    mlp_indicator_block(S.Nelem, tk.e0, glob((const double*)S.q[src]), glob(S.B), glob(S.VdmInv), W, lo, hi, nullptr,
                        nullptr, MS);
}

// ---------------------------------------------------------------------------------------------
// training (FP32, weights and activations rounded to BF16 in the forward pass as in inference)
// ---------------------------------------------------------------------------------------------
constexpr int ACT = NIN + 4 * NH + 1;    // per-sample record: xq[72], a1q, dz1, a2q, dz2, dz3

// 32 threads per sample (thread j owns hidden unit j), 8 samples per block of 256
__global__ void __launch_bounds__(256) k_train_fwd_bwd(int n, const float* __restrict__ feat,
                                                       const float* __restrict__ label, const int* __restrict__ idx,
                                                       float pos_weight, const float* __restrict__ Wq,
                                                       float* __restrict__ act, float* __restrict__ loss) {
    __shared__ float sW[ESDG_MLP_NPARAM];
    __shared__ float sx[8][NIN], sa1[8][NH], sdz2[8][NH];
    for (int t = threadIdx.x; t < ESDG_MLP_NPARAM; t += blockDim.x) sW[t] = Wq[t];
    const int ls = threadIdx.x >> 5, j = threadIdx.x & 31;
    const int s = blockIdx.x * 8 + ls;
    const bool ok = s < n;
    const int row = ok ? (idx ? idx[s] : s) : 0;
    float* A = act + (int64_t)(ok ? s : 0) * ACT;
    for (int i = j; i < NIN; i += 32) {
        const float xi = ok ? rbf(feat[(int64_t)row * NIN + i]) : 0.0f;
        sx[ls][i] = xi;
        if (ok) A[i] = xi;
    }
    __syncthreads();
    float z1 = 0.0f;
    for (int i = 0; i < NIN; ++i) z1 += sx[ls][i] * sW[OFF_W1 + i * NH + j];
    z1 += sW[OFF_B1 + j];
    const float a1 = rbf(fmaxf(z1, 0.0f));
    sa1[ls][j] = a1;
    __syncthreads();
    float z2 = 0.0f;
    for (int i = 0; i < NH; ++i) z2 += sa1[ls][i] * sW[OFF_W2 + i * NH + j];
    z2 += sW[OFF_B2 + j];
    const float a2 = rbf(fmaxf(z2, 0.0f));
    // z3 = sum_j a2_j w3_j over the 32 threads of the sample, in index order through LDS
    sdz2[ls][j] = a2 * sW[OFF_W3 + j];
    __syncthreads();
    float z3 = 0.0f;
    for (int i = 0; i < NH; ++i) z3 += sdz2[ls][i];
    z3 += sW[OFF_B3];
    const float y = ok ? label[row] : 0.0f;
    const float p = 1.0f / (1.0f + expf(-z3));
    const float dz3 = ((1.0f - y) * p - pos_weight * y * (1.0f - p)) / (float)n;
    __syncthreads();
    const float dz2 = (z2 > 0.0f) ? sW[OFF_W3 + j] * dz3 : 0.0f;
    sdz2[ls][j] = dz2;
    __syncthreads();
    float da = 0.0f;
    for (int k = 0; k < NH; ++k) da += sW[OFF_W2 + j * NH + k] * sdz2[ls][k];
    if (ok) {
        A[NIN + j] = a1;
        A[NIN + NH + j] = (z1 > 0.0f) ? da : 0.0f;
        A[NIN + 2 * NH + j] = a2;
        A[NIN + 3 * NH + j] = dz2;
        if (j == 0) {
            A[NIN + 4 * NH] = dz3;
            // -(w y log p + (1 - y) log(1 - p)) with log p = -softplus(-z3), log(1-p) = -softplus(z3)
            const float sp_neg = fmaxf(-z3, 0.0f) + log1pf(expf(-fabsf(z3)));
            const float sp_pos = fmaxf(z3, 0.0f) + log1pf(expf(-fabsf(z3)));
            loss[s] = pos_weight * y * sp_neg + (1.0f - y) * sp_pos;
        }
    }
}

// gradient entry per thread, summed over the samples in order
__global__ void __launch_bounds__(256) k_train_grad(int n, const float* __restrict__ act, const float* __restrict__ loss,
                                                    float* __restrict__ grad, float* __restrict__ loss_sum) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > ESDG_MLP_NPARAM) return;
    float g = 0.0f;
    if (p == ESDG_MLP_NPARAM) {
        for (int s = 0; s < n; ++s) g += loss[s];
        *loss_sum = g / (float)n;
        return;
    }
    int ia, ib;   // g = sum_s act[s][ia] * act[s][ib]  (ia < 0: factor 1)
    if (p < OFF_B1) { ia = p / NH; ib = NIN + NH + p % NH; }                      // W1: xq_i dz1_j
    else if (p < OFF_W2) { ia = -1; ib = NIN + NH + (p - OFF_B1); }              // b1: dz1_j
    else if (p < OFF_B2) { const int q = p - OFF_W2; ia = NIN + q / NH; ib = NIN + 3 * NH + q % NH; }  // W2: a1_i dz2_j
    else if (p < OFF_W3) { ia = -1; ib = NIN + 3 * NH + (p - OFF_B2); }          // b2: dz2_j
    else if (p < OFF_B3) { ia = NIN + 2 * NH + (p - OFF_W3); ib = NIN + 4 * NH; }  // w3: a2_j dz3
    else { ia = -1; ib = NIN + 4 * NH; }                                         // b3: dz3
    for (int s = 0; s < n; ++s) {
        const float* A = act + (int64_t)s * ACT;
        g += (ia < 0) ? A[ib] : A[ia] * A[ib];
    }
    grad[p] = g;
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
                       float lr, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ESDG_MLP_NPARAM) return;
    const float gi = g[p];
    const float mi = 0.9f * m[p] + 0.1f * gi;
    const float vi = 0.999f * v[p] + 0.001f * gi * gi;
    m[p] = mi;
    v[p] = vi;
    w[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// master -> BF16-rounded FP32 copy (training forward) and the inference layouts
__global__ void k_quantize(const float* __restrict__ w, float* __restrict__ Wq, uint16_t* __restrict__ W1T,
                           uint16_t* __restrict__ W2T, float* __restrict__ w3, float* __restrict__ bias) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < ESDG_MLP_NPARAM) {
        const bool is_bias = (p >= OFF_B1 && p < OFF_W2) || (p >= OFF_B2 && p < OFF_W3) || p == OFF_B3;
        Wq[p] = is_bias ? w[p] : rbf(w[p]);
    }
    if (p < NH * KP) {
        const int j = p / KP, k = p % KP;
        W1T[p] = (k < NIN) ? f2bf(w[OFF_W1 + k * NH + j]) : 0;
    }
    if (p < NH * NH) {
        const int j = p / NH, k = p % NH;
        W2T[p] = f2bf(w[OFF_W2 + k * NH + j]);
    }
    if (p < NH) w3[p] = rbf(w[OFF_W3 + p]);
    if (p < NH) { bias[p] = w[OFF_B1 + p]; bias[NH + p] = w[OFF_B2 + p]; }
    if (p == 0) bias[2 * NH] = w[OFF_B3];
}

}  // namespace esdg

using namespace esdg;

struct esdg_mlp {
    esdg_ctx* ctx = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr, *Wq = nullptr;
    uint16_t *W1T = nullptr, *W2T = nullptr;
    float *w3 = nullptr, *bias = nullptr, *loss = nullptr;
    float *act = nullptr, *sloss = nullptr;
    int cap = 0;
    int t = 0;
    std::vector<void*> owned;

    MlpDev dev() const { return MlpDev{W1T, W2T, w3, bias}; }
    int alloc(void** p, size_t bytes) {
        int rc = esdg_malloc(ctx, p, bytes);
        if (!rc) owned.push_back(*p);
        return rc;
    }
    int quantize() {
        constexpr int n = (NH * KP > ESDG_MLP_NPARAM) ? NH * KP : ESDG_MLP_NPARAM;
        k_quantize<<<(n + 255) / 256, 256, 0, ctx->stream>>>(w, Wq, W1T, W2T, w3, bias);
        return launch_check(ctx, "mlp quantize");
    }
    int ensure(int n) {
        if (n <= cap) return ESDG_OK;
        if (act) { esdg_free(ctx, act); act = nullptr; }
        if (sloss) { esdg_free(ctx, sloss); sloss = nullptr; }
        int rc = esdg_malloc(ctx, (void**)&act, (size_t)n * ACT * sizeof(float));
        if (!rc) rc = esdg_malloc(ctx, (void**)&sloss, (size_t)n * sizeof(float));
        if (!rc) cap = n;
        return rc;
    }
    ~esdg_mlp() {
        if (act) esdg_free(ctx, act);
        if (sloss) esdg_free(ctx, sloss);
        for (void* p : owned) esdg_free(ctx, p);
    }
};

namespace esdg {
int mlp_classify(esdg_ctx* ctx, const esdg_mlp* m, int Nelem, const double* q, const double* B, const double* VdmInv,
                 float lo, float hi, float* p, signed char* estat) {
    KernelScope ks(ctx, "mlp_indicator");
    k_mlp_indicator<<<blocks_for(Nelem, FEPB), 256, 0, ctx->stream>>>(Nelem, q, B, VdmInv, m->dev(), lo, hi, p, estat,
                                                                       MaskedSensor{});
    return launch_check(ctx, "mlp_indicator");
}

int mlp_ensemble_indicator(esdg_ctx* ctx, const esdg_mlp* m, const StageScn* scns, const StageTask* tasks, int ntask,
                           int src, const int* mark, int stage, float lo, float hi, int audit,
                           unsigned long long* counters) {
    if (ntask == 0) return ESDG_OK;
    KernelScope ks(ctx, "ens_indicator");
    k_ens_indicator<<<ntask, 256, 0, ctx->stream>>>(scns, tasks, src, mark, stage, m->dev(), lo, hi, audit, counters);
    return launch_check(ctx, "ens_indicator");
}

int mlp_classify_sensor(esdg_ctx* ctx, const esdg_mlp* m, int Nelem, const double* q, const double* B,
                        const double* VdmInv, float lo, float hi, signed char* estat, int audit,
                        unsigned long long* counters, double* visc, double* visc_plot, const double* clock,
                        double T, int clock_mode, hipStream_t stream) {
    if (ctx->p.ngl != 8) return set_error(ctx, ESDG_ERR_ARG, "the learned indicator needs N = 7");
    MaskedSensor MS;
    MS.clock = clock;
    MS.T = T;
    MS.clock_mode = clock_mode;
    MS.visc = visc;
    MS.visc_plot = visc_plot;
    MS.sc = SensorConsts{ctx->p.ZeroTOL, ctx->p.eps0, ctx->p.sigmaMin, ctx->p.sigmaMax};
    MS.audit = audit;
    MS.counters = counters;
    KernelScope ks(ctx, "mlp_indicator");
    hipStream_t st = stream ? stream : ctx->stream;
    k_mlp_indicator<<<blocks_for(Nelem, FEPB), 256, 0, st>>>(Nelem, q, B, VdmInv, m->dev(), lo, hi, nullptr, estat, MS);
    return launch_check(ctx, "mlp_indicator");
}
}  // namespace esdg

extern "C" {

int esdg_mlp_create(esdg_ctx* ctx, uint64_t seed, esdg_mlp** out) {
    ESDG_REQUIRE(ctx, ctx && out, "esdg_mlp_create: null argument");
    *out = nullptr;
    esdg_mlp* M = new (std::nothrow) esdg_mlp();
    if (!M) return set_error(ctx, ESDG_ERR_NOMEM, "esdg_mlp_create: out of host memory");
    M->ctx = ctx;
    const size_t P = ESDG_MLP_NPARAM * sizeof(float);
    int rc = M->alloc((void**)&M->w, P);
    if (!rc) rc = M->alloc((void**)&M->m, P);
    if (!rc) rc = M->alloc((void**)&M->v, P);
    if (!rc) rc = M->alloc((void**)&M->g, P);
    if (!rc) rc = M->alloc((void**)&M->Wq, P);
    if (!rc) rc = M->alloc((void**)&M->W1T, NH * KP * sizeof(uint16_t));
    if (!rc) rc = M->alloc((void**)&M->W2T, NH * NH * sizeof(uint16_t));
    if (!rc) rc = M->alloc((void**)&M->w3, NH * sizeof(float));
    if (!rc) rc = M->alloc((void**)&M->bias, (2 * NH + 1) * sizeof(float));
    if (!rc) rc = M->alloc((void**)&M->loss, sizeof(float));
    if (rc) {
        delete M;
        return rc;
    }
    // He-uniform initial weights from SplitMix64(seed, index), zero biases
    std::vector<float> h(ESDG_MLP_NPARAM, 0.0f);
    auto u01 = [&](uint64_t i) {
        uint64_t z = seed ^ (0x9E3779B97F4A7C15ull * (i + 1));
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int count, int fan_in) {
        const double a = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; ++i) h[off + i] = (float)((2.0 * u01(off + i) - 1.0) * a);
    };
    fill(OFF_W1, NIN * NH, NIN);
    fill(OFF_W2, NH * NH, NH);
    fill(OFF_W3, NH, NH);
    rc = esdg_mlp_set_params(M, h.data());
    if (rc) {
        delete M;
        return rc;
    }
    *out = M;
    return ESDG_OK;
}

int esdg_mlp_destroy(esdg_mlp* m) {
    if (!m) return ESDG_OK;
    (void)esdg_finish(m->ctx);
    delete m;
    return ESDG_OK;
}

int esdg_mlp_set_params(esdg_mlp* M, const float* hp) {
    if (!M || !hp) return set_error(M ? M->ctx : nullptr, ESDG_ERR_ARG, "esdg_mlp_set_params: null argument");
    const size_t P = ESDG_MLP_NPARAM * sizeof(float);
    int rc = esdg_memcpy(M->ctx, M->w, hp, P, ESDG_H2D);
    if (!rc) rc = hip_check(M->ctx, hipMemsetAsync(M->m, 0, P, M->ctx->stream), "mlp moments");
    if (!rc) rc = hip_check(M->ctx, hipMemsetAsync(M->v, 0, P, M->ctx->stream), "mlp moments");
    M->t = 0;
    if (!rc) rc = M->quantize();
    if (!rc) rc = esdg_finish(M->ctx);
    return rc;
}

int esdg_mlp_get_params(esdg_mlp* M, float* hp) {
    if (!M || !hp) return set_error(M ? M->ctx : nullptr, ESDG_ERR_ARG, "esdg_mlp_get_params: null argument");
    return esdg_memcpy(M->ctx, hp, M->w, ESDG_MLP_NPARAM * sizeof(float), ESDG_D2H);
}

int esdg_mlp_features(esdg_ctx* ctx, int Nelem, const double* q, const double* B, const double* VdmInv, float* feat) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem >= 0 && q && B && VdmInv && feat, "esdg_mlp_features: bad argument");
    ESDG_REQUIRE(ctx, ctx->p.ngl == 8, "esdg_mlp_features: the indicator is defined for N = 7 (ngl = 8)");
    if (Nelem == 0) return ESDG_OK;
    KernelScope ks(ctx, "mlp_features");
    k_mlp_features<<<blocks_for(Nelem, FEPB), 256, 0, ctx->stream>>>(Nelem, q, B, VdmInv, feat);
    return launch_check(ctx, "mlp_features");
}

int esdg_mlp_infer(esdg_mlp* M, int n, const float* feat, float* p) {
    if (!M) return set_error(nullptr, ESDG_ERR_ARG, "esdg_mlp_infer: null network");
    ESDG_REQUIRE(M->ctx, n >= 0 && feat && p, "esdg_mlp_infer: bad argument");
    if (n == 0) return ESDG_OK;
    KernelScope ks(M->ctx, "mlp_infer");
    const int blocks = (int)std::min<int64_t>((n + 63) / 64, 4096);
    k_mlp_infer<<<blocks, 256, 0, M->ctx->stream>>>(n, feat, M->dev(), p);
    return launch_check(M->ctx, "mlp_infer");
}

int esdg_mlp_indicator(esdg_mlp* M, int Nelem, const double* q, const double* B, const double* VdmInv, float* p) {
    if (!M) return set_error(nullptr, ESDG_ERR_ARG, "esdg_mlp_indicator: null network");
    ESDG_REQUIRE(M->ctx, Nelem >= 0 && q && B && VdmInv && p, "esdg_mlp_indicator: bad argument");
    ESDG_REQUIRE(M->ctx, M->ctx->p.ngl == 8, "esdg_mlp_indicator: the indicator is defined for N = 7 (ngl = 8)");
    if (Nelem == 0) return ESDG_OK;
    return mlp_classify(M->ctx, M, Nelem, q, B, VdmInv, 0.0f, 1.0f, p, nullptr);
}

int esdg_mlp_loss_grad(esdg_mlp* M, int n, const float* feat, const float* label, const int* idx, float pos_weight,
                       float* grad, float* loss) {
    if (!M) return set_error(nullptr, ESDG_ERR_ARG, "esdg_mlp_loss_grad: null network");
    esdg_ctx* ctx = M->ctx;
    ESDG_REQUIRE(ctx, n > 0 && feat && label && grad && loss, "esdg_mlp_loss_grad: bad argument");
    int rc = M->ensure(n);
    if (rc) return rc;
    {
        KernelScope ks(ctx, "mlp_fwd_bwd");
        k_train_fwd_bwd<<<(n + 7) / 8, 256, 0, ctx->stream>>>(n, feat, label, idx, pos_weight, M->Wq, M->act,
                                                               M->sloss);
        rc = launch_check(ctx, "mlp_fwd_bwd");
        if (rc) return rc;
    }
    KernelScope ks(ctx, "mlp_grad");
    k_train_grad<<<(ESDG_MLP_NPARAM + 1 + 255) / 256, 256, 0, ctx->stream>>>(n, M->act, M->sloss, grad, loss);
    return launch_check(ctx, "mlp_grad");
}

int esdg_mlp_train_step(esdg_mlp* M, int n, const float* feat, const float* label, const int* idx, float pos_weight,
                        float lr, float* host_loss) {
    int rc = esdg_mlp_loss_grad(M, n, feat, label, idx, pos_weight, M ? M->g : nullptr, M ? M->loss : nullptr);
    if (rc) return rc;
    esdg_ctx* ctx = M->ctx;
    M->t += 1;
    const float c1 = 1.0f - std::pow(0.9f, (float)M->t), c2 = 1.0f - std::pow(0.999f, (float)M->t);
    k_adam<<<(ESDG_MLP_NPARAM + 255) / 256, 256, 0, ctx->stream>>>(M->w, M->m, M->v, M->g, lr, c1, c2);
    rc = launch_check(ctx, "mlp adam");
    if (!rc) rc = M->quantize();
    if (!rc && host_loss) rc = esdg_memcpy(ctx, host_loss, M->loss, sizeof(float), ESDG_D2H);
    return rc;
}

}  // extern "C"
