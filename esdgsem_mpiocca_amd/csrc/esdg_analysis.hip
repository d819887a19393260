// esdg_analysis.hip -- the per-step analysis kernels of the time loop and the solution integrals.
//
//   esdg_arrival_times      okl/Analysis/ArrivalTimes.okl:16-62 (called after every step,
//                           src/deviceclass.cpp:1132-1135, buffer set to 1000.0 at :324-333)
//   esdg_maximum_elevation  okl/Analysis/MaximumElevation.okl:16-55 (:1138-1141, buffer zeroed at
//                           :337-342)
//   esdg_total_mass_dev     basis::calcTotalMass, src/basis.cpp:1088-1111, on device arrays
//   esdg_total_entropy_dev  basis::calcTotalEntropy + calcEntropyPointwise, src/basis.cpp:1004-1055
//
// The analysis kernels are element-wise and exact.  The integrals are device reductions: each
// element's sum in the reference's node order (eta outer, xi inner), then a fixed-shape tree over
// the elements, so the result does not depend on timing; it differs from the reference's single
// serial sum over all elements by rounding only.
#include <algorithm>
#include <cmath>

#include "esdg_internal.h"

namespace esdg {

__global__ void __launch_bounds__(256) k_arrival_times(int64_t n, int ngl2, double t, double h_zero,
                                                       const double* __restrict__ q, const double* __restrict__ b,
                                                       double* __restrict__ arr) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        if (arr[x] == 1000.0) {
            const int64_t e = x / ngl2, k = x - e * ngl2;
            if (fabs((q[e * ngl2 * ESDG_NEQ + k] + b[x]) - h_zero) > 0.00005) arr[x] = t;
        }
    }
}

__global__ void __launch_bounds__(256) k_maximum_elevation(int64_t n, int ngl2, double h_zero,
                                                           const double* __restrict__ q,
                                                           const double* __restrict__ b,
                                                           double* __restrict__ mx) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = x / ngl2, k = x - e * ngl2;
        const double free_surface = q[e * ngl2 * ESDG_NEQ + k] + b[x] - h_zero;
        mx[x] = fmax(mx[x], free_surface);
    }
}

// calcEntropyPointwise: u = v = 0 where h <= 1e-12
__device__ __forceinline__ double entropy_pointwise(double g, double h, double hu, double hv, double b) {
    double u = 0.0, v = 0.0;
    if (h > 1e-12) {
        u = hu / h;
        v = hv / h;
    }
    return 0.5 * h * (u * u + v * v) + 0.5 * g * h * h + g * h * b;
}

// Pass 1: a fixed grid of RED_BLOCKS blocks; block k sums the elements [k*chunk, (k+1)*chunk),
// thread t taking elements k*chunk + t, + 256, ... in order, then a fixed tree over the block.
constexpr int RED_BLOCKS = 1024;

template <bool ENTROPY>
__global__ void __launch_bounds__(256) k_integral_partial(int Nelem, int ngl, double g, const double* __restrict__ q,
                                                          const double* __restrict__ b, const double* __restrict__ Jac,
                                                          const double* __restrict__ w, double* __restrict__ part) {
    __shared__ double s[256];
    const int ngl2 = ngl * ngl;
    const int chunk = (Nelem + RED_BLOCKS - 1) / RED_BLOCKS;
    const int e0 = blockIdx.x * chunk, e1 = min(Nelem, e0 + chunk);
    double acc = 0.0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += 256) {
        double se = 0.0;
        for (int j = 0; j < ngl; ++j)
            for (int i = 0; i < ngl; ++i) {
                const int64_t xid = (int64_t)e * ngl2 + j * ngl + i;
                const int64_t id = (int64_t)e * ngl2 * ESDG_NEQ + j * ngl + i;
                const double v = ENTROPY ? entropy_pointwise(g, q[id], q[id + ngl2], q[id + 2 * ngl2], b[xid]) : q[id];
                se += v / Jac[xid] * w[i] * w[j];
            }
        acc += se;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Surface elevation h + b - h_zero at the gauge nodes (deviceclass.cpp:850: the state index of
// global node n is (n / ngl^2) (Neq - 1) ngl^2 + n).
__global__ void __launch_bounds__(64) k_gauge_sample(int ngauges, int ngl2, double h_zero, const int* __restrict__ ids,
                                                     const double* __restrict__ q, const double* __restrict__ b,
                                                     double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ngauges) return;
    const int64_t n = ids[k];
    const int64_t id = (n / ngl2) * (ESDG_NEQ - 1) * ngl2 + n;
    out[k] = q[id] + b[n] - h_zero;
}

// Pass 2: one block, a fixed tree over the RED_BLOCKS partial sums; the result to out[0].
__global__ void __launch_bounds__(256) k_integral_final(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double s[256];
    double a = 0.0;
    for (int k = threadIdx.x; k < RED_BLOCKS; k += 256) a += part[k];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

static unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 8192); }

template <bool ENTROPY>
static int integral(esdg_ctx* ctx, int Nelem, const double* q, const double* B, const double* Jac, const double* GLw,
                    double* host_out, const char* name) {
    static_assert(RED_BLOCKS + 1 <= 4096, "the reduction uses the context scratch");
    KernelScope ks(ctx, name);
    double* part = ctx->d_scratch;
    k_integral_partial<ENTROPY><<<RED_BLOCKS, 256, 0, ctx->stream>>>(Nelem, ctx->p.ngl, ctx->p.g, q, B, Jac, GLw, part);
    k_integral_final<<<1, 256, 0, ctx->stream>>>(part, part + RED_BLOCKS);
    int rc = launch_check(ctx, name);
    if (rc) return rc;
    return esdg_memcpy(ctx, host_out, part + RED_BLOCKS, sizeof(double), ESDG_D2H);
}

}  // namespace esdg

using namespace esdg;

extern "C" {

int esdg_arrival_times(esdg_ctx* ctx, int Nelem, double t, const double* q, const double* B, double* ArrivalTimings) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem >= 0 && q && B && ArrivalTimings, "esdg_arrival_times: bad argument");
    if (Nelem == 0) return ESDG_OK;
    KernelScope ks(ctx, "arrival_times");
    const int ngl2 = ctx->p.ngl * ctx->p.ngl;
    const int64_t n = (int64_t)Nelem * ngl2;
    k_arrival_times<<<grid_for(n), 256, 0, ctx->stream>>>(n, ngl2, t, ctx->p.h_zero, q, B, ArrivalTimings);
    return launch_check(ctx, "arrival_times");
}

int esdg_maximum_elevation(esdg_ctx* ctx, int Nelem, double t, const double* q, const double* B,
                           double* MaximumElevation) {
    (void)t;   // passed by the reference's call, unused by its kernel
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem >= 0 && q && B && MaximumElevation, "esdg_maximum_elevation: bad argument");
    if (Nelem == 0) return ESDG_OK;
    KernelScope ks(ctx, "maximum_elevation");
    const int ngl2 = ctx->p.ngl * ctx->p.ngl;
    const int64_t n = (int64_t)Nelem * ngl2;
    k_maximum_elevation<<<grid_for(n), 256, 0, ctx->stream>>>(n, ngl2, ctx->p.h_zero, q, B, MaximumElevation);
    return launch_check(ctx, "maximum_elevation");
}

int esdg_gauge_sample(esdg_ctx* ctx, int ngauges, const int* node_ids, const double* q, const double* B, double* out) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, ngauges >= 0 && (ngauges == 0 || (node_ids && q && B && out)), "esdg_gauge_sample: bad argument");
    if (ngauges == 0) return ESDG_OK;
    KernelScope ks(ctx, "gauge_sample");
    k_gauge_sample<<<(ngauges + 63) / 64, 64, 0, ctx->stream>>>(ngauges, ctx->p.ngl * ctx->p.ngl, ctx->p.h_zero,
                                                                node_ids, q, B, out);
    return launch_check(ctx, "gauge_sample");
}

int esdg_total_mass_dev(esdg_ctx* ctx, int Nelem, const double* q, const double* Jac, const double* GLw, double* host_mass) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem >= 0 && q && Jac && GLw && host_mass, "esdg_total_mass_dev: bad argument");
    return integral<false>(ctx, Nelem, q, nullptr, Jac, GLw, host_mass, "total_mass");
}

int esdg_total_entropy_dev(esdg_ctx* ctx, int Nelem, const double* q, const double* B, const double* Jac,
                           const double* GLw, double* host_entropy) {
    ESDG_REQUIRE(ctx, ctx, "null context");
    ESDG_REQUIRE(ctx, Nelem >= 0 && q && B && Jac && GLw && host_entropy, "esdg_total_entropy_dev: bad argument");
    return integral<true>(ctx, Nelem, q, B, Jac, GLw, host_entropy, "total_entropy");
}

}  // extern "C"
