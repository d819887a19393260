// esdg_internal.h -- shared definitions of the gfx950 implementation behind include/esdg.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "esdg.h"

#define ESDG_NEQ 3
#define ESDG_MAX_NGL 10

struct esdg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for work that runs beside the main stream's, forked from
    // and joined back into it by the two events inside one call; created by esdg_create (null if
    // that failed: the work then runs on the main stream)
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    esdg_params p{};
    std::string err;
    size_t bytes = 0;
    std::map<void*, size_t> allocs;
    double* d_scratch = nullptr;      // small device scratch (reduction partials)
    size_t scratch_count = 0;
    // profiling: per kernel name, (start,stop) event pairs not yet summed, and sums
    bool profile = false;
    struct Prof { std::vector<std::pair<hipEvent_t, hipEvent_t>> pending; int64_t launches = 0; double ms = 0; };
    std::map<std::string, Prof> prof;
    std::vector<hipEvent_t> event_pool;
};

namespace esdg {

int set_error(esdg_ctx* ctx, int code, const std::string& msg);
int hip_check(esdg_ctx* ctx, hipError_t e, const char* what);

// Records HIP events around one kernel launch when profiling is on.
class KernelScope {
  public:
    KernelScope(esdg_ctx* ctx, const char* name);
    ~KernelScope();
  private:
    esdg_ctx* ctx_;
    const char* name_;
    hipEvent_t start_ = nullptr;
};

// Checks the launch and reports through the context.
int launch_check(esdg_ctx* ctx, const char* name);

// Element / face counts are int32 (indices fit); byte offsets are 64-bit.
inline unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// The manufactured solutions h = H0 + cos(kx) sin(ky) cos t - b, u = 1/2, v = 3/2 selected by the
// boundary kinds of the convergence tests (src/deviceclass.cpp:363-387): Dirichlet_ConvTest.okl and
// S_ConvTest.okl (H0 = 8, k = 1), ...ConvTest3 (20, 1), ...ConvTest4 (20, 1e-3), ...ConvTest5 (20000, 1e-6).
__host__ __device__ inline bool is_manufactured(int bc) {
    return bc == ESDG_BC_CONVERGENCE || (bc >= ESDG_BC_CONV_TEST3 && bc <= ESDG_BC_CONV_TEST5);
}
struct Manufactured { double H0, k; };
__host__ __device__ inline Manufactured manufactured_of(int bc) {
    switch (bc) {
        case ESDG_BC_CONV_TEST3: return {20.0, 1.0};
        case ESDG_BC_CONV_TEST4: return {20.0, 0.001};
        case ESDG_BC_CONV_TEST5: return {20000.0, 0.000001};
        default: return {8.0, 1.0};
    }
}

}  // namespace esdg

#define ESDG_REQUIRE(ctx, cond, msg)                                        \
    do {                                                                    \
        if (!(cond)) return esdg::set_error((ctx), ESDG_ERR_ARG, (msg));    \
    } while (0)

// Dispatch on the runtime ngl to an ahead-of-time compiled template instance.
#define ESDG_NGL_DISPATCH(ngl, FN, ...)              \
    switch (ngl) {                                   \
        case 2: FN<2>(__VA_ARGS__); break;           \
        case 3: FN<3>(__VA_ARGS__); break;           \
        case 4: FN<4>(__VA_ARGS__); break;           \
        case 5: FN<5>(__VA_ARGS__); break;           \
        case 6: FN<6>(__VA_ARGS__); break;           \
        case 7: FN<7>(__VA_ARGS__); break;           \
        case 8: FN<8>(__VA_ARGS__); break;           \
        case 9: FN<9>(__VA_ARGS__); break;           \
        case 10: FN<10>(__VA_ARGS__); break;         \
        default: break;                              \
    }
