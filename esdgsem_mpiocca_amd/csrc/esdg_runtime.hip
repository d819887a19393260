// esdg_runtime.hip -- context, memory, errors, per-kernel event timing and the device max
// reduction behind include/esdg.h.  Replaces the OCCA device/memory layer the reference uses
// (src/deviceclass.cpp:36-57, 227-342) with one HIP stream per context.
#include <cstring>
#include <mutex>

#include "esdg_internal.h"

namespace esdg {

static thread_local std::string g_last_error_noctx;

int set_error(esdg_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    else g_last_error_noctx = msg;
    return code;
}

int hip_check(esdg_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return ESDG_OK;
    return set_error(ctx, ESDG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int launch_check(esdg_ctx* ctx, const char* name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(ctx, ESDG_ERR_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(e));
    return ESDG_OK;
}

static hipEvent_t take_event(esdg_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

KernelScope::KernelScope(esdg_ctx* ctx, const char* name) : ctx_(ctx), name_(name) {
    if (ctx_->profile) {
        start_ = take_event(ctx_);
        if (start_) (void)hipEventRecord(start_, ctx_->stream);
    }
}

KernelScope::~KernelScope() {
    if (!ctx_->profile || !start_) return;
    hipEvent_t stop = take_event(ctx_);
    if (!stop) return;
    (void)hipEventRecord(stop, ctx_->stream);
    auto& pr = ctx_->prof[name_];
    pr.pending.emplace_back(start_, stop);
    pr.launches += 1;
}

__global__ void fill_kernel(double* __restrict__ d, double v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = v;
}

// Block-level max of a grid-stride range; partial results in part[blockIdx.x].
__global__ void __launch_bounds__(256) max_partial_kernel(const double* __restrict__ x, int n, double* __restrict__ part) {
    __shared__ double s[256];
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = fmax(m, x[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// The allocation-free variant of the max reduction (esdg_reduce_max_noalloc) keeps its block
// partials in this file-scope array instead of the context's scratch: one partial per 256 values,
// so it serves n <= ESDG_NOALLOC_MAX_N.  Each device holds its own copy until the process ends.
// In host code the array's name is not a device address: kernels get the address that
// hipGetSymbolAddress returned for the device (noalloc_scratch below), never the name.
constexpr int NOALLOC_PARTIALS = ESDG_NOALLOC_MAX_N / 256;
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu;
static double* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device (by esdg_create
// for its device, else by the first esdg_reduce_max_noalloc call there).  Null on failure, with
// the message in *err.
static double* noalloc_scratch(std::string* err) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        *err = std::string("no current HIP device: ") + hipGetErrorString(e);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        g_noalloc_ptr[dev] = static_cast<double*>(p);
    }
    return g_noalloc_ptr[dev];
}

}  // namespace esdg

using namespace esdg;

extern "C" {

int esdg_create(esdg_ctx** out, int hip_device, const esdg_params* params) {
    if (!out || !params) return set_error(nullptr, ESDG_ERR_ARG, "esdg_create: null argument");
    *out = nullptr;
    if (params->ngl < 2 || params->ngl > ESDG_MAX_NGL)
        return set_error(nullptr, ESDG_ERR_ARG, "esdg_create: ngl must be in [2,10]");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return set_error(nullptr, ESDG_ERR_HIP, std::string("esdg_create: no HIP device: ") + hipGetErrorString(e));
    if (hip_device < 0 || hip_device >= ndev) return set_error(nullptr, ESDG_ERR_ARG, "esdg_create: bad device index");
    esdg_ctx* ctx = new (std::nothrow) esdg_ctx();
    if (!ctx) return set_error(nullptr, ESDG_ERR_NOMEM, "esdg_create: out of host memory");
    ctx->device = hip_device;
    ctx->p = *params;
    if (hipSetDevice(hip_device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return set_error(nullptr, ESDG_ERR_HIP, "esdg_create: stream creation failed");
    }
    ctx->scratch_count = 4096;
    if (hipMalloc(&ctx->d_scratch, ctx->scratch_count * sizeof(double)) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return set_error(nullptr, ESDG_ERR_NOMEM, "esdg_create: scratch allocation failed");
    }
    // The second stream and its fork/join events, created once here.  If any of them cannot be
    // created, work meant for the second stream runs on the main stream in order (same results);
    // the failure's error state is cleared so that no later launch check reports it.
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming) != hipSuccess) {
        for (hipEvent_t* ev : {&ctx->fork, &ctx->join})
            if (*ev) { (void)hipEventDestroy(*ev); *ev = nullptr; }
        if (ctx->side) (void)hipStreamDestroy(ctx->side);
        ctx->side = nullptr;
        (void)hipGetLastError();
    }
    // the allocation-free entry's scratch address on this device, looked up at set-up
    std::string serr;
    if (!noalloc_scratch(&serr)) {
        (void)hipGetLastError();
        for (hipEvent_t ev : {ctx->fork, ctx->join})
            if (ev) (void)hipEventDestroy(ev);
        if (ctx->side) (void)hipStreamDestroy(ctx->side);
        (void)hipFree(ctx->d_scratch);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return set_error(nullptr, ESDG_ERR_HIP, "esdg_create: " + serr);
    }
    *out = ctx;
    return ESDG_OK;
}

int esdg_destroy(esdg_ctx* ctx) {
    if (!ctx) return ESDG_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->allocs) (void)hipFree(kv.first);
    (void)hipFree(ctx->d_scratch);
    for (auto& kv : ctx->prof)
        for (auto& pe : kv.second.pending) { (void)hipEventDestroy(pe.first); (void)hipEventDestroy(pe.second); }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    for (hipEvent_t e : {ctx->fork, ctx->join})
        if (e) (void)hipEventDestroy(e);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ESDG_OK;
}

int esdg_malloc(esdg_ctx* ctx, void** d, size_t bytes) {
    ESDG_REQUIRE(ctx, ctx && d, "esdg_malloc: null argument");
    *d = nullptr;
    if (bytes == 0) bytes = 256;
    void* p = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess || hipMalloc(&p, bytes) != hipSuccess)
        return set_error(ctx, ESDG_ERR_NOMEM, "esdg_malloc: hipMalloc of " + std::to_string(bytes) + " bytes failed");
    int rc = hip_check(ctx, hipMemsetAsync(p, 0, bytes, ctx->stream), "esdg_malloc: zero fill");
    if (rc == ESDG_OK) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "esdg_malloc: sync");
    if (rc != ESDG_OK) { (void)hipFree(p); return rc; }
    ctx->allocs[p] = bytes;
    ctx->bytes += bytes;
    *d = p;
    return ESDG_OK;
}

int esdg_free(esdg_ctx* ctx, void* d) {
    ESDG_REQUIRE(ctx, ctx, "esdg_free: null context");
    if (!d) return ESDG_OK;
    auto it = ctx->allocs.find(d);
    ESDG_REQUIRE(ctx, it != ctx->allocs.end(), "esdg_free: pointer not allocated by this context");
    (void)hipStreamSynchronize(ctx->stream);
    ctx->bytes -= it->second;
    ctx->allocs.erase(it);
    return hip_check(ctx, hipFree(d), "esdg_free");
}

int esdg_memcpy(esdg_ctx* ctx, void* dst, const void* src, size_t bytes, int kind) {
    ESDG_REQUIRE(ctx, ctx && dst && src, "esdg_memcpy: null argument");
    hipMemcpyKind k = kind == ESDG_H2D ? hipMemcpyHostToDevice : kind == ESDG_D2H ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    ESDG_REQUIRE(ctx, kind >= 0 && kind <= 2, "esdg_memcpy: bad copy kind");
    int rc = hip_check(ctx, hipMemcpyAsync(dst, src, bytes, k, ctx->stream), "esdg_memcpy");
    if (rc != ESDG_OK) return rc;
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "esdg_memcpy: sync");
}

int esdg_fill(esdg_ctx* ctx, double* d, double value, size_t count) {
    ESDG_REQUIRE(ctx, ctx && d, "esdg_fill: null argument");
    if (count == 0) return ESDG_OK;
    KernelScope ks(ctx, "fill");
    unsigned nb = (unsigned)std::min<size_t>((count + 255) / 256, 8192);
    fill_kernel<<<nb, 256, 0, ctx->stream>>>(d, value, count);
    return launch_check(ctx, "fill");
}

int esdg_finish(esdg_ctx* ctx) {
    ESDG_REQUIRE(ctx, ctx, "esdg_finish: null context");
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "esdg_finish");
}

size_t esdg_bytes_allocated(const esdg_ctx* ctx) { return ctx ? ctx->bytes : 0; }

const char* esdg_last_error(const esdg_ctx* ctx) { return ctx ? ctx->err.c_str() : g_last_error_noctx.c_str(); }

void* esdg_stream(esdg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int esdg_profile_enable(esdg_ctx* ctx, int on) {
    ESDG_REQUIRE(ctx, ctx, "esdg_profile_enable: null context");
    ctx->profile = on != 0;
    return ESDG_OK;
}

int esdg_profile_reset(esdg_ctx* ctx) {
    ESDG_REQUIRE(ctx, ctx, "esdg_profile_reset: null context");
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->prof)
        for (auto& pe : kv.second.pending) { ctx->event_pool.push_back(pe.first); ctx->event_pool.push_back(pe.second); }
    ctx->prof.clear();
    return ESDG_OK;
}

int esdg_profile_query(esdg_ctx* ctx, const char* name, int64_t* launches, double* total_ms) {
    ESDG_REQUIRE(ctx, ctx && name && launches && total_ms, "esdg_profile_query: null argument");
    int rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "esdg_profile_query: sync");
    if (rc != ESDG_OK) return rc;
    auto it = ctx->prof.find(name);
    if (it == ctx->prof.end()) { *launches = 0; *total_ms = 0.0; return ESDG_OK; }
    auto& pr = it->second;
    for (auto& pe : pr.pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pe.first, pe.second) == hipSuccess) pr.ms += ms;
        ctx->event_pool.push_back(pe.first);
        ctx->event_pool.push_back(pe.second);
    }
    pr.pending.clear();
    *launches = pr.launches;
    *total_ms = pr.ms;
    return ESDG_OK;
}

int esdg_reduce_max(esdg_ctx* ctx, int n, const double* x, double* d_out) {
    ESDG_REQUIRE(ctx, ctx && x && d_out && n >= 0, "esdg_reduce_max: bad argument");
    KernelScope ks(ctx, "reduce_max");
    int nb = (int)std::min<int64_t>(std::max<int64_t>(1, ((int64_t)n + 255) / 256), 1024);
    max_partial_kernel<<<nb, 256, 0, ctx->stream>>>(x, n, ctx->d_scratch);
    max_partial_kernel<<<1, 256, 0, ctx->stream>>>(ctx->d_scratch, nb, d_out);
    return launch_check(ctx, "reduce_max");
}

int esdg_reduce_max_noalloc(void* stream, int n, const double* x, double* d_out) {
    if (!x || !d_out || n < 0) return set_error(nullptr, ESDG_ERR_ARG, "esdg_reduce_max_noalloc: bad argument");
    if (n > ESDG_NOALLOC_MAX_N)
        return set_error(nullptr, ESDG_ERR_ARG,
                         "esdg_reduce_max_noalloc: n = " + std::to_string(n) + " exceeds the maximum " +
                             std::to_string(ESDG_NOALLOC_MAX_N) + " of its static scratch");
    std::string err;
    double* part = noalloc_scratch(&err);
    if (!part) return set_error(nullptr, ESDG_ERR_HIP, "esdg_reduce_max_noalloc: " + err);
    // every partial the second pass reads is written by the first (nb blocks, one partial each),
    // so nothing left by an earlier call is read and nothing needs clearing
    const int nb = std::max(1, (n + 255) / 256);
    hipStream_t s = static_cast<hipStream_t>(stream);
    max_partial_kernel<<<nb, 256, 0, s>>>(x, n, part);
    max_partial_kernel<<<1, 256, 0, s>>>(part, nb, d_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(nullptr, ESDG_ERR_HIP, std::string("esdg_reduce_max_noalloc: launch: ") + hipGetErrorString(e));
    return ESDG_OK;
}

}  // extern "C"
