// sfem_internal.hpp -- shared internals of the gfx950 assemble-and-solve library:
// context, caching device allocator, error plumbing, wave64 reduction helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "smearfem.h"

namespace sfem {

// ---------------------------------------------------------------- errors --
struct Error {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error{code, msg}; }

#define SFEM_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            ::sfem::fail(e_ == hipErrorOutOfMemory ? SFEM_ERR_MEMORY : SFEM_ERR_DEVICE, \
                         std::string(#call) + ": " + hipGetErrorString(e_));             \
    } while (0)

#define SFEM_LAUNCH_CHECK() SFEM_HIP(hipGetLastError())

// ----------------------------------------------------- caching allocator --
// Device memory is recycled by size class so repeated assemble/solve calls do
// not pay hipMalloc/hipFree.  All users of one context share its stream, so a
// block returned to the cache can be handed out again in stream order.
class Allocator {
public:
    explicit Allocator(int device) : device_(device) {}
    ~Allocator() { release_all(); }
    void* alloc(size_t bytes);
    void free(void* p);
    void release_cached();
    void release_all();
    int64_t bytes_in_use() const { return in_use_; }
    int64_t bytes_held() const { return held_; }
    void set_limit(int64_t limit) { limit_ = limit; }

private:
    static size_t size_class(size_t bytes);
    int device_;
    std::mutex mu_;
    std::multimap<size_t, void*> cache_;
    std::map<void*, size_t> live_;
    int64_t in_use_ = 0, held_ = 0, limit_ = -1;
};

}  // namespace sfem

// A context lives until sfem_ctx_destroy has been called and every object
// made on it (mesh, matrix, bc, system) is destroyed: each object holds a
// reference, so any destroy order is safe.
struct sfem_ctx {
    std::atomic<int> refs{1};
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    sfem::Allocator* alloc = nullptr;
    // Sliced-SpMV launch shape (measured on MI355X, 256x64x64 Hex8, FP64:
    // 392 us with both defaults vs 494 us with neither).  Overrides at
    // creation: SFEM_SPMV_NT=0/1, SFEM_SELL_WAVE_SLICES=n (0 = spread the
    // slices over at most 8 blocks per CU).
    bool tune_nt = true;        // non-temporal matrix loads
    int tune_wave_slices = 1;   // slices per wavefront
    // Grid barrier of the persistent solvers: workspace, wait limit (100 MHz
    // ticks; SFEM_BARRIER_TIMEOUT_US) and whether the per-XCD grouped form may
    // be used (SFEM_BARRIER_GROUPED=0 forces the per-block form).
    unsigned* bar_ws = nullptr;
    unsigned long long bar_timeout_ticks = 10ull * 100000000ull;
    bool bar_grouped = true;
    bool bar_last_grouped = false;  // form the last launch used
    // Device address of the no-allocation solve's workspace array on this
    // context's device (hipGetSymbolAddress at creation).
    void* scratch = nullptr;
    // Streams for independent work beside `stream` (default priority, made
    // on first use; joined back to `stream` by events before its results
    // are read or its buffers released).
    static constexpr int kSideStreams = 4;
    hipStream_t side[kSideStreams] = {};
};

namespace sfem {

void ctx_retain(sfem_ctx* ctx);
void ctx_release(sfem_ctx* ctx);
hipStream_t side_stream(sfem_ctx* ctx, int i);

// RAII device buffer backed by the context allocator.
template <typename T>
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(sfem_ctx* ctx, size_t n) { reset(ctx, n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ctx_(o.ctx_), p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); ctx_ = o.ctx_; p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void reset(sfem_ctx* ctx, size_t n) {
        release();
        ctx_ = ctx;
        n_ = n;
        p_ = n ? static_cast<T*>(ctx->alloc->alloc(n * sizeof(T))) : nullptr;
    }
    void release() {
        if (p_ && ctx_) ctx_->alloc->free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

private:
    sfem_ctx* ctx_ = nullptr;
    T* p_ = nullptr;
    size_t n_ = 0;
};

// host <-> device helpers on the context stream
void h2d(sfem_ctx* ctx, void* dst, const void* src, size_t bytes);
void d2h(sfem_ctx* ctx, void* dst, const void* src, size_t bytes);  // synchronous
void memset_d(sfem_ctx* ctx, void* dst, int v, size_t bytes);
void sync(sfem_ctx* ctx);

// Runs f, turning internal errors into a status code plus ctx->last_error.
template <typename F>
inline int guard(sfem_ctx* ctx, F&& f) {
    try {
        f();
        if (ctx) ctx->last_error.clear();
        return SFEM_OK;
    } catch (const Error& e) {
        if (ctx) ctx->last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_error = "host allocation failed";
        return SFEM_ERR_MEMORY;
    } catch (...) {
        if (ctx) ctx->last_error = "unexpected internal error";
        return SFEM_ERR_DEVICE;
    }
}


inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return static_cast<unsigned>(g);
}

// ------------------------------------------------- device-side helpers --
// Sums over a 64-lane wavefront by butterfly shuffles: every lane ends with
// the same bits, and the combination order is fixed, so results do not
// depend on timing.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block-wide sum of one value for a 256-thread block (4 wavefronts).  The
// result is valid in every thread.  `sh` needs 4 slots.
template <typename T>
__device__ __forceinline__ T block_sum_256(T v, T* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    T r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const unsigned lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

}  // namespace sfem
