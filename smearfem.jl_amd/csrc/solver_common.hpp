// solver_common.hpp -- device building blocks shared by the single-system
// PCG (pcg.hip) and the batched persistent PCG (batch.hip): coherent loads,
// the last-block reduction handshake, and the CSR / node-block sliced row
// engines.
#pragma once

#include "sfem_internal.hpp"

namespace sfem {
namespace {

constexpr double kEps64 = 2.220446049250313e-16;  // FP64 unit roundoff x 2

// Global-address-space view of a device pointer.  Pointers read from memory
// (the batch descriptors) are generic to the compiler, which then emits flat
// loads; accessing through these views gives global loads and stores.
#define SFEM_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const SFEM_GLOBAL T* glb(const T* p) { return (const SFEM_GLOBAL T*)p; }
template <typename T>
__device__ __forceinline__ SFEM_GLOBAL T* glb(T* p) { return (SFEM_GLOBAL T*)p; }

constexpr int kThreads = 256;

// ------------------------------------------------------------ helpers --
__device__ __forceinline__ double ld_coherent(const double* p) {
    unsigned long long v =
        __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __longlong_as_double((long long)v);
}

// Agent-coherent store of a block's partial result: it bypasses the XCD's L2
// (per-XCD L2s are not coherent with each other) so the last block, on any
// XCD, reads it with ld_coherent.
__device__ __forceinline__ void st_coherent(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until this wave's outstanding vector memory operations, stores
// included, have completed (GFX9 s_waitcnt: vmcnt(0), expcnt and lgkmcnt
// left at their maxima).
__device__ __forceinline__ void wait_vm_complete() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// True in every thread of the block that arrives last.  Thread 0 must have
// written this block's partials with st_coherent before the call; it waits for
// those stores to complete before its counter increment.  Unlike
// __threadfence() this does not write back and invalidate the whole L2, which
// would evict the vector every other block is still gathering from.
// The arrivals are spread over 8 counters: block b arrives on counter b % 8,
// the last block of each of those groups then on the top counter.  One
// agent-scope counter for every block serialises the arrivals (about 12 ns
// each on MI355X, 50 us for 4,096 blocks); spread over 8 counters 256 bytes
// apart they cost about 1.3 ns each.  The last block resets every counter.
__device__ __forceinline__ bool last_block_arrive(Arrivals* a) {
    __shared__ int is_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        wait_vm_complete();
        const unsigned G = gridDim.x, g = blockIdx.x % 8u;
        const unsigned groups = G < 8u ? G : 8u, in_group = (G - g + 7u) / 8u;
        bool last = __hip_atomic_fetch_add(&a->c[64 * g], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                    in_group - 1u;
        if (last)
            last = __hip_atomic_fetch_add(&a->c[64 * 8], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                   groups - 1u;
        if (last)
            for (int k = 0; k < 9; ++k)
                __hip_atomic_store(&a->c[64 * k], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        is_last = last;
    }
    __syncthreads();
    return is_last;
}

// Fixed-order sums of nparts partial W-tuples (stride W) by one block:
// thread t adds parts t, t + 256, t + 512, ... in that order (component by
// component), then the block sums.  The coherent loads go out U parts at a
// time (each one a round trip past the L2; at most 8 values in flight, so
// the kernels that end in this reduction keep their register budget), the
// additions stay in order.
template <int W>
__device__ __forceinline__ void thread_partials(const double* partials, int nparts, double (&s)[W]) {
    constexpr int U = W >= 8 ? 1 : 8 / W;
#pragma unroll
    for (int k = 0; k < W; ++k) s[k] = 0.0;
    int i = threadIdx.x;
    for (; i + (U - 1) * kThreads < nparts; i += U * kThreads) {
        double v[U][W];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < W; ++k) v[u][k] = ld_coherent(partials + (int64_t)W * (i + u * kThreads) + k);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int k = 0; k < W; ++k) s[k] += v[u][k];
    }
    for (; i < nparts; i += kThreads)
#pragma unroll
        for (int k = 0; k < W; ++k) s[k] += ld_coherent(partials + (int64_t)W * i + k);
}

// Fixed-order sum of nparts partial quadruples (stride 4) by one block.
__device__ __forceinline__ void sum_partials4(const double* partials, int nparts, double out[4], double* sh) {
    double s[4];
    thread_partials<4>(partials, nparts, s);
    out[0] = block_sum_256(s[0], sh);
    out[1] = block_sum_256(s[1], sh + 4);
    out[2] = block_sum_256(s[2], sh);
    out[3] = block_sum_256(s[3], sh + 4);
}

// Fixed-order sum of nparts partial triples (stride 3) by one block.
__device__ __forceinline__ void sum_partials3(const double* partials, int nparts, double* a, double* b, double* c,
                                              double* sh) {
    double s[3];
    thread_partials<3>(partials, nparts, s);
    *a = block_sum_256(s[0], sh);
    *b = block_sum_256(s[1], sh + 4);
    *c = block_sum_256(s[2], sh);
}

// Logical block id: consecutive logical blocks land on the same XCD (the
// hardware deals workgroups round-robin over the 8 XCDs), so each XCD's L2
// sees a contiguous part of the rows and of the gathered vector.
__device__ __forceinline__ int64_t logical_block() {
    const int64_t g = gridDim.x, b = blockIdx.x;
    if (g % 8) return b;
    return (b % 8) * (g / 8) + b / 8;
}

template <typename T>
__device__ __forceinline__ T fmaT(T a, T b, T c) { return fma(a, b, c); }

// ------------------------------------------------------- operator rows --
// CSR, G lanes per row.  Returns the row sum in every lane of the group.
template <int G, typename TV, typename ACC, typename Gather>
__device__ __forceinline__ ACC csr_row(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                       const TV* __restrict__ val, int64_t row, int lane, Gather gx) {
    ACC s = 0;
    const int32_t e = rowptr[row + 1];
#pragma unroll 4  // the loads of four strides in flight; the same order of operations
    for (int32_t k = rowptr[row] + lane; k < e; k += G) s = fmaT<ACC>((ACC)val[k], gx(col[k]), s);
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// csr_row plus sum_k |val_k x_k| (the scale of the rounding error of the row
// sum, for the attainable-residual estimate).
template <int G, typename Gather>
__device__ __forceinline__ double csr_row_abs(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                              const double* __restrict__ val, int64_t row, int lane, Gather gx,
                                              double* aout) {
    double s = 0.0, a = 0.0;
    const int32_t e = rowptr[row + 1];
#pragma unroll 4
    for (int32_t k = rowptr[row] + lane; k < e; k += G) {
        const double v = val[k], x = gx(col[k]);
        s = fma(v, x, s);
        a = fma(fabs(v), fabs(x), a);
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        a += __shfl_xor(a, off, 64);
    }
    *aout = a;
    return s;
}

// Matrix streams are read once per SpMV: with NT the loads carry the
// non-temporal hint so they do not evict the gathered vector from the caches.
template <bool NT, typename V>
__device__ __forceinline__ V ld_stream(const V* p) {
    if constexpr (NT) return __builtin_nontemporal_load(glb(p));
    else return *glb(p);
}

// Operator value sources of the sliced format.  Vals<double> / Vals<float>
// read one plane array; Vals<SplitVal> reads the FP32 rounding of each FP64
// value plus a 16-bit correction and rebuilds v = hi + lo * ulp(hi) * 2^-16
// exactly in FP64 (|v' - v| <= 2^-39 |v|): 6 bytes per value.
struct SplitVal {};

// `raw` loads a value as stored and `decode` turns it into the arithmetic
// type, so a caller can issue many loads before it uses any of them.
template <typename TV>
struct Vals {
    const TV* v;
    using Raw = TV;
    template <bool NT>
    __device__ __forceinline__ Raw raw(int64_t i) const { return ld_stream<NT>(v + i); }
    __device__ __forceinline__ TV decode(Raw r) const { return r; }
    template <bool NT>
    __device__ __forceinline__ TV get(int64_t i) const { return ld_stream<NT>(v + i); }
};

template <>
struct Vals<SplitVal> {
    const float* hi;
    const int16_t* lo;
    struct Raw {
        float h;
        int l;
    };
    template <bool NT>
    __device__ __forceinline__ Raw raw(int64_t i) const {
        return Raw{ld_stream<NT>(hi + i), (int)ld_stream<NT>(lo + i)};
    }
    __device__ __forceinline__ double decode(Raw r) const {
        // ulp(h) * 2^-16 = 2^(exponent(h) - 23 - 16); below the normal range
        // (|h| < 2^-87, where lo is 0 by construction) the scale is 0
        const float scale = __int_as_float(max((__float_as_int(r.h) & 0x7F800000) - (39 << 23), 0));
        return fma((double)r.l, (double)scale, (double)r.h);
    }
    template <bool NT>
    __device__ __forceinline__ double get(int64_t i) const { return decode(raw<NT>(i)); }
};

// Node-block sliced format: this lane's block row of slice `sl`.
template <int BS, typename TV, typename ACC, bool NT = false, typename Gather>
__device__ __forceinline__ void sell_row(const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w,
                                         const int32_t* __restrict__ scol, Vals<TV> vals, int64_t sl, int lane,
                                         Gather gx, ACC out[BS]) {
    constexpr int B2 = BS * BS;
    const int64_t base = glb(sl_ptr)[sl];
    const int w = glb(sl_w)[sl];
#pragma unroll
    for (int j = 0; j < BS; ++j) out[j] = 0;
#pragma unroll 2
    for (int k = 0; k < w; ++k) {
        const int64_t slot = base + (int64_t)k * 64;
        const int64_t c = ld_stream<NT>(scol + slot + lane);
        ACC xv[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) xv[i] = gx(c * BS + i);
        const int64_t v = slot * B2 + lane;
#pragma unroll
        for (int j = 0; j < BS; ++j)
#pragma unroll
            for (int i = 0; i < BS; ++i)
                out[j] = fmaT<ACC>((ACC)vals.template get<NT>(v + (j * BS + i) * 64), xv[i], out[j]);
    }
}

// sell_row with explicit staging, for kernels whose register budget is set
// elsewhere (the persistent solvers): per pair of block columns, the column
// ids and all value loads are issued before any of them is used, then the
// gathers, then the arithmetic.  Scheduling barriers keep the compiler from
// interleaving loads with their uses, which would leave one load in flight
// per wavefront.  Same arithmetic, in the same order, as sell_row.
template <int BS, typename TV, typename ACC, bool NT = false, typename Gather>
__device__ __forceinline__ void sell_row_staged(const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w,
                                                const int32_t* __restrict__ scol, Vals<TV> vals, int64_t sl,
                                                int lane, Gather gx, ACC out[BS]) {
    constexpr int B2 = BS * BS;
    using Raw = typename Vals<TV>::Raw;
    const int64_t base = glb(sl_ptr)[sl];
    const int w = glb(sl_w)[sl];
#pragma unroll
    for (int j = 0; j < BS; ++j) out[j] = 0;
    int k = 0;
    for (; k + 2 <= w; k += 2) {
        const int64_t s0 = base + (int64_t)k * 64, s1 = s0 + 64;
        const int64_t c0 = ld_stream<NT>(scol + s0 + lane), c1 = ld_stream<NT>(scol + s1 + lane);
        Raw r0[B2], r1[B2];
#pragma unroll
        for (int m = 0; m < B2; ++m) r0[m] = vals.template raw<NT>(s0 * B2 + lane + m * 64);
#pragma unroll
        for (int m = 0; m < B2; ++m) r1[m] = vals.template raw<NT>(s1 * B2 + lane + m * 64);
        __builtin_amdgcn_sched_barrier(0);
        ACC x0[BS], x1[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) x0[i] = gx(c0 * BS + i);
#pragma unroll
        for (int i = 0; i < BS; ++i) x1[i] = gx(c1 * BS + i);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < BS; ++j)
#pragma unroll
            for (int i = 0; i < BS; ++i) out[j] = fmaT<ACC>((ACC)vals.decode(r0[j * BS + i]), x0[i], out[j]);
#pragma unroll
        for (int j = 0; j < BS; ++j)
#pragma unroll
            for (int i = 0; i < BS; ++i) out[j] = fmaT<ACC>((ACC)vals.decode(r1[j * BS + i]), x1[i], out[j]);
    }
    if (k < w) {
        const int64_t s0 = base + (int64_t)k * 64;
        const int64_t c0 = ld_stream<NT>(scol + s0 + lane);
        Raw r0[B2];
#pragma unroll
        for (int m = 0; m < B2; ++m) r0[m] = vals.template raw<NT>(s0 * B2 + lane + m * 64);
        ACC x0[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) x0[i] = gx(c0 * BS + i);
#pragma unroll
        for (int j = 0; j < BS; ++j)
#pragma unroll
            for (int i = 0; i < BS; ++i) out[j] = fmaT<ACC>((ACC)vals.decode(r0[j * BS + i]), x0[i], out[j]);
    }
}

// sell_row in FP64 plus aout[j] = sum |M_ji x_i| over the row's blocks,
// staged like sell_row_staged (pairs of block columns; the arithmetic order
// per row is k, then j, then i).
template <int BS, typename Gather>
__device__ __forceinline__ void sell_row_abs(const int64_t* __restrict__ sl_ptr, const int32_t* __restrict__ sl_w,
                                             const int32_t* __restrict__ scol, const double* __restrict__ sval,
                                             int64_t sl, int lane, Gather gx, double out[BS], double aout[BS]) {
    constexpr int B2 = BS * BS;
    const int64_t base = glb(sl_ptr)[sl];
    const int w = glb(sl_w)[sl];
#pragma unroll
    for (int j = 0; j < BS; ++j) out[j] = aout[j] = 0.0;
    auto step = [&](const double (&m)[B2], const double (&xv)[BS]) {
#pragma unroll
        for (int j = 0; j < BS; ++j)
#pragma unroll
            for (int i = 0; i < BS; ++i) {
                out[j] = fma(m[j * BS + i], xv[i], out[j]);
                aout[j] = fma(fabs(m[j * BS + i]), fabs(xv[i]), aout[j]);
            }
    };
    int k = 0;
    for (; k + 2 <= w; k += 2) {
        const int64_t s0 = base + (int64_t)k * 64, s1 = s0 + 64;
        const int64_t c0 = ld_stream<true>(scol + s0 + lane), c1 = ld_stream<true>(scol + s1 + lane);
        double m0[B2], m1[B2];
#pragma unroll
        for (int m = 0; m < B2; ++m) m0[m] = ld_stream<true>(sval + s0 * B2 + lane + m * 64);
#pragma unroll
        for (int m = 0; m < B2; ++m) m1[m] = ld_stream<true>(sval + s1 * B2 + lane + m * 64);
        __builtin_amdgcn_sched_barrier(0);
        double x0[BS], x1[BS];
#pragma unroll
        for (int i = 0; i < BS; ++i) x0[i] = gx(c0 * BS + i);
#pragma unroll
        for (int i = 0; i < BS; ++i) x1[i] = gx(c1 * BS + i);
        __builtin_amdgcn_sched_barrier(0);
        step(m0, x0);
        step(m1, x1);
    }
    if (k < w) {
        const int64_t s0 = base + (int64_t)k * 64;
        const int64_t c0 = ld_stream<true>(scol + s0 + lane);
        double m0[B2], x0[BS];
#pragma unroll
        for (int m = 0; m < B2; ++m) m0[m] = ld_stream<true>(sval + s0 * B2 + lane + m * 64);
#pragma unroll
        for (int i = 0; i < BS; ++i) x0[i] = gx(c0 * BS + i);
        step(m0, x0);
    }
}

}  // namespace
}  // namespace sfem
