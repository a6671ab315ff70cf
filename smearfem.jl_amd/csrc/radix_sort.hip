// radix_sort.hip -- device-wide exclusive scans and a stable LSD radix sort of
// (uint64 key, uint32 value) pairs for gfx950.
//
// The sort stands in for the counting sort inside Julia's sparse(I,J,V)
// (called at smearFEM src/fem.jl:253): stability keeps equal keys in input
// (COO index) order, which is the order the duplicates are summed in.
//
// Each 8-bit pass is reduce-then-scan:
//   hist    : per-tile digit counts, stored digit-major [256][ntiles]
//   scan    : one exclusive scan of that array = global start of (digit, tile)
//   scatter : stable rank inside the tile with 64-lane ballots (no block-wide
//             barriers in the ranking loop), LDS staging, then stores that are
//             contiguous within each digit run.
#include "sfem_objects.hpp"

namespace sfem {

namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 16;                    // per thread
constexpr int kScanTile = kBlock * kScanItems;    // 4096 values per block

// Exclusive scan across a 256-thread block of one value per thread; returns
// the exclusive prefix and writes the block total to *total.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* sh, T* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    T wbase = 0;
    for (int i = 0; i < w; ++i) wbase += sh[i];
    *total = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return wbase + inc - v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void scan_reduce_kernel(const T* __restrict__ in, int64_t n,
                                                             T* __restrict__ partial) {
    __shared__ T sh[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    T s = 0;
#pragma unroll 4
    for (int it = 0; it < kScanItems; ++it) {
        const int64_t i = base + (int64_t)it * kBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    const T tot = block_sum_256(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Each thread owns kScanItems consecutive values; `offsets` (may be null) adds
// the scanned block offset.
template <typename T>
__global__ __launch_bounds__(kBlock) void scan_tile_kernel(const T* in, T* out, int64_t n,
                                                           const T* __restrict__ offsets) {
    __shared__ T sh[4];
    __shared__ T stile[kScanTile];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    for (int it = 0; it < kScanItems; ++it) {  // coalesced load to LDS
        const int64_t i = base + (int64_t)it * kBlock + threadIdx.x;
        stile[it * kBlock + threadIdx.x] = i < n ? in[i] : T(0);
    }
    __syncthreads();
    T v[kScanItems];
    T s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = stile[threadIdx.x * kScanItems + k];
        s += v[k];
    }
    T total;
    T run = block_exclusive_scan(s, sh, &total);
    if (offsets) run += offsets[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        stile[threadIdx.x * kScanItems + k] = run;
        run += v[k];
    }
    __syncthreads();
    for (int it = 0; it < kScanItems; ++it) {
        const int64_t i = base + (int64_t)it * kBlock + threadIdx.x;
        if (i < n) out[i] = stile[it * kBlock + threadIdx.x];
    }
}

// Scans in stream order without a host sync; the total goes to d_total when
// it is non-null.  in == out is allowed.
template <typename T>
void scan_dev(sfem_ctx* ctx, const T* in, T* out, int64_t n, T* d_total) {
    if (n <= 0) return;
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 1) {
        if (d_total) {
            hipLaunchKernelGGL(scan_reduce_kernel<T>, dim3(1), dim3(kBlock), 0, ctx->stream, in, n, d_total);
            SFEM_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(scan_tile_kernel<T>, dim3(1), dim3(kBlock), 0, ctx->stream, in, out, n,
                           (const T*)nullptr);
        SFEM_LAUNCH_CHECK();
        return;
    }
    DevBuf<T> partial(ctx, (size_t)nb);
    hipLaunchKernelGGL(scan_reduce_kernel<T>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, in, n,
                       partial.get());
    SFEM_LAUNCH_CHECK();
    scan_dev<T>(ctx, partial.get(), partial.get(), nb, d_total);
    hipLaunchKernelGGL(scan_tile_kernel<T>, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, in, out, n,
                       (const T*)partial.get());
    SFEM_LAUNCH_CHECK();
}

template <typename T>
T scan_impl(sfem_ctx* ctx, const T* in, T* out, int64_t n) {
    if (n <= 0) return 0;
    DevBuf<T> tot(ctx, 1);
    scan_dev<T>(ctx, in, out, n, tot.get());
    T h = 0;
    d2h(ctx, &h, tot.get(), sizeof(T));
    return h;
}

// ---------------------------------------------------------------- radix --
constexpr int kRadixItems = 16;
constexpr int kWaves = kBlock / 64;
constexpr int kWaveTile = 64 * kRadixItems;     // 1024 keys per wavefront
constexpr int kRadixTile = kBlock * kRadixItems; // 4096 keys per block

__global__ __launch_bounds__(kBlock) void radix_hist_kernel(const uint64_t* __restrict__ keys,
                                                            int64_t n, int shift,
                                                            uint32_t* __restrict__ hist,
                                                            int64_t ntiles) {
    __shared__ uint32_t h[kWaves][256];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < kWaves * 256; i += kBlock) (&h[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRadixTile + (int64_t)w * kWaveTile;
#pragma unroll 4
    for (int it = 0; it < kRadixItems; ++it) {
        const int64_t i = base + it * 64 + lane;
        if (i < n) atomicAdd(&h[w][(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    const int d = threadIdx.x;
    hist[(int64_t)d * ntiles + blockIdx.x] = (h[0][d] + h[1][d]) + (h[2][d] + h[3][d]);
}

__global__ __launch_bounds__(kBlock) void radix_scatter_kernel(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, const uint32_t* __restrict__ offs,
    int64_t ntiles) {
    __shared__ uint32_t wcnt[kWaves][256];
    __shared__ uint32_t bstart[256];
    __shared__ uint32_t gstart[256];
    __shared__ uint32_t ssum[4];
    __shared__ uint64_t skey[kRadixTile];
    __shared__ uint32_t sval[kRadixTile];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t tile0 = (int64_t)blockIdx.x * kRadixTile;
    for (int i = tid; i < kWaves * 256; i += kBlock) (&wcnt[0][0])[i] = 0;
    gstart[tid] = offs[(int64_t)tid * ntiles + blockIdx.x];
    __syncthreads();

    uint64_t k[kRadixItems];
    uint32_t v[kRadixItems];
    uint32_t rank[kRadixItems];
    const int64_t base = tile0 + (int64_t)w * kWaveTile;
#pragma unroll
    for (int it = 0; it < kRadixItems; ++it) {
        const int64_t i = base + it * 64 + lane;
        const bool ok = i < n;
        k[it] = ok ? kin[i] : 0ull;
        v[it] = ok ? vin[i] : 0u;
    }
    // Stable rank within the wavefront's chunk: items are visited in input
    // order (it, lane), lanes with the same digit are found by 8 ballots.
#pragma unroll
    for (int it = 0; it < kRadixItems; ++it) {
        const bool ok = base + it * 64 + lane < n;
        const uint32_t d = (uint32_t)(k[it] >> shift) & 255u;
        uint64_t m = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t below = (uint32_t)__popcll(m & lanemask_lt());
        const uint32_t cnt = (uint32_t)__popcll(m);
        uint32_t before = 0;
        if (ok) before = wcnt[w][d];
        rank[it] = before + below;
        if (ok && below == 0) wcnt[w][d] = before + cnt;
    }
    __syncthreads();
    {  // per digit: wavefront prefix, then the block-local digit starts
        const int d = tid;
        uint32_t run = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ++ww) {
            const uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = run;
            run += c;
        }
        uint32_t total;
        bstart[d] = block_exclusive_scan(run, ssum, &total);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kRadixItems; ++it) {
        if (base + it * 64 + lane < n) {
            const uint32_t d = (uint32_t)(k[it] >> shift) & 255u;
            const uint32_t pos = bstart[d] + wcnt[w][d] + rank[it];
            skey[pos] = k[it];
            sval[pos] = v[it];
        }
    }
    __syncthreads();
    const int cnt = (int)min((int64_t)kRadixTile, n - tile0);
    for (int i = tid; i < cnt; i += kBlock) {
        const uint64_t key = skey[i];
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        const int64_t o = (int64_t)gstart[d] + (i - (int64_t)bstart[d]);
        kout[o] = key;
        vout[o] = sval[i];
    }
}

}  // namespace

uint64_t exclusive_scan_u32(sfem_ctx* ctx, const uint32_t* in, uint32_t* out, int64_t n) {
    return scan_impl<uint32_t>(ctx, in, out, n);
}

int64_t exclusive_scan_i64(sfem_ctx* ctx, const int64_t* in, int64_t* out, int64_t n) {
    return scan_impl<int64_t>(ctx, in, out, n);
}

void exclusive_scan_u32_dev(sfem_ctx* ctx, const uint32_t* in, uint32_t* out, int64_t n,
                            uint32_t* d_total) {
    scan_dev<uint32_t>(ctx, in, out, n, d_total);
}

void radix_sort_pairs(sfem_ctx* ctx, uint64_t** keys, uint64_t* keys_alt, uint32_t** vals,
                      uint32_t* vals_alt, int64_t n, int key_bits) {
    if (n <= 1 || key_bits <= 0) return;
    if (n >= (int64_t(1) << 32)) fail(SFEM_ERR_UNSUPPORTED, "radix sort: 2^32 or more pairs");
    const int64_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    DevBuf<uint32_t> hist(ctx, (size_t)(256 * ntiles));
    uint64_t* kin = *keys;
    uint64_t* kout = keys_alt;
    uint32_t* vin = *vals;
    uint32_t* vout = vals_alt;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)ntiles), dim3(kBlock), 0, ctx->stream, kin, n,
                           shift, hist.get(), ntiles);
        SFEM_LAUNCH_CHECK();
        scan_dev<uint32_t>(ctx, hist.get(), hist.get(), 256 * ntiles, nullptr);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3((unsigned)ntiles), dim3(kBlock), 0, ctx->stream, kin,
                           vin, kout, vout, n, shift, hist.get(), ntiles);
        SFEM_LAUNCH_CHECK();
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    *keys = kin;
    *vals = vin;
}

}  // namespace sfem
