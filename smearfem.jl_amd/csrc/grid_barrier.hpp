// Grid-wide barrier for the persistent solvers.
//
// The launch is cooperative, so every block of the grid is resident and a
// counter barrier cannot wait for a block that never runs.  MI355X has 8 XCDs
// with one L2 each; the hardware dispatches block b to XCD b % 8.  A barrier
// must make every block's stores visible in memory before any block reads
// past it:
//
//   grouped (the dispatch rule holds for every block of the launch): each wave
//     waits for its own stores to reach the L2, the block arrives on its XCD's
//     counter, and the last block of each XCD writes that L2 back once (one
//     agent-scope release per XCD instead of one per wave) before arriving on
//     the top counter;
//   per-block (fallback): thread 0 of every block writes the L2 back before it
//     arrives.
//
// The last block overall publishes the barrier's epoch; waiters acquire at
// agent scope (invalidating stale L1/L2 lines) once they see it.  Counters
// only grow within a launch (block b's group target for barrier k is
// k x group size), so nothing is reset between barriers; the host zeroes the
// workspace before each launch.  Waiting is bounded by the 100 MHz wall clock:
// a wait that outlasts the limit sets the abort flag, and every block returns
// from the kernel at its next barrier, so a broken barrier ends the launch
// with an error instead of a hung device.
#pragma once

#include <hip/hip_runtime.h>

namespace sfem {

constexpr int kBarXcds = 8;
constexpr int kBarLine = 64;  // words between counters (256 B)
constexpr int kBarTop = kBarXcds * kBarLine;
constexpr int kBarGen = (kBarXcds + 1) * kBarLine;
constexpr int kBarFlags = (kBarXcds + 2) * kBarLine;  // [0] a block off its XCD, [1] aborted
constexpr int kBarWords = (kBarXcds + 3) * kBarLine;

// XCC_ID hardware register, bits [3:0]: the XCD this wave runs on.
__device__ __forceinline__ unsigned xcd_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u; }

struct GridBarrier {
    unsigned* ws;
    unsigned long long timeout;  // wall-clock ticks (100 MHz)
    unsigned epoch;
    bool grouped;

    // First barrier of a launch: checks the block-to-XCD rule, then decides
    // (uniformly, from memory every block reads after the barrier) whether
    // the grouped form may be used.  Returns false if the launch aborted.
    __device__ bool start(bool allow_grouped) {
        epoch = 0;
        grouped = false;
        if (threadIdx.x == 0 && xcd_id() != blockIdx.x % kBarXcds)
            __hip_atomic_fetch_or(ws + kBarFlags, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!sync()) return false;
        __shared__ int s_grouped;
        if (threadIdx.x == 0)
            s_grouped = allow_grouped &&
                        __hip_atomic_load(ws + kBarFlags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
        __syncthreads();
        grouped = s_grouped != 0;
        return true;
    }

    __device__ bool sync() {
        ++epoch;
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores are in the L2
        __syncthreads();
        __shared__ int s_ok;
        if (threadIdx.x == 0) {
            unsigned* flags = ws + kBarFlags;
            bool ok = __hip_atomic_load(flags + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
            if (ok) {
                const unsigned G = gridDim.x;
                const unsigned ng = G < (unsigned)kBarXcds ? G : (unsigned)kBarXcds;
                const unsigned grp = blockIdx.x % ng;
                const unsigned gsize = (G - 1 - grp) / ng + 1;
                if (!grouped) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                const unsigned c =
                    __hip_atomic_fetch_add(ws + grp * kBarLine, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (c + 1 == epoch * gsize) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");  // this XCD's L2 to memory
                    const unsigned t =
                        __hip_atomic_fetch_add(ws + kBarTop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (t + 1 == epoch * ng) {
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                        __hip_atomic_store(ws + kBarGen, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                const unsigned long long t0 = wall_clock64();
                while (__hip_atomic_load(ws + kBarGen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
                    __builtin_amdgcn_s_sleep(2);
                    if (__hip_atomic_load(flags + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                        ok = false;
                        break;
                    }
                    if (wall_clock64() - t0 > timeout) {
                        __hip_atomic_store(flags + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = false;
                        break;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            s_ok = ok ? 1 : 0;
        }
        __syncthreads();
        return s_ok != 0;
    }
};

}  // namespace sfem
