// Device helpers shared by the HIP kernels: wavefront lane exchange (DPP),
// an LDS-only work-group barrier and raw buffer loads / stores whose
// out-of-range offset makes a lane's access a no-op.
#pragma once

#include <hip/hip_runtime.h>

namespace frh {

__device__ inline double wave_from_prev_lane(double x)  // lane i <- lane i-1, lane 0 <- 0
{
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xf, 0xf, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ inline double wave_from_next_lane(double x)  // lane i <- lane i+1, lane 63 <- 0
{
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xf, 0xf, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Raw buffer access (CDNA buffer resource: base, no stride, size in bytes).
// A lane whose byte offset is at or beyond the size reads 0 and its store is
// dropped, without a memory access and without a branch around it.
constexpr unsigned kOutOfRange = 0x80000000u;

__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}

// Cache policy of a raw buffer access: 0, or kNT (non-temporal: data that
// streams through once is not kept in the caches).
constexpr int kNT = 2;

template <int AUX = 0>
__device__ inline double buf_load_f64(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, AUX));
}

// two consecutive float64 values in one 16-byte load
struct F64x2 {
    double lo, hi;
};

__device__ inline F64x2 buf_load_f64x2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(F64x2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}

template <int AUX = 0>
__device__ inline void buf_store_f64(double x, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    using u2 = __attribute__((ext_vector_type(2))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x), r, (int)voff, (int)soff, AUX);
}

// two consecutive float64 values in one 16-byte store
__device__ inline void buf_store_f64x2(F64x2 x, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    using u4 = __attribute__((ext_vector_type(4))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, x), r, (int)voff, (int)soff, 0);
}

// byte offset of cell (r, d - r) of the skewed plane, or kOutOfRange when
// the lane has no cell on diagonal d
__device__ inline unsigned slot_offset(bool row_ok, int d, int r, int w, int hs)
{
    const bool ok = row_ok && (unsigned)(d - r) < (unsigned)w;
    return ok ? ((unsigned)d * (unsigned)hs + (unsigned)r) * 8u : kOutOfRange;
}

// Chains of work-groups.  Consecutive sweeps of a pass (or their adjoints)
// can run as one work-group per sweep and frame, each stage consuming its
// predecessor's output while that is still being produced: the producer
// publishes how many steps it has completed in an agent-scope flag (a
// release store every few steps, after its waves' stores have completed),
// the consumer waits for the steps it needs before it loads their results,
// and acquires.  Stages are numbered stage-major, so every producer has a
// lower work-group index than its consumers and is dispatched first; frames
// are padded to a multiple of 8 so that a frame's stages share an XCD
// (work-group index modulo 8) and its L2.  A wait is bounded: one that gives
// up sets *err and proceeds (its results are then invalid), so a launch
// always ends.
constexpr int kChainFlagStride = 64;  // ints: one flag per 256 bytes
constexpr unsigned kChainSpin = 1u << 18;

// Steps between two publications: each costs the producer a write-back of
// its L2, so fewer are faster, but each adds to the lag between stages;
// measured at 32 frames (512 x 512: 32 best, 8 takes 1.3x as long).
__host__ __device__ constexpr int chain_publish_interval(int nd)
{
    return nd / 32 < 4 ? 4 : nd / 32 > 32 ? 32 : nd / 32;
}

// frames padded to a multiple of 8 (the XCD count)
__host__ __device__ constexpr int chain_tpad(int T) { return (T + 7) / 8 * 8; }

// Waits until the producer's flag is at least need (pre: the flag as
// loaded one step earlier) and acquires; returns the flag value seen.
__device__ inline int chain_acquire(int *flag, int need, int pre, int *err)
{
    int v = pre;
    if (v < need) {
        unsigned it = 0;
        for (;;) {
            v = __builtin_amdgcn_readfirstlane(__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (v >= need || ++it >= kChainSpin) break;
            __builtin_amdgcn_s_sleep(2);
        }
        if (v < need) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            v = 0x7fffffff;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return v;
}

// s_waitcnt vmcnt(0): this wave's stores have been performed
__device__ inline void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

}  // namespace frh
