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

// byte offset of cell (r, d - r) of the skewed plane, or kOutOfRange when
// the lane has no cell on diagonal d
__device__ inline unsigned slot_offset(bool row_ok, int d, int r, int w, int hs)
{
    const bool ok = row_ok && (unsigned)(d - r) < (unsigned)w;
    return ok ? ((unsigned)d * (unsigned)hs + (unsigned)r) * 8u : kOutOfRange;
}

}  // namespace frh
