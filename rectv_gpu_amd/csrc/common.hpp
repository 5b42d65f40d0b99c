// Shared device/host helpers for the rectv HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace rtv {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define RTV_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::rtv::Error(e_ == hipErrorOutOfMemory ? -3 : -2,                       \
                               std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

#define RTV_REQUIRE(cond, msg)                                                            \
    do {                                                                                  \
        if (!(cond)) throw ::rtv::Error(-1, (msg));                                       \
    } while (0)

__host__ __device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__host__ __device__ inline float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap the hardware block index so that consecutive *logical* indices land on
// the same XCD; neighbouring tiles/rows then share that XCD's L2.
__device__ inline int xcd_remap(int p, int total) {
    const int nx = 8;
    int q = total / nx, r = total % nx;
    int xcd = p % nx, slot = p / nx;
    return (xcd < r) ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() also waits
// for every outstanding global load and store of the wave (vmcnt(0) on this
// architecture), which would stall on data that was prefetched into registers
// for a later phase; the LDS-scoped fences wait for LDS operations (lgkmcnt)
// alone.  For kernels whose waves exchange data through LDS only.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// v + (v of lane ^ 16), within each 32-lane half (ds_swizzle, no address VGPRs)
__device__ __forceinline__ float add_xor16(float v) {
    return v + __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x401F));
}
// v + (v of lane ^ 32); bp = 4 * (lane ^ 32), computed once by the caller
__device__ __forceinline__ float add_xor32(float v, int bp) {
    return v + __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(bp, __builtin_bit_cast(int, v)));
}

// Raw buffer access: a wave-uniform base in a buffer descriptor (SGPRs), a
// 32-bit per-lane byte offset and a wave-uniform byte offset.  Keeps per-lane
// address state at one VGPR where a loop walks many uniform sub-arrays.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p) {
    // dword3 0x00020000: untyped 32-bit access on gfx9-family parts; no range limit
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, -1, 0x00020000);
}
// (the b32 / b64 builtins move raw bits: reinterpret, never convert)
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
// (two 32-bit loads: the 64-bit raw buffer load builtin of this toolchain
// reads only the first dword)
__device__ __forceinline__ float2 buf_ld2(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return make_float2(buf_ld(r, voff, soff), buf_ld(r, voff + 4, soff));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
// Non-temporal stores of a kernel's output that no later access of the same kernel
// reads.  They pay off where whole cache lines are written (partial-line streams
// lose the merging in L2), so each call site is chosen by measurement.
typedef float v2f_nt __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt(float2 *p, float2 v) {
    __builtin_nontemporal_store((v2f_nt){v.x, v.y}, (v2f_nt *)p);
}
// Streaming (non-temporal) forms for data read or written once per kernel, so that
// the caches keep the data that neighbouring lanes and waves read again.
constexpr int CP_NT = 2;  // cache-policy bit NT of the buffer instructions
__device__ __forceinline__ float buf_ld_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, CP_NT));
}
__device__ __forceinline__ float2 buf_ld2_nt(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    return make_float2(buf_ld_nt(r, voff, soff), buf_ld_nt(r, voff + 4, soff));
}
__device__ __forceinline__ void buf_st_nt(float v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, CP_NT);
}

}  // namespace rtv
