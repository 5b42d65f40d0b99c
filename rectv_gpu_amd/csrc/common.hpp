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

}  // namespace rtv
