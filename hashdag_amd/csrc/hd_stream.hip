// Streaming kernels that measure the achievable HBM bandwidth on the device the tracer runs on:
// a copy (read n bytes, write n bytes) and a read-only pass (read n bytes, fold them into one
// word per block so the loads are not dead). Each thread keeps four 16-byte loads in flight per
// step (blocks of 256 threads own contiguous runs of 4 x 4 KiB); the loads and stores are
// non-temporal, as the data is touched once.
#include <hip/hip_runtime.h>

#include "hd_ctx.h"

namespace {
constexpr uint32_t kUnroll = 4;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_stream_copy(const u32x4* __restrict__ src, u32x4* __restrict__ dst,
                                                     uint64_t n16) {
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x * kUnroll;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x * kUnroll + threadIdx.x; base < n16; base += step) {
        u32x4 v[kUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            const uint64_t i = base + (uint64_t)u * blockDim.x;
            if (i < n16) v[u] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            const uint64_t i = base + (uint64_t)u * blockDim.x;
            if (i < n16) __builtin_nontemporal_store(v[u], dst + i);
        }
    }
}

__global__ __launch_bounds__(256) void k_stream_read(const u32x4* __restrict__ src, uint64_t n16,
                                                     uint32_t* __restrict__ out) {
    __shared__ uint32_t sAcc[4];
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x * kUnroll;
    uint32_t acc = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x * kUnroll + threadIdx.x; base < n16; base += step) {
        u32x4 v[kUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) {
            const uint64_t i = base + (uint64_t)u * blockDim.x;
            v[u] = i < n16 ? __builtin_nontemporal_load(src + i) : u32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (uint32_t u = 0; u < kUnroll; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc ^= __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63u) == 0) sAcc[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = sAcc[0] ^ sAcc[1] ^ sAcc[2] ^ sAcc[3];
}

constexpr uint32_t kStreamBlocks = 256 * 16;  // 16 blocks per CU on 256 CUs
}  // namespace

extern "C" hd_status hd_stream_copy(const void* d_src, void* d_dst, uint64_t bytes, void* stream) {
    if (!d_src || !d_dst || bytes % 16) return hd::fail(nullptr, HD_ERR_INVALID, "stream copy needs 16-byte multiples");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    k_stream_copy<<<kStreamBlocks, 256, 0, s>>>(reinterpret_cast<const u32x4*>(d_src), reinterpret_cast<u32x4*>(d_dst),
                                               bytes / 16);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hd::hip_fail(nullptr, e, "stream copy");
    return HD_OK;
}

extern "C" hd_status hd_stream_read(const void* d_src, uint64_t bytes, void* d_out, void* stream) {
    if (!d_src || !d_out || bytes % 16) return hd::fail(nullptr, HD_ERR_INVALID, "stream read needs 16-byte multiples");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    k_stream_read<<<kStreamBlocks, 256, 0, s>>>(reinterpret_cast<const u32x4*>(d_src), bytes / 16,
                                               reinterpret_cast<uint32_t*>(d_out));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hd::hip_fail(nullptr, e, "stream read");
    return HD_OK;
}
