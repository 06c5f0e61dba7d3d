// Streaming copy used to measure the achievable HBM bandwidth on the device the tracer
// runs on (read n words, write n words, 16-byte vector accesses, grid-stride).
#include <hip/hip_runtime.h>

#include "hd_ctx.h"

namespace {
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                     uint64_t n16) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" hd_status hd_stream_copy(const void* d_src, void* d_dst, uint64_t bytes, void* stream) {
    if (!d_src || !d_dst || bytes % 16) return hd::fail(nullptr, HD_ERR_INVALID, "stream copy needs 16-byte multiples");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const uint64_t n16 = bytes / 16;
    const uint32_t blocks = 256 * 32;  // 32 blocks per CU on 256 CUs
    k_stream_copy<<<blocks, 256, 0, s>>>(reinterpret_cast<const uint4*>(d_src), reinterpret_cast<uint4*>(d_dst), n16);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hd::hip_fail(nullptr, e, "stream copy");
    return HD_OK;
}
