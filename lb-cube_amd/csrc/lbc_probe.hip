// Streaming copy (non-temporal loads and stores, 16 bytes per lane: the fastest
// plain copy measured on MI355X) used as the measured HBM ceiling next to the
// spec peak: the bench reports the fluid kernels' achieved bandwidth against
// both, and the HBM byte counters (FETCH_SIZE / WRITE_SIZE) are scaled by what
// they read for this copy of known size.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/lbcube.h"

namespace lbc {

typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void stream_copy_kernel(const v4f* __restrict__ src, v4f* __restrict__ dst,
                                                          uint64_t n4) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace lbc

extern "C" int lbc_stream_copy(const void* src, void* dst, uint64_t bytes, void* stream) {
    if (src == nullptr || dst == nullptr) return hipErrorInvalidDevicePointer;
    if (bytes == 0 || bytes % 16 != 0 || src == dst) return hipErrorInvalidValue;
    const uint64_t n4 = bytes / 16;
    const uint64_t blocks = (n4 + 255) / 256;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(lbc::stream_copy_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const lbc::v4f*>(src), static_cast<lbc::v4f*>(dst),
                       n4);
    return hipGetLastError();
}
