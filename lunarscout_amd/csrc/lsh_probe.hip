// lsh_probe.hip -- measurement kernels: a streaming copy and a 2x2 gather.
//
// They give the march kernel's roofline its reference points on the same
// device: the copy kernel is the plain HBM stream of known size that scales
// rocprofv3's FETCH_SIZE / WRITE_SIZE counters into bytes, and the gather
// kernel measures the rate of the march's memory access (one 2x2 float32 tap
// = two 8-byte row reads at a random cell of a float32 grid) when it hits a
// window of a chosen size: L2-resident, MALL-resident or HBM-sized.

#include "lsh_internal.h"

namespace lsh {
namespace {

__global__ void __launch_bounds__(256) copy_kernel(const float4* __restrict__ src,
                                                   float4* __restrict__ dst, int64_t count) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += stride)
    dst[i] = src[i];
}

// Each thread takes `taps` 2x2 taps at cells drawn by a 32-bit LCG over a
// width x height window (cells in [0, width-2] x [0, height-2], so every tap
// stays inside) and writes the maximum, which keeps every load live.
__global__ void __launch_bounds__(256) gather_kernel(const float* __restrict__ window,
                                                     int32_t width, int32_t height, int32_t taps,
                                                     float* __restrict__ out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t state = static_cast<uint32_t>(t) * 2654435761u + 12345u;
  const uint32_t cw = static_cast<uint32_t>(width - 1), ch = static_cast<uint32_t>(height - 1);
  float m = -3.0e38f;
  for (int k = 0; k < taps; ++k) {
    state = state * 1664525u + 1013904223u;
    const uint32_t x = static_cast<uint32_t>((static_cast<uint64_t>(state >> 8) * cw) >> 24);
    state = state * 1664525u + 1013904223u;
    const uint32_t y = static_cast<uint32_t>((static_cast<uint64_t>(state >> 8) * ch) >> 24);
    const float* r0 = window + static_cast<int64_t>(y) * width + x;
    const float* r1 = r0 + width;
    m = fmaxf(m, fmaxf(fmaxf(r0[0], r0[1]), fmaxf(r1[0], r1[1])));
  }
  out[t] = m;
}

}  // namespace

hipError_t launch_probe_copy(const float* src, float* dst, int64_t bytes, hipStream_t stream) {
  const int64_t count = bytes / 16;
  int blocks = static_cast<int>(count / 256 + 1 < 65536 ? count / 256 + 1 : 65536);
  hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, stream,
                     reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), count);
  return hipGetLastError();
}

hipError_t launch_probe_gather(const float* window, int32_t width, int32_t height, int32_t taps,
                               int32_t threads, float* out, hipStream_t stream) {
  hipLaunchKernelGGL(gather_kernel, dim3(threads / 256), dim3(256), 0, stream, window, width,
                     height, taps, out);
  return hipGetLastError();
}

}  // namespace lsh
