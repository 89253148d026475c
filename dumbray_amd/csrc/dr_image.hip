// Frame output: the device BlockBasedFrameBuffer<8,16> to row-major RGBA8 bytes with
// Images::translateChannelFloatToByte (Images.cu:7-19): v = 255 * channel, 255 above 255,
// 0 below 0 (and for NaN), else truncated.  One thread per pixel; a row of a block is 128
// contiguous bytes of the frame, the byte rows are written coalesced.
#include "dr_common.h"

namespace dr {

__device__ __forceinline__ unsigned char channel_byte(float c) {
  const float v = 255.0f * c;
  if (v > 255.0f) return 255;
  if (v < 0.0f) return 0;
  if (v != v) return 0;
  return (unsigned char)v;
}

__global__ __launch_bounds__(256) void k_frame_to_rgba8(const float4 *__restrict__ frame, int width, int height,
                                                       int blocks_w, uchar4 *__restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= width || y >= height) return;
  const size_t src = (size_t)128 * ((y / 16) * blocks_w + x / 8) + (y % 16) * 8 + (x % 8);
  const float4 c = frame[src];
  out[(size_t)y * width + x] = make_uchar4(channel_byte(c.x), channel_byte(c.y), channel_byte(c.z), channel_byte(c.w));
}

}  // namespace dr

extern "C" int dr_frame_to_rgba8(const float *frame, int width, int height, uint8_t *out, void *stream) {
  if (!frame || !out || width <= 0 || height <= 0 || height > 65535)
    return dr::fail(DR_ERR_ARG, "dr_frame_to_rgba8: bad arguments");
  const int blocks_w = (width + 7) / 8;
  hipLaunchKernelGGL(dr::k_frame_to_rgba8, dim3((width + 255) / 256, height), dim3(256), 0, (hipStream_t)stream,
                     (const float4 *)frame, width, height, blocks_w, (uchar4 *)out);
  DR_HIP_CHECK(hipGetLastError());
  return DR_OK;
}
