// Capture input formats: the packed 10-bit format of the Domesday
// Duplicator (4 samples in 5 bytes, big-endian bit order), unpacked to the
// int16 layout the rest of the decoder reads ((code - 512) << 6), as
// unpack_data_4_40 / load_packed_data_4_40 do (utils.py:371-410).  The
// demodulator also reads packed input directly (LD_IN_PACKED10).
#include "internal.hpp"

namespace ldgpu {

constexpr int kUnpackThreads = 256;

// Four samples per thread: the 5-byte group g holds samples 4g .. 4g+3.
__global__ __launch_bounds__(kUnpackThreads) void k_unpack_10bit(const unsigned char* __restrict__ packed,
                                                                 long long first, long long n,
                                                                 short* __restrict__ out) {
  const long long g0 = first >> 2;
  const long long g = g0 + (long long)blockIdx.x * kUnpackThreads + threadIdx.x;
  const unsigned char* p = packed + 5 * g;
  const long long j = 4 * g;            // first sample of the group
  if (j + 3 < first || j >= first + n) return;
  const unsigned b0 = p[0], b1 = p[1], b2 = p[2], b3 = p[3], b4 = p[4];
  const unsigned u[4] = {(b0 << 2) | (b1 >> 6), ((b1 & 0x3fu) << 4) | (b2 >> 4), ((b2 & 0x0fu) << 6) | (b3 >> 2),
                         ((b3 & 0x03u) << 8) | b4};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long long o = j + i - first;
    if (o >= 0 && o < n) out[o] = (short)(((int)u[i] - 512) * 64);
  }
}

}  // namespace ldgpu

using namespace ldgpu;

extern "C" int ld_unpack_10bit(ld_handle* h, const uint8_t* packed, int64_t packed_bytes, int64_t first, int64_t n,
                               int16_t* out, void* stream) {
  if (!h || !packed || !out || first < 0 || n <= 0) {
    set_error("ld_unpack_10bit: bad arguments");
    return LD_EINVAL;
  }
  // the last group read must lie inside the packed buffer
  const int64_t last_group = (first + n - 1) >> 2;
  if ((last_group + 1) * 5 > packed_bytes) {
    set_error("ld_unpack_10bit: samples [%lld, %lld) need %lld bytes, buffer has %lld", (long long)first,
              (long long)(first + n), (long long)((last_group + 1) * 5), (long long)packed_bytes);
    return LD_ERANGE;
  }
  const int64_t groups = last_group - (first >> 2) + 1;
  const int64_t grid = (groups + kUnpackThreads - 1) / kUnpackThreads;
  if (grid > 0x7fffffffLL) {
    set_error("ld_unpack_10bit: too many samples");
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  hipLaunchKernelGGL(k_unpack_10bit, dim3((unsigned)grid), dim3(kUnpackThreads), 0, st, packed, (long long)first,
                     (long long)n, (short*)out);
  LD_HIP(hipGetLastError());
  return LD_OK;
}
