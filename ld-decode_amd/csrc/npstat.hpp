// NumPy reductions as the reference's float32 data see them, shared by the
// line-location kernels: slice bounds and numpy's pairwise float32 sum.
#pragma once
#include <hip/hip_runtime.h>

namespace ldgpu {

// Python/numpy slice bounds of [a, b) over n elements.
__device__ __forceinline__ void np_slice(long long a, long long b, long long n, long long& s, long long& e) {
  if (a < 0) a = (a + n < 0) ? 0 : a + n;
  if (a > n) a = n;
  if (b < 0) b = (b + n < 0) ? 0 : b + n;
  if (b > n) b = n;
  s = a;
  e = b > a ? b : a;
}

// numpy's pairwise float32 summation (8 accumulators per block of at most
// 128 elements, halves rounded down to a multiple of 8 above that).
template <class F>
__device__ float np_block_sum(F x, int off, int n) {
  if (n < 8) {
    float r = -0.0f;
    for (int i = 0; i < n; ++i) r = r + x(off + i);
    return r;
  }
  float r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = x(off + j);
  int i = 8;
  for (; i < n - (n % 8); i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = r[j] + x(off + i + j);
  }
  float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res = res + x(off + i);
  return res;
}

// numpy's recursion for any n: halves rounded down to a multiple of 8 until
// a block holds at most 128 elements (depth D covers n <= 128 * 2^D).
template <int D, class F>
__device__ __noinline__ float np_pairwise_sum_any(F x, int off, int n) {
  if constexpr (D == 0) {
    return np_block_sum(x, off, n);
  } else {
    if (n <= 128) return np_block_sum(x, off, n);
    int n2 = n / 2;
    n2 -= n2 % 8;
    const float left = np_pairwise_sum_any<D - 1>(x, off, n2);
    return left + np_pairwise_sum_any<D - 1>(x, off + n2, n - n2);
  }
}

// np.add.reduce of a float32 array: the reduction runs over buffers of 8192
// elements (NumPy's ufunc buffer size), each summed pairwise, the buffer
// sums added in order.
template <class F>
__device__ float np_sum_f32(F x, long long n) {
  float r = -0.0f;
  for (long long i = 0; i < n; i += 8192) {
    const int m = (int)(n - i < 8192 ? n - i : 8192);
    auto chunk = [&x, i](int j) { return x(i + j); };
    r = r + np_pairwise_sum_any<6>(chunk, 0, m);
  }
  return r;
}

}  // namespace ldgpu
