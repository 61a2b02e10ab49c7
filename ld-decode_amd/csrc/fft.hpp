// Workgroup-resident FP32 FFTs for gfx950 (CDNA4).
//
// One workgroup of T threads transforms N complex points.  Each thread keeps
// P = N / T points in registers in "natural" layout: v[m] holds element
// tid + T*m.  A transform is a sequence of Stockham autosort passes of radix
// R; a pass gathers R points at stride N/R into registers, applies the
// inter-pass twiddles, runs an in-register radix-R DFT (compile-time
// constants), and scatters the R results through LDS to the positions the
// next pass reads.  The first pass reads registers (the natural layout is
// exactly what it needs) and the last pass writes registers, so an N-point
// transform with three passes makes two LDS round trips.
//
// LDS holds real and imaginary parts in separate float arrays with one pad
// word every 32 words, which makes the strided scatters of every pass used
// here conflict-free.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ldgpu {

// ---------------------------------------------------------------- complex ops
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// ------------------------------------------------- compile-time twiddles
constexpr double kPi = 3.141592653589793238462643383279502884;

// Taylor series on [-pi, pi]; 30 terms reach double rounding, far below the
// float32 rounding the constants are stored at.
constexpr double c_sin(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int i = 1; i < 30; ++i) {
    term *= -x2 / double((2 * i) * (2 * i + 1));
    sum += term;
  }
  return sum;
}
constexpr double c_cos(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int i = 1; i < 30; ++i) {
    term *= -x2 / double((2 * i - 1) * (2 * i));
    sum += term;
  }
  return sum;
}
// exp(DIR * 2*pi*i*K/N) with the angle reduced to [-pi, pi]
template <int N, int K, int DIR>
struct Tw {
  static constexpr int kr = ((K % N) + N) % N;
  static constexpr int ks = (2 * kr > N) ? kr - N : kr;
  static constexpr double ang = 2.0 * kPi * double(ks) / double(N);
  static constexpr float re = float(c_cos(ang));
  static constexpr float im = float(DIR * c_sin(ang));
};

// d * exp(DIR*2*pi*i*K/N), trivial rotations done without multiplies
template <int N, int K, int DIR>
__device__ __forceinline__ float2 rot_const(float2 d) {
  constexpr int k = ((K % N) + N) % N;
  if constexpr (k == 0) {
    return d;
  } else if constexpr (4 * k == N) {  // * (DIR i)
    return make_float2(-DIR * d.y, DIR * d.x);
  } else if constexpr (2 * k == N) {
    return make_float2(-d.x, -d.y);
  } else if constexpr (4 * k == 3 * N) {  // * (-DIR i)
    return make_float2(DIR * d.y, -DIR * d.x);
  } else {
    constexpr float c = Tw<N, K, DIR>::re, s = Tw<N, K, DIR>::im;
    return make_float2(d.x * c - d.y * s, d.x * s + d.y * c);
  }
}

// ------------------------------------------------ in-register radix-R DFT
// Radix-2 decimation in frequency on v[0..N): leaves X[k] at v[bitrev(k)].
template <int N, int DIR, int K = 0>
__device__ __forceinline__ void dif_layer(float2* v) {
  if constexpr (K < N / 2) {
    float2 a = v[K], b = v[K + N / 2];
    v[K] = cadd(a, b);
    v[K + N / 2] = rot_const<N, K, DIR>(csub(a, b));
    dif_layer<N, DIR, K + 1>(v);
  }
}
template <int N, int DIR>
__device__ __forceinline__ void dif_rec(float2* v) {
  if constexpr (N > 1) {
    dif_layer<N, DIR>(v);
    dif_rec<N / 2, DIR>(v);
    dif_rec<N / 2, DIR>(v + N / 2);
  }
}
constexpr int bitrev(int x, int n) {
  int r = 0;
  for (int b = 1; b < n; b <<= 1) {
    r = (r << 1) | (x & 1);
    x >>= 1;
  }
  return r;
}

// --------------------------------------------------------------- LDS layout
__device__ __forceinline__ int pidx(int i) { return i + (i >> 5); }
template <int N>
struct LdsSize {
  static constexpr int floats = N + N / 32;  // per real/imag array
};

// ------------------------------------------------------------ Stockham pass
// Pass of radix R after NS points have been combined.  TW points at a table
// of exp(-2*pi*i*t/L) for t < L (L a multiple of N); DIR = -1 forward, +1
// inverse (conjugated twiddles).  FROM_LDS: inputs come from LDS (else from
// registers in natural layout).  TO_LDS: outputs go to LDS (else, only legal
// when NS*R == N, to registers in natural layout).
template <int N, int T, int R, int NS, int DIR, bool FROM_LDS, bool TO_LDS, int L>
__device__ __forceinline__ void stockham_pass(float2 (&v)[N / T], float* __restrict__ lre,
                                              float* __restrict__ lim,
                                              const float2* __restrict__ tw) {
  constexpr int P = N / T;
  constexpr int Q = P / R;  // butterflies per thread
  static_assert(P % R == 0, "radix must divide points per thread");
  static_assert(TO_LDS || NS * R == N, "only the last pass may end in registers");
  const int tid = threadIdx.x;
  if constexpr (FROM_LDS) {
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int i = tid + T * q + r * (N / R);
        v[q + r * Q] = make_float2(lre[pidx(i)], lim[pidx(i)]);
      }
    __syncthreads();
  } else if constexpr (TO_LDS) {
    __syncthreads();  // LDS may still be read by a previous stage
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = tid + T * q;
    const int k = j & (NS - 1);
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[q + r * Q];
    if constexpr (NS > 1) {
      constexpr int step = L / (NS * R);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w = tw[r * k * step];
        if constexpr (DIR > 0) w.y = -w.y;
        u[r] = cmul(u[r], w);
      }
    }
    dif_rec<R, DIR>(u);
    if constexpr (TO_LDS) {
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int o = base + r * NS;
        float2 x = u[bitrev(r, R)];
        lre[pidx(o)] = x.x;
        lim[pidx(o)] = x.y;
      }
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * Q] = u[bitrev(r, R)];
    }
  }
  if constexpr (TO_LDS) __syncthreads();
}

// Full transforms used by the demodulator.  On entry v[m] = x[tid + T*m];
// on exit v[m] = X[tid + T*m] (unnormalised).  LDS is free again on return.
template <int N, int DIR, int L>
struct WgFft;

// 16384 points, 512 threads: radix 32 x 32 x 16
template <int DIR, int L>
struct WgFft<16384, DIR, L> {
  static constexpr int T = 512;
  static __device__ __forceinline__ void run(float2 (&v)[32], float* lre, float* lim,
                                             const float2* tw) {
    stockham_pass<16384, 512, 32, 1, DIR, false, true, L>(v, lre, lim, tw);
    stockham_pass<16384, 512, 32, 32, DIR, true, true, L>(v, lre, lim, tw);
    stockham_pass<16384, 512, 16, 1024, DIR, true, false, L>(v, lre, lim, tw);
  }
};

// 8192 points, 256 threads: radix 32 x 16 x 16
template <int DIR, int L>
struct WgFft<8192, DIR, L> {
  static constexpr int T = 256;
  static __device__ __forceinline__ void run(float2 (&v)[32], float* lre, float* lim,
                                             const float2* tw) {
    stockham_pass<8192, 256, 32, 1, DIR, false, true, L>(v, lre, lim, tw);
    stockham_pass<8192, 256, 16, 32, DIR, true, true, L>(v, lre, lim, tw);
    stockham_pass<8192, 256, 16, 512, DIR, true, false, L>(v, lre, lim, tw);
  }
};

// 1024 points, 64 threads (one wavefront): radix 16 x 16 x 4
template <int DIR, int L>
struct WgFft<1024, DIR, L> {
  static constexpr int T = 64;
  static __device__ __forceinline__ void run(float2 (&v)[16], float* lre, float* lim,
                                             const float2* tw) {
    stockham_pass<1024, 64, 16, 1, DIR, false, true, L>(v, lre, lim, tw);
    stockham_pass<1024, 64, 16, 16, DIR, true, true, L>(v, lre, lim, tw);
    stockham_pass<1024, 64, 4, 256, DIR, true, false, L>(v, lre, lim, tw);
  }
};

}  // namespace ldgpu
