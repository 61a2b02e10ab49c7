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

// ------------------------------------------------------------ buffer access
// Raw buffer descriptors (gfx9 family word 3 = 0x00020000): one SGPR
// descriptor per array, a per-thread VGPR byte offset and a scalar byte
// offset per element group.  Out-of-range loads return 0 and out-of-range
// stores are dropped by the hardware range check (num_records).
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 bload2(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ float bload1(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
// Stores take the cache policy as a template argument: 0 (default) or
// kStream, the non-temporal hint ("nt" on gfx950), for outputs this kernel
// never reads again, so that they do not push the filter tables out of L2.
constexpr int kStream = 2;
template <int AUX = 0>
__device__ __forceinline__ void bstore1(rsrc_t r, float v, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ void bstore2(rsrc_t r, float2 v, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                        r, (int)voff, (int)soff, AUX);
}

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
// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int I>
struct IntC {
  static constexpr int value = I;
};
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(IntC<B>{});
    static_for<B + 1, E>(f);
  }
}

// static_for with a scheduling fence after every G iterations: bounds how far
// the scheduler hoists independent loads (register pressure) in long
// element-wise loops
template <int B, int E, int G, class F>
__device__ __forceinline__ void static_for_chunked(F&& f) {
  if constexpr (B < E) {
    f(IntC<B>{});
    if constexpr ((B + 1) % G == 0) __builtin_amdgcn_sched_barrier(0);
    static_for_chunked<B + 1, E, G>(f);
  }
}

constexpr int ilog2(int x) {
  int r = 0;
  while ((1 << (r + 1)) <= x) ++r;
  return r;
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
// pidx(base + C) for a compile-time C, written so that the per-thread part
// (pidx(base)) is shared and C contributes an immediate offset whenever
// (base & 31) + (C & 31) < 32 is provable (always when C % 32 == 0).
template <int C>
__device__ __forceinline__ int pidx_off(int pbase, int base) {
  return pbase + C + (C >> 5) + (((base & 31) + (C & 31)) >> 5);
}
template <int N>
struct LdsSize {
  static constexpr int floats = N + N / 32 + 32;  // per real/imag array (+ one wrap slot at index N)
};

// Per-thread LDS base for a (real, imaginary) pair of float arrays, kept in a
// register the optimiser cannot see through, so every access is the base
// plus an immediate offset (< 64 KiB) instead of a separately materialised
// address.  Element offsets passed to ld/st are compile-time constants.
typedef __attribute__((address_space(3))) float lds_float;
__device__ __forceinline__ uint32_t lds_addr(const float* p) {
  return (uint32_t)(uintptr_t)(const lds_float*)p;
}
struct LdsPair {
  uint32_t re, im;
  __device__ __forceinline__ LdsPair(const float* lre, const float* lim, int elem) {
    re = lds_addr(lre) + 4u * (uint32_t)elem;
    im = lds_addr(lim) + 4u * (uint32_t)elem;
    asm volatile("" : "+v"(re), "+v"(im));
  }
  template <int C>
  __device__ __forceinline__ float2 ld() const {
    return make_float2(((const lds_float*)(uintptr_t)re)[C], ((const lds_float*)(uintptr_t)im)[C]);
  }
  template <int C>
  __device__ __forceinline__ void st(float2 v) const {
    ((lds_float*)(uintptr_t)re)[C] = v.x;
    ((lds_float*)(uintptr_t)im)[C] = v.y;
  }
};
// padded offset of element (base + C) relative to padded base, when
// (base & 31) + (C & 31) < 32 (true for every access pattern used here)
template <int C>
struct POff {
  static constexpr int value = C + (C >> 5);
};

// ------------------------------------------------------------ barriers
// Block-wide (all T threads of the workgroup take part) or wavefront-local
// (T == 64 and other waves may be doing something else).
enum { kSyncBlock = 0, kSyncWave = 1 };
template <int SYNC>
__device__ __forceinline__ void lds_barrier() {
  if constexpr (SYNC == kSyncBlock) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// ------------------------------------------------------------ Stockham pass
// Twiddle sources may hold a pass's full rows, W^{r k} for every r < R
// (member template row<NS>(k, r), defined only for those passes): the pass
// then reads each twiddle instead of forming the powers of the base ones.
template <class TWF, int NS, class = void>
struct HasTwRow : std::false_type {};
template <class TWF, int NS>
struct HasTwRow<TWF, NS, std::void_t<decltype(std::declval<const TWF&>().template row<NS>(0, 0))>>
    : std::true_type {};

// Pass of radix R after NS points have been combined.  TW is this pass's
// twiddle rows: TW[k*log2(R) + i] = exp(-2*pi*i*2^i*k/(NS*R)) for k < NS
// (unused when NS == 1); DIR = -1 forward, +1 inverse (conjugated twiddles).  FROM_LDS: inputs come from LDS (else from registers in natural
// layout).  TO_LDS: outputs go to LDS (else, only legal when NS*R == N, to
// registers in natural layout).
template <int N, int T, int R, int NS, int DIR, bool FROM_LDS, bool TO_LDS, int SYNC, class TWF>
__device__ __forceinline__ void stockham_pass(float2 (&v)[N / T], float* __restrict__ lre,
                                              float* __restrict__ lim, const TWF& twf, int tid_in) {
  constexpr int P = N / T;
  constexpr int Q = P / R;  // butterflies per thread
  static_assert(P % R == 0, "radix must divide points per thread");
  static_assert(TO_LDS || NS * R == N, "only the last pass may end in registers");
  const int tid = tid_in & (T - 1);
  static_assert(T % 32 == 0 && (N / R) % 32 == 0, "LDS offsets assume 32-aligned strides");
  if constexpr (FROM_LDS) {
    const LdsPair b(lre, lim, pidx(tid));
    static_for<0, Q>([&](auto qc) {
      static_for<0, R>([&](auto rc) {
        constexpr int q = decltype(qc)::value, r = decltype(rc)::value;
        v[q + r * Q] = b.template ld<POff<T * q + r * (N / R)>::value>();
      });
    });
    lds_barrier<SYNC>();
  } else if constexpr (TO_LDS) {
    lds_barrier<SYNC>();  // LDS may still be read by a previous stage
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = tid + T * q;
    const int k = j & (NS - 1);
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[r] = v[q + r * Q];
    if constexpr (NS > 1 && HasTwRow<TWF, NS>::value) {
      static_for_chunked<1, R, 8>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        float2 w = twf.template row<NS>(k, r);
        if constexpr (DIR > 0) w.y = -w.y;
        u[r] = cmul(u[r], w);
      });
    } else if constexpr (NS > 1) {
      // w_r = W^{r k}: the log2(R) powers W^{2^i k} come from the table, the
      // others are products over the set bits of r (at most log2(R)-1
      // multiplies deep: a few float32 ulp)
      constexpr int LR = ilog2(R);
      float2 base[LR];
#pragma unroll
      for (int i = 0; i < LR; ++i) {
        base[i] = twf.template get<NS>(k, i);
        if constexpr (DIR > 0) base[i].y = -base[i].y;
      }
      float2 w[R];
      static_for_chunked<1, R, 8>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        constexpr int hb = ilog2(r);
        if constexpr (r == (1 << hb))
          w[r] = base[hb];
        else
          w[r] = cmul(w[r - (1 << hb)], base[hb]);
        u[r] = cmul(u[r], w[r]);
      });
    }
    dif_rec<R, DIR>(u);
    if constexpr (TO_LDS) {
      const int base = (j / NS) * NS * R + k;
      const LdsPair w(lre, lim, pidx(base));
      static_for<0, R>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        w.template st<POff<r * NS>::value>(u[bitrev(r, R)]);
      });
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q + r * Q] = u[bitrev(r, R)];
    }
  }
  if constexpr (TO_LDS) lds_barrier<SYNC>();
}

// Twiddle rows of every pass with NS > 1, concatenated (global table).
// Offsets (in float2) of each transform's passes; the host builds the table
// with the same layout: row k of a pass holds exp(-2 pi i 2^i k / (NS R)),
// i < log2(R).
struct TwLayout {
  // 16384 points: passes (R=32, NS=32) and (R=16, NS=1024)
  static constexpr int k16k_p2 = 0;
  static constexpr int k16k_p3 = k16k_p2 + 32 * 5;
  // 1024 points: passes (R=16, NS=16) and (R=4, NS=256)
  static constexpr int k1k_p2 = k16k_p3 + 1024 * 4;
  static constexpr int k1k_p3 = k1k_p2 + 16 * 4;
  // 2048 points (audio slices of 65536-sample blocks): passes (R=32, NS=32)
  // and (R=2, NS=1024)
  static constexpr int k2k_p2 = k1k_p3 + 256 * 2;
  static constexpr int k2k_p3 = k2k_p2 + 32 * 5;
  // 512 points (audio slices of 16384-sample blocks): passes (R=8, NS=8)
  // and (R=8, NS=64)
  static constexpr int k512_p2 = k2k_p3 + 1024 * 1;
  static constexpr int k512_p3 = k512_p2 + 8 * 3;
  // 256 points (audio slices of 8192-sample blocks): passes (R=4, NS=4),
  // (R=4, NS=16) and (R=4, NS=64)
  static constexpr int k256_p2 = k512_p3 + 64 * 3;
  static constexpr int k256_p3 = k256_p2 + 4 * 2;
  static constexpr int k256_p4 = k256_p3 + 16 * 2;
  static constexpr int total = k256_p4 + 64 * 2;
};

// Twiddle source reading the global rows through a buffer descriptor.
struct TwGlobal {
  rsrc_t tw;
  int off_p2, off_p3;  // row tables of the second and third pass
  int ns_p2, lr_p2, lr_p3;
  int off_p4 = 0, ns_p3 = 0, lr_p4 = 0;  // a fourth pass (ns_p3 = 0: none)
  template <int NS>
  __device__ __forceinline__ float2 get(int k, int i) const {
    const bool p2 = NS == ns_p2, p3 = !p2 && (ns_p3 == 0 || NS == ns_p3);
    const int off = p2 ? off_p2 : (p3 ? off_p3 : off_p4);
    const int lr = p2 ? lr_p2 : (p3 ? lr_p3 : lr_p4);
    return bload2(tw, (uint32_t)(off + k * lr + i) * 8u, 0u);
  }
};

// Full transforms used by the demodulator.  On entry v[m] = x[tid + T*m];
// on exit v[m] = X[tid + T*m] (unnormalised).  LDS is free again on return.
template <int N, int DIR, int SYNC = kSyncBlock>
struct WgFft;

// 16384 points, 512 threads: radix 32 x 32 x 16
template <int DIR, int SYNC>
struct WgFft<16384, DIR, SYNC> {
  static constexpr int T = 512;
  template <class TWF>
  // tid: the thread index (callers may pass an opaque copy to keep address
  // arithmetic from being hoisted out of their loops)
  static __device__ __forceinline__ void run(float2 (&v)[32], float* lre, float* lim, const TWF& twf,
                                             int tid = threadIdx.x) {
    stockham_pass<16384, 512, 32, 1, DIR, false, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<16384, 512, 32, 32, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<16384, 512, 16, 1024, DIR, true, false, SYNC>(v, lre, lim, twf, tid);
  }
};

// 16384 points, 1024 threads (16 per thread): radix 16 x 16 x 16 x 4
template <int DIR, int SYNC = kSyncBlock>
struct WgFft16kT1024 {
  static constexpr int T = 1024;
  template <class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[16], float* lre, float* lim, const TWF& twf,
                                             int tid = threadIdx.x) {
    stockham_pass<16384, 1024, 16, 1, DIR, false, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<16384, 1024, 16, 16, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<16384, 1024, 16, 256, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<16384, 1024, 4, 4096, DIR, true, false, SYNC>(v, lre, lim, twf, tid);
  }
};

// 512 points, 64 threads (one wavefront): radix 8 x 8 x 8
template <int DIR, int SYNC>
struct WgFft<512, DIR, SYNC> {
  static constexpr int T = 64;
  template <class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[8], float* lre, float* lim, const TWF& twf,
                                             int tid = threadIdx.x) {
    stockham_pass<512, 64, 8, 1, DIR, false, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<512, 64, 8, 8, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<512, 64, 8, 64, DIR, true, false, SYNC>(v, lre, lim, twf, tid);
  }
};

// 256 points, 64 threads (one wavefront): radix 4 x 4 x 4 x 4
template <int DIR, int SYNC>
struct WgFft<256, DIR, SYNC> {
  static constexpr int T = 64;
  template <class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[4], float* lre, float* lim, const TWF& twf,
                                             int tid = threadIdx.x) {
    stockham_pass<256, 64, 4, 1, DIR, false, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<256, 64, 4, 4, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<256, 64, 4, 16, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<256, 64, 4, 64, DIR, true, false, SYNC>(v, lre, lim, twf, tid);
  }
};

// 4096 points, 256 threads (16 per thread): radix 16 x 16 x 16
template <int DIR, int SYNC = kSyncBlock>
struct WgFft4kT256 {
  static constexpr int T = 256;
  template <class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[16], float* lre, float* lim, const TWF& twf,
                                             int tid = threadIdx.x) {
    stockham_pass<4096, 256, 16, 1, DIR, false, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<4096, 256, 16, 16, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<4096, 256, 16, 256, DIR, true, false, SYNC>(v, lre, lim, twf, tid);
  }
};

// 8192 points, 512 threads (16 per thread): radix 16 x 16 x 16 x 2
template <int DIR, int SYNC = kSyncBlock>
struct WgFft8kT512 {
  static constexpr int T = 512;
  template <class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[16], float* lre, float* lim, const TWF& twf,
                                             int tid = threadIdx.x) {
    stockham_pass<8192, 512, 16, 1, DIR, false, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<8192, 512, 16, 16, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<8192, 512, 16, 256, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<8192, 512, 2, 4096, DIR, true, false, SYNC>(v, lre, lim, twf, tid);
  }
};

// 2048 points, 64 threads (one wavefront): radix 32 x 32 x 2
template <int DIR, int SYNC>
struct WgFft<2048, DIR, SYNC> {
  static constexpr int T = 64;
  template <class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[32], float* lre, float* lim, const TWF& twf,
                                             int tid = threadIdx.x) {
    stockham_pass<2048, 64, 32, 1, DIR, false, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<2048, 64, 32, 32, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<2048, 64, 2, 1024, DIR, true, false, SYNC>(v, lre, lim, twf, tid);
  }
};

// 1024 points, 64 threads (one wavefront): radix 16 x 16 x 4
template <int DIR, int SYNC>
struct WgFft<1024, DIR, SYNC> {
  static constexpr int T = 64;
  template <class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[16], float* lre, float* lim, const TWF& twf,
                                             int tid = threadIdx.x) {
    stockham_pass<1024, 64, 16, 1, DIR, false, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<1024, 64, 16, 16, DIR, true, true, SYNC>(v, lre, lim, twf, tid);
    stockham_pass<1024, 64, 4, 256, DIR, true, false, SYNC>(v, lre, lim, twf, tid);
  }
};

// ------------------------------------------------ split-BF16 MFMA radix 64
// A radix-64 Stockham pass as a real matrix product on the matrix cores:
// for the C = M/64 columns j of a pass (inputs a[j + r C], r < 64, twiddled
// by W_{64 NS}^{r (j mod NS)} in FP32), the DFT-64 is
//   [Y_re; Y_im] (128 x C) = [[F_re, -F_im]; [F_im, F_re]] (128 x 128) [U_re; U_im]
// with F[s][r] = exp(-2 pi i r s / 64).  Both operands are split into
// KT bf16 terms and the product summed over the term pairs that matter,
// accumulated in FP32 by v_mfma_f32_32x32x16_bf16:
//   KT = 2: x = hi + lo (to about 2^-17 relative); G_hi U_hi + G_hi U_lo +
//           G_lo U_hi (the lo*lo term, about 2^-18, dropped): 3 products;
//   KT = 3: x = hi + mid + lo (exact for an FP32 U: 3 x 8 significand
//           bits); the six products of order >= 2^-16 (hi*hi, hi*mid,
//           mid*hi, hi*lo, lo*hi, mid*mid; the dropped mid*lo, lo*mid are
//           about 2^-24): FP32-level error.
// The G fragments come precomputed in the instruction's lane layout from a
// global table (L1/L2 resident, 96 KiB: three terms).
//
// Lane layout of v_mfma_f32_32x32x16_bf16 (checked on gfx950): lane l holds
// A[l % 32][8 (l / 32) + e] and B[8 (l / 32) + e][l % 32] for e < 8, and the
// result D[8 (q / 4) + 4 (l / 32) + q % 4][l % 32] in accumulator q < 16.
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2_t;
typedef __attribute__((ext_vector_type(16))) float f32x16_t;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4_t;

// float -> bf16 bits, round to nearest even (finite values)
__device__ __forceinline__ uint32_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}
// two floats -> packed bf16 pair (a in the low half), round to nearest
// even: one v_cvt_pk_bf16_f32
__device__ __forceinline__ uint32_t bf16_pack(float a, float b) {
  const bf16x2_t h = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, h);
}
// the packed pair's values as floats
__device__ __forceinline__ float bf16_lo_f(uint32_t p) { return __uint_as_float(p << 16); }
__device__ __forceinline__ float bf16_hi_f(uint32_t p) { return __uint_as_float(p & 0xffff0000u); }
// (a, b) = sum of KT packed bf16 pairs t[0..KT), each the rounded residual
// of the terms before it
template <int KT>
__device__ __forceinline__ void bf16_split_pair(float a, float b, uint32_t (&t)[KT]) {
  t[0] = bf16_pack(a, b);
  if constexpr (KT > 1) {
    a -= bf16_lo_f(t[0]);
    b -= bf16_hi_f(t[0]);
    t[1] = bf16_pack(a, b);
  }
  if constexpr (KT > 2) {
    a -= bf16_lo_f(t[1]);
    b -= bf16_hi_f(t[1]);
    t[2] = bf16_pack(a, b);
  }
}

// G fragment table layout: [row tile 0..3][k step 0..7][term 0..2][lane 0..63]
// of 8 bf16 (one u32x4 per lane).
constexpr int kMfmaTerms = 3;
constexpr int kMfmaTabEntries = 4 * 8 * kMfmaTerms * 64;

// W_{64 NS}^q for q < 64 NS as the product of a coarse and a fine table:
// q = 64 qh + ql, W^q = W_{NS}^{qh} W_{64 NS}^{ql} (NS = 64 here: both 64).
struct MfmaTw {
  float2 coarse[64];  // exp(-2 pi i qh / NS)
  float2 fine[64];    // exp(-2 pi i ql / (64 NS))
};

// acc += G[rows of tile] x U over one k step pair, for KT-term operands:
// a_re / a_im: the G fragments of the U_re and U_im rows, b_re / b_im: U's
template <int KT>
__device__ __forceinline__ f32x16_t mfma_terms(f32x16_t acc, const bf16x8_t (&a_re)[KT], const bf16x8_t (&a_im)[KT],
                                               const bf16x8_t (&b_re)[KT], const bf16x8_t (&b_im)[KT]) {
  // smallest terms first
  if constexpr (KT == 3) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_re[1], b_re[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_im[1], b_im[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_re[0], b_re[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_im[0], b_im[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_re[2], b_re[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_im[2], b_im[0], acc, 0, 0, 0);
  }
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_re[0], b_re[1], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_im[0], b_im[1], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_re[1], b_re[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_im[1], b_im[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_re[0], b_re[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_im[0], b_im[0], acc, 0, 0, 0);
  return acc;
}

template <int M, int T, int NS, int KT>
__device__ __forceinline__ void mfma64_pass(float* __restrict__ lre, float* __restrict__ lim,
                                            const u32x4_t* __restrict__ gtab, const MfmaTw* tw, int tid) {
  constexpr int C = M / 64;           // columns
  constexpr int NW = T / 64;          // wavefronts
  constexpr int NCT = C / 32;         // column tiles of 32
  static_assert(NW % 4 == 0 && NCT % (NW / 4) == 0, "tiles must divide over the wavefronts");
  static_assert(KT == 2 || KT == 3, "two or three bf16 terms");
  constexpr int CTW = NCT / (NW / 4); // column tiles per wavefront
  static_assert(NS == 1 || NS == 64, "twiddle tables cover NS = 1 and 64");
  const int w = tid >> 6, l = tid & 63;
  const int rt = w & 3;               // row tile: output rows 32 rt .. 32 rt + 31
  const int cg = w >> 2;              // column group
  const int half = l >> 5, col = l & 31;
  f32x16_t acc[CTW];
#pragma unroll
  for (int c = 0; c < CTW; ++c)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[c][q] = 0.0f;
  const u32x4_t* ga = gtab + (size_t)rt * (8 * kMfmaTerms * 64) + l;
#pragma unroll 1
  for (int ks = 0; ks < 4; ++ks) {
    // G fragments of k steps ks (U_re rows) and ks + 4 (U_im rows)
    bf16x8_t are[KT], aim[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      are[k] = __builtin_bit_cast(bf16x8_t, ga[(ks * kMfmaTerms + k) * 64]);
      aim[k] = __builtin_bit_cast(bf16x8_t, ga[((ks + 4) * kMfmaTerms + k) * 64]);
    }
    const int r0 = 16 * ks + 8 * half;
#pragma unroll
    for (int c = 0; c < CTW; ++c) {
      const int j = 32 * (cg * CTW + c) + col;
      u32x4_t bre[KT], bim[KT];
#pragma unroll
      for (int e2 = 0; e2 < 4; ++e2) {
        float2 u[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int r = r0 + 2 * e2 + t;
          const int idx = pidx(j + r * C);
          u[t] = make_float2(lre[idx], lim[idx]);
          if constexpr (NS > 1) {
            const int q = (r * (j & (NS - 1))) & (64 * NS - 1);
            u[t] = cmul(u[t], cmul(tw->coarse[q >> 6], tw->fine[q & 63]));
          }
        }
        uint32_t pr[KT], pi[KT];
        bf16_split_pair<KT>(u[0].x, u[1].x, pr);
        bf16_split_pair<KT>(u[0].y, u[1].y, pi);
#pragma unroll
        for (int k = 0; k < KT; ++k) {
          bre[k][e2] = pr[k];
          bim[k][e2] = pi[k];
        }
      }
      bf16x8_t br[KT], bi[KT];
#pragma unroll
      for (int k = 0; k < KT; ++k) {
        br[k] = __builtin_bit_cast(bf16x8_t, bre[k]);
        bi[k] = __builtin_bit_cast(bf16x8_t, bim[k]);
      }
      acc[c] = mfma_terms<KT>(acc[c], are, aim, br, bi);
    }
  }
  __syncthreads();  // every input read before any output lands on it
#pragma unroll
  for (int c = 0; c < CTW; ++c) {
    const int j = 32 * (cg * CTW + c) + col;
    const int base = (j / NS) * NS * 64 + (j & (NS - 1));
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int row = 32 * rt + 8 * (q >> 2) + 4 * half + (q & 3);  // 0..127
      const int sidx = row & 63;
      float* dst = row < 64 ? lre : lim;
      dst[pidx(base + sidx * NS)] = acc[c][q];
    }
  }
  __syncthreads();
}

// 4096 points, 256 threads, forward: radix 64 x 64 on the matrix cores
// (split BF16); the result read back from LDS in natural order.
struct WgFft4kMfma {
  static constexpr int T = 256;
  template <int KT, class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[16], float* lre, float* lim, const TWF&,
                                             const u32x4_t* gtab, const MfmaTw* tw, int tid = threadIdx.x) {
    __syncthreads();
    const LdsPair own(lre, lim, pidx(tid));
    static_for<0, 16>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      own.template st<POff<256 * m>::value>(v[m]);
    });
    __syncthreads();
    mfma64_pass<4096, 256, 1, KT>(lre, lim, gtab, tw, tid);
    mfma64_pass<4096, 256, 64, KT>(lre, lim, gtab, tw, tid);
    static_for<0, 16>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      v[m] = own.template ld<POff<256 * m>::value>();
    });
    __syncthreads();
  }
};

// 8192 points, 512 threads, forward: radix 64 x 64 on the matrix cores
// (split BF16), then radix 2 in FP32.
struct WgFft8kMfma {
  static constexpr int T = 512;
  template <int KT, class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[16], float* lre, float* lim, const TWF& twf,
                                             const u32x4_t* gtab, const MfmaTw* tw, int tid = threadIdx.x) {
    __syncthreads();
    const LdsPair own(lre, lim, pidx(tid));
    static_for<0, 16>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      own.template st<POff<512 * m>::value>(v[m]);
    });
    __syncthreads();
    mfma64_pass<8192, 512, 1, KT>(lre, lim, gtab, tw, tid);
    mfma64_pass<8192, 512, 64, KT>(lre, lim, gtab, tw, tid);
    stockham_pass<8192, 512, 2, 4096, -1, true, false, kSyncBlock>(v, lre, lim, twf, tid);
  }
};

// 16384 points, 1024 threads, forward: radix 64 x 64 on the matrix cores
// (split BF16), then radix 4 in FP32.  On entry v[m] = x[tid + T m]; on exit
// v[m] = X[tid + T m].  TWF serves the last pass (NS = 4096).
struct WgFft16kMfma {
  static constexpr int T = 1024;
  template <int KT, class TWF>
  static __device__ __forceinline__ void run(float2 (&v)[16], float* lre, float* lim, const TWF& twf,
                                             const u32x4_t* gtab, const MfmaTw* tw, int tid = threadIdx.x) {
    __syncthreads();  // LDS may still be read by a previous stage
    const LdsPair own(lre, lim, pidx(tid));
    static_for<0, 16>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      own.template st<POff<1024 * m>::value>(v[m]);
    });
    __syncthreads();
    mfma64_pass<16384, 1024, 1, KT>(lre, lim, gtab, tw, tid);
    mfma64_pass<16384, 1024, 64, KT>(lre, lim, gtab, tw, tid);
    stockham_pass<16384, 1024, 4, 4096, -1, true, false, kSyncBlock>(v, lre, lim, twf, tid);
  }
};

}  // namespace ldgpu
