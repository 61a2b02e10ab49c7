// RF demodulation kernel (FP32) and the handle / demod part of the C ABI.
//
// k_demod_block: one workgroup of kT threads per block of N = 32768 real
// samples (M = 16384 complex points) computes the reference's
// RFDecode.demodblock_cpu (core.py:755-879) end to end:
//
//   1. X = rFFT(block)                                        (core.py:761)
//   2. rfhpf = irFFT(X * Frfhpf), cut/rolled by video_rot     (core.py:769-772)
//   3. analytic z = iFFT(X * RFVideo * MTF^l), even and odd   (core.py:800-805)
//      samples as two M-point inverse transforms
//   4. demod_raw = arg(z[n] conj z[n-1]) * fs / 2pi, wrapped  (utils.py:898-926)
//   5. D = rFFT(clip(demod_raw) - dc)                         (core.py:809)
//   6. demod, demod_05, demod_burst[, demod_pilot] =          (core.py:811-845)
//      irFFT(D * H) + dc * H(0), rolled and cut
//   7. audio stage 1: sliced X bins * filt1 -> 1024-pt iFFT   (core.py:854-875)
//      -> instantaneous frequency, one wavefront per channel
//
// Real transforms of length N use one M-point complex transform plus the
// even/odd split step.  Every intermediate stays in registers and LDS: HBM
// sees the input block and the requested outputs only.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "fft.hpp"
#include "internal.hpp"

namespace ldgpu {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

int hip_fail(hipError_t e, const char* what) {
  set_error("%s failed: %s", what, hipGetErrorString(e));
  return LD_EHIP;
}

constexpr int kN = 32768;      // block length handled by these kernels
constexpr int kM = kN / 2;     // complex points per transform
// Threads per workgroup for the M-point FFT: 1024 (16 points per thread,
// radix 16 x 16 x 16 x 4, four wavefronts per SIMD to hide LDS and barrier
// latency) or 512 (32 points, radix 32 x 32 x 16, two per SIMD).
constexpr int kT = 1024;
constexpr int kP = kM / kT;    // points per thread
constexpr int kNA = 1024;      // audio slice bins at N = 32768
constexpr float kTwoPi = 6.283185307179586f;

using FwdM = std::conditional_t<kT == 512, WgFft<kM, -1>, WgFft16kT1024<-1>>;
using InvA = WgFft<kNA, +1, kSyncWave>;

template <int M>
struct LdsT;

// Pass twiddles of the M-point transform, kept in LDS for the whole kernel
// (copied once from the global table).  Row k of a pass after NS points holds
// exp(-2 pi i 2^i k / (NS R)), i < log2 R; long rows are products of two
// short tables (k = F h + l).
template <int T>
struct TwLdsT;
template <>
struct TwLdsT<512> {   // passes (R 32, NS 32), (R 16, NS 1024 = 32 h + l)
  float2 p2[32 * 5];
  float2 a[32 * 4];
  float2 b[32 * 4];
  static constexpr int count = 32 * 5 + 32 * 4 + 32 * 4;
  template <int NS>
  __device__ __forceinline__ float2 get(int k, int i) const {
    if constexpr (NS == 32) {
      return p2[k * 5 + i];
    } else {
      static_assert(NS == 1024, "M-point transform passes");
      return cmul(a[(k >> 5) * 4 + i], b[(k & 31) * 4 + i]);
    }
  }
  static void fill(float2* t, float2 (*cexp)(double, double)) {
    for (int k = 0; k < 32; ++k)
      for (int i = 0; i < 5; ++i) *t++ = cexp((double)(1 << i) * k, 1024);
    for (int h = 0; h < 32; ++h)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * 32 * h, 16384);
    for (int l = 0; l < 32; ++l)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * l, 16384);
  }
};
template <>
struct TwLdsT<1024> {  // passes (R 16, NS 16), (R 16, NS 256 = 16 h + l), (R 4, NS 4096 = 64 h + l)
  float2 p16[16 * 4];
  float2 a256[16 * 4];
  float2 b256[16 * 4];
  float2 a4k[64 * 2];
  float2 b4k[64 * 2];
  static constexpr int count = 16 * 4 * 3 + 64 * 2 * 2;
  template <int NS>
  __device__ __forceinline__ float2 get(int k, int i) const {
    if constexpr (NS == 16) {
      return p16[k * 4 + i];
    } else if constexpr (NS == 256) {
      return cmul(a256[(k >> 4) * 4 + i], b256[(k & 15) * 4 + i]);
    } else {
      static_assert(NS == 4096, "M-point transform passes");
      return cmul(a4k[(k >> 6) * 2 + i], b4k[(k & 63) * 2 + i]);
    }
  }
  static void fill(float2* t, float2 (*cexp)(double, double)) {
    for (int k = 0; k < 16; ++k)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * k, 256);
    for (int h = 0; h < 16; ++h)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * 16 * h, 4096);
    for (int l = 0; l < 16; ++l)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * l, 4096);
    for (int h = 0; h < 64; ++h)
      for (int i = 0; i < 2; ++i) *t++ = cexp((double)(1 << i) * 64 * h, 16384);
    for (int l = 0; l < 64; ++l)
      for (int i = 0; i < 2; ++i) *t++ = cexp((double)(1 << i) * l, 16384);
  }
};
using TwLds = TwLdsT<kT>;
constexpr int kTwLdsCount = TwLds::count;

template <class TW>
struct TwSrcLdsT {
  const TW* t;
  template <int NS>
  __device__ __forceinline__ float2 get(int k, int i) const {
    return t->template get<NS>(k, i);
  }
  template <int NS, std::enable_if_t<HasTwRow<TW, NS>::value, int> = 0>
  __device__ __forceinline__ float2 row(int k, int r) const {
    return t->template row<NS>(k, r);
  }
};
using TwSrcLds = TwSrcLdsT<TwLds>;

// The 8192-point, 512-thread transform's pass twiddles: passes (R 16, NS 16),
// (R 16, NS 256 = 16 h + l), (R 2, NS 4096 = 64 h + l).
struct TwLds8k {
  float2 p16[16 * 4];
  float2 a256[16 * 4];
  float2 b256[16 * 4];
  float2 a4k[64];
  float2 b4k[64];
  static constexpr int count = 16 * 4 * 3 + 64 * 2;
  template <int NS>
  __device__ __forceinline__ float2 get(int k, int i) const {
    if constexpr (NS == 16) {
      return p16[k * 4 + i];
    } else if constexpr (NS == 256) {
      return cmul(a256[(k >> 4) * 4 + i], b256[(k & 15) * 4 + i]);
    } else {
      static_assert(NS == 4096, "M-point transform passes");
      return cmul(a4k[k >> 6], b4k[k & 63]);
    }
  }
  static void fill(float2* t, float2 (*cexp)(double, double)) {
    for (int k = 0; k < 16; ++k)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * k, 256);
    for (int h = 0; h < 16; ++h)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * 16 * h, 4096);
    for (int l = 0; l < 16; ++l)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * l, 4096);
    for (int h = 0; h < 64; ++h) *t++ = cexp(64.0 * h, 8192);
    for (int l = 0; l < 64; ++l) *t++ = cexp((double)l, 8192);
  }
};
static_assert(TwLds8k::count <= TwLds::count, "the handle table holds either image");

// The 4096-point, 256-thread transform's pass twiddles: passes (R 16, NS 16),
// its full rows W_256^{r k} (r, k < 16) stored by r (a read's lanes take
// consecutive k: no bank conflicts), and (R 16, NS 256 = 16 h + l).
struct TwLds4k {
  float2 p16[16 * 16];
  float2 a256[16 * 4];
  float2 b256[16 * 4];
  static constexpr int count = 16 * 16 + 16 * 4 * 2;
  template <int NS>
  __device__ __forceinline__ float2 get(int k, int i) const {
    if constexpr (NS == 16) {
      return p16[(1 << i) * 16 + k];
    } else {
      static_assert(NS == 256, "M-point transform passes");
      return cmul(a256[(k >> 4) * 4 + i], b256[(k & 15) * 4 + i]);
    }
  }
  template <int NS, std::enable_if_t<NS == 16, int> = 0>
  __device__ __forceinline__ float2 row(int k, int r) const {
    return p16[r * 16 + k];
  }
  static void fill(float2* t, float2 (*cexp)(double, double)) {
    for (int r = 0; r < 16; ++r)
      for (int k = 0; k < 16; ++k) *t++ = cexp((double)r * k, 256);
    for (int h = 0; h < 16; ++h)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * 16 * h, 4096);
    for (int l = 0; l < 16; ++l)
      for (int i = 0; i < 4; ++i) *t++ = cexp((double)(1 << i) * l, 4096);
  }
};

static_assert(TwLds4k::count <= TwLds::count, "the handle table holds either image");

template <int M>
struct alignas(16) LdsT {
  float re[LdsSize<M>::floats];
  float im[LdsSize<M>::floats];
};

// Block-length configurations of the single-level demodulator: N real
// samples per block, one M = N / 2 point complex transform by T threads of P
// points, audio slices of NA bins.
template <int N_>
struct DemodCfg;
template <>
struct DemodCfg<32768> {
  static constexpr int N = 32768, M = 16384, T = 1024, P = 16, NA = 1024;
  static constexpr int NotchMax = 508;  // bins of the spur notch window (148 used at 40 MSPS)
  static constexpr int Wg = 1;  // workgroups per CU (LDS: 156 KB each)
  static constexpr int WavesPerSimd = Wg * T / 64 / 4;  // __launch_bounds__'s second argument
  using Fwd = WgFft16kT1024<-1>;
  using FwdMfma = WgFft16kMfma;
  using Tw = TwLdsT<1024>;
  using Inv = WgFft<1024, +1, kSyncWave>;
  static constexpr uint32_t HalfBytes = (M + 1) * 8, StrideB = T * 8;
  static __device__ __forceinline__ TwGlobal audio_tw(rsrc_t r) {
    return TwGlobal{r, TwLayout::k1k_p2, TwLayout::k1k_p3, 16, 4, 2};
  }
};
template <>
struct DemodCfg<16384> {
  static constexpr int N = 16384, M = 8192, T = 512, P = 16, NA = 512;
  static constexpr int NotchMax = 508;
  static constexpr int Wg = 2;  // two workgroups per CU (LDS: 80 KB each; 128 VGPRs)
  static constexpr int WavesPerSimd = Wg * T / 64 / 4;
  using Fwd = WgFft8kT512<-1>;
  using FwdMfma = WgFft8kMfma;
  using Tw = TwLds8k;
  using Inv = WgFft<512, +1, kSyncWave>;
  static constexpr uint32_t HalfBytes = (M + 1) * 8, StrideB = T * 8;
  static __device__ __forceinline__ TwGlobal audio_tw(rsrc_t r) {
    return TwGlobal{r, TwLayout::k512_p2, TwLayout::k512_p3, 8, 3, 3};
  }
};
using Cfg32 = DemodCfg<32768>;
using Lds = LdsT<kM>;
using Cfg16 = DemodCfg<16384>;
template <>
struct DemodCfg<8192> {
  static constexpr int N = 8192, M = 4096, T = 256, P = 16, NA = 256;
  using Fwd = WgFft4kT256<-1>;
  using FwdMfma = WgFft4kMfma;
  using Tw = TwLds4k;
  using Inv = WgFft<256, +1, kSyncWave>;
  static constexpr uint32_t HalfBytes = (M + 1) * 8, StrideB = T * 8;
  static constexpr int NotchMax = 128;  // 37 bins used at 40 MSPS
  static constexpr int Wg = 4;          // four workgroups per CU (LDS: 40 KB each; 128 VGPRs)
  static constexpr int WavesPerSimd = Wg * T / 64 / 4;
  static __device__ __forceinline__ TwGlobal audio_tw(rsrc_t r) {
    TwGlobal g{r, TwLayout::k256_p2, TwLayout::k256_p3, 4, 2, 2};
    g.off_p4 = TwLayout::k256_p4;
    g.ns_p3 = 16;
    g.lr_p4 = 2;
    return g;
  }
};
using Cfg8 = DemodCfg<8192>;

// W_N^(tid + T m) from the thread's split[tid] = exp(-2 pi i tid / N) and a
// compile-time constant exp(-2 pi i m T / N)
template <class C, int M_>
__device__ __forceinline__ float2 split_tw(float2 wt) {
  return cmul(wt, make_float2(Tw<C::N / C::T, M_, -1>::re, Tw<C::N / C::T, M_, -1>::im));
}
// Elements m and m + 8 of a thread are 8 T bins apart: with N / T = 32 and
// P = 16, W_N^(T (m + 8)) = W_32^(m + 8) = -i W_32^m, so the twiddle of
// element m + 8 is the one of element m times -i (an exact rotation).
template <class C>
__device__ __forceinline__ float2 split_tw_half(float2 w) {
  static_assert(C::N / C::T == 32 && C::P == 16, "pairs m, m + 8 need N / T = 32, P = 16");
  return make_float2(w.y, -w.x);
}

// Everything one fused launch needs.  Filters are half spectra (bins 0..M)
// with the inverse-transform scale folded in.  The c2r outputs (rfhpf and the
// post filters) are described uniformly.
struct C2rOut {
  const float2* h;
  float* out;
  float dc;
  int shift, cut, len;
};

struct DemodArgs {
  const void* in;
  int64_t first, step;
  int in_bytes;            // bytes per input sample (2: int16, 4: float32, 0: packed 10-bit)
  uint32_t in_total;       // packed input: bytes of the whole input buffer
  const float2* tw_rows;   // Stockham pass twiddles of the audio transform (TwLayout)
  const float2* tw_lds;    // TwLds image, copied to LDS by every workgroup
  const float2* tw_split;  // split[t] = exp(-2 pi i t / N), t < kT
  const u32x4_t* mfma_g;   // split-BF16 MFMA variant: DFT-64 fragments (kMfmaTabEntries)
  const MfmaTw* mfma_tw;   // and its twiddle factor tables, copied to LDS
  const float2* h_rfv;
  const float2* h_rfv_odd; // RFVideo times exp(+2 pi i k / N): the odd analytic samples' table
  C2rOut hpf;              // out == nullptr: skip
  C2rOut post[4];
  int npost;
  float* demod_raw;
  int vcut, vlen;
  int64_t vstride;
  float2* ws_even;         // per block M float2: even analytic samples
  const float2* h_audio[2];
  int lowbin[2];
  float rate[2], low_freq[2];
  float* out_audio[2];
  int acut, alen;
  int64_t astride;
  float fs, clip_lo, clip_hi, dc;
  int notch_lo, notch_hi;  // spur notch window, empty: off
  const float2* h_efm;     // EFM equaliser (Hermitian half spectrum, c2r scale)
  int16_t* efm;            // EFM output (cut and strided as the video outputs)
  uint32_t* below;         // in-pulse bits of one post output (video_stride / 32 words per block)
  float below_highf;       // largest float not above the threshold
  int below_post;          // index of that output in post[], -1: off
};

constexpr int kNotchMax = 508;  // bins in the notch window

constexpr uint32_t kHalfBytes = (kM + 1) * 8;  // a half-spectrum table
constexpr uint32_t kStrideB = kT * 8;           // bytes between a thread's elements m and m+1

// ---------------------------------------------------------------- helpers
// After a forward M-point FFT of y[m] = x[2m] + i x[2m+1] (v[m] = Y[k],
// k = tid+T*m) form TWICE the half spectrum X[k] = E + W_N^k O of the real
// sequence x in x[m], and 2 X[M] (real) in xM on the thread holding k = 0:
// the halvings of E and O are left to the tables this spectrum meets (the
// c2r tables carry a factor 1/4, see c2r_pre; the analytic signal, the
// notch decision and the audio slices only see phases or ratios), which
// saves a multiply per component.  Scaling by 2 is exact in float32, so the
// outputs are those of the halved form.
template <class C>
__device__ __forceinline__ void r2c_split(const float2 (&v)[C::P], float2 (&x)[C::P], float& xM, LdsT<C::M>& s, float2 wt,
                                          int tid) {
  const LdsPair own(s.re, s.im, pidx(tid)), mir(s.re, s.im, pidx(-tid));
  static_for<0, C::P>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    own.template st<POff<C::T * m>::value>(v[m]);
  });
  if (tid == 0) {  // Y[M] = Y[0] (periodic), read by k = 0
    s.re[pidx(C::M)] = v[0].x;
    s.im[pidx(C::M)] = v[0].y;
  }
  __syncthreads();
  // elements m and m + 8 together, sharing one twiddle product
  static_for<0, C::P / 2>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const float2 w = split_tw<C, m>(wt);
    const float2 wh = split_tw_half<C>(w);
    {
      float2 p = v[m];
      float2 q = mir.template ld<POff<C::M - C::T * m>::value>();
      float2 E = make_float2(p.x + q.x, p.y - q.y);
      float2 O = make_float2(p.y + q.y, q.x - p.x);
      x[m] = cadd(E, cmul(w, O));
      if (m == 0 && tid == 0) xM = 2.0f * (p.x - p.y);
    }
    {
      constexpr int m2 = m + C::P / 2;
      float2 p = v[m2];
      float2 q = mir.template ld<POff<C::M - C::T * m2>::value>();
      float2 E = make_float2(p.x + q.x, p.y - q.y);
      float2 O = make_float2(p.y + q.y, q.x - p.x);
      x[m2] = cadd(E, cmul(wh, O));
    }
    // bound how far the scheduler hoists the mirror loads (register pressure)
    if constexpr ((m & 3) == 3) __builtin_amdgcn_sched_barrier(0);
  });
  __syncthreads();
}

// Input of a real inverse transform of Z = X * H (Hermitian half spectrum;
// x holds 2 X (r2c_split) and the table H / (4 M), so that x H = Z / (2 M):
// E and O come out without their halvings; its bins tid + T m arrive in v,
// bin M in hM), conjugated so that a forward FFT produces the conjugate of
// the inverse: v[m] = conj(E + i O).
template <class C>
__device__ __forceinline__ void c2r_pre(const float2 (&x)[C::P], float xM, float2 hM, float2 (&v)[C::P], LdsT<C::M>& s,
                                        float2 wt, int tid) {
  const LdsPair own(s.re, s.im, pidx(tid)), mir(s.re, s.im, pidx(-tid));
  static_for<0, C::P>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    float2 z = cmul(x[m], v[m]);
    v[m] = z;
    own.template st<POff<C::T * m>::value>(z);
  });
  if (tid == 0) {  // Z[M], read by k = 0
    float2 zM = cscale(hM, xM);
    s.re[pidx(C::M)] = zM.x;
    s.im[pidx(C::M)] = zM.y;
  }
  __syncthreads();
  static_for<0, C::P / 2>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const float2 w = split_tw<C, m>(wt);
    const float2 wh = split_tw_half<C>(w);
    {
      float2 B = cconj(mir.template ld<POff<C::M - C::T * m>::value>());
      float2 A = v[m];
      float2 E = make_float2(A.x + B.x, A.y + B.y);
      float2 O = cmulc(make_float2(A.x - B.x, A.y - B.y), w);
      v[m] = make_float2(E.x - O.y, -(E.y + O.x));
    }
    {
      constexpr int m2 = m + C::P / 2;
      float2 B = cconj(mir.template ld<POff<C::M - C::T * m2>::value>());
      float2 A = v[m2];
      float2 E = make_float2(A.x + B.x, A.y + B.y);
      float2 O = cmulc(make_float2(A.x - B.x, A.y - B.y), wh);
      v[m2] = make_float2(E.x - O.y, -(E.y + O.x));
    }
    if constexpr ((m & 3) == 3) __builtin_amdgcn_sched_barrier(0);
  });
  // the FFT's first pass synchronises before it writes LDS
}

// A transform's output v goes to LDS at the thread's own positions (the
// transform's last pass synchronises after its final LDS reads), which
// frees v to receive the next phase's filter before this phase's global
// stores are issued: the in-order memory counter then lets the next loads
// complete without waiting for the stores, and the next transform runs
// while the stores drain.
template <class C>
__device__ __forceinline__ void stage_own(const float2 (&v)[C::P], LdsT<C::M>& s, int tid) {
  const LdsPair own(s.re, s.im, pidx(tid));
  static_for<0, C::P>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    own.template st<POff<C::T * m>::value>(v[m]);
  });
}

// Bits of a paired in-pulse word (even samples in bits 0..15, odd in
// 16..31) that hold samples 0 .. left-1 of the word, left < 32.
__device__ __forceinline__ uint32_t below_mask_of(uint32_t left) {
  const uint32_t ne = (left + 1) >> 1, no = left >> 1;  // even / odd samples kept
  return ((1u << ne) - 1u) | (((1u << no) - 1u) << 16);
}

// Write y (from a forward FFT of conjugated input: y[2m] = v.x, y[2m+1] =
// -v.y, staged in LDS) as out[i] = y[(i + cut + shift) mod N] + dc for
// i < len; stores with i >= len fall outside the descriptor and are dropped.
//
// words != nullptr (offset a multiple of 32): also the in-pulse bits of the
// output, findpulses' y <= high as y <= highf (highf: the largest float not
// above high).  A wavefront's elements m cover 128 consecutive y; per 32 of
// them (one output word) bits 0..15 are the even samples' and 16..31 the
// odd samples' flags, from two ballots; lanes 0..3 store the four words.
template <class C>
__device__ __forceinline__ void c2r_store(const C2rOut& o, int64_t base, int tid, LdsT<C::M>& s,
                                          uint32_t* words = nullptr, float highf = 0.0f) {
  const rsrc_t r = make_rsrc(o.out + base, (uint32_t)o.len * 4u);
  const LdsPair own(s.re, s.im, pidx(tid));
  const int off = o.cut + o.shift;
  const float dc = o.dc;
  if (((off | o.len) & 1) == 0) {
    // even offset and length: y[2m], y[2m+1] land on an aligned pair that
    // is wholly inside or wholly outside [0, len): one 8-byte store
    const int q = tid & 3;
    static_for<0, C::P>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const float2 v = own.template ld<POff<C::T * m>::value>();
      int n = 2 * (tid + C::T * m);
      uint32_t i0 = (uint32_t)((n - off) & (C::N - 1));
      const float y0 = v.x + dc, y1 = dc - v.y;
      bstore2<kStream>(r, make_float2(y0, y1), i0 * 4u, 0u);
      if (words != nullptr) {
        const unsigned long long e0 = __ballot(y0 <= highf), e1 = __ballot(y1 <= highf);
        const uint32_t w = ((uint32_t)(e0 >> (16 * q)) & 0xffffu) | ((uint32_t)(e1 >> (16 * q)) << 16);
        const uint32_t j0 = (uint32_t)((2 * (C::T * m + (tid & ~63)) + 32 * q - off) & (C::N - 1));
        if ((tid & 63) < 4 && j0 < (uint32_t)o.len) {
          const uint32_t left = (uint32_t)o.len - j0;
          words[j0 >> 5] = left >= 32 ? w : (w & below_mask_of(left));
        }
      }
    });
    return;
  }
  static_for<0, C::P>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const float2 v = own.template ld<POff<C::T * m>::value>();
    int n = 2 * (tid + C::T * m);
    uint32_t i0 = (uint32_t)((n - off) & (C::N - 1));
    uint32_t i1 = (uint32_t)((n + 1 - off) & (C::N - 1));
    bstore1<kStream>(r, v.x + dc, i0 * 4u, 0u);
    bstore1<kStream>(r, dc - v.y, i1 * 4u, 0u);
  });
}

// atan2(y, x) in [-pi, pi] for finite y, x (the complex products of the
// discriminator), about 2 float32 ulp: the octant reduction a = min / max
// (hardware reciprocal refined by one Newton step), atan(a) on [0, 1] as
// a + a s Q(s), s = a^2, with Q a degree-7 polynomial fitted to
// (atan(a) / a - 1) / s by iteratively reweighted least squares on the
// relative error (max error 1.7e-8 before float32 rounding), then the
// reflections pi/2 - t (|y| > |x|), pi - t (x < 0) and the sign of y.
// (0, 0) gives 0, as numpy's angle(0) does.
__device__ __forceinline__ float disc_atan2(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  float r = __builtin_amdgcn_rcpf(mx);
  r = fmaf(fmaf(-mx, r, 1.0f), r, r);
  const float a = mx > 0.0f ? mn * r : 0.0f;
  const float s2 = a * a;
  float q = 0.002920095343142748f;
  q = fmaf(q, s2, -0.016365554183721542f);
  q = fmaf(q, s2, 0.04320802912116051f);
  q = fmaf(q, s2, -0.07551892846822739f);
  q = fmaf(q, s2, 0.10665856301784515f);
  q = fmaf(q, s2, -0.1421101838350296f);
  q = fmaf(q, s2, 0.19993768632411957f);
  q = fmaf(q, s2, -0.33333152532577515f);
  float t = fmaf(a * s2, q, a);
  t = ay > ax ? 1.5707963267948966f - t : t;
  t = x < 0.0f ? 3.141592653589793f - t : t;
  return copysignf(t, y);
}

// Hz of the phase increment d (radians, in (-pi, pi]) wrapped to [0, 2pi):
// utils.py:921-926
__device__ __forceinline__ float inc_to_hz(float d, float hz_per_rad, float fs) {
  float hz = d * hz_per_rad;
  return d < 0.0f ? hz + fs : hz;
}

// Analogue audio stage 1 for one channel by one wavefront (utils.py:898-926,
// core.py:857-866): slice bins, 1024-point inverse FFT, instantaneous
// frequency at the decimated rate plus the slice's low frequency.
template <class C>
__device__ void audio_wave(const DemodArgs& a, int c, const float2* __restrict__ slice, float* lre, float* lim,
                           int b, const TwGlobal& twr) {
  const int lane = threadIdx.x & 63;
  const rsrc_t F = make_rsrc(a.h_audio[c], C::NA * 8u);
  float2 v[C::NA / 64];
#pragma unroll
  for (int m = 0; m < C::NA / 64; ++m) {
    int k = lane + 64 * m;
    float2 f = bload2(F, lane * 8u, m * 512u);
    float2 z = make_float2(0.0f, 0.0f);
    if (k < C::NA / 2)
      z = cmul(slice[k], f);
    else if (k == C::NA / 2)  // bin N - lowbin - nbins/2 of the full spectrum
      z = cmul(cconj(slice[C::NA / 2]), f);
    v[m] = z;
  }
  C::Inv::run(v, lre, lim, twr);
#pragma unroll
  for (int m = 0; m < C::NA / 64; ++m) {
    int n = lane + 64 * m;
    lre[pidx(n)] = v[m].x;
    lim[pidx(n)] = v[m].y;
  }
  lds_barrier<kSyncWave>();
  const float rate = a.rate[c], lf = a.low_freq[c];
  const float hz_per_rad = rate / kTwoPi;
  const rsrc_t out = make_rsrc(a.out_audio[c] + (int64_t)b * a.astride, (uint32_t)a.alen * 4u);
#pragma unroll
  for (int m = 0; m < C::NA / 64; ++m) {
    int n = lane + 64 * m;
    float hz = 0.0f;
    if (n > 0) {
      int pn = n - 1;
      float2 p = cmulc(v[m], make_float2(lre[pidx(pn)], lim[pidx(pn)]));
      hz = inc_to_hz(disc_atan2(p.y, p.x), hz_per_rad, rate);
    }
    int i = n - a.acut;  // i < 0 wraps to a huge offset: dropped
    bstore1<kStream>(out, hz + lf, (uint32_t)i * 4u, 0u);
  }
}

// PAL LD-V4300D spur notch (core.py:774-798), decision, by one wavefront:
// from the magnitudes of window bins [lo, hi) in mag, the window's mean +
// 3 std (population, in double) flags the bins above it; flag[i] belongs to
// bin lo - 2 + i (two zero guards each side).  Bins that are flagged or next
// to a flagged bin are zeroed in the analytic-signal and audio inputs (their
// mirror bins are implied by the half-spectrum form).
__device__ __forceinline__ void notch_decide(const float* mag, int* flag, double* red, int lo, int hi) {
  int lane = threadIdx.x, n = hi - lo;
  // opaque: keeps the compiler from hoisting this rarely used work into the
  // kernel prologue, where its registers would be held (and spilled)
  asm volatile("" : "+v"(lane), "+v"(n));
  double sum = 0.0;
  for (int i = lane; i < n; i += 64) sum += mag[i];
  red[lane] = sum;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  sum = 0.0;
  for (int l = 0; l < 64; ++l) sum += red[l];
  const double mean = sum / (double)n;
  double var = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double d = mag[i] - mean;
    var += d * d;
  }
  __builtin_amdgcn_wave_barrier();
  red[64 + lane] = var;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  var = 0.0;
  for (int l = 0; l < 64; ++l) var += red[64 + l];
  const double thr = mean + 3.0 * sqrt(var / (double)n);
  for (int i = lane; i < n + 4; i += 64) {
    const int w = i - 2;
    flag[i] = (w >= 0 && w < n) ? ((double)mag[w] > thr) : 0;
  }
}

// Whether spectrum bin k is zeroed by the notch (flags from notch_decide).
__device__ __forceinline__ bool notch_zero(const int* flag, int lo, int hi, int k) {
  if (k < lo - 1 || k > hi) return false;
  const int i = k - lo + 2;
  return (flag[i - 1] | flag[i] | flag[i + 1]) != 0;
}

// The thread's bins tid + T m of a half-spectrum table (bin M of every
// table sits in LDS, see bin_m).
template <class C>
__device__ __forceinline__ void load_half(const float2* table, float2 (&h)[C::P], int tid) {
  const rsrc_t H = make_rsrc(table, C::HalfBytes);
  const uint32_t vo = tid * 8u;
#pragma unroll
  for (int m = 0; m < C::P; ++m) h[m] = bload2(H, vo, m * C::StrideB);
}

// ---------------------------------------------------------------- kernel
// One workgroup demodulates one block end to end (RFDecode.demodblock_cpu,
// core.py:755-879); only the input and the requested outputs touch HBM
// (plus the even analytic samples, which wait in a per-block workspace).
// The work is a sequence of phases around ONE forward FFT instance (inverse
// transforms run on conjugated data), which keeps the kernel's code small
// enough for the instruction cache:
//   0        block -> X                         (rFFT)
//   1        X * Frfhpf -> rfhpf                (c2r, optional)
//   2, 3     X * RFVideo -> analytic even / odd (inverse)
//   4        inst. frequency, clip -> D         (rFFT)
//   5..8     D * post filter -> video outputs   (c2r)
enum { kPhIn = 0, kPhHpf = 1, kPhEven = 2, kPhOdd = 3, kPhDemod = 4, kPhPost = 5, kPhEfm = -1 };

// The EFM output (core.py:847-851) from a transform staged as c2r_store
// reads it: np.int16(np.clip(y, -32768, 32767)) (truncation toward zero)
// for i < len; pairs on even offsets go out as one 32-bit store.
template <class C>
__device__ __forceinline__ void efm_store(int16_t* outp, int cut, int len, int tid, LdsT<C::M>& s) {
  const rsrc_t r = make_rsrc(outp, (uint32_t)len * 2u);
  const LdsPair own(s.re, s.im, pidx(tid));
  auto q = [](float y) -> uint32_t {
    return (uint32_t)(uint16_t)(int16_t)(int)fminf(fmaxf(y, -32768.0f), 32767.0f);
  };
  const bool pairs = ((cut | len) & 1) == 0;
  static_for<0, C::P>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const float2 v = own.template ld<POff<C::T * m>::value>();
    const int n = 2 * (tid + C::T * m);
    const uint32_t i0 = (uint32_t)((n - cut) & (C::N - 1));
    if (pairs) {
      __builtin_amdgcn_raw_buffer_store_b32(q(v.x) | (q(-v.y) << 16), r, (int)(i0 * 2u), 0, kStream);
    } else {
      const uint32_t i1 = (uint32_t)((n + 1 - cut) & (C::N - 1));
      __builtin_amdgcn_raw_buffer_store_b16((uint16_t)q(v.x), r, (int)(i0 * 2u), 0, kStream);
      __builtin_amdgcn_raw_buffer_store_b16((uint16_t)q(-v.y), r, (int)(i1 * 2u), 0, kStream);
    }
  });
}

// kPacked: the input is the packed 10-bit format; kEfm: the EFM output is
// produced too, as one more c2r phase of the input spectrum between the RF
// high-pass and the analytic signal (separate instantiations, so that the
// plain kernel's register allocation is not touched).
//
// s0: the block's first input sample; bo: its output slot (outputs at bo
// times the strides); bw: its workspace slot.
//
// kMfma = 2 or 3: the transforms' radix-64 stages run as split-BF16 matrix
// products of that many bf16 terms per operand on the matrix cores
// (C::FwdMfma) instead of FP32 radix-16 passes; 0: FP32 passes.
template <class C, bool kPacked, bool kEfm, int kMfma = 0>
__device__ __forceinline__ void demod_block(const DemodArgs& a, int64_t s0, int bo, int bw) {
  __shared__ LdsT<C::M> s;
  __shared__ MfmaTw mtw;
  __shared__ float2 aslice[2][C::NA / 2 + 1];
  __shared__ typename C::Tw twl;
  __shared__ int notch_flag[C::NotchMax + 4];
  __shared__ uint32_t notch_mask[C::T];  // bit m: bin tid + T m zeroed by the spur notch
  // bin M of every filter table: 0 rfhpf, 1 RFVideo, 2.. post filters, 6 EFM
  __shared__ float2 bin_m[2 + 4 + 1];
  const int tid = threadIdx.x;
  const rsrc_t zev = make_rsrc(a.ws_even + (int64_t)bw * C::M, C::M * 8u);
  const uint32_t vo = tid * 8u;
  // v: the transform's data; between phases it carries the next phase's
  // filter (bins tid + T m), loaded after the transform and before that
  // phase's output stores (see stage_own)
  float2 v[C::P], x[C::P];
  if constexpr (!kPacked) {
    // the block's samples (core.py:761: y[j] = x[2j] + i x[2j+1]), loaded
    // before the table setup so that the loads overlap it; int16 samples
    // stay raw bits in v until phase 0 converts them (a conversion here
    // would wait for the loads)
    const rsrc_t in = make_rsrc((const char*)a.in + s0 * a.in_bytes, (uint32_t)(C::N * a.in_bytes));
    if (a.in_bytes == 4) {
#pragma unroll
      for (int m = 0; m < C::P; ++m) v[m] = bload2(in, vo, m * C::StrideB);
    } else if ((s0 & 1) == 0) {
#pragma unroll
      for (int m = 0; m < C::P; ++m) v[m].x = bload1(in, tid * 4u, m * (C::T * 4u));  // both samples
    } else {  // odd start: 16-bit loads
#pragma unroll
      for (int m = 0; m < C::P; ++m) {
        uint32_t o = tid * 4u + m * (C::T * 4u);
        v[m].x = __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(in, (int)o, 0, 0));
        v[m].y = __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(in, (int)o + 2, 0, 0));
      }
    }
  }
  {  // twiddle tables into LDS
    const rsrc_t src = make_rsrc(a.tw_lds, C::Tw::count * 8u);
    float2* dst = reinterpret_cast<float2*>(&twl);
    for (int i = tid; i < C::Tw::count; i += C::T) dst[i] = bload2(src, (uint32_t)i * 8u, 0u);
    if (tid < 2 + a.npost) {
      const float2* t = tid == 0 ? a.hpf.h : (tid == 1 ? a.h_rfv : a.post[tid - 2].h);
      bin_m[tid] = t ? bload2(make_rsrc(t, C::HalfBytes), 0u, C::M * 8u) : make_float2(0.0f, 0.0f);
    }
    if constexpr (kEfm) {
      if (tid == C::T - 1) bin_m[6] = bload2(make_rsrc(a.h_efm, C::HalfBytes), 0u, C::M * 8u);
    }
    if constexpr (kMfma) {
      const rsrc_t mt = make_rsrc(a.mfma_tw, sizeof(MfmaTw));
      if (tid < (int)(sizeof(MfmaTw) / 8)) reinterpret_cast<float2*>(&mtw)[tid] = bload2(mt, (uint32_t)tid * 8u, 0u);
    }
  }
  __syncthreads();
  const TwSrcLdsT<typename C::Tw> twf{&twl};
  const float2 wt0 = bload2(make_rsrc(a.tw_split, C::T * 8u), vo, 0u);
  float xM = 0.0f;
  const int last = kPhPost + a.npost;
  const bool has_hpf = a.hpf.out != nullptr;

  if constexpr (kPacked) {
    // packed 10-bit input (utils.py:371-390): sample j is the big-endian
    // 16-bit word at byte 5 (j >> 2) + (j & 3), shifted right by
    // 6 - 2 (j & 3); a sample pair always lies in the 4 bytes from its first
    // byte.  The block's bytes are staged in LDS (the exchange area is free
    // before the first transform) as dwords from the one holding its first
    // byte, then each thread decodes its 32 pairs into v, as phase 0 of the
    // int16 path loads them.
    const rsrc_t pk = make_rsrc(a.in, a.in_total);
    const uint32_t q5 = (uint32_t)(5 * (s0 >> 2));
    const uint32_t r = (uint32_t)(s0 & 3);
    const uint32_t A0 = q5 & ~3u;
    const uint32_t lastb = q5 + 5u * ((r + (uint32_t)C::N - 1u) >> 2) + ((r + (uint32_t)C::N - 1u) & 3u) + 1u;
    const int nwords = (int)((lastb - A0) / 4u + 1u);
    uint32_t* w32 = reinterpret_cast<uint32_t*>(s.re);
    for (int i = tid; i < nwords; i += C::T) w32[i] = __builtin_bit_cast(uint32_t, bload1(pk, A0 + 4u * (uint32_t)i, 0u));
    __syncthreads();
    static_for<0, C::P>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const uint32_t jr = r + 2u * (uint32_t)(tid + C::T * m);
      const uint32_t B = q5 + 5u * (jr >> 2) + (jr & 3u) - A0;
      const uint32_t W = __builtin_amdgcn_alignbyte(w32[(B >> 2) + 1u], w32[B >> 2], B & 3u);
      const uint32_t r0 = jr & 3u;
      const uint32_t b0 = W & 0xffu, b1 = (W >> 8) & 0xffu, b2 = (W >> 16) & 0xffu, b3 = W >> 24;
      const uint32_t v0 = (((b0 << 8) | b1) >> (6u - 2u * r0)) & 0x3ffu;
      const uint32_t v1 = r0 < 3u ? ((((b1 << 8) | b2) >> (4u - 2u * r0)) & 0x3ffu) : ((((b2 << 8) | b3) >> 6) & 0x3ffu);
      // (code - 512) << 6 as the int16 path sees it, exactly
      v[m] = make_float2((float)((int)v0 - 512) * 64.0f, (float)((int)v1 - 512) * 64.0f);
    });
    __syncthreads();    // the first transform reuses the exchange area
  }

  for (int it = 0; it < last + (kEfm ? 1 : 0); ++it) {
    int ph = it;
    if constexpr (kEfm) ph = it == kPhEven ? kPhEfm : (it > kPhEven ? it - 1 : it);
    // the split twiddles derived from wt are loop-invariant; keep the
    // compiler from hoisting (and spilling) all 32 of them out of the loop
    float2 wt = wt0;
    int tidv = tid;  // opaque copy for per-element store offsets (same reason)
    asm volatile("" : "+v"(wt.x), "+v"(wt.y), "+v"(tidv));
    // ---------------------------------------------------------- pre
    if (ph == kPhIn) {
      // the block's samples are in v (loaded or decoded before the phase
      // loop); int16 samples from their raw bits
      if constexpr (!kPacked) {
        if (a.in_bytes != 4) {
          const bool pair = (s0 & 1) == 0;
#pragma unroll
          for (int m = 0; m < C::P; ++m) {
            const uint32_t w0 = __float_as_uint(v[m].x), w1 = pair ? (w0 >> 16) : __float_as_uint(v[m].y);
            v[m] = make_float2((float)(int16_t)(w0 & 0xffffu), (float)(int16_t)(w1 & 0xffffu));
          }
        }
      }
    } else if (ph == kPhHpf) {
      // RF high-pass branch (core.py:769-772)
      if (!has_hpf) continue;
      c2r_pre<C>(x, xM, bin_m[0], v, s, wt, tidv);
    } else if (kEfm && ph == kPhEfm) {
      // EFM equaliser on the input spectrum (core.py:847)
      c2r_pre<C>(x, xM, bin_m[6], v, s, wt, tidv);
    } else if (ph == kPhEven || ph == kPhOdd) {
      // analytic signal (core.py:800-805), conjugated: even samples from
      // Z[k] (+ Z[M] folded into bin 0), odd samples from Z[k] W_N^-k
      // (- Z[M] folded into bin 0).  The odd phase's table is RFVideo
      // times W_N^-k (formed in double on the host; bin 0 unchanged).
      const float2 hM = cscale(bin_m[1], xM);  // RFVideo bin M
      const bool odd = (ph == kPhOdd);
      const uint32_t zmask = a.notch_hi > a.notch_lo ? notch_mask[tidv] : 0u;
      static_for<0, C::P>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        float2 z = cmul(x[m], v[m]);
        if (zmask & (1u << m)) z = make_float2(0.0f, 0.0f);
        if (m == 0 && tid == 0) z = odd ? csub(z, hM) : cadd(z, hM);
        v[m] = cconj(z);
      });
    } else if (ph == kPhDemod) {
      // instantaneous frequency (utils.py:898-926), clip (core.py:809).
      // The workspace (even) and v (odd) hold conj(z): arg(a b*) of the
      // conjugates is minus the wanted increment.  Odd samples go to LDS
      // shifted by one so each thread reads z[2j-1] at its own index j.
      const LdsPair sh(s.re, s.im, pidx(tidv + 1)), own(s.re, s.im, pidx(tidv));
      // the even samples' loads go out first, so that their latency
      // overlaps the exchange and its barrier; every one of them comes
      // before any demod_raw store (the in-order memory counter never
      // makes a load wait for a store)
      float2 ev[C::P];
#pragma unroll
      for (int m = 0; m < C::P; ++m) ev[m] = bload2(zev, vo, m * C::StrideB);
      static_for<0, C::P>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        sh.template st<POff<C::T * m>::value>(v[m]);
      });
      __syncthreads();
      const float hz_per_rad = a.fs / kTwoPi;
      static_for<0, C::P>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        float2 e = ev[m], od = v[m];
        float2 prev = own.template ld<POff<C::T * m>::value>();
        float2 p = cmulc(e, prev);
        // sample 0 has no predecessor: 0 (only element 0 of thread 0; the
        // test on the plain thread index, as tidv is opaque to the compiler
        // and would cost a branch per element)
        float h0 = inc_to_hz(-disc_atan2(p.y, p.x), hz_per_rad, a.fs);
        if (m == 0 && tid == 0) h0 = 0.0f;
        float2 q = cmulc(od, e);
        float h1 = inc_to_hz(-disc_atan2(q.y, q.x), hz_per_rad, a.fs);
        v[m] = make_float2(h0, h1);
      });
      if (a.demod_raw != nullptr) {
        const rsrc_t raw = make_rsrc(a.demod_raw + (int64_t)bo * a.vstride, (uint32_t)a.vlen * 4u);
        const bool pairs = ((a.vcut | a.vlen) & 1) == 0;
#pragma unroll
        for (int m = 0; m < C::P; ++m) {  // i < 0 wraps to a huge offset: dropped
          const uint32_t i0 = (uint32_t)(2 * (tidv + C::T * m) - a.vcut);
          if (pairs) {
            bstore2<kStream>(raw, v[m], i0 * 4u, 0u);  // aligned pair, wholly in or out
          } else {
            bstore1<kStream>(raw, v[m].x, i0 * 4u, 0u);
            bstore1<kStream>(raw, v[m].y, i0 * 4u + 4u, 0u);
          }
        }
      }
#pragma unroll
      for (int m = 0; m < C::P; ++m)
        v[m] = make_float2(fminf(fmaxf(v[m].x, a.clip_lo), a.clip_hi) - a.dc,
                           fminf(fmaxf(v[m].y, a.clip_lo), a.clip_hi) - a.dc);
    } else {
      // video post filters (core.py:811-820)
      c2r_pre<C>(x, xM, bin_m[2 + ph - kPhPost], v, s, wt, tidv);
    }

    // ---------------------------------------------------------- transform
    if constexpr (kMfma)
      C::FwdMfma::template run<kMfma>(v, s.re, s.im, twf, a.mfma_g, &mtw, tidv);
    else
      C::Fwd::run(v, s.re, s.im, twf, tidv);

    // ---------------------------------------------------------- post
    if (ph == kPhIn || ph == kPhDemod) {
      r2c_split<C>(v, x, xM, s, wt, tidv);
      if (ph == kPhIn) {
        if (a.notch_hi > a.notch_lo) {  // PAL spur notch decision (core.py:774-798)
          // the magnitudes and the reduction use the exchange area, free
          // from the split's last barrier until the next phase's first LDS
          // store, which every thread reaches only after the decision's
          // second barrier (the flags, read after it, have their own array)
          float* notch_mag = s.re;
          double* notch_red = reinterpret_cast<double*>(s.im);
          static_assert(C::NotchMax <= LdsSize<C::M>::floats && 128 * 2 <= LdsSize<C::M>::floats,
                        "notch scratch within the exchange area");
          static_for<0, C::P>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const int k = tid + C::T * m;
            if (k >= a.notch_lo && k < a.notch_hi)
              notch_mag[k - a.notch_lo] = sqrtf(x[m].x * x[m].x + x[m].y * x[m].y);
          });
          __syncthreads();
          if (tid < 64) notch_decide(notch_mag, notch_flag, notch_red, a.notch_lo, a.notch_hi);
          __syncthreads();
          uint32_t zm = 0;
          static_for<0, C::P>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            if (notch_zero(notch_flag, a.notch_lo, a.notch_hi, tid + C::T * m)) zm |= 1u << m;
          });
          notch_mask[tid] = zm;
        }
        const uint32_t zmask = a.notch_hi > a.notch_lo ? notch_mask[tid] : 0u;
        static_for<0, 2>([&](auto cc) {
          constexpr int c = decltype(cc)::value;
          if (a.out_audio[c] != nullptr) {
#pragma unroll
            for (int m = 0; m < C::P; ++m) {
              int i = tidv + C::T * m - a.lowbin[c];
              if (i >= 0 && i <= C::NA / 2) aslice[c][i] = (zmask & (1u << m)) ? make_float2(0.0f, 0.0f) : x[m];
            }
          }
        });
      }
      // the next phase's filter: rfhpf's or RFVideo's after the input
      // transform, the first post filter's after the demod transform
      if (ph == kPhIn)
        load_half<C>(has_hpf ? a.hpf.h : (kEfm ? a.h_efm : a.h_rfv), v, tidv);
      else if (a.npost > 0)
        load_half<C>(a.post[0].h, v, tidv);
    } else if (ph == kPhEven) {
      stage_own<C>(v, s, tidv);
      load_half<C>(a.h_rfv_odd, v, tidv);  // the odd half's table
      const LdsPair own(s.re, s.im, pidx(tidv));
      static_for<0, C::P>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        bstore2(zev, own.template ld<POff<C::T * m>::value>(), vo, m * C::StrideB);
      });
    } else if (ph == kPhHpf) {
      stage_own<C>(v, s, tidv);
      load_half<C>(kEfm ? a.h_efm : a.h_rfv, v, tidv);
      c2r_store<C>(a.hpf, (int64_t)bo * a.vstride, tidv, s);
    } else if (kEfm && ph == kPhEfm) {
      stage_own<C>(v, s, tidv);
      load_half<C>(a.h_rfv, v, tidv);
      efm_store<C>(a.efm + (int64_t)bo * a.vstride, a.vcut, a.vlen, tidv, s);
    } else if (ph >= kPhPost) {
      stage_own<C>(v, s, tidv);
      if (ph + 1 < last) load_half<C>(a.post[ph + 1 - kPhPost].h, v, tidv);
      c2r_store<C>(a.post[ph - kPhPost], (int64_t)bo * a.vstride, tidv, s,
                ph - kPhPost == a.below_post ? a.below + (((int64_t)bo * a.vstride) >> 5) : nullptr, a.below_highf);
    }
  }

  // analogue audio stage 1, one wavefront per channel (core.py:854-875)
  __syncthreads();
  if (tid < 128) {
    const int c = tid >> 6;
    const TwGlobal tg = C::audio_tw(make_rsrc(a.tw_rows, TwLayout::total * 8u));
    if (a.out_audio[c] != nullptr)
      audio_wave<C>(a, c, aslice[c], s.re + c * LdsSize<C::NA>::floats, s.im + c * LdsSize<C::NA>::floats, bo, tg);
  }
}

#define LD_DEMOD_AT(a) (a), (a).first + (a).step * (int64_t)blockIdx.x, (int)blockIdx.x, (int)blockIdx.x
#define LD_DEMOD_KERNELS(C, sfx)                                                                             \
  __global__ __launch_bounds__(C::T, C::WavesPerSimd) void k_demod_block##sfx(DemodArgs a) {                                 \
    demod_block<C, false, false>(LD_DEMOD_AT(a));                                                            \
  }                                                                                                          \
  __global__ __launch_bounds__(C::T, C::WavesPerSimd) void k_demod_block##sfx##_packed(DemodArgs a) {                        \
    demod_block<C, true, false>(LD_DEMOD_AT(a));                                                             \
  }                                                                                                          \
  __global__ __launch_bounds__(C::T, C::WavesPerSimd) void k_demod_block##sfx##_efm(DemodArgs a) {                           \
    demod_block<C, false, true>(LD_DEMOD_AT(a));                                                             \
  }                                                                                                          \
  __global__ __launch_bounds__(C::T, C::WavesPerSimd) void k_demod_block##sfx##_packed_efm(DemodArgs a) {                    \
    demod_block<C, true, true>(LD_DEMOD_AT(a));                                                              \
  }                                                                                                          \
  __global__ __launch_bounds__(C::T, C::WavesPerSimd) void k_demod_block##sfx##_mfma(DemodArgs a) {                          \
    demod_block<C, false, false, 2>(LD_DEMOD_AT(a));                                                         \
  }                                                                                                          \
  __global__ __launch_bounds__(C::T, C::WavesPerSimd) void k_demod_block##sfx##_mfma3(DemodArgs a) {                         \
    demod_block<C, false, false, 3>(LD_DEMOD_AT(a));                                                         \
  }
LD_DEMOD_KERNELS(Cfg32, )
LD_DEMOD_KERNELS(Cfg16, 16)
LD_DEMOD_KERNELS(Cfg8, 8)
#undef LD_DEMOD_KERNELS
#undef LD_DEMOD_AT

// Blocks of many captures and both systems in one launch: launch block i
// demodulates the block the DEVICE table entry i describes with the
// parameters and filter tables of its system (entries with sys < 0 are
// padding and exit at once).
// An entry whose block would read outside the input or write outside its
// system's output slots is skipped as well (the table is device data the
// host cannot check before the launch).
struct DemodBatchArgs {
  DemodArgs sys[2];
  const ld_batch_block* blocks;
  int64_t input_len;
  int nslots[2];
};
static_assert(sizeof(DemodBatchArgs) <= 4096, "kernel arguments over 4 KiB");
template <class C, int kMfma>
__device__ __forceinline__ void demod_batch(const DemodBatchArgs& a) {
  const ld_batch_block d = a.blocks[blockIdx.x];
  if (d.sys < 0 || d.sys > 1 || d.first < 0 || d.first > a.input_len - C::N || d.slot < 0 ||
      d.slot >= a.nslots[d.sys])
    return;
  demod_block<C, false, false, kMfma>(a.sys[d.sys], d.first, d.slot, (int)blockIdx.x);
}
__global__ __launch_bounds__(Cfg32::T) void k_demod_batch(DemodBatchArgs a) { demod_batch<Cfg32, 0>(a); }
__global__ __launch_bounds__(Cfg32::T) void k_demod_batch_mfma(DemodBatchArgs a) { demod_batch<Cfg32, 2>(a); }
__global__ __launch_bounds__(Cfg32::T) void k_demod_batch_mfma3(DemodBatchArgs a) { demod_batch<Cfg32, 3>(a); }
__global__ __launch_bounds__(Cfg16::T, Cfg16::WavesPerSimd) void k_demod_batch16(DemodBatchArgs a) { demod_batch<Cfg16, 0>(a); }
__global__ __launch_bounds__(Cfg16::T, Cfg16::WavesPerSimd) void k_demod_batch16_mfma(DemodBatchArgs a) { demod_batch<Cfg16, 2>(a); }
__global__ __launch_bounds__(Cfg16::T, Cfg16::WavesPerSimd) void k_demod_batch16_mfma3(DemodBatchArgs a) { demod_batch<Cfg16, 3>(a); }
__global__ __launch_bounds__(Cfg8::T, Cfg8::WavesPerSimd) void k_demod_batch8(DemodBatchArgs a) { demod_batch<Cfg8, 0>(a); }
__global__ __launch_bounds__(Cfg8::T, Cfg8::WavesPerSimd) void k_demod_batch8_mfma(DemodBatchArgs a) { demod_batch<Cfg8, 2>(a); }
__global__ __launch_bounds__(Cfg8::T, Cfg8::WavesPerSimd) void k_demod_batch8_mfma3(DemodBatchArgs a) { demod_batch<Cfg8, 3>(a); }

// The demodulator's forward M-point transform alone (FP32 or split-BF16
// MFMA), one workgroup per transform: in[b M .. + M) -> out[b M .. + M).
template <int kMfma>
__device__ __forceinline__ void fft_only(const float2* in, float2* out, const DemodArgs& a) {
  __shared__ Lds s;
  __shared__ TwLds twl;
  __shared__ MfmaTw mtw;
  const int tid = threadIdx.x;
  {
    const rsrc_t src = make_rsrc(a.tw_lds, kTwLdsCount * 8u);
    float2* dst = reinterpret_cast<float2*>(&twl);
    for (int i = tid; i < kTwLdsCount; i += kT) dst[i] = bload2(src, (uint32_t)i * 8u, 0u);
    if (tid < (int)(sizeof(MfmaTw) / 8))
      reinterpret_cast<float2*>(&mtw)[tid] = bload2(make_rsrc(a.mfma_tw, sizeof(MfmaTw)), (uint32_t)tid * 8u, 0u);
  }
  __syncthreads();
  const TwSrcLds twf{&twl};
  const rsrc_t ri = make_rsrc(in + (int64_t)blockIdx.x * kM, kM * 8u);
  const rsrc_t ro = make_rsrc(out + (int64_t)blockIdx.x * kM, kM * 8u);
  float2 v[kP];
#pragma unroll
  for (int m = 0; m < kP; ++m) v[m] = bload2(ri, tid * 8u, m * kStrideB);
  if constexpr (kMfma)
    WgFft16kMfma::template run<kMfma>(v, s.re, s.im, twf, a.mfma_g, &mtw, tid);
  else
    FwdM::run(v, s.re, s.im, twf, tid);
#pragma unroll
  for (int m = 0; m < kP; ++m) bstore2(ro, v[m], tid * 8u, m * kStrideB);
}
__global__ __launch_bounds__(kT) void k_fft_fp32(const float2* in, float2* out, DemodArgs a) {
  fft_only<0>(in, out, a);
}
__global__ __launch_bounds__(kT) void k_fft_mfma(const float2* in, float2* out, DemodArgs a) {
  fft_only<2>(in, out, a);
}
__global__ __launch_bounds__(kT) void k_fft_mfma3(const float2* in, float2* out, DemodArgs a) {
  fft_only<3>(in, out, a);
}

// =================================================== 65536-sample blocks
// The demodulation of 65536-sample blocks (RFDecode(blocklen=65536)): the
// 32768-point complex transforms are one radix-2 step in registers around
// two 16384-point workgroup transforms (the 32768-sample kernel's FP32 or
// split-BF16 MFMA transform, LDS-resident), so that LDS holds 16384 complex
// points at a time.  A forward transform of data in natural order (thread
// t holds element t + T m, m < 32) decimates in frequency and ends in pair
// order (thread t holds bins 2j, 2j + 1 in v[m], v[m + 16], j = t + T m,
// m < 16); a forward transform of data in pair order decimates in time and
// ends in natural order.  Spectra go from pair order through filters back to
// natural order, so no transform needs a transpose.  The spectra (X, then D)
// and the even analytic samples wait in a per-block workspace: registers and
// LDS of one CU cannot hold a 32768-point spectrum next to the working data.
constexpr int kN64 = 65536;
constexpr int kM64 = kN64 / 2;
constexpr int kH64 = kM64 / 2;           // points per half transform (16384)
constexpr int kNA64 = kN64 / 32;         // audio slice bins (2048)
// workspace per block (float2 units): X / D spectrum (bins 0..M-1), even
// analytic samples, the two audio slices
constexpr int kWs64Spec = 0;
constexpr int kWs64Even = kM64 + 64;
constexpr int kWs64Slice = kWs64Even + kM64;
constexpr int kWs64Floats2 = kWs64Slice + 2 * (kNA64 / 2 + 1) + 62;   // multiple of 64 float2 (512 B)
static_assert(kWs64Floats2 % 64 == 0, "workspace blocks stay 512-byte aligned");

using InvA64 = WgFft<kNA64, +1, kSyncWave>;

// W_M^(tid + T m) from wt = W_M^tid and a compile-time constant
template <int M_>
__device__ __forceinline__ float2 wm64(float2 wt) {
  return cmul(wt, make_float2(Tw<kM64, kT * M_, -1>::re, Tw<kM64, kT * M_, -1>::im));
}

template <int kMfma>
__device__ __forceinline__ void half_fft(float2* hv, Lds& s, const TwSrcLds& twf, const DemodArgs& a,
                                         const MfmaTw* mtw, int tid) {
  float2 (&h)[16] = *reinterpret_cast<float2(*)[16]>(hv);
  if constexpr (kMfma)
    WgFft16kMfma::template run<kMfma>(h, s.re, s.im, twf, a.mfma_g, mtw, tid);
  else
    FwdM::run(h, s.re, s.im, twf, tid);
}

// natural order in, pair order out (radix-2 decimation in frequency)
template <int kMfma>
__device__ __forceinline__ void fft_dif64(float2 (&v)[32], Lds& s, const TwSrcLds& twf, const DemodArgs& a,
                                          const MfmaTw* mtw, float2 wt, int tid) {
  static_for<0, 16>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const float2 x0 = v[m], x1 = v[m + 16];
    v[m] = cadd(x0, x1);
    v[m + 16] = cmul(csub(x0, x1), wm64<m>(wt));
  });
  half_fft<kMfma>(v, s, twf, a, mtw, tid);
  half_fft<kMfma>(v + 16, s, twf, a, mtw, tid);
}

// pair order in, natural order out (radix-2 decimation in time)
template <int kMfma>
__device__ __forceinline__ void fft_dit64(float2 (&v)[32], Lds& s, const TwSrcLds& twf, const DemodArgs& a,
                                          const MfmaTw* mtw, float2 wt, int tid) {
  half_fft<kMfma>(v, s, twf, a, mtw, tid);
  half_fft<kMfma>(v + 16, s, twf, a, mtw, tid);
  static_for<0, 16>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const float2 e = v[m], t = cmul(v[m + 16], wm64<m>(wt));
    v[m] = cadd(e, t);
    v[m + 16] = csub(e, t);
  });
}

// W_N^k for the thread's bins k = 2j (odd = 0) or 2j + 1 (odd = 1), j = tid + T m
template <int M_, int ODD>
__device__ __forceinline__ float2 wbin64(float2 wt) {
  const float2 w = wm64<M_>(wt);
  if constexpr (ODD) return cmul(w, make_float2(Tw<kN64, 1, -1>::re, Tw<kN64, 1, -1>::im));
  return w;
}

// After a forward transform of y[n] = x[2n] + i x[2n+1] (pair order in v),
// the half spectrum X[k] of the real x, in pair order in v, and X[M] in xM.
// The mirror bins M - k come through LDS, one half at a time.
__device__ __forceinline__ void r2c_split64(float2 (&v)[32], float& xM, Lds& s, float2 wt, int tid) {
  const LdsPair own(s.re, s.im, pidx(tid));
#pragma unroll
  for (int hf = 0; hf < 2; ++hf) {
    __syncthreads();
    static_for<0, 16>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      own.template st<POff<kT * m>::value>(v[m + 16 * hf]);
    });
    if (hf == 0 && tid == 0) {  // Y[M] = Y[0] (periodic), the mirror of bin 0
      s.re[pidx(kH64)] = v[0].x;
      s.im[pidx(kH64)] = v[0].y;
    }
    __syncthreads();
    // mirror of bin 2j is A[H - j]; of bin 2j + 1, B[H - 1 - j]
    const LdsPair mir(s.re, s.im, pidx(kH64 - hf - tid));
    static_for<0, 16>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const float2 p = v[m + 16 * hf];
      const float2 q = mir.template ld<-POff<kT * m>::value>();
      const float2 E = make_float2(0.5f * (p.x + q.x), 0.5f * (p.y - q.y));
      const float2 O = make_float2(0.5f * (p.y + q.y), -0.5f * (p.x - q.x));
      if (hf == 0 && m == 0 && tid == 0) xM = p.x - p.y;
      v[m + 16 * hf] = cadd(E, cmul(hf ? wbin64<m, 1>(wt) : wbin64<m, 0>(wt), O));
      if constexpr ((m & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    });
  }
  __syncthreads();
}

// The input of a real inverse transform of Z = X H (half spectrum in pair
// order in v, Z[M] in zM), conjugated for the forward transform: conj(E + iO).
__device__ __forceinline__ void c2r_pre64(float2 (&v)[32], float2 zM, Lds& s, float2 wt, int tid) {
  const LdsPair own(s.re, s.im, pidx(tid));
#pragma unroll
  for (int hf = 0; hf < 2; ++hf) {
    __syncthreads();
    static_for<0, 16>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      own.template st<POff<kT * m>::value>(v[m + 16 * hf]);
    });
    if (hf == 0 && tid == 0) {  // Z[M], the mirror of bin 0
      s.re[pidx(kH64)] = zM.x;
      s.im[pidx(kH64)] = zM.y;
    }
    __syncthreads();
    const LdsPair mir(s.re, s.im, pidx(kH64 - hf - tid));
    static_for<0, 16>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const float2 A = v[m + 16 * hf];
      const float2 B = cconj(mir.template ld<-POff<kT * m>::value>());
      const float2 E = make_float2(0.5f * (A.x + B.x), 0.5f * (A.y + B.y));
      const float2 O = cmulc(make_float2(0.5f * (A.x - B.x), 0.5f * (A.y - B.y)),
                             hf ? wbin64<m, 1>(wt) : wbin64<m, 0>(wt));
      v[m + 16 * hf] = make_float2(E.x - O.y, -(E.y + O.x));
      if constexpr ((m & 7) == 7) __builtin_amdgcn_sched_barrier(0);
    });
  }
  // the transform's first step works in registers; its half transforms
  // synchronise before they write LDS
}

// Stores of a transform result y (natural order in v; y[2n] = v.x, y[2n+1]
// = -v.y) as out[i] = y[(i + cut + shift) mod N] + dc, i < len, with the
// in-pulse words of the output when asked (as c2r_store).
__device__ __forceinline__ void c2r_store64(const C2rOut& o, int64_t base, const float2 (&v)[32], int tid,
                                            uint32_t* words = nullptr, float highf = 0.0f) {
  const rsrc_t r = make_rsrc(o.out + base, (uint32_t)o.len * 4u);
  const int off = o.cut + o.shift;
  const float dc = o.dc;
  const bool pairs = ((off | o.len) & 1) == 0;
  const int q = tid & 3;
  static_for<0, 32>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    const int n = 2 * (tid + kT * m);
    const float y0 = v[m].x + dc, y1 = dc - v[m].y;
    const uint32_t i0 = (uint32_t)((n - off) & (kN64 - 1));
    if (pairs) {
      bstore2<kStream>(r, make_float2(y0, y1), i0 * 4u, 0u);
      if (words != nullptr) {
        const unsigned long long e0 = __ballot(y0 <= highf), e1 = __ballot(y1 <= highf);
        const uint32_t w = ((uint32_t)(e0 >> (16 * q)) & 0xffffu) | ((uint32_t)(e1 >> (16 * q)) << 16);
        const uint32_t j0 = (uint32_t)((2 * (kT * m + (tid & ~63)) + 32 * q - off) & (kN64 - 1));
        if ((tid & 63) < 4 && j0 < (uint32_t)o.len) {
          const uint32_t left = (uint32_t)o.len - j0;
          words[j0 >> 5] = left >= 32 ? w : (w & below_mask_of(left));
        }
      }
    } else {
      const uint32_t i1 = (uint32_t)((n + 1 - off) & (kN64 - 1));
      bstore1<kStream>(r, y0, i0 * 4u, 0u);
      bstore1<kStream>(r, y1, i1 * 4u, 0u);
    }
  });
}

// v = the workspace spectrum times a half spectrum table, bin by bin, in
// pair order (bins 2j, 2j + 1 in v[m], v[m + 16], j = tid + T m)
__device__ __forceinline__ void load_prod64(rsrc_t spec, const float2* table, float2 (&v)[32], int tid) {
  const rsrc_t H = make_rsrc(table, (kM64 + 1) * 8u);
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const uint32_t o = (uint32_t)(tid + kT * m) * 16u;
    v[m] = cmul(bload2(spec, o, 0u), bload2(H, o, 0u));
    v[m + 16] = cmul(bload2(spec, o + 8u, 0u), bload2(H, o + 8u, 0u));
  }
}

// Analogue audio stage 1 of a 65536-sample block, one wavefront per
// channel: the slice (in the workspace) times filt1, 2048-point inverse,
// instantaneous frequency (as audio_wave).
__device__ void audio_wave64(const DemodArgs& a, int c, const float2* slice, float* lre, float* lim, int b,
                             const TwGlobal& twr) {
  const int lane = threadIdx.x & 63;
  const rsrc_t F = make_rsrc(a.h_audio[c], kNA64 * 8u);
  const rsrc_t S = make_rsrc(slice, (kNA64 / 2 + 1) * 8u);
  float2 v[kNA64 / 64];
#pragma unroll
  for (int m = 0; m < kNA64 / 64; ++m) {
    const int k = lane + 64 * m;
    const float2 f = bload2(F, lane * 8u, m * 512u);
    float2 z = make_float2(0.0f, 0.0f);
    if (k < kNA64 / 2)
      z = cmul(bload2(S, (uint32_t)k * 8u, 0u), f);
    else if (k == kNA64 / 2)
      z = cmul(cconj(bload2(S, (uint32_t)k * 8u, 0u)), f);
    v[m] = z;
  }
  InvA64::run(v, lre, lim, twr);
#pragma unroll
  for (int m = 0; m < kNA64 / 64; ++m) {
    const int n = lane + 64 * m;
    lre[pidx(n)] = v[m].x;
    lim[pidx(n)] = v[m].y;
  }
  lds_barrier<kSyncWave>();
  const float rate = a.rate[c], lf = a.low_freq[c];
  const float hz_per_rad = rate / kTwoPi;
  const rsrc_t out = make_rsrc(a.out_audio[c] + (int64_t)b * a.astride, (uint32_t)a.alen * 4u);
#pragma unroll
  for (int m = 0; m < kNA64 / 64; ++m) {
    const int n = lane + 64 * m;
    float hz = 0.0f;
    if (n > 0) {
      const float2 p = cmulc(v[m], make_float2(lre[pidx(n - 1)], lim[pidx(n - 1)]));
      hz = inc_to_hz(disc_atan2(p.y, p.x), hz_per_rad, rate);
    }
    bstore1<kStream>(out, hz + lf, (uint32_t)(n - a.acut) * 4u, 0u);
  }
}

template <int kMfma>
__device__ __forceinline__ void demod_block64(const DemodArgs& a, int64_t s0, int bo, int bw) {
  __shared__ Lds s;
  __shared__ TwLds twl;
  __shared__ MfmaTw mtw;
  __shared__ float notch_mag[kNotchMax];
  __shared__ int notch_flag[kNotchMax + 4];
  __shared__ double notch_red[128];
  __shared__ float2 bin_m[2 + 4];  // bin M of every filter table: 0 rfhpf, 1 RFVideo, 2.. post filters
  __shared__ float spec_m;         // X[M], then D[M]
  const int tid = threadIdx.x;
  float2* ws = a.ws_even + (int64_t)bw * kWs64Floats2;
  const rsrc_t wspec = make_rsrc(ws + kWs64Spec, kM64 * 8u);
  const rsrc_t weven = make_rsrc(ws + kWs64Even, kM64 * 8u);
  float2* slices = ws + kWs64Slice;
  {
    const rsrc_t src = make_rsrc(a.tw_lds, kTwLdsCount * 8u);
    float2* dst = reinterpret_cast<float2*>(&twl);
    for (int i = tid; i < kTwLdsCount; i += kT) dst[i] = bload2(src, (uint32_t)i * 8u, 0u);
    if (tid < 2 + a.npost) {
      const float2* t = tid == 0 ? a.hpf.h : (tid == 1 ? a.h_rfv : a.post[tid - 2].h);
      bin_m[tid] = t ? bload2(make_rsrc(t, (kM64 + 1) * 8u), 0u, kM64 * 8u) : make_float2(0.0f, 0.0f);
    }
    if constexpr (kMfma) {
      const rsrc_t mt = make_rsrc(a.mfma_tw, sizeof(MfmaTw));
      if (tid >= 64 && tid < 64 + (int)(sizeof(MfmaTw) / 8))
        reinterpret_cast<float2*>(&mtw)[tid - 64] = bload2(mt, (uint32_t)(tid - 64) * 8u, 0u);
    }
  }
  __syncthreads();
  const TwSrcLds twf{&twl};
  const float2 wt = bload2(make_rsrc(a.tw_split, kT * 8u), tid * 8u, 0u);   // W_M^tid
  const bool has_hpf = a.hpf.out != nullptr;
  const bool notch = a.notch_hi > a.notch_lo;
  float2 v[32];
  float xM = 0.0f;
  uint32_t zmask = 0u;   // bit m: v[m]'s bin zeroed by the spur notch

  // ------------------------------------------------ X = rFFT(block)
  {
    const rsrc_t in = make_rsrc((const char*)a.in + s0 * a.in_bytes, (uint32_t)(kN64 * a.in_bytes));
    if (a.in_bytes == 4) {
#pragma unroll
      for (int m = 0; m < 32; ++m) v[m] = bload2(in, tid * 8u, m * (kT * 8u));
    } else if ((s0 & 1) == 0) {
#pragma unroll
      for (int m = 0; m < 32; ++m) {
        const uint32_t w = __builtin_bit_cast(uint32_t, bload1(in, tid * 4u, m * (kT * 4u)));
        v[m] = make_float2((float)(int16_t)(w & 0xffffu), (float)(int16_t)(w >> 16));
      }
    } else {
#pragma unroll
      for (int m = 0; m < 32; ++m) {
        const uint32_t o = tid * 4u + m * (kT * 4u);
        const int16_t lo = (int16_t)__builtin_amdgcn_raw_buffer_load_b16(in, (int)o, 0, 0);
        const int16_t hi = (int16_t)__builtin_amdgcn_raw_buffer_load_b16(in, (int)o + 2, 0, 0);
        v[m] = make_float2((float)lo, (float)hi);
      }
    }
  }
  fft_dif64<kMfma>(v, s, twf, a, &mtw, wt, tid);
  r2c_split64(v, xM, s, wt, tid);
  if (notch) {  // PAL spur notch decision (core.py:774-798)
    static_for<0, 32>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const int k = 2 * (tid + kT * (m & 15)) + (m >> 4);
      if (k >= a.notch_lo && k < a.notch_hi) notch_mag[k - a.notch_lo] = sqrtf(v[m].x * v[m].x + v[m].y * v[m].y);
    });
    __syncthreads();
    if (tid < 64) notch_decide(notch_mag, notch_flag, notch_red, a.notch_lo, a.notch_hi);
    __syncthreads();
    static_for<0, 32>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      if (notch_zero(notch_flag, a.notch_lo, a.notch_hi, 2 * (tid + kT * (m & 15)) + (m >> 4))) zmask |= 1u << m;
    });
  }
  // the spectrum to the workspace (bins 2j, 2j + 1 side by side), X[M], the
  // audio slices (notched bins zeroed)
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const uint32_t o = (uint32_t)(tid + kT * m) * 16u;
    bstore2(wspec, v[m], o, 0u);
    bstore2(wspec, v[m + 16], o + 8u, 0u);
  }
  if (tid == 0) spec_m = xM;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (a.out_audio[c] == nullptr) continue;
    const rsrc_t sl = make_rsrc(slices + c * (kNA64 / 2 + 1), (kNA64 / 2 + 1) * 8u);
    static_for<0, 32>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      const int i = 2 * (tid + kT * (m & 15)) + (m >> 4) - a.lowbin[c];
      if (i >= 0 && i <= kNA64 / 2) bstore2(sl, (zmask & (1u << m)) ? make_float2(0.0f, 0.0f) : v[m], (uint32_t)i * 8u, 0u);
    });
  }
  __syncthreads();
  xM = spec_m;

  // ------------------------------------------------ rfhpf (core.py:769-772)
  if (has_hpf) {
    load_prod64(wspec, a.hpf.h, v, tid);
    c2r_pre64(v, cscale(bin_m[0], xM), s, wt, tid);
    fft_dit64<kMfma>(v, s, twf, a, &mtw, wt, tid);
    c2r_store64(a.hpf, (int64_t)bo * a.vstride, v, tid);
  }

  // ------------------------------------------------ analytic signal (core.py:800-805)
  const float2 hM = cscale(bin_m[1], xM);
#pragma unroll
  for (int odd = 0; odd < 2; ++odd) {
    load_prod64(wspec, a.h_rfv, v, tid);
    static_for<0, 32>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      float2 z = v[m];
      if (zmask & (1u << m)) z = make_float2(0.0f, 0.0f);
      if (m == 0 && tid == 0) z = odd ? csub(z, hM) : cadd(z, hM);   // bin 0: Z[M] folded in
      else if (odd) z = cmulc(z, wbin64<(m & 15), (m >> 4)>(wt));    // odd samples: Z[k] W_N^-k
      v[m] = cconj(z);
    });
    fft_dit64<kMfma>(v, s, twf, a, &mtw, wt, tid);
    if (!odd) {
#pragma unroll
      for (int m = 0; m < 32; ++m) bstore2(weven, v[m], tid * 8u, m * (kT * 8u));
    }
  }

  // ------------------------------------------------ instantaneous frequency, clip (utils.py:898-926, core.py:809)
  {
    // z_odd[n - 1] for every n: the odd samples shifted by one through LDS,
    // half a block at a time (the sample before the second half's first is
    // carried over by the thread that holds it)
    const float hz_per_rad = a.fs / kTwoPi;
    const float2 last0 = v[15];     // z_odd[H - 1] on thread T - 1
#pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
      __syncthreads();
      const LdsPair sh(s.re, s.im, pidx(tid + 1));
      static_for<0, 16>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        sh.template st<POff<kT * m>::value>(v[m + 16 * hf]);
      });
      if (hf == 1 && tid == kT - 1) {
        s.re[pidx(0)] = last0.x;
        s.im[pidx(0)] = last0.y;
      }
      __syncthreads();
      const LdsPair own(s.re, s.im, pidx(tid));
      static_for<0, 16>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const int j = tid + kT * (m + 16 * hf);
        const float2 prev = own.template ld<POff<kT * m>::value>();
        const float2 e = bload2(weven, tid * 8u, (m + 16 * hf) * (kT * 8u)), od = v[m + 16 * hf];
        const float2 p = cmulc(e, prev);
        const float h0 = (j > 0) ? inc_to_hz(-disc_atan2(p.y, p.x), hz_per_rad, a.fs) : 0.0f;
        const float2 q = cmulc(od, e);
        const float h1 = inc_to_hz(-disc_atan2(q.y, q.x), hz_per_rad, a.fs);
        v[m + 16 * hf] = make_float2(h0, h1);
      });
    }
    if (a.demod_raw != nullptr) {
      const rsrc_t raw = make_rsrc(a.demod_raw + (int64_t)bo * a.vstride, (uint32_t)a.vlen * 4u);
      const bool pairs = ((a.vcut | a.vlen) & 1) == 0;
#pragma unroll
      for (int m = 0; m < 32; ++m) {
        const uint32_t i0 = (uint32_t)(2 * (tid + kT * m) - a.vcut);
        if (pairs) {
          bstore2<kStream>(raw, v[m], i0 * 4u, 0u);
        } else {
          bstore1<kStream>(raw, v[m].x, i0 * 4u, 0u);
          bstore1<kStream>(raw, v[m].y, i0 * 4u + 4u, 0u);
        }
      }
    }
#pragma unroll
    for (int m = 0; m < 32; ++m)
      v[m] = make_float2(fminf(fmaxf(v[m].x, a.clip_lo), a.clip_hi) - a.dc,
                         fminf(fmaxf(v[m].y, a.clip_lo), a.clip_hi) - a.dc);
  }
  fft_dif64<kMfma>(v, s, twf, a, &mtw, wt, tid);
  r2c_split64(v, xM, s, wt, tid);     // D and D[M]
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const uint32_t o = (uint32_t)(tid + kT * m) * 16u;
    bstore2(wspec, v[m], o, 0u);
    bstore2(wspec, v[m + 16], o + 8u, 0u);
  }
  if (tid == 0) spec_m = xM;
  __syncthreads();
  xM = spec_m;

  // ------------------------------------------------ video post filters (core.py:811-820)
  for (int p = 0; p < a.npost; ++p) {
    const C2rOut& o = a.post[p];
    load_prod64(wspec, o.h, v, tid);
    c2r_pre64(v, cscale(bin_m[2 + p], xM), s, wt, tid);
    fft_dit64<kMfma>(v, s, twf, a, &mtw, wt, tid);
    c2r_store64(o, (int64_t)bo * a.vstride, v, tid,
                p == a.below_post ? a.below + (((int64_t)bo * a.vstride) >> 5) : nullptr, a.below_highf);
  }

  // ------------------------------------------------ audio stage 1 (core.py:854-875)
  __syncthreads();
  if (tid < 128) {
    const int c = tid >> 6;
    const TwGlobal tg{make_rsrc(a.tw_rows, TwLayout::total * 8u), TwLayout::k2k_p2, TwLayout::k2k_p3, 32, 5, 1};
    if (a.out_audio[c] != nullptr)
      audio_wave64(a, c, slices + c * (kNA64 / 2 + 1), s.re + c * LdsSize<kNA64>::floats,
                   s.im + c * LdsSize<kNA64>::floats, bo, tg);
  }
}

__global__ __launch_bounds__(kT) void k_demod_block64(DemodArgs a) {
  demod_block64<0>(a, a.first + a.step * (int64_t)blockIdx.x, (int)blockIdx.x, (int)blockIdx.x);
}
__global__ __launch_bounds__(kT) void k_demod_block64_mfma(DemodArgs a) {
  demod_block64<2>(a, a.first + a.step * (int64_t)blockIdx.x, (int)blockIdx.x, (int)blockIdx.x);
}
__global__ __launch_bounds__(kT) void k_demod_block64_mfma3(DemodArgs a) {
  demod_block64<3>(a, a.first + a.step * (int64_t)blockIdx.x, (int)blockIdx.x, (int)blockIdx.x);
}
template <int kMfma>
__device__ __forceinline__ void demod_batch64(const DemodBatchArgs& a) {
  const ld_batch_block d = a.blocks[blockIdx.x];
  if (d.sys < 0 || d.sys > 1 || d.first < 0 || d.first > a.input_len - kN64 || d.slot < 0 ||
      d.slot >= a.nslots[d.sys])
    return;
  demod_block64<kMfma>(a.sys[d.sys], d.first, d.slot, (int)blockIdx.x);
}
__global__ __launch_bounds__(kT) void k_demod_batch64(DemodBatchArgs a) { demod_batch64<0>(a); }
__global__ __launch_bounds__(kT) void k_demod_batch64_mfma(DemodBatchArgs a) { demod_batch64<2>(a); }
__global__ __launch_bounds__(kT) void k_demod_batch64_mfma3(DemodBatchArgs a) { demod_batch64<3>(a); }

// ------------------------------------------------------- audio stage 2
// One blocklen segment of a field read's stage-1 audio: y = the filtered
// segment (index 0 = its first sample); output sample out_base + i takes
// y[i] for i in [r0, r1) and [r2, r3) (the samples no later segment writes).
struct Audio2Seg {
  int64_t in_off;    // stitched stage-1 index of the segment's first sample
  int64_t out_base;  // stage-2 index that y[0] lands on
  int32_t len;       // samples read (the rest of the block is zero)
  int32_t f32;       // round to float32 (the reference's long-read array)
  int32_t r0, r1, r2, r3;
};

struct Audio2Args {
  const float* in[2];           // stitched stage-1 audio, left / right
  const Audio2Seg* segs;
  const float2* h[2];           // stage-2 half spectra (bins 0..M), scale 1/M
  double center[2];
  double* out[2];
  const float2* tw_lds;
  const float2* tw_split;
};

constexpr float kClipLow = 500000.0f;  // findpeaks threshold, core.py:894

// runfilter_audio_phase2 (core.py:884-915) for one (segment, channel):
// raw = stage-1 audio - carrier (float32); the left channel's findpeaks
// clips in the segment (utils.py:1207-1217: t[i] > t[last] and t[i+1] > t[i]
// on the copy t with values under 500 kHz zeroed; clip l = i - 1) zero
// raw[l-8, l+8) in both channels; the zero-padded block is filtered over N
// points as the demodulator's c2r outputs are (one forward M-point
// transform, the real split, the filter product, the conjugated inverse)
// and the carrier is added back in float64.
__global__ __launch_bounds__(kT) void k_audio2_block(Audio2Args a) {
  __shared__ Lds s;
  __shared__ TwLds twl;
  const int tid = threadIdx.x;
  const int g = blockIdx.x >> 1, c = blockIdx.x & 1;
  const Audio2Seg sg = a.segs[g];
  const int len = sg.len;
  const uint32_t vo = tid * 8u;
  {
    const rsrc_t src = make_rsrc(a.tw_lds, kTwLdsCount * 8u);
    float2* dst = reinterpret_cast<float2*>(&twl);
    for (int i = tid; i < kTwLdsCount; i += kT) dst[i] = bload2(src, (uint32_t)i * 8u, 0u);
  }
  const rsrc_t lin = make_rsrc(a.in[0] + sg.in_off, (uint32_t)len * 4u);
  const float cl = (float)a.center[0];
  // clip candidates: flag[i] for i < len - 1 (bytes over the exchange area)
  uint8_t* flag = reinterpret_cast<uint8_t*>(&s);
  float tlast = bload1(lin, (uint32_t)(len - 1) * 4u, 0u) - cl;
  tlast = tlast < kClipLow ? 0.0f : tlast;
  int any = 0;
  for (int i = tid; i < kN; i += kT) {
    uint8_t fl = 0;
    if (i + 1 < len) {
      float t0 = bload1(lin, (uint32_t)i * 4u, 0u) - cl, t1 = bload1(lin, (uint32_t)(i + 1) * 4u, 0u) - cl;
      t0 = t0 < kClipLow ? 0.0f : t0;
      t1 = t1 < kClipLow ? 0.0f : t1;
      fl = (t0 > tlast) && (t1 > t0);
    }
    flag[i] = fl;
    any |= fl;
  }
  any = __syncthreads_or(any);
  // raw samples of this channel as pairs (2n, 2n + 1), n = tid + T m
  const rsrc_t in = make_rsrc(a.in[c] + sg.in_off, (uint32_t)len * 4u);
  const float cc = (float)a.center[c];
  float2 v[kP], x[kP];
#pragma unroll
  for (int m = 0; m < kP; ++m) {
    const int j = 2 * (tid + kT * m);
    float r0 = j < len ? bload1(in, (uint32_t)j * 4u, 0u) - cc : 0.0f;
    float r1 = j + 1 < len ? bload1(in, (uint32_t)(j + 1) * 4u, 0u) - cc : 0.0f;
    if (any) {
      // sample q is zeroed when a flagged i lies in [q - 6, q + 9]
      bool z0 = false, z1 = false;
      for (int i = j - 6; i <= j + 10; ++i) {
        if (i < 0 || i >= len - 1 || !flag[i]) continue;
        z0 |= i <= j + 9;
        z1 |= i >= j - 5;
      }
      if (z0) r0 = 0.0f;
      if (z1) r1 = 0.0f;
    }
    v[m] = make_float2(r0, r1);
  }
  __syncthreads();   // the flags share LDS with the transform
  const TwSrcLds twf{&twl};
  const float2 wt = bload2(make_rsrc(a.tw_split, kT * 8u), vo, 0u);
  float xM = 0.0f;
  FwdM::run(v, s.re, s.im, twf, tid);
  r2c_split<Cfg32>(v, x, xM, s, wt, tid);
  load_half<Cfg32>(a.h[c], v, tid);
  const float2 hM = bload2(make_rsrc(a.h[c], kHalfBytes), 0u, kM * 8u);
  c2r_pre<Cfg32>(x, xM, hM, v, s, wt, tid);
  FwdM::run(v, s.re, s.im, twf, tid);
  double* out = a.out[c] + sg.out_base;
  const double center = a.center[c];
  auto put = [&](int i, float y) {
    if ((i >= sg.r0 && i < sg.r1) || (i >= sg.r2 && i < sg.r3)) {
      const double o = (double)y + center;
      out[i] = sg.f32 ? (double)(float)o : o;
    }
  };
#pragma unroll
  for (int m = 0; m < kP; ++m) {
    const int j = 2 * (tid + kT * m);
    put(j, v[m].x);
    put(j + 1, -v[m].y);
  }
}

constexpr size_t kMaxRetired = 32;  // retired tables kept before a synchronising free

// Scratch of the no-allocation demod entry point: the per-block workspace
// (even analytic samples) of up to kScratchBytes, one copy per GPU that
// loads the library, held until the process ends.  Host code uses the
// address hipGetSymbolAddress returns on each device, never the array's name.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// The scratch address on `device` (current), looked up once per device.
static int scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) {
    set_error("device %d outside [0, %d)", device, kMaxDevices);
    return LD_EINVAL;
  }
  if (g_scratch_ptr[device] == nullptr) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return LD_OK;
}

int replace_table(ld_handle* h, void** slot, const void* host, size_t bytes) {
  LD_HIP(hipSetDevice(h->device));
  void* fresh = nullptr;
  LD_HIP(hipMalloc(&fresh, bytes));
  hipError_t e = hipMemcpy(fresh, host, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(fresh);
    return hip_fail(e, "table upload");
  }
  if (*slot != nullptr) h->retired.push_back(*slot);
  *slot = fresh;
  if (h->retired.size() > kMaxRetired) free_retired(h);
  return LD_OK;
}

void free_retired(ld_handle* h) {
  if (h->retired.empty()) return;
  (void)hipSetDevice(h->device);
  (void)hipDeviceSynchronize();
  for (void* p : h->retired) (void)hipFree(p);
  h->retired.clear();
}

}  // namespace ldgpu

using namespace ldgpu;

// ==================================================================== C ABI
extern "C" {

int ld_abi_version(void) { return LDGPU_ABI_VERSION; }

const char* ld_last_error(void) { return g_err.c_str(); }

int ld_create(int device, int blocklen, ld_handle** out) {
  if (!out) {
    set_error("ld_create: out is NULL");
    return LD_EINVAL;
  }
  *out = nullptr;
  if (blocklen != Cfg8::N && blocklen != Cfg16::N && blocklen != kN && blocklen != kN64) {
    set_error("ld_create: blocklen %d not supported (%d, %d, %d or %d)", blocklen, Cfg8::N, Cfg16::N, kN, kN64);
    return LD_EINVAL;
  }
  LD_HIP(hipSetDevice(device));
  void* scr = nullptr;
  if (int rc = scratch_address(device, &scr)) return rc;
  ld_handle* h = new ld_handle();
  h->device = device;
  h->scratch = scr;
  h->blocklen = blocklen;
  // [0, TwLayout::total): Stockham pass rows (TwLayout), then the TwLds
  // image (pass rows and factor tables), then the split twiddles
  const int M = 0;
  std::vector<float2> tw(TwLayout::total + kTwLdsCount + kT);
  float2 (*cexp)(double, double) = [](double num, double den) {
    double a = -2.0 * M_PI * num / den;
    return make_float2((float)cos(a), (float)sin(a));
  };
  auto rows = [&](int off, int R, int NS) {
    int lr = 0;
    while ((2 << lr) <= R) ++lr;
    for (int k = 0; k < NS; ++k)
      for (int i = 0; i < lr; ++i) tw[M + off + k * lr + i] = cexp((double)(1 << i) * k, (double)NS * R);
  };
  rows(TwLayout::k16k_p2, 32, 32);
  rows(TwLayout::k16k_p3, 16, 1024);
  rows(TwLayout::k1k_p2, 16, 16);
  rows(TwLayout::k1k_p3, 4, 256);
  rows(TwLayout::k2k_p2, 32, 32);
  rows(TwLayout::k2k_p3, 2, 1024);
  rows(TwLayout::k512_p2, 8, 8);
  rows(TwLayout::k512_p3, 8, 64);
  rows(TwLayout::k256_p2, 4, 4);
  rows(TwLayout::k256_p3, 4, 16);
  rows(TwLayout::k256_p4, 4, 64);
  // split[t] = exp(-2 pi i t / 32768), t < 1024: W_N^t of the 32768-sample
  // kernel's real split, W_M^t of the 65536-sample kernel's radix-2 step and
  // its split; at 16384 samples split[t] = exp(-2 pi i t / 16384), t < 512
  if (blocklen == Cfg16::N) {
    TwLds8k::fill(tw.data() + TwLayout::total, cexp);
    for (int t2 = 0; t2 < Cfg16::T; ++t2) tw[TwLayout::total + kTwLdsCount + t2] = cexp(t2, Cfg16::N);
  } else if (blocklen == Cfg8::N) {
    TwLds4k::fill(tw.data() + TwLayout::total, cexp);
    for (int t2 = 0; t2 < Cfg8::T; ++t2) tw[TwLayout::total + kTwLdsCount + t2] = cexp(t2, Cfg8::N);
  } else {
    TwLds::fill(tw.data() + TwLayout::total, cexp);
    for (int t2 = 0; t2 < kT; ++t2) tw[TwLayout::total + kTwLdsCount + t2] = cexp(t2, kN);
  }
  hipError_t e = hipMalloc(&h->twiddle, sizeof(float2) * tw.size());
  if (e == hipSuccess) e = hipMemcpy(h->twiddle, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice);
  // split-BF16 MFMA variants: the DFT-64 matrix [[F_re, -F_im]; [F_im, F_re]]
  // (F = exp(-2 pi i r s / 64)) as three bf16 terms (hi, mid, lo) in the lane layout
  // of v_mfma_f32_32x32x16_bf16, and the NS = 64 twiddle factor tables
  std::vector<uint32_t> gt((size_t)kMfmaTabEntries * 4);
  auto bf16 = [](float x) -> uint32_t {
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
  };
  auto bf2f = [](uint32_t b) -> double {
    const uint32_t u = b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
  };
  for (int rt = 0; rt < 4; ++rt)
    for (int ks = 0; ks < 8; ++ks)
      for (int term = 0; term < kMfmaTerms; ++term)
        for (int l = 0; l < 64; ++l) {
          uint32_t part[8];
          for (int el = 0; el < 8; ++el) {
            const int row = 32 * rt + (l & 31), col = 16 * ks + 8 * (l >> 5) + el;
            const int sidx = row & 63, r = col & 63;
            const double ang = 2.0 * M_PI * (double)((r * sidx) & 63) / 64.0;
            const double c = cos(ang), sn = sin(ang);
            double g = row < 64 ? (col < 64 ? c : sn) : (col < 64 ? -sn : c);
            // term k: the bf16 rounding of what terms 0..k-1 leave of g
            uint32_t t = bf16((float)g);
            for (int k = 1; k <= term; ++k) {
              g -= bf2f(t);
              t = bf16((float)g);
            }
            part[el] = t;
          }
          uint32_t* dst = gt.data() + ((size_t)((rt * 8 + ks) * kMfmaTerms + term) * 64 + l) * 4;
          for (int w = 0; w < 4; ++w) dst[w] = part[2 * w] | (part[2 * w + 1] << 16);
        }
  MfmaTw mt;
  for (int q = 0; q < 64; ++q) {
    mt.coarse[q] = cexp(q, 64);
    mt.fine[q] = cexp(q, 4096);
  }
  if (e == hipSuccess) e = hipMalloc(&h->mfma_g, gt.size() * 4);
  if (e == hipSuccess) e = hipMemcpy(h->mfma_g, gt.data(), gt.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&h->mfma_tw, sizeof(MfmaTw));
  if (e == hipSuccess) e = hipMemcpy(h->mfma_tw, &mt, sizeof(MfmaTw), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    int rc = hip_fail(e, "ld_create twiddle upload");
    ld_destroy(h);
    return rc;
  }
  *out = h;
  return LD_OK;
}

void ld_destroy(ld_handle* h) {
  if (!h) return;
  free_retired(h);
  hipSetDevice(h->device);
  hipFree(h->twiddle);
  hipFree(h->mfma_g);
  hipFree(h->mfma_tw);
  hipFree(h->sinc_lut);
  for (auto& f : h->filt) hipFree(f);
  for (auto& f : h->filt_rolled) hipFree(f);
  for (auto& f : h->audio_filt) hipFree(f);
  for (auto& f : h->audio2_filt) hipFree(f);
  delete h;
}

static int upload_complex(ld_handle* h, float2** dst, const double* reim, int n, double scale) {
  std::vector<float2> tmp(n);
  for (int i = 0; i < n; ++i) tmp[i] = make_float2((float)(reim[2 * i] * scale), (float)(reim[2 * i + 1] * scale));
  return replace_table(h, reinterpret_cast<void**>(dst), tmp.data(), sizeof(float2) * n);
}

// The c2r tables of the single-level kernels (8192 .. 32768 samples) and of
// audio stage 2 carry a factor 1/4: r2c_split leaves its spectrum doubled and
// c2r_pre its halves undivided (exact powers of two).  RFVideo, whose product
// only feeds phases, and the 65536-sample kernels' tables are unscaled.
static double c2r_table_scale(const ld_handle* h, int id) {
  return (id == LD_FILT_RFVIDEO || h->blocklen == kN64) ? 1.0 : 0.25;
}

int ld_set_filter(ld_handle* h, int id, const double* reim, int nbins, double scale, double dc_gain,
                  int shift) {
  if (!h || !reim || id < 0 || id >= LD_NFILT) {
    set_error("ld_set_filter: bad handle, table or id %d", id);
    return LD_EINVAL;
  }
  if (nbins != h->blocklen / 2 + 1) {
    set_error("ld_set_filter: nbins %d != blocklen/2+1 = %d", nbins, h->blocklen / 2 + 1);
    return LD_EINVAL;
  }
  if (shift < -h->blocklen || shift > h->blocklen) {
    set_error("ld_set_filter: shift %d out of range", shift);
    return LD_EINVAL;
  }
  scale *= c2r_table_scale(h, id);
  int rc = upload_complex(h, &h->filt[id], reim, nbins, scale);
  if (rc) return rc;
  // the one-sample roll, in double: bin k times exp(+2 pi i k / N)
  std::vector<double> rolled(2 * (size_t)nbins);
  for (int k = 0; k < nbins; ++k) {
    const double a = 2.0 * M_PI * (double)k / (double)h->blocklen, c = cos(a), sn = sin(a);
    rolled[2 * k] = reim[2 * k] * c - reim[2 * k + 1] * sn;
    rolled[2 * k + 1] = reim[2 * k] * sn + reim[2 * k + 1] * c;
  }
  rc = upload_complex(h, &h->filt_rolled[id], rolled.data(), nbins, scale);
  if (rc) return rc;
  h->filt_dc[id] = dc_gain;
  h->filt_shift[id] = shift;
  return LD_OK;
}

// A c2r output's table and roll: out[i] = y[(i + cut + shift) mod N].  Where
// cut + shift is odd, the rolled table (y advanced by one sample) with
// shift - 1 gives the same output at an even offset, so that the kernels
// store aligned 8-byte pairs.
static void c2r_table(const ld_handle* h, int id, int cut, int shift, const float2*& table, int& out_shift) {
  if (((cut + shift) & 1) != 0 && h->filt_rolled[id] != nullptr) {
    table = h->filt_rolled[id];
    out_shift = shift - 1;
  } else {
    table = h->filt[id];
    out_shift = shift;
  }
}

int ld_set_audio2(ld_handle* h, int ch, const double* reim, int nbins, double center_freq) {
  if (!h || !reim || ch < 0 || ch > 1) {
    set_error("ld_set_audio2: bad handle, table or channel %d", ch);
    return LD_EINVAL;
  }
  if (nbins != h->blocklen / 2 + 1) {
    set_error("ld_set_audio2: nbins %d != blocklen/2+1 = %d", nbins, h->blocklen / 2 + 1);
    return LD_EINVAL;
  }
  // k_audio2_block's r2c_split / c2r_pre pair: the factor 1/4 (see c2r_table_scale)
  int rc = upload_complex(h, &h->audio2_filt[ch], reim, nbins, 0.25 / (h->blocklen / 2));
  if (rc) return rc;
  h->audio2_center[ch] = center_freq;
  return LD_OK;
}

// The segments of one field read as audio_phase2 (core.py:917-950) cuts
// it: one zero-padded block up to blocklen samples; longer reads: blocks at
// 0 and every blocklen - 512 samples while below len - (blocklen - 512)
// (written from their sample 512 on), and the block at len - blocklen - 1
// written one sample late; each output sample belongs to the last block
// that writes it.
static void audio2_segments(const ld_audio_field& fd, std::vector<Audio2Seg>& out) {
  const int64_t n = fd.audio_len, N = kN, askip = 512, sjump = kN - 512;
  if (n <= N) {
    out.push_back(Audio2Seg{fd.audio_offset, fd.a2_offset, (int32_t)n, 0, 0, (int32_t)n, 0, 0});
    return;
  }
  struct W { int64_t in, base, len, w0, w1; };   // writes output [w0, w1) = y[w0 - base ..]
  std::vector<W> ws;
  ws.push_back({0, 0, N, 0, N});
  for (int64_t smp = sjump; smp < n - sjump; smp += sjump) {
    const int64_t m = std::min(N, n - smp);
    ws.push_back({smp, smp, m, smp + askip, smp + m});
  }
  ws.push_back({n - N - 1, n - N, N, n - (N - askip), n});
  // ownership: subtract the union of later writes
  std::vector<std::pair<int64_t, int64_t>> later;
  std::vector<Audio2Seg> segs;
  for (int j = (int)ws.size() - 1; j >= 0; --j) {
    std::vector<std::pair<int64_t, int64_t>> runs{{ws[j].w0, ws[j].w1}};
    for (auto& L : later) {
      std::vector<std::pair<int64_t, int64_t>> next;
      for (auto& r : runs) {
        if (L.second <= r.first || L.first >= r.second) {
          next.push_back(r);
          continue;
        }
        if (r.first < L.first) next.push_back({r.first, L.first});
        if (L.second < r.second) next.push_back({L.second, r.second});
      }
      runs.swap(next);
    }
    later.push_back({ws[j].w0, ws[j].w1});
    if (runs.empty()) continue;
    Audio2Seg g{fd.audio_offset + ws[j].in, fd.a2_offset + ws[j].base, (int32_t)ws[j].len, 1, 0, 0, 0, 0};
    g.r0 = (int32_t)(runs[0].first - ws[j].base);
    g.r1 = (int32_t)(runs[0].second - ws[j].base);
    if (runs.size() > 1) {
      g.r2 = (int32_t)(runs[1].first - ws[j].base);
      g.r3 = (int32_t)(runs[1].second - ws[j].base);
    }
    segs.push_back(g);
  }
  out.insert(out.end(), segs.rbegin(), segs.rend());
}

size_t ld_audio2_workspace_bytes(const ld_handle*, const ld_audio_field* fields, int nfields) {
  if (!fields || nfields <= 0) return 0;
  size_t nseg = 0;
  for (int f = 0; f < nfields; ++f) {
    const int64_t n = fields[f].audio_len;
    nseg += n <= kN ? 1 : (size_t)((n + kN - 513) / (kN - 512) + 2);
  }
  return (sizeof(Audio2Seg) * nseg + 255) & ~(size_t)255;
}

int ld_audio2_fields(ld_handle* h, const float* left, const float* right, int64_t total,
                     const ld_audio_field* fields, int nfields, double* out_left, double* out_right,
                     int64_t out_total, void* workspace, size_t ws_bytes, void* stream) {
  if (!h || !left || !right || !fields || nfields <= 0 || !out_left || !out_right || !workspace) {
    set_error("ld_audio2_fields: bad arguments");
    return LD_EINVAL;
  }
  if (h->blocklen != kN) {
    set_error("ld_audio2_fields: audio stage 2 runs on %d-sample handles", kN);
    return LD_EINVAL;
  }
  if (!h->audio2_filt[0] || !h->audio2_filt[1]) {
    set_error("ld_audio2_fields: stage-2 filters not set (ld_set_audio2)");
    return LD_ESTATE;
  }
  std::vector<Audio2Seg> segs;
  for (int f = 0; f < nfields; ++f) {
    const ld_audio_field& fd = fields[f];
    if (fd.audio_len < 2 || fd.audio_len > (int64_t)1 << 30) {
      set_error("ld_audio2_fields: field %d: %lld audio samples", f, (long long)fd.audio_len);
      return LD_EINVAL;
    }
    if (fd.audio_offset < 0 || fd.audio_offset + fd.audio_len > total || fd.a2_offset < 0 ||
        fd.a2_offset + fd.audio_len > out_total) {
      set_error("ld_audio2_fields: field %d outside the arrays", f);
      return LD_ERANGE;
    }
    audio2_segments(fd, segs);
  }
  if (ws_bytes < sizeof(Audio2Seg) * segs.size()) {
    set_error("ld_audio2_fields: workspace %zu bytes < %zu needed", ws_bytes, sizeof(Audio2Seg) * segs.size());
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  LD_HIP(hipMemcpyAsync(workspace, segs.data(), sizeof(Audio2Seg) * segs.size(), hipMemcpyHostToDevice, st));
  Audio2Args a{};
  a.in[0] = left;
  a.in[1] = right;
  a.segs = (const Audio2Seg*)workspace;
  a.h[0] = h->audio2_filt[0];
  a.h[1] = h->audio2_filt[1];
  a.center[0] = h->audio2_center[0];
  a.center[1] = h->audio2_center[1];
  a.out[0] = out_left;
  a.out[1] = out_right;
  a.tw_lds = h->twiddle + TwLayout::total;
  a.tw_split = h->twiddle + TwLayout::total + kTwLdsCount;
  hipLaunchKernelGGL(k_audio2_block, dim3(2 * (unsigned)segs.size()), dim3(kT), 0, st, a);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

int ld_set_audio(ld_handle* h, int ch, const double* reim, int nbins, int lowbin, double a1_freq,
                 double low_freq) {
  if (!h || !reim || ch < 0 || ch > 1) {
    set_error("ld_set_audio: bad handle, table or channel %d", ch);
    return LD_EINVAL;
  }
  if (nbins != h->blocklen / 32) {
    set_error("ld_set_audio: nbins %d not supported (only %d at blocklen %d)", nbins, h->blocklen / 32, h->blocklen);
    return LD_EINVAL;
  }
  if (lowbin < 1 || lowbin + nbins / 2 >= h->blocklen / 2) {
    set_error("ld_set_audio: lowbin %d outside the half spectrum", lowbin);
    return LD_EINVAL;
  }
  int rc = upload_complex(h, &h->audio_filt[ch], reim, nbins, 1.0 / nbins);
  if (rc) return rc;
  h->audio_nbins = nbins;
  h->audio_lowbin[ch] = lowbin;
  h->audio_rate[ch] = a1_freq;
  h->audio_low_freq[ch] = low_freq;
  return LD_OK;
}

int ld_set_demod_params(ld_handle* h, double fs, double clip_lo, double clip_hi, double dc_offset,
                        int video_rot) {
  if (!h || !(fs > 0) || !(clip_hi > clip_lo)) {
    set_error("ld_set_demod_params: bad arguments");
    return LD_EINVAL;
  }
  if (video_rot <= -h->cut_end || video_rot >= h->cut_begin) {
    set_error("ld_set_demod_params: video_rot %d outside (-%d, %d)", video_rot, h->cut_end, h->cut_begin);
    return LD_EINVAL;
  }
  h->fs = fs;
  h->clip_lo = clip_lo;
  h->clip_hi = clip_hi;
  h->dc_offset = dc_offset;
  h->video_rot = video_rot;
  h->params_set = true;
  return LD_OK;
}

int ld_set_notch(ld_handle* h, int lo, int hi) {
  if (!h) {
    set_error("ld_set_notch: null handle");
    return LD_EINVAL;
  }
  const int wmax = h->blocklen == Cfg8::N ? Cfg8::NotchMax : kNotchMax;
  if (!(lo == 0 && hi == 0) && !(lo >= 1 && lo < hi && hi < h->blocklen / 2 && hi - lo <= wmax)) {
    set_error("ld_set_notch: window [%d, %d) outside [1, %d) or wider than %d bins", lo, hi, h->blocklen / 2, wmax);
    return LD_EINVAL;
  }
  h->notch_lo = lo;
  h->notch_hi = hi;
  return LD_OK;
}

size_t ld_demod_workspace_bytes(const ld_handle* h, int nblocks) {
  if (!h || nblocks <= 0) return 0;
  // even analytic samples of every block (M float2)
  if (h->blocklen == kN64) return (size_t)nblocks * (size_t)kWs64Floats2 * sizeof(float2);
  return (size_t)nblocks * (size_t)(h->blocklen / 2) * sizeof(float2);
}

// The launch arguments of one handle's demodulation (shared by the plain
// and the batched entry points); `who` names the caller in errors.
// nblocks: the number of output slots the strides are checked for.
static int build_demod_args(ld_handle* h, const void* input, int dtype, int64_t input_len, int nblocks, int cut,
                            const ld_demod_outputs* out, void* workspace, DemodArgs& a, const char* who) {
  if (dtype != LD_IN_INT16 && dtype != LD_IN_FLOAT32 && dtype != LD_IN_PACKED10) {
    set_error("%s: unknown input dtype %d", who, dtype);
    return LD_EINVAL;
  }
  if (!h->params_set || !h->filt[LD_FILT_RFVIDEO] || !h->filt[LD_FILT_RFHPF] || !h->filt[LD_FILT_VIDEO] ||
      !h->filt[LD_FILT_VIDEO05] || !h->filt[LD_FILT_BURST]) {
    set_error("%s: demod parameters or video filters not uploaded", who);
    return LD_ESTATE;
  }
  if (out->demod_pilot && !h->filt[LD_FILT_PILOT]) {
    set_error("%s: demod_pilot requested but no pilot filter uploaded", who);
    return LD_ESTATE;
  }
  if ((out->audio_left && !h->audio_filt[0]) || (out->audio_right && !h->audio_filt[1])) {
    set_error("%s: audio requested but audio filters not uploaded", who);
    return LD_ESTATE;
  }
  const int N = h->blocklen;
  const int S = N - h->cut_begin - h->cut_end;
  const int vcut = cut ? h->cut_begin : 0;
  const int vlen = cut ? S : N;
  const int NA = h->audio_nbins ? h->audio_nbins : h->blocklen / 32;
  const int fdiv = N / NA;
  const int acut = cut ? h->cut_begin / fdiv : 0;
  const int alen = cut ? NA - acut - (h->cut_end + fdiv - 1) / fdiv : NA;
  if (out->video_stride < 0 || out->audio_stride < 0 ||
      (nblocks > 1 && ((out->video_stride < vlen) || ((out->audio_left || out->audio_right) && out->audio_stride < alen)))) {
    set_error("%s: output strides shorter than one block's output", who);
    return LD_EINVAL;
  }
  a = DemodArgs{};
  a.in = input;
  a.tw_rows = h->twiddle;
  a.tw_lds = h->twiddle + TwLayout::total;
  a.tw_split = h->twiddle + TwLayout::total + kTwLdsCount;
  a.mfma_g = (const u32x4_t*)h->mfma_g;
  a.mfma_tw = (const MfmaTw*)h->mfma_tw;
  a.h_rfv = h->filt[LD_FILT_RFVIDEO];
  a.h_rfv_odd = h->filt_rolled[LD_FILT_RFVIDEO];
  c2r_table(h, LD_FILT_RFHPF, h->cut_begin, -h->video_rot, a.hpf.h, a.hpf.shift);
  a.hpf.out = out->rfhpf;
  a.hpf.dc = 0.0f;
  a.hpf.cut = h->cut_begin;
  a.hpf.len = S;
  const int ids[4] = {LD_FILT_VIDEO, LD_FILT_VIDEO05, LD_FILT_BURST, LD_FILT_PILOT};
  float* outs[4] = {out->demod, out->demod_05, out->demod_burst, out->demod_pilot};
  a.below_post = -1;
  if (out->demod_05_below) {
    if (!out->demod_05 || (nblocks > 1 && out->video_stride % 32 != 0)) {
      set_error("%s: demod_05_below needs demod_05 and a video stride that is a multiple of 32", who);
      return LD_EINVAL;
    }
    if ((vcut + h->filt_shift[LD_FILT_VIDEO05]) % 32 != 0 || vlen % 2 != 0) {
      set_error("%s: demod_05_below needs demod_05's output offset to be a multiple of 32", who);
      return LD_EINVAL;
    }
    a.below = out->demod_05_below;
    // y <= highf exactly when (double)y <= below_high
    float f = (float)out->below_high;
    if ((double)f > out->below_high) f = nextafterf(f, -INFINITY);
    a.below_highf = f;
  }
  for (int i = 0; i < 4; ++i) {
    if (!outs[i]) continue;
    if (i == 1 && out->demod_05_below) a.below_post = a.npost;
    C2rOut& o = a.post[a.npost++];
    c2r_table(h, ids[i], vcut, h->filt_shift[ids[i]], o.h, o.shift);
    o.out = outs[i];
    o.dc = (float)(h->dc_offset * h->filt_dc[ids[i]]);
    o.cut = vcut;
    o.len = vlen;
  }
  a.demod_raw = out->demod_raw;
  a.ws_even = (float2*)workspace;
  a.vcut = vcut;
  a.vlen = vlen;
  a.vstride = out->video_stride;
  float* aouts[2] = {out->audio_left, out->audio_right};
  for (int c = 0; c < 2; ++c) {
    a.h_audio[c] = h->audio_filt[c];
    a.lowbin[c] = h->audio_lowbin[c];
    a.rate[c] = (float)h->audio_rate[c];
    a.low_freq[c] = (float)h->audio_low_freq[c];
    a.out_audio[c] = aouts[c];
  }
  a.acut = acut;
  a.alen = alen;
  a.astride = out->audio_stride;
  a.fs = (float)h->fs;
  a.clip_lo = (float)h->clip_lo;
  a.clip_hi = (float)h->clip_hi;
  a.dc = (float)h->dc_offset;
  a.notch_lo = h->notch_lo;
  a.notch_hi = h->notch_hi;
  a.in_bytes = dtype == LD_IN_INT16 ? 2 : (dtype == LD_IN_FLOAT32 ? 4 : 0);
  a.in_total = 0;
  if (dtype == LD_IN_PACKED10) {
    const int64_t bytes = (input_len + 3) / 4 * 5;
    if (bytes > 0xfffffff0LL) {
      set_error("%s: packed input over 4 GiB (pass a shorter span)", who);
      return LD_EINVAL;
    }
    // dword loads cover the pairs' bytes: every dword holding a sample byte
    // lies below the byte count rounded up to 4 (the buffer's allocation
    // covers that round-up)
    a.in_total = (uint32_t)((bytes + 3) / 4 * 4);
  }
  a.efm = out->efm;
  a.h_efm = h->filt[LD_FILT_EFM];
  if (a.efm && !a.h_efm) {
    set_error("%s: EFM output requested without the EFM filter (LD_FILT_EFM)", who);
    return LD_ESTATE;
  }
  return LD_OK;
}

int ld_demod_blocks(ld_handle* h, const void* input, int dtype, int64_t input_len, int64_t first,
                    int64_t step, int nblocks, int cut, const ld_demod_outputs* out, void* workspace,
                    size_t ws_bytes, void* stream) {
  if (!h || !input || !out || nblocks <= 0 || !workspace) {
    set_error("ld_demod_blocks: null argument or nblocks <= 0");
    return LD_EINVAL;
  }
  const int N = h->blocklen;
  if (first < 0 || step < 0 || first + (int64_t)(nblocks - 1) * step + N > input_len) {
    set_error("ld_demod_blocks: blocks [%lld, +%d*%lld + %d) exceed input length %lld", (long long)first,
              nblocks - 1, (long long)step, N, (long long)input_len);
    return LD_ERANGE;
  }
  if (ws_bytes < ld_demod_workspace_bytes(h, nblocks)) {
    set_error("ld_demod_blocks: workspace %zu bytes < %zu needed", ws_bytes, ld_demod_workspace_bytes(h, nblocks));
    return LD_EINVAL;
  }
  DemodArgs a;
  if (int rc = build_demod_args(h, input, dtype, input_len, nblocks, cut, out, workspace, a, "ld_demod_blocks"))
    return rc;
  a.first = first;
  a.step = step;
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  if (h->blocklen == kN64) {
    if (dtype == LD_IN_PACKED10 || a.efm) {
      set_error("ld_demod_blocks: 65536-sample blocks take int16 / float32 input without EFM output");
      return LD_EINVAL;
    }
    hipLaunchKernelGGL(h->fft_variant == LD_FFT_MFMA_BF16X3   ? k_demod_block64_mfma
                       : h->fft_variant == LD_FFT_MFMA_BF16X6 ? k_demod_block64_mfma3
                                                              : k_demod_block64,
                       dim3(nblocks), dim3(kT), 0, st, a);
    LD_HIP(hipGetLastError());
    return LD_OK;
  }
  if (h->fft_variant != LD_FFT_FP32 && (dtype == LD_IN_PACKED10 || a.efm)) {
    set_error("ld_demod_blocks: the split-BF16 MFMA variant takes int16 / float32 input without EFM output");
    return LD_EINVAL;
  }
  // kernels per block length: {plain, packed, efm, packed_efm, mfma, mfma3}
  using K = void (*)(DemodArgs);
  static const K k32[6] = {k_demod_block, k_demod_block_packed, k_demod_block_efm, k_demod_block_packed_efm,
                           k_demod_block_mfma, k_demod_block_mfma3};
  static const K k16[6] = {k_demod_block16, k_demod_block16_packed, k_demod_block16_efm,
                           k_demod_block16_packed_efm, k_demod_block16_mfma, k_demod_block16_mfma3};
  static const K k8[6] = {k_demod_block8, k_demod_block8_packed, k_demod_block8_efm, k_demod_block8_packed_efm,
                          k_demod_block8_mfma, k_demod_block8_mfma3};
  const K* ks = h->blocklen == Cfg16::N ? k16 : (h->blocklen == Cfg8::N ? k8 : k32);
  const int threads = h->blocklen == Cfg16::N ? Cfg16::T : (h->blocklen == Cfg8::N ? Cfg8::T : kT);
  int which = (dtype == LD_IN_PACKED10 ? 1 : 0) + (a.efm ? 2 : 0);
  if (h->fft_variant == LD_FFT_MFMA_BF16X3) which = 4;
  if (h->fft_variant == LD_FFT_MFMA_BF16X6) which = 5;
  hipLaunchKernelGGL(ks[which], dim3(nblocks), dim3(threads), 0, st, a);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

int ld_demod_batch(ld_handle* const* sys, const void* input, int dtype, int64_t input_len,
                   const ld_batch_block* blocks, int nblocks, const int32_t* nslots, int cut,
                   const ld_demod_outputs* outs, void* workspace, size_t ws_bytes, void* stream) {
  if (!sys || !input || !blocks || nblocks <= 0 || !nslots || !outs || !workspace) {
    set_error("ld_demod_batch: null argument or nblocks <= 0");
    return LD_EINVAL;
  }
  if (dtype != LD_IN_INT16 && dtype != LD_IN_FLOAT32) {
    set_error("ld_demod_batch: input dtype %d (int16 or float32 only)", dtype);
    return LD_EINVAL;
  }
  ld_handle* h0 = sys[0] ? sys[0] : sys[1];
  if (!h0) {
    set_error("ld_demod_batch: no system handle");
    return LD_EINVAL;
  }
  DemodBatchArgs b = {};
  for (int k = 0; k < 2; ++k) {
    ld_handle* h = sys[k];
    b.nslots[k] = 0;
    if (!h) continue;
    if (h->blocklen != h0->blocklen || h->device != h0->device) {
      set_error("ld_demod_batch: system %d handle: blocklen %d or device %d differs", k, h->blocklen, h->device);
      return LD_EINVAL;
    }
    if (outs[k].efm) {
      set_error("ld_demod_batch: no EFM output in the batched launch");
      return LD_EINVAL;
    }
    if (nslots[k] < 0) {
      set_error("ld_demod_batch: system %d: %d output slots", k, nslots[k]);
      return LD_EINVAL;
    }
    const int nb = nslots[k] > 1 ? nslots[k] : 2;  // strides are checked as for several blocks
    if (int rc = build_demod_args(h, input, dtype, input_len, nb, cut, &outs[k], workspace, b.sys[k],
                                  "ld_demod_batch"))
      return rc;
    b.nslots[k] = nslots[k];
  }
  if (ws_bytes < ld_demod_workspace_bytes(h0, nblocks)) {
    set_error("ld_demod_batch: workspace %zu bytes < %zu needed", ws_bytes, ld_demod_workspace_bytes(h0, nblocks));
    return LD_EINVAL;
  }
  b.blocks = blocks;
  b.input_len = input_len;
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h0->device));
  for (int k = 0; k < 2; ++k)
    if (sys[k] && sys[k]->fft_variant != h0->fft_variant) {
      set_error("ld_demod_batch: the system handles select different transform variants");
      return LD_EINVAL;
    }
  const int v = h0->fft_variant;
  void (*kern)(DemodBatchArgs);
  int threads = kT;
  if (h0->blocklen == kN64) {
    kern = v == LD_FFT_MFMA_BF16X3 ? k_demod_batch64_mfma : v == LD_FFT_MFMA_BF16X6 ? k_demod_batch64_mfma3
                                                                                     : k_demod_batch64;
  } else if (h0->blocklen == Cfg16::N) {
    kern = v == LD_FFT_MFMA_BF16X3 ? k_demod_batch16_mfma : v == LD_FFT_MFMA_BF16X6 ? k_demod_batch16_mfma3
                                                                                     : k_demod_batch16;
    threads = Cfg16::T;
  } else if (h0->blocklen == Cfg8::N) {
    kern = v == LD_FFT_MFMA_BF16X3 ? k_demod_batch8_mfma : v == LD_FFT_MFMA_BF16X6 ? k_demod_batch8_mfma3
                                                                                   : k_demod_batch8;
    threads = Cfg8::T;
  } else {
    kern = v == LD_FFT_MFMA_BF16X3 ? k_demod_batch_mfma : v == LD_FFT_MFMA_BF16X6 ? k_demod_batch_mfma3
                                                                                   : k_demod_batch;
  }
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(threads), 0, st, b);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

int ld_set_fft_variant(ld_handle* h, int variant) {
  if (!h || (variant != LD_FFT_FP32 && variant != LD_FFT_MFMA_BF16X3 && variant != LD_FFT_MFMA_BF16X6)) {
    set_error("ld_set_fft_variant: bad handle or variant %d", variant);
    return LD_EINVAL;
  }
  h->fft_variant = variant;
  return LD_OK;
}

int ld_fft_forward(ld_handle* h, const void* in, void* out, int ntransforms, int variant, void* stream) {
  if (!h || !in || !out || ntransforms <= 0 ||
      (variant != LD_FFT_FP32 && variant != LD_FFT_MFMA_BF16X3 && variant != LD_FFT_MFMA_BF16X6)) {
    set_error("ld_fft_forward: bad arguments");
    return LD_EINVAL;
  }
  if (h->blocklen == Cfg16::N || h->blocklen == Cfg8::N) {
    set_error("ld_fft_forward: the 16384-point transform needs a 32768- or 65536-sample handle");
    return LD_EINVAL;
  }
  DemodArgs a = {};
  a.tw_lds = h->twiddle + TwLayout::total;
  a.mfma_g = (const u32x4_t*)h->mfma_g;
  a.mfma_tw = (const MfmaTw*)h->mfma_tw;
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  hipLaunchKernelGGL(variant == LD_FFT_MFMA_BF16X3   ? k_fft_mfma
                     : variant == LD_FFT_MFMA_BF16X6 ? k_fft_mfma3
                                                     : k_fft_fp32,
                     dim3(ntransforms), dim3(kT), 0, st,
                     (const float2*)in, (float2*)out, a);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

size_t ld_demod_noalloc_max_bytes(void) { return kScratchBytes; }

int ld_demod_blocks_noalloc(ld_handle* h, const void* input, int dtype, int64_t input_len, int64_t first,
                            int64_t step, int nblocks, int cut, const ld_demod_outputs* out, void* stream) {
  if (!h || nblocks <= 0) {
    set_error("ld_demod_blocks_noalloc: null handle or nblocks <= 0");
    return LD_EINVAL;
  }
  if (h->scratch == nullptr) {
    set_error("ld_demod_blocks_noalloc: the handle has no scratch address");
    return LD_ESTATE;
  }
  const size_t need = ld_demod_workspace_bytes(h, nblocks);
  if (need > kScratchBytes) {
    set_error("ld_demod_blocks_noalloc: %d blocks need %zu workspace bytes, over the %zu-byte scratch "
              "(at most %zu blocks of %d)", nblocks, need, kScratchBytes,
              kScratchBytes / ld_demod_workspace_bytes(h, 1), h->blocklen);
    return LD_EINVAL;
  }
  return ld_demod_blocks(h, input, dtype, input_len, first, step, nblocks, cut, out, h->scratch, kScratchBytes,
                         stream);
}

}  // extern "C"
