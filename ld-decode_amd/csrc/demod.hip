// RF demodulation kernels (FP32) and the handle / demod part of the C ABI.
//
// Per block of N = 32768 real samples (M = N/2 = 16384 complex points), the
// reference's RFDecode.demodblock_cpu (core.py:755-879) is computed as:
//
//   k_fwd        int16/f32 block -> X = rFFT(x), bins 0..M          (core.py:761)
//   k_c2r        X * Frfhpf -> irFFT -> cut/roll -> rfhpf            (core.py:769-772)
//   k_analytic   X * RFVideo(*MTF^l)/N -> analytic z, even and odd   (core.py:800-805)
//                samples as two M-point inverse transforms
//   k_demod      z -> instantaneous frequency (conjugate product,    (utils.py:898-926,
//                atan2) -> demod_raw; clip, remove DC, rFFT -> D      core.py:806-809)
//   k_c2r        D * {FVideo, FVideo05, FVideoBurst, FVideoPilot}    (core.py:811-845)
//                -> irFFT -> + dc*H(0) -> roll -> cut
//   k_audio      sliced X bins * filt1 -> 1024-pt iFFT -> inst. freq (core.py:854-875)
//
// Real transforms of length N use one M-point complex transform plus the
// standard split (even/odd packing) step.  Spectra live in a per-block
// workspace in HBM between kernels.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
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
constexpr int kT = 512;        // threads per workgroup for the M-point FFT
constexpr int kP = kM / kT;    // points per thread (32)
constexpr int kLdx = kM + 64;  // workspace row pitch for half spectra (float2)
constexpr int kNA = 1024;      // audio slice bins at N = 32768
constexpr float kTwoPi = 6.283185307179586f;

using FwdM = WgFft<kM, -1, kN>;
using InvM = WgFft<kM, +1, kN>;
using InvA = WgFft<kNA, +1, kN>;

struct Lds {
  float re[LdsSize<kM>::floats];
  float im[LdsSize<kM>::floats];
};

// ---------------------------------------------------------------- helpers
// Split step after an M-point FFT of y[m] = x[2m] + i x[2m+1]: thread's
// v[m] = Y[k], k = tid + T*m, becomes X[k] = E + W_N^k O.  Writes X[0..M].
__device__ __forceinline__ void r2c_split_store(float2 (&v)[kP], Lds& s, const float2* __restrict__ tw,
                                                float2* __restrict__ X) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int m = 0; m < kP; ++m) {
    int k = tid + kT * m;
    s.re[pidx(k)] = v[m].x;
    s.im[pidx(k)] = v[m].y;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < kP; ++m) {
    int k = tid + kT * m;
    int km = (kM - k) & (kM - 1);
    float2 a = v[m];
    float2 b = make_float2(s.re[pidx(km)], s.im[pidx(km)]);
    float2 E = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
    float2 O = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
    float2 w = tw[k];
    X[k] = cadd(E, cmul(w, O));
    if (k == 0) X[kM] = make_float2(a.x - a.y, 0.0f);
  }
  __syncthreads();
}

// Pre-step of an inverse real transform: from Z = X*H (bins 0..M, Hermitian
// extension implied) build Y[k] = E + i O in the thread's natural layout.
__device__ __forceinline__ void c2r_pre(float2 (&v)[kP], const float2* __restrict__ X,
                                        const float2* __restrict__ H, const float2* __restrict__ tw) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int m = 0; m < kP; ++m) {
    int k = tid + kT * m;
    int km = kM - k;
    float2 A = cmul(X[k], H[k]);
    float2 B = cconj(cmul(X[km], H[km]));
    float2 E = make_float2(0.5f * (A.x + B.x), 0.5f * (A.y + B.y));
    float2 Dd = make_float2(0.5f * (A.x - B.x), 0.5f * (A.y - B.y));
    float2 O = cmulc(Dd, tw[k]);  // * W_N^{-k}
    v[m] = make_float2(E.x - O.y, E.y + O.x);
  }
}

// Hz of the phase increment d (radians, from atan2 in (-pi, pi]) wrapped to
// [0, 2pi): utils.py:921-926
__device__ __forceinline__ float inc_to_hz(float d, float hz_per_rad, float fs) {
  float hz = d * hz_per_rad;
  return d < 0.0f ? hz + fs : hz;
}

// ---------------------------------------------------------------- kernels
template <int DT>
__global__ __launch_bounds__(kT) void k_fwd(const void* __restrict__ in, int64_t first, int64_t step,
                                           float2* __restrict__ ws_x, const float2* __restrict__ tw) {
  __shared__ Lds s;
  const int tid = threadIdx.x;
  const int64_t s0 = first + step * blockIdx.x;
  float2 v[kP];
#pragma unroll
  for (int m = 0; m < kP; ++m) {
    int64_t i = s0 + 2 * (int64_t)(tid + kT * m);
    if constexpr (DT == LD_IN_INT16) {
      const int16_t* p = static_cast<const int16_t*>(in);
      v[m] = make_float2((float)p[i], (float)p[i + 1]);
    } else {
      const float* p = static_cast<const float*>(in);
      v[m] = make_float2(p[i], p[i + 1]);
    }
  }
  FwdM::run(v, s.re, s.im, tw);
  r2c_split_store(v, s, tw, ws_x + (int64_t)blockIdx.x * kLdx);
}

// Inverse real transform of (spectrum row * H) with output roll/cut:
// out[i] = y[(i + cut_begin + shift) mod N] + dc, for i < outlen.
struct C2rArgs {
  const float2* H[4];
  float dc[4];
  int shift[4];
  float* out[4];
};

__global__ __launch_bounds__(kT) void k_c2r(const float2* __restrict__ ws_spec, C2rArgs a, int cut_begin,
                                           int outlen, int64_t out_stride, const float2* __restrict__ tw) {
  __shared__ Lds s;
  const int tid = threadIdx.x;
  const int b = blockIdx.x, f = blockIdx.y;
  float* out = a.out[f];
  if (out == nullptr) return;
  const float2* X = ws_spec + (int64_t)b * kLdx;
  float2 v[kP];
  c2r_pre(v, X, a.H[f], tw);
  InvM::run(v, s.re, s.im, tw);
  const float dc = a.dc[f];
  const int off = cut_begin + a.shift[f];
  out += (int64_t)b * out_stride;
#pragma unroll
  for (int m = 0; m < kP; ++m) {
    int n = 2 * (tid + kT * m);
    int i0 = (n - off) & (kN - 1);
    int i1 = (n + 1 - off) & (kN - 1);
    if (i0 < outlen) out[i0] = v[m].x + dc;
    if (i1 < outlen) out[i1] = v[m].y + dc;
  }
}

// Analytic signal: z[n] = (1/N) sum_{k=0..M} Z[k] e^{+2 pi i k n / N} with
// Z = X * H (H carries the 0/1/2 analytic mask and 1/N).  Even samples:
// M-point inverse transform of Z[k] (+Z[M] folded into bin 0); odd samples:
// of Z[k] W_N^{-k} (-Z[M] folded into bin 0).  blockIdx.y selects the half.
__global__ __launch_bounds__(kT) void k_analytic(const float2* __restrict__ ws_x, const float2* __restrict__ H,
                                                float2* __restrict__ ws_z, const float2* __restrict__ tw) {
  __shared__ Lds s;
  const int tid = threadIdx.x;
  const int b = blockIdx.x, half = blockIdx.y;
  const float2* X = ws_x + (int64_t)b * kLdx;
  float2 v[kP];
#pragma unroll
  for (int m = 0; m < kP; ++m) {
    int k = tid + kT * m;
    float2 z = cmul(X[k], H[k]);
    if (k == 0) {
      float2 zm = cmul(X[kM], H[kM]);
      z = half ? csub(z, zm) : cadd(z, zm);
    } else if (half) {
      z = cmulc(z, tw[k]);
    }
    v[m] = z;
  }
  InvM::run(v, s.re, s.im, tw);
  float2* Z = ws_z + ((int64_t)b * 2 + half) * kM;
#pragma unroll
  for (int m = 0; m < kP; ++m) Z[tid + kT * m] = v[m];
}

// Instantaneous frequency + clip + DC removal + forward rFFT of the result.
__global__ __launch_bounds__(kT) void k_demod(const float2* __restrict__ ws_z, float2* __restrict__ ws_d,
                                             float* __restrict__ demod_raw, int cut_begin, int outlen,
                                             int64_t out_stride, float fs, float clip_lo, float clip_hi,
                                             float dc, const float2* __restrict__ tw) {
  __shared__ Lds s;
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  const float2* ze = ws_z + (int64_t)b * 2 * kM;
  const float2* zo = ze + kM;
  const float hz_per_rad = fs / kTwoPi;
  float* raw = demod_raw ? demod_raw + (int64_t)b * out_stride : nullptr;
  float2 v[kP];
#pragma unroll
  for (int m = 0; m < kP; ++m) {
    int j = tid + kT * m;
    float2 e = ze[j], o = zo[j];
    float h0 = 0.0f;
    if (j > 0) {
      float2 p = cmulc(e, zo[j - 1]);
      h0 = inc_to_hz(atan2f(p.y, p.x), hz_per_rad, fs);
    }
    float2 q = cmulc(o, e);
    float h1 = inc_to_hz(atan2f(q.y, q.x), hz_per_rad, fs);
    if (raw) {
      int i0 = 2 * j - cut_begin, i1 = i0 + 1;
      if (i0 >= 0 && i0 < outlen) raw[i0] = h0;
      if (i1 >= 0 && i1 < outlen) raw[i1] = h1;
    }
    v[m] = make_float2(fminf(fmaxf(h0, clip_lo), clip_hi) - dc, fminf(fmaxf(h1, clip_lo), clip_hi) - dc);
  }
  FwdM::run(v, s.re, s.im, tw);
  r2c_split_store(v, s, tw, ws_d + (int64_t)b * kLdx);
}

// Analogue audio stage 1: one wavefront per (block, channel).
struct AudioArgs {
  const float2* F[2];
  int lowbin[2];
  float rate[2];
  float low_freq[2];
  float* out[2];
};

__global__ __launch_bounds__(64) void k_audio(const float2* __restrict__ ws_x, AudioArgs a, int cut_begin,
                                             int outlen, int64_t out_stride, const float2* __restrict__ tw) {
  __shared__ float sre[LdsSize<kNA>::floats], sim[LdsSize<kNA>::floats];
  const int tid = threadIdx.x;
  const int b = blockIdx.x, ch = blockIdx.y;
  float* out = a.out[ch];
  if (out == nullptr) return;
  const float2* X = ws_x + (int64_t)b * kLdx;
  const float2* F = a.F[ch];
  const int lb = a.lowbin[ch];
  float2 v[kNA / 64];
#pragma unroll
  for (int m = 0; m < kNA / 64; ++m) {
    int k = tid + 64 * m;
    float2 z = make_float2(0.0f, 0.0f);
    if (k < kNA / 2)
      z = cmul(X[lb + k], F[k]);
    else if (k == kNA / 2)  // bin N - lowbin - nbins/2 of the full spectrum
      z = cmul(cconj(X[lb + kNA / 2]), F[k]);
    v[m] = z;
  }
  InvA::run(v, sre, sim, tw);
#pragma unroll
  for (int m = 0; m < kNA / 64; ++m) {
    int n = tid + 64 * m;
    sre[pidx(n)] = v[m].x;
    sim[pidx(n)] = v[m].y;
  }
  __syncthreads();
  const float rate = a.rate[ch], lf = a.low_freq[ch];
  const float hz_per_rad = rate / kTwoPi;
  out += (int64_t)b * out_stride;
#pragma unroll
  for (int m = 0; m < kNA / 64; ++m) {
    int n = tid + 64 * m;
    float hz = 0.0f;
    if (n > 0) {
      int pn = n - 1;
      float2 p = cmulc(v[m], make_float2(sre[pidx(pn)], sim[pidx(pn)]));
      hz = inc_to_hz(atan2f(p.y, p.x), hz_per_rad, rate);
    }
    int i = n - cut_begin;
    if (i >= 0 && i < outlen) out[i] = hz + lf;
  }
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
  if (blocklen != kN) {
    set_error("ld_create: blocklen %d not supported (only %d)", blocklen, kN);
    return LD_EINVAL;
  }
  LD_HIP(hipSetDevice(device));
  ld_handle* h = new ld_handle();
  h->device = device;
  h->blocklen = blocklen;
  std::vector<float2> tw(blocklen);
  for (int t = 0; t < blocklen; ++t) {
    double a = -2.0 * M_PI * (double)t / (double)blocklen;
    tw[t] = make_float2((float)cos(a), (float)sin(a));
  }
  hipError_t e = hipMalloc(&h->twiddle, sizeof(float2) * blocklen);
  if (e == hipSuccess) e = hipMemcpy(h->twiddle, tw.data(), sizeof(float2) * blocklen, hipMemcpyHostToDevice);
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
  hipSetDevice(h->device);
  hipFree(h->twiddle);
  hipFree(h->sinc_lut);
  for (auto& f : h->filt) hipFree(f);
  for (auto& f : h->audio_filt) hipFree(f);
  delete h;
}

static int upload_complex(float2** dst, const double* reim, int n, double scale, int device) {
  std::vector<float2> tmp(n);
  for (int i = 0; i < n; ++i) tmp[i] = make_float2((float)(reim[2 * i] * scale), (float)(reim[2 * i + 1] * scale));
  LD_HIP(hipSetDevice(device));
  if (*dst == nullptr) LD_HIP(hipMalloc(dst, sizeof(float2) * n));
  LD_HIP(hipMemcpy(*dst, tmp.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return LD_OK;
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
  int rc = upload_complex(&h->filt[id], reim, nbins, scale, h->device);
  if (rc) return rc;
  h->filt_dc[id] = dc_gain;
  h->filt_shift[id] = shift;
  return LD_OK;
}

int ld_set_audio(ld_handle* h, int ch, const double* reim, int nbins, int lowbin, double a1_freq,
                 double low_freq) {
  if (!h || !reim || ch < 0 || ch > 1) {
    set_error("ld_set_audio: bad handle, table or channel %d", ch);
    return LD_EINVAL;
  }
  if (nbins != kNA) {
    set_error("ld_set_audio: nbins %d not supported (only %d at blocklen %d)", nbins, kNA, kN);
    return LD_EINVAL;
  }
  if (lowbin < 1 || lowbin + nbins / 2 > h->blocklen / 2) {
    set_error("ld_set_audio: lowbin %d outside the half spectrum", lowbin);
    return LD_EINVAL;
  }
  int rc = upload_complex(&h->audio_filt[ch], reim, nbins, 1.0 / nbins, h->device);
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

size_t ld_demod_workspace_bytes(const ld_handle* h, int nblocks) {
  if (!h || nblocks <= 0) return 0;
  // X and D half spectra (kLdx each) + analytic signal (2 * kM)
  return (size_t)nblocks * sizeof(float2) * (2 * (size_t)kLdx + 2 * (size_t)kM);
}

int ld_demod_blocks(ld_handle* h, const void* input, int dtype, int64_t input_len, int64_t first,
                    int64_t step, int nblocks, int cut, const ld_demod_outputs* out, void* workspace,
                    size_t ws_bytes, void* stream) {
  if (!h || !input || !out || !workspace || nblocks <= 0) {
    set_error("ld_demod_blocks: null argument or nblocks <= 0");
    return LD_EINVAL;
  }
  if (dtype != LD_IN_INT16 && dtype != LD_IN_FLOAT32) {
    set_error("ld_demod_blocks: unknown input dtype %d", dtype);
    return LD_EINVAL;
  }
  if (!h->params_set || !h->filt[LD_FILT_RFVIDEO] || !h->filt[LD_FILT_RFHPF] || !h->filt[LD_FILT_VIDEO] ||
      !h->filt[LD_FILT_VIDEO05] || !h->filt[LD_FILT_BURST]) {
    set_error("ld_demod_blocks: demod parameters or video filters not uploaded");
    return LD_ESTATE;
  }
  if (out->demod_pilot && !h->filt[LD_FILT_PILOT]) {
    set_error("ld_demod_blocks: demod_pilot requested but no pilot filter uploaded");
    return LD_ESTATE;
  }
  if ((out->audio_left && !h->audio_filt[0]) || (out->audio_right && !h->audio_filt[1])) {
    set_error("ld_demod_blocks: audio requested but audio filters not uploaded");
    return LD_ESTATE;
  }
  const int N = h->blocklen;
  if (first < 0 || step < 0 || first + (int64_t)(nblocks - 1) * step + N > input_len) {
    set_error("ld_demod_blocks: blocks [%lld, +%d*%lld + %d) exceed input length %lld", (long long)first,
              nblocks - 1, (long long)step, N, (long long)input_len);
    return LD_ERANGE;
  }
  size_t need = ld_demod_workspace_bytes(h, nblocks);
  if (ws_bytes < need) {
    set_error("ld_demod_blocks: workspace %zu bytes < %zu needed", ws_bytes, need);
    return LD_EINVAL;
  }
  const int S = N - h->cut_begin - h->cut_end;
  const int vcut = cut ? h->cut_begin : 0;
  const int vlen = cut ? S : N;
  const int NA = h->audio_nbins ? h->audio_nbins : kNA;
  const int fdiv = N / NA;
  const int acut = cut ? h->cut_begin / fdiv : 0;
  const int alen = cut ? NA - acut - (h->cut_end + fdiv - 1) / fdiv : NA;
  if (out->video_stride < 0 || out->audio_stride < 0 || (nblocks > 1 && out->video_stride < vlen &&
      out->video_stride != 0)) {
    set_error("ld_demod_blocks: bad output strides");
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  float2* ws_x = (float2*)workspace;
  float2* ws_d = ws_x + (size_t)nblocks * kLdx;
  float2* ws_z = ws_d + (size_t)nblocks * kLdx;
  const float2* tw = h->twiddle;
  if (dtype == LD_IN_INT16)
    hipLaunchKernelGGL(k_fwd<LD_IN_INT16>, dim3(nblocks), dim3(kT), 0, st, input, first, step, ws_x, tw);
  else
    hipLaunchKernelGGL(k_fwd<LD_IN_FLOAT32>, dim3(nblocks), dim3(kT), 0, st, input, first, step, ws_x, tw);
  LD_HIP(hipGetLastError());
  if (out->rfhpf) {
    C2rArgs a = {};
    a.H[0] = h->filt[LD_FILT_RFHPF];
    a.dc[0] = 0.0f;
    a.shift[0] = -h->video_rot;
    a.out[0] = out->rfhpf;
    hipLaunchKernelGGL(k_c2r, dim3(nblocks, 1), dim3(kT), 0, st, ws_x, a, h->cut_begin, S, out->video_stride, tw);
    LD_HIP(hipGetLastError());
  }
  if (out->audio_left || out->audio_right) {
    AudioArgs a = {};
    for (int c = 0; c < 2; ++c) {
      a.F[c] = h->audio_filt[c];
      a.lowbin[c] = h->audio_lowbin[c];
      a.rate[c] = (float)h->audio_rate[c];
      a.low_freq[c] = (float)h->audio_low_freq[c];
    }
    a.out[0] = out->audio_left;
    a.out[1] = out->audio_right;
    hipLaunchKernelGGL(k_audio, dim3(nblocks, 2), dim3(64), 0, st, ws_x, a, acut, alen, out->audio_stride, tw);
    LD_HIP(hipGetLastError());
  }
  const bool need_video = out->demod || out->demod_raw || out->demod_05 || out->demod_burst || out->demod_pilot;
  if (!need_video) return LD_OK;
  hipLaunchKernelGGL(k_analytic, dim3(nblocks, 2), dim3(kT), 0, st, ws_x, (const float2*)h->filt[LD_FILT_RFVIDEO],
                     ws_z, tw);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_demod, dim3(nblocks), dim3(kT), 0, st, ws_z, ws_d, out->demod_raw, vcut, vlen,
                     out->video_stride, (float)h->fs, (float)h->clip_lo, (float)h->clip_hi, (float)h->dc_offset, tw);
  LD_HIP(hipGetLastError());
  C2rArgs a = {};
  const int ids[4] = {LD_FILT_VIDEO, LD_FILT_VIDEO05, LD_FILT_BURST, LD_FILT_PILOT};
  float* outs[4] = {out->demod, out->demod_05, out->demod_burst, out->demod_pilot};
  int nf = 0;
  for (int i = 0; i < 4; ++i) {
    if (!outs[i]) continue;
    a.H[nf] = h->filt[ids[i]];
    a.dc[nf] = (float)(h->dc_offset * h->filt_dc[ids[i]]);
    a.shift[nf] = h->filt_shift[ids[i]];
    a.out[nf] = outs[i];
    ++nf;
  }
  if (nf) {
    hipLaunchKernelGGL(k_c2r, dim3(nblocks, nf), dim3(kT), 0, st, ws_d, a, vcut, vlen, out->video_stride, tw);
    LD_HIP(hipGetLastError());
  }
  return LD_OK;
}

}  // extern "C"
