// Analogue audio output rate (downscale_audio, core.py:1423-1578), batched
// over fields: one workgroup per field computes the tick positions and the
// wow factors of _downscale_audio_compute_locs_and_swow (the 1.5% wow clamp
// is a chain over the ticks: one thread runs it, and only when some step of
// the computed wow exceeds 1.5%), then
// every tick's output pair as _downscale_audio_to_output does: the in-order
// float64 mean of the stage-2 audio under it, times the wow, minus the
// carrier, rescaled to int16 with round-half-even and clipped, negated.
#include <math.h>

#include "internal.hpp"

#pragma clang fp contract(off)

namespace ldgpu {

constexpr int kDsThreads = 1024;
constexpr int kDsMaxTicks = 4096;   // ticks per field (nout + 1)
constexpr int kDsBatch = 8;         // stage-2 samples loaded together per mean

struct DsArgs {
  const double* a2[2];
  int64_t a2_total;
  const ld_audio_field* fields;
  const double* ll;
  ld_audio_out_params p;
  int16_t* out;
  int32_t* failed;
};

// dsa_rescale_and_clip (utils.py:1328-1332): int(np.round(x * 32767 / 371081))
// clipped to +-32766 (rint rounds half to even, as np.round)
__device__ __forceinline__ int dsa_rescale(double x) {
  const double v = rint(x * 32767.0 / 371081.0);
  return v < -32766.0 ? -32766 : (v > 32766.0 ? 32766 : (int)v);
}

// In-order float64 sum of x[s..e) (numba's np.mean before the division),
// kDsBatch loads in flight at a time.
__device__ __forceinline__ double ordered_sum(const double* __restrict__ x, long long s, long long e) {
  double acc = 0.0;
  for (long long k = s; k < e; k += kDsBatch) {
    double v[kDsBatch];
#pragma unroll
    for (int q = 0; q < kDsBatch; ++q) v[q] = k + q < e ? x[k + q] : 0.0;
#pragma unroll
    for (int q = 0; q < kDsBatch; ++q)
      if (k + q < e) acc += v[q];
  }
  return acc;
}

__global__ __launch_bounds__(kDsThreads) void k_audio_downscale(DsArgs a) {
  __shared__ double locs[kDsMaxTicks];
  __shared__ double swow[kDsMaxTicks];
  __shared__ int bad;
  const int f = blockIdx.x;
  const ld_audio_field fd = a.fields[f];
  const double* li = a.ll + fd.lineloc_offset;
  const int nl = fd.nlines;
  const int nt = fd.nout + 1;
  const double step = 1.0 / a.p.freq;
  if (threadIdx.x == 0) bad = 0;
  // ticks: numba's arange, start + i * step
  for (int i = threadIdx.x; i < nt; i += kDsThreads) {
    const double t = fd.timeoffset + (double)i * step;
    const double linenum = ((t * 1000000.0) / a.p.line_period) + 1.0;
    double cur, nxt;
    if (linenum < 0.0) {
      const double v = li[0] + (a.p.linelen * linenum);
      cur = (double)(long long)v;          // int(): toward zero
      nxt = cur + a.p.linelen;
    } else if ((double)nl > linenum + 2.0) {
      const int il = (int)linenum;
      cur = li[il];
      nxt = li[il + 1];
    } else {
      cur = li[nl - 2];
      nxt = cur + a.p.linelen;
    }
    double loc = cur;
    loc += (nxt - cur) * (linenum - floor(linenum));
    double w = (nxt - cur) / a.p.linelen;
    w = (w - 1.0) + 1.0;
    swow[i] = w;
    locs[i] = loc / a.p.scale;
  }
  __syncthreads();
  // the wow clamp: swow[i] = swow[i-1] when they differ by more than 1.5%.
  // Without a jump that large between the computed values it never fires
  // (each previous value is then its own computed value): the chain runs
  // only when some jump exceeds 1.5%.
  int jump = 0;
  for (int i = 1 + threadIdx.x; i < nt; i += kDsThreads) jump |= fabs(swow[i] - swow[i - 1]) > 0.015;
  if (__syncthreads_or(jump) && threadIdx.x == 0) {
    double prev = swow[0];
    for (int i0 = 1; i0 < nt; i0 += 16) {
      double r[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) r[k] = i0 + k < nt ? swow[i0 + k] : 0.0;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        if (i0 + k < nt) {
          if (fabs(r[k] - prev) > 0.015) r[k] = prev;
          prev = r[k];
        }
      }
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if (i0 + k < nt) swow[i0 + k] = r[k];
    }
  }
  __syncthreads();
  const double* L = a.a2[0] + fd.a2_offset;
  const double* R = a.a2[1] + fd.a2_offset;
  const long long len = fd.audio_len;
  int16_t* out = a.out + 2 * fd.out_offset;
  int mybad = 0;
  for (int i = threadIdx.x; i < fd.nout; i += kDsThreads) {
    const long long s = (long long)locs[i], e = (long long)locs[i + 1];
    int16_t ol = 0, orr = 0;
    if (e > s && e < len && s >= 0) {
      const double n = (double)(e - s);
      const double vl = ((ordered_sum(L, s, e) / n) * swow[i]) - a.p.lfreq;
      const double vr = ((ordered_sum(R, s, e) / n) * swow[i]) - a.p.rfreq;
      ol = (int16_t)(-dsa_rescale(vl));
      orr = (int16_t)(-dsa_rescale(vr));
    } else {
      mybad = 1;
    }
    out[2 * i] = ol;
    out[2 * i + 1] = orr;
  }
  if (__syncthreads_or(mybad) && threadIdx.x == 0) bad = 1;
  if (threadIdx.x == 0) a.failed[f] = bad;
}

}  // namespace ldgpu

using namespace ldgpu;

extern "C" int ld_audio_downscale(ld_handle* h, const double* a2_left, const double* a2_right, int64_t a2_total,
                                  const ld_audio_field* fields, int nfields, const double* linelocs,
                                  int64_t nlinelocs, const ld_audio_out_params* prm, int16_t* out,
                                  int64_t out_pairs, int32_t* failed, void* workspace, size_t ws_bytes,
                                  void* stream) {
  if (!h || !a2_left || !a2_right || !fields || nfields <= 0 || !linelocs || !prm || !out || !failed ||
      !workspace) {
    set_error("ld_audio_downscale: bad arguments");
    return LD_EINVAL;
  }
  if (ws_bytes < sizeof(ld_audio_field) * (size_t)nfields) {
    set_error("ld_audio_downscale: workspace too small");
    return LD_EINVAL;
  }
  if (!(prm->freq > 0.0) || !(prm->line_period > 0.0) || !(prm->linelen > 0.0) || !(prm->scale > 0.0)) {
    set_error("ld_audio_downscale: rate, line period, line length and scale must be positive");
    return LD_EINVAL;
  }
  for (int f = 0; f < nfields; ++f) {
    const ld_audio_field& fd = fields[f];
    if (fd.nout < 0 || fd.nout + 1 > kDsMaxTicks) {
      set_error("ld_audio_downscale: field %d: %d output samples (at most %d)", f, fd.nout, kDsMaxTicks - 1);
      return LD_EINVAL;
    }
    if (fd.nlines < 2 || fd.lineloc_offset < 0 || fd.lineloc_offset + fd.nlines > nlinelocs) {
      set_error("ld_audio_downscale: field %d: line range", f);
      return LD_EINVAL;
    }
    if (fd.a2_offset < 0 || fd.audio_len < 0 || fd.a2_offset + fd.audio_len > a2_total || fd.out_offset < 0 ||
        fd.out_offset + fd.nout > out_pairs) {
      set_error("ld_audio_downscale: field %d outside the arrays", f);
      return LD_ERANGE;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  LD_HIP(hipMemcpyAsync(workspace, fields, sizeof(ld_audio_field) * nfields, hipMemcpyHostToDevice, st));
  DsArgs a{};
  a.a2[0] = a2_left;
  a.a2[1] = a2_right;
  a.a2_total = a2_total;
  a.fields = (const ld_audio_field*)workspace;
  a.ll = linelocs;
  a.p = *prm;
  a.out = out;
  a.failed = failed;
  hipLaunchKernelGGL(k_audio_downscale, dim3(nfields), dim3(kDsThreads), 0, st, a);
  LD_HIP(hipGetLastError());
  return LD_OK;
}
