// Field time-base correction and FP64 hsync refinement, plus their C ABI.
//
//   k_tbc_lines   one workgroup per field: per-line slopes of the line map,
//                 weighted median / MAD outlier clamp of the wow factors, and
//                 the EMA start value of every line (utils.py:131-149)
//   k_tbc_pixels  one thread per output pixel: FP64 source position from the
//                 piecewise-linear line map (core.py:2591-2623), 16-tap
//                 sinc-LUT interpolation (utils.py:151-183), level adjust, and
//                 the Hz -> uint16 code map (utils.py:1061-1075)
//   k_hsync       one wavefront per line: sub-sample hsync edge search on
//                 demod_05 (core.py:2484-2551, calczc utils.py:795-854)
#include <math.h>

#include <vector>

#include "internal.hpp"

// Positions, level adjust and the output code map follow the reference's
// operation order with one rounding per operation; the compiler must not fuse
// multiplies and adds here (the sinc weight interpolation uses an explicit fmaf).
#pragma clang fp contract(off)

namespace ldgpu {

constexpr int kLinesMax = 1024;  // line locations per field handled by k_tbc_lines
constexpr int kLineThreads = 256;

// Workspace row per field: level (double), EMA start (double), first pixel (int64)
struct LineRow {
  double level[kLinesMax];
  double ema_start[kLinesMax];
  long long first_px[kLinesMax + 1];
};

// Segment of the k=1 spline holding x: knots at j*inlinelen, right-continuous
// (scipy's interval search picks t[j] <= x < t[j+1]).
__device__ __forceinline__ int segment_of(double x, double inl, int nseg_max) {
  int j = (int)floor(x / inl);
  if ((double)(j + 1) * inl <= x) ++j;
  if ((double)j * inl > x) --j;
  return j < 0 ? 0 : (j > nseg_max ? nseg_max : j);
}

// Bitonic sort of n <= kLinesMax (key, weight) pairs in shared memory.
__device__ void block_sort(double* key, int* wgt, int n) {
  int npow = 1;
  while (npow < n) npow <<= 1;
  for (int i = n + threadIdx.x; i < npow; i += blockDim.x) {
    key[i] = INFINITY;
    wgt[i] = 0;
  }
  __syncthreads();
  for (int k = 2; k <= npow; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < npow; i += blockDim.x) {
        int ixj = i ^ j;
        if (ixj > i) {
          bool up = (i & k) == 0;
          double a = key[i], b = key[ixj];
          if ((a > b) == up) {
            key[i] = b;
            key[ixj] = a;
            int t = wgt[i];
            wgt[i] = wgt[ixj];
            wgt[ixj] = t;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Median of the multiset {key[i] repeated wgt[i] times} (sorted keys): the
// mean of the elements at ranks (L-1)/2 and L/2, as numpy's median.
__device__ double weighted_median(const double* key, const int* wgt, int n) {
  long long L = 0;
  for (int i = 0; i < n; ++i) L += wgt[i];
  long long r0 = (L - 1) / 2, r1 = L / 2;
  double v0 = 0, v1 = 0;
  long long c = 0;
  bool got0 = false;
  for (int i = 0; i < n; ++i) {
    long long nc = c + wgt[i];
    if (!got0 && r0 < nc) {
      v0 = key[i];
      got0 = true;
    }
    if (r1 < nc) {
      v1 = key[i];
      break;
    }
    c = nc;
  }
  return (L % 2) ? v1 : (v0 + v1) / 2.0;
}

__global__ __launch_bounds__(kLineThreads) void k_tbc_lines(const ld_tbc_field* __restrict__ fields,
                                                           const double* __restrict__ linelocs,
                                                           LineRow* __restrict__ rows, double smoothing,
                                                           double threshold) {
  __shared__ double key[kLinesMax];
  __shared__ int wgt[kLinesMax];
  __shared__ double slope[kLinesMax];
  __shared__ int cnt[kLinesMax];
  __shared__ double med_s, mad_s;
  const ld_tbc_field fd = fields[blockIdx.x];
  const double* ll = linelocs + fd.lineloc_offset;
  LineRow* row = rows + blockIdx.x;
  const int W = fd.outwidth;
  const long long npix = (long long)(fd.linesout + fd.lineoffset + 1) * W;  // pixel_locs length
  const double inl = fd.inlinelen;
  const double outscale = inl / (double)W;
  const int nseg = fd.nlines - 1;
  // first pixel of every segment (monotone in p; boundaries sit next to j*W)
  for (int j = threadIdx.x; j <= nseg; j += blockDim.x) {
    long long b;
    if (j == 0) {
      b = 0;
    } else if (j == nseg) {
      b = npix;
    } else {
      b = (long long)j * W - 2;
      if (b < 0) b = 0;
      while (b < npix && segment_of((double)b * outscale, inl, nseg - 1) < j) ++b;
    }
    row->first_px[j] = b < npix ? b : npix;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nseg; j += blockDim.x) {
    // derivative of the k=1 B-spline on [t_j, t_j+1): c_j*(-w) + c_{j+1}*w, w = 1/dt
    double w = 1.0 / (((double)(j + 1) * inl) - ((double)j * inl));
    double s = ll[j] * (-w) + ll[j + 1] * w;
    slope[j] = s;
    key[j] = s;
    int c = (int)(row->first_px[j + 1] - row->first_px[j]);
    cnt[j] = c;
    wgt[j] = c;
  }
  __syncthreads();
  block_sort(key, wgt, nseg);
  if (threadIdx.x == 0) med_s = weighted_median(key, wgt, nseg);
  __syncthreads();
  const double med = med_s;
  for (int j = threadIdx.x; j < nseg; j += blockDim.x) {
    key[j] = fabs(slope[j] - med);
    wgt[j] = cnt[j];
  }
  __syncthreads();
  block_sort(key, wgt, nseg);
  if (threadIdx.x == 0) mad_s = weighted_median(key, wgt, nseg);
  __syncthreads();
  const double mad = mad_s;
  const double thr = mad > 0 ? threshold * mad : 0.001;
  for (int j = threadIdx.x; j < nseg; j += blockDim.x)
    row->level[j] = fabs(slope[j] - med) > thr ? med : slope[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    // EMA value at the pixel before each line's first pixel (utils.py:142-149,
    // y_i = a*l_i + (1-a)*y_{i-1}, y_0 = l_0), in closed form per line.
    double y = row->level[0];
    row->ema_start[0] = y;
    if (smoothing > 0) {
      const double alpha = 1.0 / (smoothing * W);
      const double keep = 1.0 - alpha;
      for (int j = 0; j < nseg; ++j) {
        row->ema_start[j] = y;
        long long c = row->first_px[j + 1] - row->first_px[j];
        if (j == 0) c -= 1;  // pixel 0 is not smoothed
        if (c > 0) y = row->level[j] + pow(keep, (double)c) * (y - row->level[j]);
      }
    }
  }
}

struct TbcPixelArgs {
  const float* demod;
  const ld_tbc_field* fields;
  const double* linelocs;
  const LineRow* rows;
  const float* lut;
  float* out_f32;
  uint16_t* out_u16;
  double smoothing;
  double lscale, loffset;  // hz_to_output: x*lscale + loffset
  int coord_f32;
};

__global__ __launch_bounds__(256) void k_tbc_pixels(TbcPixelArgs a) {
  const ld_tbc_field fd = a.fields[blockIdx.y];
  const int W = fd.outwidth;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nout = (long long)fd.linesout * W;
  if (q >= nout) return;
  const LineRow* row = a.rows + blockIdx.y;
  const double* ll = a.linelocs + fd.lineloc_offset;
  const long long p = q + (long long)W * (fd.lineoffset + 1);
  const double inl = fd.inlinelen;
  const double x = (double)p * (inl / (double)W);
  const int nseg = fd.nlines - 1;
  const int j = segment_of(x, inl, nseg - 1);
  // value of the k=1 B-spline (scipy's de Boor order of operations)
  const double t0 = (double)j * inl, t1 = (double)(j + 1) * inl;
  const double w = 1.0 / (t1 - t0);
  const double loc = ll[j] * (w * (t1 - x)) + ll[j + 1] * (w * (x - t0));
  double level = row->level[j];
  if (a.smoothing > 0) {
    const long long t = p - row->first_px[j];
    if (!(j == 0)) {
      const double keep = 1.0 - 1.0 / (a.smoothing * W);
      level = level + pow(keep, (double)(t + 1)) * (row->ema_start[j] - level);
    }
  }
  long long ci;
  int ph0;
  float alpha;
  if (a.coord_f32) {
    const float c = (float)loc;
    ci = (long long)c;  // truncation toward zero, as int()
    const float frac = c - (float)ci;
    const float pp = frac * 65536.0f;
    ph0 = (int)pp;
    alpha = pp - (float)ph0;
  } else {
    ci = (long long)loc;
    const double frac = loc - (double)ci;
    const double pp = frac * 65536.0;
    ph0 = (int)pp;
    alpha = (float)(pp - (double)ph0);
  }
  // negative coordinates have no LUT row (the reference reads out of bounds):
  // use phase 0; their taps fall outside the buffer and read 0
  if (ph0 < 0 || ph0 >= kSincPhases) {
    ph0 = 0;
    alpha = 0.0f;
  }
  const float4* w0 = reinterpret_cast<const float4*>(a.lut + (size_t)ph0 * kSincTaps);
  const float4* w1 = reinterpret_cast<const float4*>(a.lut + (size_t)(ph0 + 1) * kSincTaps);
  const float* src = a.demod + fd.demod_offset;
  const long long start = ci - (kSincTaps / 2 - 1);
  double acc = 0.0;
#pragma unroll
  for (int g = 0; g < kSincTaps / 4; ++g) {
    float4 u = w0[g], v = w1[g];
    float ws[4] = {u.x, u.y, u.z, u.w};
    float we[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      long long idx = start + 4 * g + e;
      float s = (idx >= 0 && idx < fd.demod_len) ? src[idx] : 0.0f;
      float wt = fmaf(alpha, we[e] - ws[e], ws[e]);
      acc += (double)(s * wt);
    }
  }
  const float r = (float)(level * acc);
  const long long o = fd.out_offset + q;
  if (a.out_f32) a.out_f32[o] = r;
  if (a.out_u16) {
    double vv = (double)r * a.lscale + a.loffset + 0.5;
    vv = vv < 0.0 ? 0.0 : (vv > 65535.0 ? 65535.0 : vv);
    a.out_u16[o] = (uint16_t)vv;
  }
}

// ------------------------------------------------------------- hsync refine
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// calczc (utils.py:795-841) by one wavefront: first crossing of `target` in
// data[s .. s+count] with s = int(start); edge from data[s] (rising if below).
__device__ bool wave_calczc(const float* d, long long len, double start, double target, int count,
                            double* out) {
  const long long s = (long long)start;  // int() truncates toward zero
  if (s < 0 || s >= len) return false;
  const bool rising = (double)d[s] < target;
  long long n = (long long)count + 1;
  if (s + n > len) n = len - s;
  for (long long base = 1; base < n; base += 64) {
    long long i = base + lane_id();
    bool hit = false;
    if (i < n) {
      double a = d[s + i - 1], b = d[s + i];
      hit = rising ? (a < target && b >= target) : (a > target && b <= target);
    }
    unsigned long long m = __ballot(hit);
    if (m) {
      long long first = base + (__ffsll((long long)m) - 1);
      long long x = s + first;
      double a = (double)d[x - 1] - target, b = (double)d[x] - target;
      double y = (b - a) != 0 ? -a / (-a + b) : 0.0;
      *out = (double)(x - 1) + y;
      return true;
    }
  }
  return false;
}

__device__ float wave_sort64(float v) {
  const int lane = lane_id();
  for (int k = 2; k <= 64; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      float o = __shfl_xor(v, j);
      bool up = (lane & k) == 0;
      bool lower = (lane & j) == 0;
      v = (lower == up) ? fminf(v, o) : fmaxf(v, o);
    }
  return v;
}

// median of d[b, e) (e - b <= 64), numba's np.median: 0.5 * (a + b) with a + b in float32
__device__ double wave_median(const float* d, long long len, long long b, long long e) {
  if (b < 0) b = 0;
  if (e > len) e = len;
  int n = (int)(e - b);
  if (n <= 0) return 0.0;
  if (n > 64) n = 64;
  float v = lane_id() < n ? d[b + lane_id()] : INFINITY;
  v = wave_sort64(v);
  if (n & 1) return (double)__shfl(v, n / 2);
  float lo = __shfl(v, n / 2 - 1), hi = __shfl(v, n / 2);
  return 0.5 * (double)(lo + hi);
}

__global__ __launch_bounds__(64) void k_hsync(const float* __restrict__ demod05,
                                             const ld_tbc_field* __restrict__ fields,
                                             const double* __restrict__ ll1_all,
                                             const uint8_t* __restrict__ bad_in_all, double freq, double sync_mid,
                                             double lo_hz, double hi_hz, int pal, double* __restrict__ ll2_all,
                                             uint8_t* __restrict__ bad_out_all) {
  const ld_tbc_field fd = fields[blockIdx.y];
  const int i = blockIdx.x;
  if (i >= fd.nlines) return;
  const float* d = demod05 + fd.demod_offset;
  const long long len = fd.demod_len;
  const long long li = fd.lineloc_offset + i;
  const double l1 = ll1_all[li];
  bool bad = bad_in_all ? bad_in_all[li] != 0 : false;
  double res = l1;
  if ((i >= 3 && i <= 6) || (pal && i >= 1 && i <= 2)) {
    bad = true;
  } else {
    const double start = l1 - freq;
    double zc;
    bool found = wave_calczc(d, len, start, sync_mid, (int)(freq * 2), &zc);
    if (found && !bad) {
      long long b0 = (long long)(zc - freq * 0.75), e0 = (long long)(zc + freq * 8);
      if (b0 < 0) b0 = 0;
      if (e0 > len) e0 = len;
      float mn = INFINITY, mx = -INFINITY;
      for (long long k = b0 + lane_id(); k < e0; k += 64) {
        float v = d[k];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
      }
      for (int o = 32; o > 0; o >>= 1) {
        mn = fminf(mn, __shfl_xor(mn, o));
        mx = fmaxf(mx, __shfl_xor(mx, o));
      }
      if ((double)mn < lo_hz || (double)mx > hi_hz) {
        bad = true;
      } else {
        double porch = wave_median(d, len, (long long)(zc + freq * 8), (long long)(zc + freq * 9));
        double sync = wave_median(d, len, (long long)(zc + freq * 1), (long long)(zc + freq * 2.5));
        double zc2;
        bool f2 = wave_calczc(d, len, start, (porch + sync) / 2, 400, &zc2);
        if (f2 && fabs(zc2 - zc) < freq / 2)
          res = zc2;
        else
          bad = true;
      }
    } else {
      bad = true;
    }
  }
  if (bad) res = l1;
  if (lane_id() == 0) {
    ll2_all[li] = res;
    bad_out_all[li] = bad ? 1 : 0;
  }
}

}  // namespace ldgpu

using namespace ldgpu;

extern "C" {

int ld_set_sinc_lut(ld_handle* h, const float* host_lut, int rows, int taps) {
  if (!h || !host_lut || rows != kSincPhases + 1 || taps != kSincTaps) {
    set_error("ld_set_sinc_lut: expected a %d x %d float32 table", kSincPhases + 1, kSincTaps);
    return LD_EINVAL;
  }
  LD_HIP(hipSetDevice(h->device));
  size_t bytes = sizeof(float) * (size_t)rows * taps;
  if (!h->sinc_lut) LD_HIP(hipMalloc(&h->sinc_lut, bytes));
  LD_HIP(hipMemcpy(h->sinc_lut, host_lut, bytes, hipMemcpyHostToDevice));
  return LD_OK;
}

static size_t tbc_fields_bytes(int nfields) {
  return ((sizeof(ld_tbc_field) * (size_t)nfields + 255) / 256) * 256;
}

size_t ld_tbc_workspace_bytes(const ld_handle* h, int nfields) {
  if (!h || nfields <= 0) return 0;
  return sizeof(LineRow) * (size_t)nfields + tbc_fields_bytes(nfields);
}

int ld_tbc_fields(ld_handle* h, const float* demod, int64_t demod_total, const ld_tbc_field* fields,
                  int nfields, const double* linelocs, int64_t nlinelocs, int coord_f32, double smoothing,
                  double threshold, const ld_level_params* level, float* out_f32, uint16_t* out_u16,
                  int64_t out_total, void* workspace, size_t ws_bytes, void* stream) {
  if (!h || !demod || !fields || !linelocs || nfields <= 0 || !workspace) {
    set_error("ld_tbc_fields: null argument or empty batch");
    return LD_EINVAL;
  }
  int max_pixels_per_field = 0;
  for (int f = 0; f < nfields; ++f) {
    const ld_tbc_field& fd = fields[f];
    if (fd.nlines < 2 || fd.nlines > kLinesMax || fd.outwidth <= 0 || fd.linesout <= 0 || fd.lineoffset < 0 ||
        !(fd.inlinelen > 0) || fd.lineloc_offset < 0 || fd.lineloc_offset + fd.nlines > nlinelocs) {
      set_error("ld_tbc_fields: field %d: bad geometry (nlines %d, max %d) or line range", f, fd.nlines, kLinesMax);
      return LD_EINVAL;
    }
    if (fd.demod_offset < 0 || fd.demod_len < 0 || fd.demod_offset + fd.demod_len > demod_total) {
      set_error("ld_tbc_fields: field %d: demod span outside the buffer", f);
      return LD_ERANGE;
    }
    long long npx = (long long)fd.linesout * fd.outwidth;
    if (fd.out_offset < 0 || fd.out_offset + npx > out_total) {
      set_error("ld_tbc_fields: field %d: output span outside the buffer", f);
      return LD_ERANGE;
    }
    if (npx > max_pixels_per_field) max_pixels_per_field = (int)npx;
  }
  if (!out_f32 && !out_u16) {
    set_error("ld_tbc_fields: no output requested");
    return LD_EINVAL;
  }
  if (out_u16 && !level) {
    set_error("ld_tbc_fields: uint16 output needs level parameters");
    return LD_EINVAL;
  }
  if (!h->sinc_lut) {
    set_error("ld_tbc_fields: sinc LUT not uploaded");
    return LD_ESTATE;
  }
  if (ws_bytes < ld_tbc_workspace_bytes(h, nfields)) {
    set_error("ld_tbc_fields: workspace %zu bytes < %zu needed", ws_bytes, ld_tbc_workspace_bytes(h, nfields));
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  ld_tbc_field* dfields = (ld_tbc_field*)workspace;
  LineRow* rows = (LineRow*)((char*)workspace + tbc_fields_bytes(nfields));
  LD_HIP(hipMemcpyAsync(dfields, fields, sizeof(ld_tbc_field) * (size_t)nfields, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_tbc_lines, dim3(nfields), dim3(kLineThreads), 0, st, dfields, linelocs, rows, smoothing,
                     threshold);
  LD_HIP(hipGetLastError());
  TbcPixelArgs a;
  a.demod = demod;
  a.fields = dfields;
  a.linelocs = linelocs;
  a.rows = rows;
  a.lut = h->sinc_lut;
  a.out_f32 = out_f32;
  a.out_u16 = out_u16;
  a.smoothing = smoothing;
  a.coord_f32 = coord_f32;
  if (level) {
    a.lscale = level->out_scale / level->hz_ire;
    a.loffset = level->output_zero - level->vsync_ire * level->out_scale - level->ire0 * a.lscale;
  } else {
    a.lscale = a.loffset = 0.0;
  }
  dim3 grid((max_pixels_per_field + 255) / 256, nfields);
  hipLaunchKernelGGL(k_tbc_pixels, grid, dim3(256), 0, st, a);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

// Validate host field descriptors and return the largest line count (or an error code).
static int check_fields(const ld_tbc_field* fields, int nfields, int64_t demod_total, int64_t nlinelocs,
                        const char* who) {
  int maxl = 0;
  for (int f = 0; f < nfields; ++f) {
    const ld_tbc_field& fd = fields[f];
    if (fd.nlines < 1 || fd.nlines > kLinesMax || fd.lineloc_offset < 0 || fd.lineloc_offset + fd.nlines > nlinelocs) {
      set_error("%s: field %d: bad line range (nlines %d, max %d)", who, f, fd.nlines, kLinesMax);
      return LD_EINVAL;
    }
    if (fd.demod_offset < 0 || fd.demod_len <= 0 || fd.demod_offset + fd.demod_len > demod_total) {
      set_error("%s: field %d: demod span outside the buffer", who, f);
      return LD_ERANGE;
    }
    if (fd.nlines > maxl) maxl = fd.nlines;
  }
  return maxl;
}

size_t ld_hsync_workspace_bytes(const ld_handle* h, int nfields) {
  if (!h || nfields <= 0) return 0;
  return tbc_fields_bytes(nfields);
}

int ld_refine_hsync(ld_handle* h, const float* demod_05, int64_t demod_total, const ld_tbc_field* fields,
                    int nfields, const double* linelocs1, const uint8_t* linebad_in, int64_t nlinelocs,
                    double freq, double sync_mid_hz, double lo_hz, double hi_hz, int pal, double* linelocs2,
                    uint8_t* linebad_out, void* workspace, size_t ws_bytes, void* stream) {
  if (!h || !demod_05 || !fields || nfields <= 0 || !linelocs1 || !linelocs2 || !linebad_out || !workspace ||
      !(freq > 0)) {
    set_error("ld_refine_hsync: bad arguments");
    return LD_EINVAL;
  }
  int maxl = check_fields(fields, nfields, demod_total, nlinelocs, "ld_refine_hsync");
  if (maxl < 0) return maxl;
  if (ws_bytes < tbc_fields_bytes(nfields)) {
    set_error("ld_refine_hsync: workspace %zu bytes < %zu needed", ws_bytes, tbc_fields_bytes(nfields));
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  ld_tbc_field* dfields = (ld_tbc_field*)workspace;
  LD_HIP(hipMemcpyAsync(dfields, fields, sizeof(ld_tbc_field) * (size_t)nfields, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_hsync, dim3(maxl, nfields), dim3(64), 0, st, demod_05, dfields, linelocs1, linebad_in, freq,
                     sync_mid_hz, lo_hz, hi_hz, pal, linelocs2, linebad_out);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

}  // extern "C"
