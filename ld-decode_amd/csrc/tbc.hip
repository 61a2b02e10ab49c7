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

#include <cstring>
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

constexpr int kTbcThreads = 256;       // pixels per workgroup (one chunk of a line)
constexpr int kTbcMaxWidth = 2048;
constexpr int kTbcChunksMax = kTbcMaxWidth / kTbcThreads;
constexpr int kChunkStage = 1024;      // staged source samples per chunk (256 pixels at up to 3.9 samples each)

// Everything the pixel kernel needs for one output line, written by
// k_tbc_lines so that a line's workgroups start with a single record load.
struct LineRec {
  double ll[4];          // knots jB .. jB+3
  double lev[3], ema[3]; // per-segment level and EMA start, segments jB .. jB+2
  double w[3];           // 1 / (t_{j+1} - t_j) of those segments
  long long first[3];    // first pixel of those segments
  int bnd[4];            // first pixel of segments jB .. jB+3 relative to p0 (the line's pixel t
                         // lies in segment jB + k for bnd[k] <= t < bnd[k+1])
  double tk[4];          // knot times (jB + k) * inlinelen
  // the segment holding most of the line's pixels (jB + 1 unless empty):
  // its pixel range relative to p0 and its values, at fixed offsets
  int mb0, mb1, mj;
  double m_ll0, m_ll1, m_lev, m_w, m_t0, m_t1;
  long long p0;          // pixel index of the line's first output pixel
  double p0d;            // the same as a double
  long long demod_offset, demod_len, out_offset;
  long long lineloc_offset;
  double inl, rinl, outscale;
  int jB, nseg, W, field;
  // per 256-pixel chunk: first staged source sample and staged count (0:
  // the chunk's taps are read from global memory)
  long long slo[kTbcChunksMax];
  int nst[kTbcChunksMax];
};

// Segment of the k=1 spline holding x: knots at j*inlinelen, right-continuous
// (scipy's interval search picks t[j] <= x < t[j+1]).
__device__ __forceinline__ int segment_of(double x, double inl, int nseg_max) {
  int j = (int)floor(x / inl);
  if ((double)(j + 1) * inl <= x) ++j;
  if ((double)j * inl > x) --j;
  return j < 0 ? 0 : (j > nseg_max ? nseg_max : j);
}

// Sort n <= kLinesMax (key, weight) pairs in shared memory by rank: each
// element's position is the number of elements ordered before it (ties by
// index), so one pass of n^2/blockDim comparisons and two barriers.
__device__ void block_sort(double* key, int* wgt, double* tkey, int* twgt, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double ki = key[i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double kj = key[j];
      r += (kj < ki) || (kj == ki && j < i);
    }
    tkey[r] = ki;
    twgt[r] = wgt[i];
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    key[i] = tkey[i];
    wgt[i] = twgt[i];
  }
  __syncthreads();
}

// Median of the multiset {key[i] repeated wgt[i] times} over sorted keys:
// the mean of the elements at ranks (L-1)/2 and L/2 (numpy's median), with
// L the total weight.  Inclusive prefix sums of the weights (block scan in
// `scan`) locate both ranks in parallel.  All threads return the result.
__device__ double weighted_median(const double* key, const int* wgt, int* scan, int n, double* out) {
  constexpr int PER = kLinesMax / kLineThreads;  // elements per thread
  const int t = threadIdx.x;
  int local[PER];
  int run = 0;
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int i = t * PER + e;
    run += (i < n) ? wgt[i] : 0;
    local[e] = run;
  }
  scan[t] = run;
  __syncthreads();
  for (int d = 1; d < kLineThreads; d <<= 1) {  // Hillis-Steele over thread totals
    int v = (t >= d) ? scan[t - d] : 0;
    __syncthreads();
    scan[t] += v;
    __syncthreads();
  }
  const int before = (t > 0) ? scan[t - 1] : 0;
  const long long L = scan[kLineThreads - 1];
  const long long r0 = (L - 1) / 2, r1 = L / 2;
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int i = t * PER + e;
    if (i >= n) break;
    const long long hi = before + local[e], lo = hi - wgt[i];
    if (r0 >= lo && r0 < hi) out[0] = key[i];
    if (r1 >= lo && r1 < hi) out[1] = key[i];
  }
  __syncthreads();
  const double m = (L % 2) ? out[1] : (out[0] + out[1]) / 2.0;
  __syncthreads();
  return m;
}

// Source position of pixel p of a line (the spline value of pixel_prep).
__device__ __forceinline__ double line_loc(const double* __restrict__ linelocs, const LineRec& rc, long long p) {
  const double x = (double)p * rc.outscale;
  int j = (int)(x * rc.rinl);
  if ((double)(j + 1) * rc.inl <= x) ++j;
  if ((double)j * rc.inl > x) --j;
  j = j < 0 ? 0 : (j > rc.nseg - 1 ? rc.nseg - 1 : j);
  const int jl = j - rc.jB;
  double llj, llj1, w;
  if (jl >= 0 && jl <= 2) {
    llj = rc.ll[jl];
    llj1 = rc.ll[jl + 1];
    w = rc.w[jl];
  } else {
    const double* ll = linelocs + rc.lineloc_offset;
    llj = ll[j];
    llj1 = ll[j + 1];
    w = 1.0 / (((double)(j + 1) * rc.inl) - ((double)j * rc.inl));
  }
  const double t0 = (double)j * rc.inl, t1 = (double)(j + 1) * rc.inl;
  return llj * (w * (t1 - x)) + llj1 * (w * (x - t0));
}

// Staged source span of pixels c0 .. c1 of a line: from the positions of
// the first and last pixel, with the taps' reach either side.
__device__ __forceinline__ void chunk_span(const double* __restrict__ linelocs, const LineRec& rc, int c0, int c1,
                                           long long& slo, int& nst) {
  const double la = line_loc(linelocs, rc, rc.p0 + c0), lb = line_loc(linelocs, rc, rc.p0 + c1);
  const double lo = fmin(la, lb), hi = fmax(la, lb);
  const bool finite = lo == lo && hi == hi && fabs(lo) < 1e15 && fabs(hi) < 1e15;
  slo = finite ? (long long)floor(lo) - (kSincTaps / 2 + 2) : 0;
  const long long shi = finite ? (long long)floor(hi) + (kSincTaps / 2 + 4) : 0;
  nst = (finite && shi - slo <= kChunkStage) ? (int)(shi - slo) : 0;
}

__global__ __launch_bounds__(kLineThreads) void k_tbc_lines(const ld_tbc_field* __restrict__ fields,
                                                           const double* __restrict__ linelocs,
                                                           LineRow* __restrict__ rows, double smoothing,
                                                           double threshold, LineRec* __restrict__ recs,
                                                           const int* __restrict__ rec_base) {
  __shared__ double key[kLinesMax];
  __shared__ int wgt[kLinesMax];
  __shared__ double slope[kLinesMax];
  __shared__ int cnt[kLinesMax];
  __shared__ double tkey[kLinesMax];
  __shared__ int twgt[kLinesMax];
  __shared__ int scan[kLineThreads];
  __shared__ double sel[2];
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
  block_sort(key, wgt, tkey, twgt, nseg);
  const double med = weighted_median(key, wgt, scan, nseg, sel);
  for (int j = threadIdx.x; j < nseg; j += blockDim.x) {
    key[j] = fabs(slope[j] - med);
    wgt[j] = cnt[j];
  }
  __syncthreads();
  block_sort(key, wgt, tkey, twgt, nseg);
  const double mad = weighted_median(key, wgt, scan, nseg, sel);
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
  __syncthreads();
  // one record per output line
  LineRec* out = recs + rec_base[blockIdx.x];
  for (int line = threadIdx.x; line < fd.linesout; line += blockDim.x) {
    LineRec r;
    const long long p0 = (long long)(line + fd.lineoffset + 1) * W;
    int jB = (int)(p0 / W) - 1;
    jB = jB > nseg - 3 ? nseg - 3 : jB;
    jB = jB < 0 ? 0 : jB;
    for (int i = 0; i < 4; ++i) r.ll[i] = ll[jB + i < nseg ? jB + i : nseg];
    for (int i = 0; i < 3; ++i) {
      const int js = jB + i < nseg ? jB + i : nseg - 1;
      r.lev[i] = row->level[js];
      r.ema[i] = row->ema_start[js];
      r.first[i] = row->first_px[js];
      r.w[i] = 1.0 / (((double)(js + 1) * inl) - ((double)js * inl));
    }
    r.p0 = p0;
    r.p0d = (double)p0;
    for (int i = 0; i < 4; ++i) {
      // segments past the last are empty (their bound is the pixel count)
      r.bnd[i] = (int)(row->first_px[jB + i <= nseg ? jB + i : nseg] - p0);
      r.tk[i] = (double)(jB + i) * inl;
    }
    const int k = r.bnd[1] < r.bnd[2] ? 1 : 0;
    r.mb0 = r.bnd[k];
    r.mb1 = r.bnd[k + 1];
    r.mj = jB + k;
    r.m_ll0 = r.ll[k];
    r.m_ll1 = r.ll[k + 1];
    r.m_lev = r.lev[k];
    r.m_w = r.w[k];
    r.m_t0 = r.tk[k];
    r.m_t1 = r.tk[k + 1];
    r.demod_offset = fd.demod_offset;
    r.demod_len = fd.demod_len;
    r.out_offset = fd.out_offset + (long long)line * W;
    r.lineloc_offset = fd.lineloc_offset;
    r.inl = inl;
    r.rinl = 1.0 / inl;
    r.outscale = outscale;
    r.jB = jB;
    r.nseg = nseg;
    r.W = W;
    r.field = blockIdx.x;
    for (int c = 0; c < kTbcChunksMax; ++c) {
      const int c0 = c * kTbcThreads;
      r.slo[c] = 0;
      r.nst[c] = 0;
      if (c0 < W) chunk_span(linelocs, r, c0, (W - c0 < kTbcThreads ? W : c0 + kTbcThreads) - 1, r.slo[c], r.nst[c]);
    }
    out[line] = r;
  }
}

struct TbcPixelArgs {
  const LineRec* recs;
  const float* demod;
  const double* linelocs;
  const LineRow* rows;
  const float* lut;
  float* out_f32;
  uint16_t* out_u16;
  double smoothing;
  double lscale, loffset;  // hz_to_output: x*lscale + loffset
  int coord_f32;
  int lut_folded;          // read rows > 32768 as mirrored rows of the lower half
  int chunks;              // workgroups per output line
};

constexpr int kStagePerThread = kChunkStage / kTbcThreads;

typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ __forceinline__ float4 lut_load(rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}

__device__ __forceinline__ float comp(const float4& f, int c) {
  return c == 0 ? f.x : (c == 1 ? f.y : (c == 2 ? f.z : f.w));
}

// Per-pixel state between the LUT row addresses and the tap sum.
struct PixelPrep {
  double start;        // first tap's sample index (field-relative), an integer
  double level;
  float alpha;
  uint32_t fold_mask;  // all ones in a folded row (groups and taps reversed)
  uint32_t ou, ov;     // byte offsets of the ws / we groups, see pixel_prep
};
// LUT offset past the table: buffer loads there return zeros
constexpr uint32_t kLutNone = 0x7fffff00u;

// A wave-uniform double, marked as such (its loads then stay scalar).
__device__ __forceinline__ double wave_uniform(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Source position, level factor and LUT rows of pixel t of the line.
__device__ __forceinline__ void pixel_prep(const TbcPixelArgs& a, const LineRec& rc, rsrc_t lut, int t,
                                           PixelPrep& q) {
  const int W = rc.W;
  const double x = (rc.p0d + (double)t) * rc.outscale;  // p0 + t is exact in a double
  // segment of the k=1 spline holding x (knots at j*inlinelen, right-
  // continuous): inside the three cached segments from the line's segment
  // bounds (pixel indices k_tbc_lines found with the same search), else
  // searched here
  int j;
  double llj, llj1, lev, ema, w, t0, t1;
  long long first;
  // the main segment's values, loaded with the record's other fields (both
  // paths below read them, so their loads are not deferred to a branch)
  const int mj = rc.mj;
  const double m_ll0 = wave_uniform(rc.m_ll0), m_ll1 = wave_uniform(rc.m_ll1);
  const double m_lev = wave_uniform(rc.m_lev), m_w = wave_uniform(rc.m_w);
  const double m_t0 = wave_uniform(rc.m_t0), m_t1 = wave_uniform(rc.m_t1);
  if (__ballot(t < rc.mb0 || t >= rc.mb1) == 0) {
    // the wavefront's pixels all lie in the line's main segment (all
    // wavefronts but those holding a segment boundary): no per-pixel
    // memory round trip
    j = mj;
    llj = m_ll0;
    llj1 = m_ll1;
    lev = m_lev;
    w = m_w;
    t0 = m_t0;
    t1 = m_t1;
    ema = rc.ema[j - rc.jB];
    first = rc.first[j - rc.jB];
  } else {
    int jl = -1;
    if (t >= rc.bnd[0] && t < rc.bnd[3]) jl = (t >= rc.bnd[1] ? 1 : 0) + (t >= rc.bnd[2] ? 1 : 0);
    if (jl >= 0 && rc.jB + jl == mj) {
      j = mj;
      llj = m_ll0;
      llj1 = m_ll1;
      lev = m_lev;
      ema = rc.ema[jl];
      first = rc.first[jl];
      w = m_w;
    } else if (jl >= 0) {
      j = rc.jB + jl;
      llj = rc.ll[jl];
      llj1 = rc.ll[jl + 1];
      lev = rc.lev[jl];
      ema = rc.ema[jl];
      first = rc.first[jl];
      w = rc.w[jl];
    } else {
      // the guess x*(1/inl) is within one of the segment
      j = (int)(x * rc.rinl);
      if ((double)(j + 1) * rc.inl <= x) ++j;
      if ((double)j * rc.inl > x) --j;
      j = j < 0 ? 0 : (j > rc.nseg - 1 ? rc.nseg - 1 : j);
      const double* ll = a.linelocs + rc.lineloc_offset;
      const LineRow* row = a.rows + rc.field;
      llj = ll[j];
      llj1 = ll[j + 1];
      lev = row->level[j];
      ema = row->ema_start[j];
      first = row->first_px[j];
      w = 1.0 / (((double)(j + 1) * rc.inl) - ((double)j * rc.inl));
    }
    t0 = (double)j * rc.inl;
    t1 = (double)(j + 1) * rc.inl;
  }
  // value of the k=1 B-spline (scipy's de Boor order of operations)
  const double loc = llj * (w * (t1 - x)) + llj1 * (w * (x - t0));
  q.level = lev;
  if (a.smoothing > 0 && j != 0) {
    const double keep = 1.0 - 1.0 / (a.smoothing * W);
    q.level = lev + pow(keep, (double)(rc.p0 + t - first + 1)) * (ema - lev);
  }
  // int(coord) truncates toward zero; the integer is kept as a float /
  // double (trunc), which equals the converted integer wherever the
  // coordinate has a fraction, and the coordinate itself where it has none
  double ci;
  int ph0;
  float alpha;
  if (a.coord_f32) {
    const float c = (float)loc;
    const float tc = truncf(c);
    const float frac = c - tc;
    const float pp = frac * 65536.0f;
    ph0 = (int)pp;
    alpha = pp - (float)ph0;
    ci = (double)tc;
  } else {
    ci = trunc(loc);
    const double frac = loc - ci;
    const double pp = frac * 65536.0;
    ph0 = (int)pp;
    alpha = (float)(pp - (double)ph0);
  }
  // negative coordinates have no LUT row (the reference reads out of
  // bounds): use phase 0; their taps fall outside the buffer and read 0
  if (ph0 < 0 || ph0 >= kSincPhases) {
    ph0 = 0;
    alpha = 0.0f;
  }
  q.alpha = alpha;
  q.start = ci - (double)(kSincTaps / 2 - 1);
  // Weights ws = row ph0, we = row ph0+1.  In a folded table rows above
  // 32768 are read as tap-reversed rows of the lower half: ws = reversed row
  // 65536-ph0, we = reversed row 65535-ph0 (for ph0 = 65535 the last row
  // repeats row 65535, i.e. ws).  The float4 groups are fetched in reversed
  // group order here and their components reversed in pixel_finish.
  const bool fold = a.lut_folded && ph0 > kSincPhases / 2 - 1;
  const int b = fold ? kSincPhases - 1 - ph0 : ph0;
  const uint32_t rowA = (uint32_t)b * (kSincTaps * 4u), rowB = rowA + kSincTaps * 4u;
  const bool last = ph0 == kSincPhases - 1;
  const uint32_t ou = fold ? rowB + 48u : rowA;
  const uint32_t ov = fold ? (last ? rowB : rowA) + 48u : rowB;
  q.ou = ou;
  q.ov = ov;
  q.fold_mask = fold ? 0xffffffffu : 0u;
}

// Value of lane g of the thread's 4-lane group (DPP quad broadcast).
template <int G>
__device__ __forceinline__ int quad_bcast(int v) {
  return __builtin_amdgcn_mov_dpp(v, G * 0x55, 0xf, 0xf, true);
}

// The sinc weights of the 4 pixels of each 4-lane group are loaded
// together: lane l of the group fetches tap group l (taps 4l .. 4l+3) of
// every pixel of the group, so one load instruction touches one 64-byte row
// per 4 lanes instead of one per lane.  quad_fetch issues the loads;
// quad_store forms wt = ws + alpha (we - ws) for those taps (the same fmaf
// as a per-pixel evaluation) and writes them to LDS at wts_slot.
// Every lane of the wavefront takes part.
//
// wts_slot: float4 slot of tap group c of pixel P.  Pixel blocks are
// permuted (P ^ (P/4 mod 4)) and the groups inside a block rotated by P, so
// that both the stores (lane l of quad m: group l of pixel 4m+g) and the
// loads (lane P: group k of its own pixel) spread over all banks.
__device__ __forceinline__ int wts_slot(int P, int c) { return 4 * (P ^ ((P >> 2) & 3)) + ((c + P) & 3); }
struct QuadRows {
  float4 u[4], v[4];
  uint32_t fold_mask[4];
  float alpha[4];
};
__device__ __forceinline__ void quad_fetch(rsrc_t lut, const PixelPrep& q, QuadRows& r, int tid) {
  const uint32_t l16 = (uint32_t)(tid & 3) * 16u;
  auto fetch = [&](auto gc) {
    constexpr int g = decltype(gc)::value;
    const uint32_t ou = (uint32_t)quad_bcast<g>((int)q.ou), ov = (uint32_t)quad_bcast<g>((int)q.ov);
    const uint32_t m = (uint32_t)quad_bcast<g>((int)q.fold_mask);
    const uint32_t d = (l16 ^ m) - m;  // +16 l, or -16 l in a folded row
    r.fold_mask[g] = m;
    r.alpha[g] = __int_as_float(quad_bcast<g>(__float_as_int(q.alpha)));
    r.u[g] = lut_load(lut, ou + d);
    r.v[g] = lut_load(lut, ov + d);
  };
  fetch(std::integral_constant<int, 0>());
  fetch(std::integral_constant<int, 1>());
  fetch(std::integral_constant<int, 2>());
  fetch(std::integral_constant<int, 3>());
}
__device__ __forceinline__ float bit_select(uint32_t m, float a, float b) {
  return __uint_as_float((__float_as_uint(a) & m) | (__float_as_uint(b) & ~m));
}
__device__ __forceinline__ void quad_store(const QuadRows& r, float4* wts, int tid) {
  const int l = tid & 3;
  const int qb = tid & ~3;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    float w[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const float ws = comp(r.u[g], c), we = comp(r.v[g], c);
      w[c] = fmaf(r.alpha[g], we - ws, ws);
    }
    // a folded row holds the taps of the group in reverse order (bitwise
    // selects: one instruction per component, no branch)
    const uint32_t m = r.fold_mask[g];
    const float4 o = make_float4(bit_select(m, w[3], w[0]), bit_select(m, w[2], w[1]), bit_select(m, w[1], w[2]),
                                 bit_select(m, w[0], w[3]));
    wts[wts_slot(qb + g, l)] = o;
  }
}

// 16-tap sum, level factor and output codes of pixel t.
__device__ __forceinline__ void pixel_finish(const TbcPixelArgs& a, const LineRec& rc, const float* stage,
                                             const float* src, double slo, int nst, int t, const PixelPrep& q,
                                             const float4* wts, int c0) {
  const double rel = q.start - slo;  // integers below 2^53: exact
  float sv[kSincTaps];
  if (rel >= 0.0 && rel <= (double)(nst - kSincTaps)) {
    const int r0 = (int)rel;
#pragma unroll
    for (int e = 0; e < kSincTaps; ++e) sv[e] = stage[r0 + e];
  } else {
    // taps outside the staged span, from the buffer (0 outside it; a
    // coordinate too large for an int64 has every tap outside)
    const bool finite = fabs(q.start) < 4.0e18;
    const long long s0 = finite ? (long long)q.start : 0;
#pragma unroll
    for (int e = 0; e < kSincTaps; ++e) {
      const long long idx = s0 + e;
      sv[e] = (finite && idx >= 0 && idx < rc.demod_len) ? src[idx] : 0.0f;
    }
  }
  float4 wq[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) wq[g] = wts[wts_slot(t - c0, g)];
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < kSincTaps; ++e) acc += (double)(sv[e] * comp(wq[e >> 2], e & 3));
  const float r = (float)(q.level * acc);
  const long long o = rc.out_offset + t;
  if (a.out_f32) a.out_f32[o] = r;
  if (a.out_u16) {
    double vv = (double)r * a.lscale + a.loffset + 0.5;
    vv = vv < 0.0 ? 0.0 : (vv > 65535.0 ? 65535.0 : vv);
    a.out_u16[o] = (uint16_t)vv;
  }
}

// One workgroup per 256-pixel chunk of an output line (blockIdx.x = line *
// chunks + chunk over the flat line records).  The chunk's source span
// follows from the positions of its first and last pixel; the LUT loads of
// every pixel and the staging loads of the span are in flight together, and
// after one barrier every thread sums its taps from LDS.  Pixels whose taps
// leave the span (decreasing line locations, spans over kChunkStage) read
// them from global memory.
__global__ __launch_bounds__(kTbcThreads) void k_tbc_pixels(TbcPixelArgs a) {
  // staged samples at LDS offset 0, so that a pixel's 16 reads share one
  // address with the tap index in the instruction's offset field
  __shared__ float4 lds[kChunkStage / 4 + kTbcThreads * 4];
  float* stage = reinterpret_cast<float*>(lds);
  float4* wts = lds + kChunkStage / 4;
  const int line = blockIdx.x / a.chunks, chunk = blockIdx.x - line * a.chunks;
  const LineRec& rc = a.recs[line];
  const int W = rc.W;
  const int c0 = chunk * kTbcThreads;
  if (c0 >= W) return;
  const int c1 = (W - c0 < kTbcThreads ? W : c0 + kTbcThreads) - 1;
  const float* src = a.demod + rc.demod_offset;
  const int tid = threadIdx.x;
  const int t = c0 + tid;
  const rsrc_t lut = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(a.lut), (short)0,
                                                       (int)((kSincPhases + 1) * kSincTaps * 4), 0x00020000);
  // staging span of the chunk (uniform, from the line record)
  const long long slo = rc.slo[chunk];
  const int nst = rc.nst[chunk];
  // LUT loads first: the in-order load counter then lets the weights be
  // formed while the staging loads issued after them are in flight
  PixelPrep q;
  q.start = 0.0;
  q.level = 0.0;
  q.alpha = 0.0f;
  q.fold_mask = 0u;
  q.ou = q.ov = kLutNone;
  if (t <= c1) pixel_prep(a, rc, lut, t, q);
  QuadRows rows;
  quad_fetch(lut, q, rows, tid);
  // staged samples i with 0 <= slo + i < demod_len and i < nst (bounds
  // formed once per workgroup in int64, compared per element in int32)
  const long long below = -slo, above = rc.demod_len - slo;
  const int ilo = below <= 0 ? 0 : (below >= nst ? nst : (int)below);
  const int ihi = above <= 0 ? 0 : (above >= nst ? nst : (int)above);
  float sreg[kStagePerThread];
#pragma unroll
  for (int k = 0; k < kStagePerThread; ++k) {
    const int i = tid + k * kTbcThreads;
    sreg[k] = (i >= ilo && i < ihi) ? src[slo + i] : 0.0f;
  }
  quad_store(rows, wts, tid);
  // zeros past nst are never read (a pixel reads the stage only when its
  // 16 taps lie below nst)
#pragma unroll
  for (int k = 0; k < kStagePerThread; ++k) stage[tid + k * kTbcThreads] = sreg[k];
  __syncthreads();
  if (t <= c1) pixel_finish(a, rc, stage, src, (double)slo, nst, t, q, wts, c0);
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
  size_t bytes = sizeof(float) * (size_t)rows * taps;
  int rc = replace_table(h, reinterpret_cast<void**>(&h->sinc_lut), host_lut, bytes);
  if (rc) return rc;
  // A Kaiser-windowed sinc table is mirror-symmetric: the weights of phase
  // p/65536 at distances (7-t)+p/65536 equal those of phase 1-p/65536 at the
  // negated distances.  When every row holds that bitwise (and the last row
  // repeats row 65535), the pixel kernel reads only rows 0..32768, which
  // halves the table's cache footprint.
  int folded = 1;
  const uint32_t* u = reinterpret_cast<const uint32_t*>(host_lut);
  for (int p = 1; p < kSincPhases && folded; ++p)
    for (int t = 0; t < kSincTaps; ++t)
      if (u[(size_t)p * kSincTaps + t] != u[(size_t)(kSincPhases - p) * kSincTaps + (kSincTaps - 1 - t)]) {
        folded = 0;
        break;
      }
  if (std::memcmp(u + (size_t)kSincPhases * kSincTaps, u + (size_t)(kSincPhases - 1) * kSincTaps,
                  sizeof(float) * kSincTaps) != 0)
    folded = 0;
  h->sinc_folded = folded;
  return LD_OK;
}

static size_t tbc_fields_bytes(int nfields) {
  // field descriptors followed by the per-field line-record base indices
  return ((sizeof(ld_tbc_field) * (size_t)nfields + sizeof(int) * ((size_t)nfields + 1) + 255) / 256) * 256;
}
constexpr int kMaxLinesOut = 1024;  // output lines per field

size_t ld_tbc_workspace_bytes(const ld_handle* h, int nfields) {
  if (!h || nfields <= 0) return 0;
  return tbc_fields_bytes(nfields) + sizeof(LineRow) * (size_t)nfields +
         sizeof(LineRec) * (size_t)nfields * kMaxLinesOut;
}

int ld_tbc_fields(ld_handle* h, const float* demod, int64_t demod_total, const ld_tbc_field* fields,
                  int nfields, const double* linelocs, int64_t nlinelocs, int coord_f32, double smoothing,
                  double threshold, const ld_level_params* level, float* out_f32, uint16_t* out_u16,
                  int64_t out_total, void* workspace, size_t ws_bytes, void* stream) {
  if (!h || !demod || !fields || !linelocs || nfields <= 0 || !workspace) {
    set_error("ld_tbc_fields: null argument or empty batch");
    return LD_EINVAL;
  }
  for (int f = 0; f < nfields; ++f) {
    const ld_tbc_field& fd = fields[f];
    if (fd.nlines < 2 || fd.nlines > kLinesMax || fd.outwidth <= 0 || fd.outwidth > kTbcMaxWidth ||
        fd.linesout <= 0 || fd.linesout > kMaxLinesOut || fd.lineoffset < 0 || fd.lineoffset > kLinesMax ||
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
  // field descriptors and each field's first line-record index, one upload
  std::vector<char> hdr(tbc_fields_bytes(nfields), 0);
  std::memcpy(hdr.data(), fields, sizeof(ld_tbc_field) * (size_t)nfields);
  int* base = (int*)(hdr.data() + sizeof(ld_tbc_field) * (size_t)nfields);
  base[0] = 0;
  for (int f = 0; f < nfields; ++f) base[f + 1] = base[f] + fields[f].linesout;
  const int total_lines = base[nfields];
  ld_tbc_field* dfields = (ld_tbc_field*)workspace;
  const int* dbase = (const int*)((char*)workspace + sizeof(ld_tbc_field) * (size_t)nfields);
  LineRow* rows = (LineRow*)((char*)workspace + tbc_fields_bytes(nfields));
  LineRec* recs = (LineRec*)((char*)rows + sizeof(LineRow) * (size_t)nfields);
  // a pageable source is staged before hipMemcpyAsync returns, so the
  // temporary host buffer may go out of scope
  LD_HIP(hipMemcpyAsync(dfields, hdr.data(), hdr.size(), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_tbc_lines, dim3(nfields), dim3(kLineThreads), 0, st, dfields, linelocs, rows, smoothing,
                     threshold, recs, dbase);
  LD_HIP(hipGetLastError());
  TbcPixelArgs a;
  a.recs = recs;
  a.demod = demod;
  a.linelocs = linelocs;
  a.rows = rows;
  a.lut = h->sinc_lut;
  a.lut_folded = h->sinc_folded;
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
  int max_width = 0;
  for (int f = 0; f < nfields; ++f) max_width = fields[f].outwidth > max_width ? fields[f].outwidth : max_width;
  a.chunks = (max_width + kTbcThreads - 1) / kTbcThreads;
  hipLaunchKernelGGL(k_tbc_pixels, dim3(total_lines * a.chunks), dim3(kTbcThreads), 0, st, a);
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
