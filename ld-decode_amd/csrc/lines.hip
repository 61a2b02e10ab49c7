// Line-location refinements after the hsync refine, batched over fields:
//   NTSC  FieldNTSC.refine_linelocs_burst (core.py:3391-3440) with
//         compute_burst_offsets (:3361-3388), compute_line_bursts
//         (:2999-3047) and clb_findbursts (utils.py:1339-1368), then
//         fix_badlines (:2566-2589) and the 117.25 degree subcarrier shift
//         (:3491-3495)
//   PAL   FieldPAL.refine_linelocs_pilot (core.py:3054-3093) twice, then
//         fix_badlines (:3221-3226)
// Kernels:
//   k_burst_scan    one thread per (field, line): burst area statistics and
//                   the zero-crossing scan, reduced to 16 cycle positions
//   k_burst_chain   one workgroup walks the fields in order: each field's
//                   phase-adjust median seeds the next field (core.py:3366-
//                   3369); lines in parallel within a field
//   k_burst_finish  one workgroup per field: sample adjustments, their
//                   median and outlier fallback, fix_badlines, shift
//   k_pilot_pass    one workgroup per field: one pilot alignment pass
//   k_pilot_finish  one workgroup per field: fix_badlines
// Positions are FP64; the demod data are float32 and handled with the
// reference's float32 statistics (numba's sequential float32 mean, numpy's
// pairwise float32 sums).
#include <math.h>

#include <cstring>
#include <vector>

#include "internal.hpp"

#pragma clang fp contract(off)

namespace ldgpu {

constexpr int kBurstLines = 266;  // compute_burst_offsets line range
constexpr int kPilotLines = 323;  // refine_linelocs_pilot line range
constexpr int kRefThreads = 512;  // lines per field handled by a workgroup
constexpr int kZc = 16;           // clb_findbursts slots

struct BurstLine {
  double base[kZc];  // (bstart + zcs[k] - s_rem) / zcburstdiv, all 16 slots
  double adj;        // phase_adjust / 2 (k_burst_chain)
  int count;         // zero crossings found
  int rising;        // bit k: isrising[k]
  int valid;         // compute_line_bursts returned a value
  int rise;          // its `rising` result (k_burst_chain)
};

struct FieldState {
  int field14;
  int nvalid;
};

// ---------------------------------------------------------------- helpers
// Python/numpy slice bounds of [a, b) over n elements.
__device__ __forceinline__ void np_slice(long long a, long long b, long long n, long long& s, long long& e) {
  if (a < 0) a = (a + n < 0) ? 0 : a + n;
  if (a > n) a = n;
  if (b < 0) b = (b + n < 0) ? 0 : b + n;
  if (b > n) b = n;
  s = a;
  e = b > a ? b : a;
}

// numba's np.mean of a float32 array: float32 accumulation in order, the
// division in float64.
template <class F>
__device__ double nb_mean_f32(F x, int n) {
  float c = 0.0f;
  for (int i = 0; i < n; ++i) c = c + x(i);
  return (double)c / (double)n;
}

// numpy's pairwise float32 summation (8 accumulators per block of at most
// 128, halves rounded down to a multiple of 8 above that).
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

template <class F>
__device__ float np_pairwise_sum(F x, int n) {
  if (n <= 128) return np_block_sum(x, 0, n);
  // explicit stack over the recursion of numpy's pairwise_sum
  int so[24], sn[24], state[24];
  float part[24];
  int sp = 0;
  so[0] = 0;
  sn[0] = n;
  state[0] = 0;
  float ret = 0.0f;
  while (sp >= 0) {
    const int o = so[sp], m = sn[sp];
    if (m <= 128) {
      const float v = np_block_sum(x, o, m);
      --sp;
      if (sp < 0) {
        ret = v;
        break;
      }
      if (state[sp] == 1) {
        part[sp] = v;
        state[sp] = 2;
        int n2 = sn[sp] / 2;
        n2 -= n2 % 8;
        ++sp;
        so[sp] = so[sp - 1] + n2;
        sn[sp] = sn[sp - 1] - n2;
        state[sp] = 0;
      } else {
        // second half done: combine and pop until a parent needs its second half
        float acc = part[sp] + v;
        --sp;
        while (sp >= 0 && state[sp] == 2) {
          acc = part[sp] + acc;
          --sp;
        }
        if (sp < 0) {
          ret = acc;
          break;
        }
        part[sp] = acc;
        state[sp] = 2;
        int n2 = sn[sp] / 2;
        n2 -= n2 % 8;
        ++sp;
        so[sp] = so[sp - 1] + n2;
        sn[sp] = sn[sp - 1] - n2;
        state[sp] = 0;
      }
    } else {
      int n2 = m / 2;
      n2 -= n2 % 8;
      state[sp] = 1;
      ++sp;
      so[sp] = o;
      sn[sp] = n2;
      state[sp] = 0;
    }
  }
  return ret;
}

// np.mean of a float32 array: pairwise float32 sum, float32 division.
template <class F>
__device__ float np_mean_f32(F x, int n) {
  return np_pairwise_sum(x, n) / (float)n;
}

// calczc_do (utils.py:812-841) with target 0, edge 0, count 16 over x[0, n):
// the first crossing in [start, start + 17), linearly interpolated.
template <class F>
__device__ bool calczc0(F x, int n, int start, double& zc) {
  const int stop = start + 17 < n ? start + 17 : n;
  const bool rising = x(start) < 0.0f;
  for (int i = start + 1; i < stop; ++i) {
    const float p = x(i - 1), c = x(i);
    if (rising ? (p < 0.0f && c >= 0.0f) : (p > 0.0f && c <= 0.0f)) {
      const double a = (double)p, b = (double)c;
      const double y = (b - a != 0.0) ? -a / (-a + b) : 0.0;
      zc = (double)(i - 1) + y;
      return true;
    }
  }
  return false;
}

// Field.get_linelen (core.py:1656-1679) with lineoffset 0.
__device__ __forceinline__ double get_linelen(const double* ll, int line, int linecount, double linelen) {
  double length;
  if (line >= linecount)
    length = (ll[line] - ll[line - 1]) / 1;
  else if (line > 0)
    length = (ll[line + 1] - ll[line - 1]) / 2;
  else
    length = (ll[line + 1] - ll[line]) / 1;
  return length <= 0 ? linelen : length;
}

// compare-exchange network sort of 16 values (ascending)
__device__ __forceinline__ void sort16(double (&v)[kZc]) {
#pragma unroll
  for (int k = 2; k <= kZc; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < kZc; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          const double a = v[i], b = v[l];
          const bool sw = up ? (a > b) : (a < b);
          v[i] = sw ? b : a;
          v[l] = sw ? a : b;
        }
      }
    }
  }
}

// Median of the valid entries of val[0, n) (flag[i] != 0) held in LDS,
// numpy's definition (mean of the two middle values for an even count);
// NaN when none is valid.  All threads of the block call it.
__device__ double block_median(const double* val, const int* flag, int n, int* cnt, double* mid) {
  const int t = threadIdx.x;
  if (t == 0) *cnt = 0;
  __syncthreads();
  if (t < n && flag[t]) atomicAdd(cnt, 1);
  __syncthreads();
  const int m = *cnt;
  if (t < n && flag[t]) {
    const double v = val[t];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      if (!flag[j]) continue;
      const double w = val[j];
      r += (w < v) || (w == v && j < t);
    }
    if (r == (m - 1) / 2) mid[0] = v;
    if (r == m / 2) mid[1] = v;
  }
  __syncthreads();
  const double med = m == 0 ? (double)NAN : ((m % 2) ? mid[1] : (mid[0] + mid[1]) / 2.0);
  __syncthreads();
  return med;
}

// Field.compute_deriv_error (core.py:2553-2564) of line l.
__device__ __forceinline__ bool deriv_error(const double* ll, int l, int n) {
  bool e = false;
  if (l >= 1 && l + 1 < n) e |= fabs((ll[l + 1] - ll[l]) - (ll[l] - ll[l - 1])) > 4;
  if (l >= 2) e |= fabs((ll[l] - ll[l - 1]) - (ll[l - 1] - ll[l - 2])) > 4;
  return e;
}

// Field.fix_badlines (core.py:2566-2589) on one field held in LDS: ll
// (NaN already replaced from the backup), bad (deriv error included).
// Each bad line is interpolated between its nearest good neighbours.
__device__ void fix_badlines_block(double* ll, const int* bad, int n, int firstcheck, double* out) {
  const int l = threadIdx.x;
  if (l < n) {
    double v = ll[l];
    if (bad[l]) {
      int pg = l - 1;
      while (pg >= 0 && bad[pg]) --pg;
      int ng = l + 1;
      while (ng < n && bad[ng]) ++ng;
      if (pg >= firstcheck && ng < n) {
        const double gap = (ll[ng] - ll[pg]) / (double)(ng - pg);
        v = (gap * (double)(l - pg)) + ll[pg];
      }
    }
    out[l] = v;
  }
}

struct RefineArgs {
  const ld_line_field* fields;
  int nfields;
  const double* ll2;
  const uint8_t* bad_in;
  double* ll_out;
  uint8_t* bad_out;
  double* phase_out;
  int* fpid_out;
  BurstLine* lines;
  FieldState* state;
  double freq, linelen, spl, fsc_mhz, pilot_mhz, hz_ire, shift, pilot_len;
  double prev_phase;
};

// ---------------------------------------------------------------- NTSC
__global__ __launch_bounds__(64) void k_burst_scan(RefineArgs a, const float* __restrict__ demod,
                                                   const float* __restrict__ dburst) {
  const int l = blockIdx.x * 64 + threadIdx.x;
  const int f = blockIdx.y;
  if (l >= kBurstLines) return;
  const ld_line_field fd = a.fields[f];
  const double* ll = a.ll2 + fd.lineloc_offset;
  BurstLine& out = a.lines[(size_t)f * kBurstLines + l];
  out.valid = 0;
  out.count = 0;
  out.rising = 0;
  const double loc = ll[l];
  if (!(fabs(loc) < 9.0e15)) return;
  const long long s = (long long)loc;
  const double s_rem = loc - (double)s;
  const double lfreq = a.spl * get_linelen(ll, l, fd.linecount, a.linelen);
  const double fsc_inv = 1 / a.fsc_mhz;
  const long long bstart = (long long)((21 * fsc_inv) * lfreq);
  const long long bend = (long long)((28 * fsc_inv) * lfreq);
  long long s0, s1;
  np_slice(s + bstart, s + bend, fd.demod_len, s0, s1);
  const int n = (int)(s1 - s0);
  if (n == 0) return;
  const float* pb = dburst + fd.demod_offset + s0;
  const float* pd = demod + fd.demod_offset + s0;
  const float mb = (float)nb_mean_f32([&](int i) { return pb[i]; }, n);
  auto bf = [&](int i) { return pb[i] - mb; };
  const float m2 = np_mean_f32(bf, n);
  const float thr = sqrtf(np_mean_f32([&](int i) {
    const float d = bf(i) - m2;
    return d * d;
  }, n));
  const float md = (float)nb_mean_f32([&](int i) { return pd[i]; }, n);
  float amax = 0.0f;
  for (int i = 0; i < n; ++i) amax = fmaxf(amax, fabsf(pd[i] - md));
  if ((double)amax > 30 * a.hz_ire) return;
  const double zcburstdiv = (lfreq * fsc_inv) / 2;
  float zcs[kZc];
#pragma unroll
  for (int k = 0; k < kZc; ++k) zcs[k] = 0.0f;
  int count = 0, rising = 0;
  int j = 0;
  const int end = n - 1;
  while (j < end && count < kZc) {
    if (fabsf(bf(j)) > thr) {
      double zc;
      if (calczc0(bf, n, j, zc)) {
        if (bf(j) < 0.0f) rising |= 1 << count;
        // zcs is a float32 array in the reference
#pragma unroll
        for (int k = 0; k < kZc; ++k)
          if (k == count) zcs[k] = (float)zc;
        ++count;
        j = (int)zc + 1;
      } else {
        break;
      }
    } else {
      ++j;
    }
  }
#pragma unroll
  for (int k = 0; k < kZc; ++k) out.base[k] = (((double)bstart + (double)zcs[k]) - s_rem) / zcburstdiv;
  out.count = count;
  out.rising = rising;
  out.valid = 1;
}

// The two clb_findbursts passes of one line from its stored cycle bases,
// starting at phase_adjust = pa; returns the final phase_adjust and the
// rising count of the second pass.
__device__ double burst_passes(const BurstLine& bl, double pa, int& rising_count) {
  rising_count = 0;
  if (bl.count == 0) return pa;
  for (int pass = 0; pass < 2; ++pass) {
    double d[kZc];
    int odd = 0;
#pragma unroll
    for (int k = 0; k < kZc; ++k) {
      const double cyc = bl.base[k] + pa;
      const double rs = cyc + .5;
      const int r = (rs > -2147483648.0 && rs < 2147483648.0) ? (int)rs : 0;  // astype(np.int32)
      d[k] = (double)r - cyc;
      odd |= ((r % 2) != 0) << k;
    }
    sort16(d);
    pa += (d[7] + d[8]) / 2.0;
    if (pass == 1) rising_count = __builtin_popcount((unsigned)(bl.rising ^ odd) & 0xffffu);
  }
  return pa;
}

__global__ __launch_bounds__(kRefThreads) void k_burst_chain(RefineArgs a) {
  __shared__ double adj[kRefThreads];
  __shared__ int valid[kRefThreads];
  __shared__ int cnt, rsum;
  __shared__ double mid[2];
  const int l = threadIdx.x;
  double prev = a.prev_phase;
  for (int f = 0; f < a.nfields; ++f) {
    int v = 0, rise = 0;
    double ad = 0.0;
    if (l < kBurstLines) {
      BurstLine& bl = a.lines[(size_t)f * kBurstLines + l];
      v = bl.valid;
      if (v) {
        int rc;
        const double pa = burst_passes(bl, -prev, rc);
        rise = rc > (bl.count / 2.0);
        ad = (-pa) / 2;
        bl.adj = ad;
        bl.rise = rise;
      }
    }
    adj[l] = ad;
    valid[l] = v;
    if (l == 0) rsum = 0;
    __syncthreads();
    if (v && rise && (l % 2) == 0) atomicAdd(&rsum, 1);
    const double med = block_median(adj, valid, kBurstLines, &cnt, mid) * 2;
    if (l == 0) {
      a.phase_out[f] = med;
      a.state[f].field14 = rsum > (cnt / 4);
      a.state[f].nvalid = cnt;
    }
    // the next field starts from this field's median (its own median is
    // still 0 when its lines run, core.py:3366-3369)
    prev = med;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kRefThreads) void k_burst_finish(RefineArgs a) {
  __shared__ double ll[kRefThreads], ll2[kRefThreads], sadj[kRefThreads], tmp[kRefThreads];
  __shared__ int bad[kRefThreads], use[kRefThreads];
  __shared__ int cnt;
  __shared__ double mid[2];
  const int f = blockIdx.x;
  const ld_line_field fd = a.fields[f];
  const int n = fd.nlines;
  const int l = threadIdx.x;
  const double* g2 = a.ll2 + fd.lineloc_offset;
  if (l < n) {
    ll2[l] = g2[l];
    bad[l] = a.bad_in[fd.lineloc_offset + l] != 0;
  }
  __syncthreads();
  // Field.process: linebad |= deriv error of linelocs2
  int b = 0;
  if (l < n) b = bad[l] || deriv_error(ll2, l, n);
  __syncthreads();
  const BurstLine* bl = a.lines + (size_t)f * kBurstLines;
  if (l < n) {
    if (l >= 1 && l < kBurstLines && !bl[l].valid) b = 1;
    bad[l] = b;
  }
  // sample adjustments of the good lines (core.py:3405-3415)
  int u = 0;
  double s = 0.0;
  if (l < kBurstLines && l < n && !(isnan(ll2[l]) || b) && bl[l].valid) {
    const double lfreq = a.spl * get_linelen(ll2, l, fd.linecount, a.linelen);
    s = bl[l].adj * lfreq * (1 / a.fsc_mhz);
    u = 1;
  }
  sadj[l] = s;
  use[l] = u;
  __syncthreads();
  const double amed = block_median(sadj, use, kBurstLines < n ? kBurstLines : n, &cnt, mid);
  const int nadj = cnt;
  if (l < n) ll[l] = ll2[l];
  __syncthreads();
  if (l == 0) {
    if (nadj > 0) {
      double lastvalid = amed;
      const int m = kBurstLines < n ? kBurstLines : n;
      for (int k = 0; k < m; ++k) {
        if (use[k] && (sadj[k] - amed >= -2) && (sadj[k] - amed <= 2)) {
          ll[k] += sadj[k];
          lastvalid = sadj[k];
        } else {
          ll[k] += lastvalid;
        }
      }
      const int f14 = a.state[f].field14;
      const int first = fd.first_field != 0;
      a.fpid_out[f] = first ? (f14 ? 1 : 3) : (f14 ? 4 : 2);
    } else {
      a.fpid_out[f] = 1;
    }
  }
  __syncthreads();
  // fix_badlines(linelocs3, linelocs2) (core.py:2566-2589, :3491)
  int b2 = 0;
  if (l < n) b2 = bad[l] || deriv_error(ll, l, n);
  __syncthreads();
  if (l < n) {
    bad[l] = b2;
    if (isnan(ll[l])) ll[l] = ll2[l];
  }
  __syncthreads();
  fix_badlines_block(ll, bad, n, 1, tmp);
  if (l < n) {
    a.ll_out[fd.lineloc_offset + l] = tmp[l] - a.shift;
    a.bad_out[fd.lineloc_offset + l] = (uint8_t)bad[l];
  }
}

// ---------------------------------------------------------------- PAL
// One refine_linelocs_pilot pass (core.py:3054-3093): ll_in -> ll_out.
__global__ __launch_bounds__(kRefThreads) void k_pilot_pass(RefineArgs a, const float* __restrict__ dpilot,
                                                            const double* __restrict__ ll_in,
                                                            double* __restrict__ ll_out) {
  __shared__ double zc[kRefThreads];
  __shared__ int ok[kRefThreads];
  __shared__ double sre[kRefThreads], sim[kRefThreads];
  const int f = blockIdx.x;
  const ld_line_field fd = a.fields[f];
  const int n = fd.nlines;
  const double* ll = ll_in + fd.lineloc_offset;
  const int l = threadIdx.x;
  double plen = 0.0;
  int good = 0;
  double z = 0.0;
  if (l < kPilotLines) {
    double adjfreq = a.freq;
    if (l > 1) adjfreq /= (ll[l] - ll[l - 1]) / a.linelen;
    plen = (adjfreq / a.pilot_mhz) / 2;
    const double begin = ll[l] + 0.0;
    if (fabs(begin) < 9.0e15) {
      const long long st = (long long)begin;
      const long long sp = (long long)(begin + a.pilot_len + 1);
      long long s0, s1;
      np_slice(st, sp, fd.demod_len, s0, s1);
      const int m = (int)(s1 - s0);
      const double lsoffset = ll[l] - (double)st;
      const float* p = dpilot + fd.demod_offset + s0;
      if (m > 0) {
        int peak = 0;
        float best = fabsf(p[0]);
        for (int i = 1; i < m; ++i) {
          const float v = fabsf(p[i]);
          if (v > best) {
            best = v;
            peak = i;
          }
        }
        double zb;
        if (calczc0([&](int i) { return p[i]; }, m, peak, zb)) {
          z = (zb - lsoffset) / plen;
          good = 1;
        }
      }
    }
  }
  zc[l] = z;
  ok[l] = good;
  __syncthreads();
  // a line without a crossing takes the previous line's value (or 0)
  if (l < kPilotLines && !good) {
    int k = l - 1;
    while (k >= 0 && !ok[k]) --k;
    z = k >= 0 ? zc[k] : 0.0;
  }
  __syncthreads();
  zc[l] = z;
  // angular mean (utils.py:1288-1302, core.py:3081-3084)
  double re = 0.0, im = 0.0;
  if (l < kPilotLines) {
    const double x2 = z - (double)(int)z;
    double sn, cs;
    sincos(x2 * M_PI * 2, &sn, &cs);
    re = cs;
    im = sn;
  }
  sre[l] = re;
  sim[l] = im;
  __syncthreads();
  for (int o = kRefThreads / 2; o > 0; o >>= 1) {
    if (l < o) {
      sre[l] += sre[l + o];
      sim[l] += sim[l + o];
    }
    __syncthreads();
  }
  double am = atan2(sim[0] / kPilotLines, sre[0] / kPilotLines) / (M_PI * 2);
  if (am < 0) am = 1 + am;
  if (l < n) {
    double v = ll[l];
    if (l < kPilotLines) {
      double d = (z - floor(z)) - am;
      if (d < -0.5)
        d += 1;
      else if (d > 0.5)
        d -= 1;
      v += (d * plen) * 1;
    }
    ll_out[fd.lineloc_offset + l] = v;
  }
}

// fix_badlines after the pilot passes (core.py:3226); linebad carries the
// hsync result plus the deriv errors of linelocs2 (Field.process) and of
// the refined locations.
__global__ __launch_bounds__(kRefThreads) void k_pilot_finish(RefineArgs a, const double* __restrict__ ll3) {
  __shared__ double ll[kRefThreads], ll2[kRefThreads], tmp[kRefThreads];
  __shared__ int bad[kRefThreads];
  const int f = blockIdx.x;
  const ld_line_field fd = a.fields[f];
  const int n = fd.nlines;
  const int l = threadIdx.x;
  if (l < n) {
    ll2[l] = a.ll2[fd.lineloc_offset + l];
    ll[l] = ll3[fd.lineloc_offset + l];
  }
  __syncthreads();
  int b = 0;
  if (l < n) b = (a.bad_in[fd.lineloc_offset + l] != 0) || deriv_error(ll2, l, n) || deriv_error(ll, l, n);
  if (l < n) bad[l] = b;
  __syncthreads();
  fix_badlines_block(ll, bad, n, 0, tmp);
  if (l < n) {
    a.ll_out[fd.lineloc_offset + l] = tmp[l];
    a.bad_out[fd.lineloc_offset + l] = (uint8_t)bad[l];
  }
}

}  // namespace ldgpu

using namespace ldgpu;

// ---------------------------------------------------------------- host
static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

static int check_line_fields(const ld_line_field* fields, int nfields, int64_t demod_total, int64_t nlinelocs,
                             int min_lines, const char* who) {
  if (!fields || nfields <= 0) {
    set_error("%s: no fields", who);
    return LD_EINVAL;
  }
  for (int f = 0; f < nfields; ++f) {
    const ld_line_field& fd = fields[f];
    if (fd.nlines < min_lines || fd.nlines > kRefThreads || fd.lineloc_offset < 0 ||
        fd.lineloc_offset + fd.nlines > nlinelocs) {
      set_error("%s: field %d: %d line locations (need %d..%d) or range outside the arrays", who, f, fd.nlines,
                min_lines, kRefThreads);
      return LD_EINVAL;
    }
    if (fd.demod_offset < 0 || fd.demod_len < 0 || fd.demod_offset + fd.demod_len > demod_total) {
      set_error("%s: field %d: demod span outside the buffer", who, f);
      return LD_ERANGE;
    }
    if (fd.linecount < 1) {
      set_error("%s: field %d: linecount %d", who, f, fd.linecount);
      return LD_EINVAL;
    }
  }
  return LD_OK;
}

size_t ld_refine_workspace_bytes(const ld_handle* h, int nfields, int64_t nlinelocs) {
  if (!h || nfields <= 0 || nlinelocs < 0) return 0;
  return align256(sizeof(ld_line_field) * (size_t)nfields) + align256(sizeof(FieldState) * (size_t)nfields) +
         align256(sizeof(BurstLine) * (size_t)nfields * kBurstLines) + 2 * align256(sizeof(double) * (size_t)nlinelocs);
}

static RefineArgs refine_args(ld_handle* h, const ld_line_params* p, int nfields, void* ws, int64_t nlinelocs) {
  RefineArgs a = {};
  char* w = (char*)ws;
  a.fields = (const ld_line_field*)w;
  w += align256(sizeof(ld_line_field) * (size_t)nfields);
  a.state = (FieldState*)w;
  w += align256(sizeof(FieldState) * (size_t)nfields);
  a.lines = (BurstLine*)w;
  a.nfields = nfields;
  a.freq = p->freq_mhz;
  a.linelen = p->linelen;
  a.spl = a.freq / p->linelen;
  a.fsc_mhz = p->fsc_mhz;
  a.pilot_mhz = p->pilot_mhz;
  a.hz_ire = p->hz_ire;
  a.shift = (117.25 / 360) / p->fsc_mhz * a.freq;
  a.pilot_len = 6 * (a.spl * p->linelen);
  (void)nlinelocs;
  return a;
}

int ld_refine_burst_ntsc(ld_handle* h, const float* demod, const float* demod_burst, int64_t demod_total,
                         const ld_line_field* fields, int nfields, const ld_line_params* params,
                         const double* linelocs2, const uint8_t* linebad_in, int64_t nlinelocs,
                         double prev_phase_adjust, double* linelocs_out, uint8_t* linebad_out,
                         double* phase_adjust_out, int32_t* field_phase_id, void* workspace, size_t ws_bytes,
                         void* stream) {
  if (!h || !demod || !demod_burst || !params || !linelocs2 || !linebad_in || !linelocs_out || !linebad_out ||
      !phase_adjust_out || !field_phase_id || !workspace) {
    set_error("ld_refine_burst_ntsc: null argument");
    return LD_EINVAL;
  }
  int rc = check_line_fields(fields, nfields, demod_total, nlinelocs, kBurstLines + 1, "ld_refine_burst_ntsc");
  if (rc != LD_OK) return rc;
  if (!(params->linelen > 0) || !(params->fsc_mhz > 0) || !(params->freq_mhz > 0)) {
    set_error("ld_refine_burst_ntsc: bad line parameters");
    return LD_EINVAL;
  }
  if (ws_bytes < ld_refine_workspace_bytes(h, nfields, nlinelocs)) {
    set_error("ld_refine_burst_ntsc: workspace %zu bytes < %zu needed", ws_bytes,
              ld_refine_workspace_bytes(h, nfields, nlinelocs));
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  RefineArgs a = refine_args(h, params, nfields, workspace, nlinelocs);
  a.ll2 = linelocs2;
  a.bad_in = linebad_in;
  a.ll_out = linelocs_out;
  a.bad_out = linebad_out;
  a.phase_out = phase_adjust_out;
  a.fpid_out = field_phase_id;
  a.prev_phase = prev_phase_adjust;
  // a pageable source is staged before hipMemcpyAsync returns
  LD_HIP(hipMemcpyAsync((void*)a.fields, fields, sizeof(ld_line_field) * (size_t)nfields, hipMemcpyHostToDevice,
                        st));
  hipLaunchKernelGGL(k_burst_scan, dim3((kBurstLines + 63) / 64, nfields), dim3(64), 0, st, a, demod, demod_burst);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_burst_chain, dim3(1), dim3(kRefThreads), 0, st, a);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_burst_finish, dim3(nfields), dim3(kRefThreads), 0, st, a);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

int ld_refine_pilot_pal(ld_handle* h, const float* demod_pilot, int64_t demod_total, const ld_line_field* fields,
                        int nfields, const ld_line_params* params, const double* linelocs2,
                        const uint8_t* linebad_in, int64_t nlinelocs, double* linelocs_out, uint8_t* linebad_out,
                        void* workspace, size_t ws_bytes, void* stream) {
  if (!h || !demod_pilot || !params || !linelocs2 || !linebad_in || !linelocs_out || !linebad_out || !workspace) {
    set_error("ld_refine_pilot_pal: null argument");
    return LD_EINVAL;
  }
  int rc = check_line_fields(fields, nfields, demod_total, nlinelocs, kPilotLines, "ld_refine_pilot_pal");
  if (rc != LD_OK) return rc;
  if (!(params->linelen > 0) || !(params->pilot_mhz > 0) || !(params->freq_mhz > 0)) {
    set_error("ld_refine_pilot_pal: bad line parameters");
    return LD_EINVAL;
  }
  if (ws_bytes < ld_refine_workspace_bytes(h, nfields, nlinelocs)) {
    set_error("ld_refine_pilot_pal: workspace %zu bytes < %zu needed", ws_bytes,
              ld_refine_workspace_bytes(h, nfields, nlinelocs));
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  RefineArgs a = refine_args(h, params, nfields, workspace, nlinelocs);
  a.ll2 = linelocs2;
  a.bad_in = linebad_in;
  a.ll_out = linelocs_out;
  a.bad_out = linebad_out;
  char* w = (char*)a.lines + align256(sizeof(BurstLine) * (size_t)nfields * kBurstLines);
  double* p1 = (double*)w;
  double* p2 = (double*)(w + align256(sizeof(double) * (size_t)nlinelocs));
  LD_HIP(hipMemcpyAsync((void*)a.fields, fields, sizeof(ld_line_field) * (size_t)nfields, hipMemcpyHostToDevice,
                        st));
  hipLaunchKernelGGL(k_pilot_pass, dim3(nfields), dim3(kRefThreads), 0, st, a, demod_pilot, linelocs2, p1);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pilot_pass, dim3(nfields), dim3(kRefThreads), 0, st, a, demod_pilot, p1, p2);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pilot_finish, dim3(nfields), dim3(kRefThreads), 0, st, a, p2);
  LD_HIP(hipGetLastError());
  return LD_OK;
}
