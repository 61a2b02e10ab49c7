// Line-location refinements after the hsync refine, batched over fields:
//   NTSC  FieldNTSC.refine_linelocs_burst (core.py:3391-3440) with
//         compute_burst_offsets (:3361-3388), compute_line_bursts
//         (:2999-3047) and clb_findbursts (utils.py:1339-1368), then
//         fix_badlines (:2566-2589) and the 117.25 degree subcarrier shift
//         (:3491-3495)
//   PAL   FieldPAL.refine_linelocs_pilot (core.py:3054-3093) twice, then
//         fix_badlines (:3221-3226)
// Kernels:
//   k_burst_scan    8 lanes per (field, line), 8 lines per wavefront: burst
//                   area statistics, the zero-crossing scan, the 16 cycle
//                   positions and the step table over them
//   k_burst_chain   one workgroup per chain walks its fields in order: each
//                   field's phase-adjust median seeds the next field
//                   (core.py:3366-3369); lines in parallel within a field
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
#include "npstat.hpp"

#pragma clang fp contract(off)

namespace ldgpu {

constexpr int kBurstLines = 266;  // compute_burst_offsets line range
constexpr int kPilotLines = 323;  // refine_linelocs_pilot line range
constexpr int kRefThreads = 512;  // lines per field handled by a workgroup
constexpr int kZc = 16;           // clb_findbursts slots

// Per (field, line) burst data.  The two clb_findbursts passes depend on
// their starting phase_adjust p0 only through which integer each slot's
// cycle position rounds to, so the result is a step function of p0 whose
// steps sit where some base[k] + p0 + 0.5 is an integer, and it moves by q
// when p0 moves by an integer q (the rounded integers all move by q; the
// rising count then flips when q is odd).  k_burst_scan tabulates one
// period: bp[] are the 16 step positions in [0, 1) sorted, tab[i] / rc[i]
// the final phase_adjust and rising count on the i-th interval between them.
struct BurstLine {
  double base[kZc];    // (bstart + zcs[k] - s_rem) / zcburstdiv, all 16 slots
  double bp[kZc];      // sorted step positions in [0, 1)
  double tab[kZc + 1]; // phase_adjust after both passes, per interval
  double adj;          // phase_adjust / 2 (k_burst_chain)
  double minbase;      // smallest base[k]
  unsigned char rc[kZc + 1];
  unsigned char pad_[7];
  int count;           // zero crossings found
  int rising;          // bit k: isrising[k]
  int valid;           // compute_line_bursts returned a value
  int rise;            // its `rising` result (k_burst_chain)
};

struct FieldState {
  int field14;
  int nvalid;
};

// ---------------------------------------------------------------- helpers
// numba's np.mean of a float32 array: float32 accumulation in order, the
// division in float64.
template <class F>
__device__ double nb_mean_f32(F x, int n) {
  float c = 0.0f;
  for (int i = 0; i < n; ++i) c = c + x(i);
  return (double)c / (double)n;
}

// The same over a 16-byte aligned LDS array: four values per read, the
// additions still strictly in order.
__device__ __forceinline__ double nb_mean_f32_lds(const float* x, int n) {
  float c = 0.0f;
  int i = 0;
#pragma unroll 4
  for (; i + 4 <= n; i += 4) {
    const float4 q = *reinterpret_cast<const float4*>(x + i);
    c = c + q.x;
    c = c + q.y;
    c = c + q.z;
    c = c + q.w;
  }
  for (; i < n; ++i) c = c + x[i];
  return (double)c / (double)n;
}

// numpy's recursion for n <= 256 (two levels at most).
template <class F>
__device__ float np_pairwise_sum(F x, int n) {
  if (n <= 128) return np_block_sum(x, 0, n);
  int n2 = n / 2;
  n2 -= n2 % 8;
  const int n3 = n - n2;
  const float left = np_block_sum(x, 0, n2);
  float right;
  if (n3 <= 128) {
    right = np_block_sum(x, n2, n3);
  } else {
    int m2 = n3 / 2;
    m2 -= m2 % 8;
    right = np_block_sum(x, n2, m2) + np_block_sum(x, n2 + m2, n3 - m2);
  }
  return left + right;
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

// Bitonic network sort of 16 values (ascending).  Each compare-exchange is
// a min and a max (two FP64 instructions); for values without NaN or -0.0,
// which is what both callers sort, the result is the same sequence as a
// compare-and-swap network's.
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
          const double lo = fmin(v[i], v[l]), hi = fmax(v[i], v[l]);
          v[i] = up ? lo : hi;
          v[l] = up ? hi : lo;
        }
      }
    }
  }
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
  // NTSC burst chains: chain c is fields [chain_first[c], chain_first[c+1])
  // and its first field's search starts from chain_seed[c]
  const int* chain_first;
  const double* chain_seed;
  int nchains;
};

// ---------------------------------------------------------------- NTSC
// Burst area of line l (compute_line_bursts, core.py:2999-3047): on return
// s0 is its first sample and n its length (0: no area), bstart / s_rem /
// zcburstdiv as the reference computes them.
struct BurstArea {
  long long s0, bstart;
  int n;
  double s_rem, zcburstdiv;
};

__device__ __forceinline__ bool burst_area(const RefineArgs& a, const ld_line_field& fd, const double* ll, int l,
                                           BurstArea& ba) {
  const double loc = ll[l];
  if (!(fabs(loc) < 9.0e15)) return false;
  const long long s = (long long)loc;
  ba.s_rem = loc - (double)s;
  const double lfreq = a.spl * get_linelen(ll, l, fd.linecount, a.linelen);
  const double fsc_inv = 1 / a.fsc_mhz;
  ba.bstart = (long long)((21 * fsc_inv) * lfreq);
  const long long bend = (long long)((28 * fsc_inv) * lfreq);
  long long s1;
  np_slice(s + ba.bstart, s + bend, fd.demod_len, ba.s0, s1);
  ba.n = (int)(s1 - ba.s0);
  ba.zcburstdiv = (lfreq * fsc_inv) / 2;
  return ba.n > 0;
}

// The two clb_findbursts passes of one line from its cycle bases (all 16
// slots), starting at phase_adjust = pa; returns the final phase_adjust and
// the rising count of the second pass.
__device__ double burst_passes(const double (&base)[kZc], int count, int rmask, double pa, int& rising_count) {
  rising_count = 0;
  if (count == 0) return pa;
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    double d[kZc];
    int odd = 0;
#pragma unroll
    for (int k = 0; k < kZc; ++k) {
      const double cyc = base[k] + pa;
      const double rs = cyc + .5;
      const int r = (rs > -2147483648.0 && rs < 2147483648.0) ? (int)rs : 0;  // astype(np.int32)
      d[k] = (double)r - cyc;
      odd |= ((r % 2) != 0) << k;
    }
    sort16(d);
    pa += (d[7] + d[8]) / 2.0;
    if (pass == 1) rising_count = __builtin_popcount((unsigned)(rmask ^ odd) & 0xffffu);
  }
  return pa;
}

constexpr int kAreaMax = 256;              // burst-area samples a line may have
constexpr int kScanLines = 8;              // lines per wavefront of k_burst_scan
constexpr int kLaneGroup = 64 / kScanLines; // lanes per line
constexpr int kAreaPitch = kAreaMax + 4;   // LDS row of one line's area (rows 16-byte aligned, 4 banks apart)

// compute_line_bursts (core.py:2999-3047) and the burst table (see
// BurstLine), kScanLines lines per wavefront with kLaneGroup lanes per
// line.  The area statistics and the zero-crossing walk are sequential per
// line (the float32 sums keep the reference's order), so every lane of a
// line's group runs them on the line's area staged in LDS; the 17 table
// intervals are then split over the group's lanes.
__global__ __launch_bounds__(64) void k_burst_scan(RefineArgs a, const float* __restrict__ demod,
                                                   const float* __restrict__ dburst) {
  __shared__ __align__(16) float sb[kScanLines][kAreaPitch];
  __shared__ __align__(16) float sd[kScanLines][kAreaPitch];
  __shared__ double sbp[kScanLines][kZc];
  const int lane = threadIdx.x;
  const int q = lane % kLaneGroup, ln = lane / kLaneGroup;
  const int l = blockIdx.x * kScanLines + ln;
  const int f = blockIdx.y;
  const ld_line_field fd = a.fields[f];
  const double* ll = a.ll2 + fd.lineloc_offset;
  const bool inrange = l < kBurstLines;
  BurstLine& out = a.lines[(size_t)f * kBurstLines + (inrange ? l : 0)];
  if (inrange && q == 0) {
    out.valid = 0;
    out.count = 0;
    out.rising = 0;
  }
  BurstArea ba{};
  const bool has = inrange && burst_area(a, fd, ll, l, ba) && ba.n <= kAreaMax;
  const int n = has ? ba.n : 0;
  // stage the line's area: lane q of the group loads samples q + 8k
  float* b = sb[ln];
  float* d = sd[ln];
  {
    const float* pb = dburst + fd.demod_offset + (has ? ba.s0 : 0);
    const float* pd = demod + fd.demod_offset + (has ? ba.s0 : 0);
    for (int k0 = q; k0 < n; k0 += 8 * kLaneGroup) {
      float vb[8], vd[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = k0 + u * kLaneGroup;
        vb[u] = i < n ? pb[i] : 0.0f;
        vd[u] = i < n ? pd[i] : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = k0 + u * kLaneGroup;
        if (i < n) {
          b[i] = vb[u];
          d[i] = vd[u];
        }
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (!has) return;
  // burstarea - nb_mean(burstarea), rounded to float32 (core.py:3022)
  const float mb = (float)nb_mean_f32_lds(b, n);
  auto bf = [&](int i) { return b[i] - mb; };
  const float m2 = np_mean_f32(bf, n);
  const float thr = sqrtf(np_mean_f32([&](int i) {
    const float e = bf(i) - m2;
    return e * e;
  }, n));
  const float md = (float)nb_mean_f32_lds(d, n);
  float amax = 0.0f;
  for (int i = 0; i < n; ++i) amax = fmaxf(amax, fabsf(d[i] - md));
  if ((double)amax > 30 * a.hz_ire) return;
  // clb_findbursts' walk: from each sample above the threshold, the next
  // crossing of 0 in its direction within 16 samples
  float zcs[kZc];
#pragma unroll
  for (int k = 0; k < kZc; ++k) zcs[k] = 0.0f;
  int count = 0, rising = 0;
  int j = 0;
  const int end = n - 1;
  while (j < end && count < kZc) {
    while (j < end && !(fabsf(bf(j)) > thr)) ++j;
    if (j >= end) break;
    double zc;
    if (!calczc0(bf, n, j, zc)) break;
    if (bf(j) < 0.0f) rising |= 1 << count;
#pragma unroll
    for (int k = 0; k < kZc; ++k)
      if (k == count) zcs[k] = (float)zc;  // zcs is a float32 array in the reference
    ++count;
    j = (int)zc + 1;
  }
  double base[kZc], t[kZc];
#pragma unroll
  for (int k = 0; k < kZc; ++k) {
    base[k] = (((double)ba.bstart + (double)zcs[k]) - ba.s_rem) / ba.zcburstdiv;
    // slot k's step position frac(-0.5 - base[k])
    const double v = -0.5 - base[k];
    t[k] = v - floor(v);
  }
#pragma unroll
  for (int k = 0; k < kZc; ++k)
    if ((k % kLaneGroup) == q) out.base[k] = base[k];
  sort16(t);
  if (q == 0) {
#pragma unroll
    for (int k = 0; k < kZc; ++k) sbp[ln][k] = t[k];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // both passes at the middle of each interval between step positions
  for (int iv = q; iv <= kZc; iv += kLaneGroup) {
    const double lo = iv == 0 ? 0.0 : sbp[ln][iv - 1];
    const double hi = iv == kZc ? 1.0 : sbp[ln][iv];
    int rc;
    const double pa = burst_passes(base, count, rising, 0.5 * (lo + hi), rc);
    out.tab[iv] = pa;
    out.rc[iv] = (unsigned char)rc;
    if (iv < kZc) out.bp[iv] = hi;
  }
  if (q == 0) {
    double mb2 = base[0];
#pragma unroll
    for (int k = 1; k < kZc; ++k) mb2 = fmin(mb2, base[k]);
    out.minbase = mb2;
    out.count = count;
    out.rising = rising;
    out.valid = 1;
  }
}

// Order-preserving 64-bit key of a double with its lowest 9 mantissa bits
// replaced by a unique index: keys of distinct entries never tie, and two
// values order by index only when they differ in those 9 bits (about 1e-13
// relative), which leaves the selected median within that of numpy's.
__device__ __forceinline__ unsigned long long rank_key(double v, int idx) {
  unsigned long long u = __double_as_longlong(v);
  u = (u >> 63) ? ~u : (u | 0x8000000000000000ull);
  return (u & ~0x1ffull) | (unsigned long long)idx;
}

// Rank of key among keys[0, n) (n even; unused entries hold keys above
// every value): 128-bit LDS reads, one 64-bit compare per entry.
__device__ __forceinline__ int rank_in(const unsigned long long* keys, int n, unsigned long long key) {
  const ulonglong2* kv = reinterpret_cast<const ulonglong2*>(keys);
  int r = 0;
#pragma unroll 8
  for (int j2 = 0; j2 < n / 2; ++j2) {
    const ulonglong2 q = kv[j2];
    r += q.x < key;
    r += q.y < key;
  }
  return r;
}

// The r-th smallest of the wavefront's keys (kk[j] of every lane, unique):
// the largest t with fewer than r + 1 keys below it, built bit by bit from
// the top; the counts are ballots, so every step is a few scalar ops.
template <int J>
__device__ __forceinline__ unsigned long long wave_select(const unsigned long long (&kk)[J], int r) {
  unsigned long long ans = 0;
  for (int b = 63; b >= 0; --b) {
    const unsigned long long t = ans | (1ull << b);
    int c = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) c += __popcll(__ballot(kk[j] < t));
    if (c <= r) ans = t;
  }
  return ans;
}

constexpr int kChainJ = (kBurstLines + 63) / 64;   // keys per lane of the selecting wavefront

// One workgroup per chain walks its fields in order; line l = thread l.  Per field:
// both burst passes from the previous field's median, then the median of
// the line adjustments, selected by the first wavefront from the keys in
// LDS (two barriers; LDS double-buffered by field parity).  The next
// field's line data load while the median forms.
__global__ __launch_bounds__(kRefThreads) void k_burst_chain(RefineArgs a) {
  __shared__ __align__(16) unsigned long long keys[2][kRefThreads];
  __shared__ double adjs[2][kRefThreads];
  __shared__ int wcnt[2][kRefThreads / 64], wrise[2][kRefThreads / 64], wnan[2][kRefThreads / 64];
  __shared__ double mid[2][2];
  const int l = threadIdx.x, w = l >> 6, lane = l & 63;
  const int c = blockIdx.x;
  const int f0 = a.chain_first[c], f1 = a.chain_first[c + 1];
  double prev = a.chain_seed[c];
  double bp[kZc], tab[kZc + 1], minbase = 0.0;
  unsigned rcw[5];
  int count = 0, rmask = 0, valid = 0;
  auto load = [&](int f) {
    valid = 0;
    if (l < kBurstLines) {
      const BurstLine& bl = a.lines[(size_t)f * kBurstLines + l];
      valid = bl.valid;
      count = bl.count;
      rmask = bl.rising;
      minbase = bl.minbase;
#pragma unroll
      for (int k = 0; k < kZc; ++k) bp[k] = bl.bp[k];
#pragma unroll
      for (int k = 0; k <= kZc; ++k) tab[k] = bl.tab[k];
      const unsigned* rw = reinterpret_cast<const unsigned*>(bl.rc);
#pragma unroll
      for (int k = 0; k < 5; ++k) rcw[k] = rw[k];
    }
  };
  load(f0);
  for (int f = f0; f < f1; ++f) {
    const int p = f & 1;
    double adj = INFINITY;
    int rise = 0;
    const int v = valid;
    if (v) {
      const double p0 = -prev;
      double pa;
      int rc = 0;
      if (count == 0) {
        pa = p0;  // clb_findbursts leaves phase_adjust unchanged
      } else if (minbase > 0.0 && minbase + p0 >= 0.0) {
        // step-table lookup: p0 = q + t, t in [0, 1).  The table holds the
        // passes in the regime where int32 truncation equals floor; pass 2
        // starts from p0 + median with median > -0.5, so every rounding
        // term of both passes stays positive only if minbase + p0 >= 0.
        const double q = floor(p0);
        const double t = p0 - q;
        int idx = 0;
#pragma unroll
        for (int k = 0; k < kZc; ++k) idx += bp[k] <= t;
        double tv = 0.0;
        int rr = 0;
#pragma unroll
        for (int i = 0; i <= kZc; ++i)
          if (i == idx) {
            tv = tab[i];
            rr = (rcw[i >> 2] >> (8 * (i & 3))) & 0xff;
          }
        pa = tv + q;
        rc = (((long long)q) & 1) ? kZc - rr : rr;
      } else {
        const BurstLine& bl = a.lines[(size_t)f * kBurstLines + l];
        double base[kZc];
#pragma unroll
        for (int k = 0; k < kZc; ++k) base[k] = bl.base[k];
        pa = burst_passes(base, count, rmask, p0, rc);
      }
      rise = rc > (count / 2.0);
      adj = (-pa) / 2;
      BurstLine& bl = a.lines[(size_t)f * kBurstLines + l];
      bl.adj = adj;
      bl.rise = rise;
    }
    const unsigned long long key = v ? rank_key(adj, l) : ~0ull;
    keys[p][l] = key;
    adjs[p][l] = adj;
    const unsigned long long vb = __ballot(v);
    const unsigned long long rb = __ballot(v && rise && (l % 2) == 0);
    const unsigned long long nb = __ballot(v && !(adj == adj));
    if (lane == 0) {
      wcnt[p][w] = __popcll(vb);
      wrise[p][w] = __popcll(rb);
      wnan[p][w] = __popcll(nb);
    }
    if (f + 1 < f1) load(f + 1);
    __syncthreads();
    int m = 0, rs = 0, nn = 0;
#pragma unroll
    for (int i = 0; i < kRefThreads / 64; ++i) {
      m += wcnt[p][i];
      rs += wrise[p][i];
      nn += wnan[p][i];
    }
    if (w == 0 && m > 0 && nn == 0) {
      // ranks (m - 1) / 2 and m / 2 of the valid keys (invalid lines hold
      // the largest key and rank last)
      unsigned long long kk[kChainJ];
#pragma unroll
      for (int j = 0; j < kChainJ; ++j) {
        const int i = lane + 64 * j;
        kk[j] = i < kBurstLines ? keys[p][i] : ~0ull;
      }
      const unsigned long long k0 = wave_select(kk, (m - 1) / 2);
      unsigned long long k1 = k0;
      if ((m % 2) == 0) {
        unsigned long long lo = ~0ull;
#pragma unroll
        for (int j = 0; j < kChainJ; ++j)
          if (kk[j] > k0 && kk[j] < lo) lo = kk[j];
        for (int o = 32; o > 0; o >>= 1) {
          const unsigned long long q = __shfl_xor(lo, o);
          lo = q < lo ? q : lo;
        }
        k1 = lo;
      }
      if (lane == 0) {
        mid[p][0] = adjs[p][k0 & 0x1ff];
        mid[p][1] = adjs[p][k1 & 0x1ff];
      }
    }
    __syncthreads();
    // numpy's median (NaN if any value is NaN or none exists), times 2
    const double med = (m == 0 || nn > 0) ? (double)NAN : ((m % 2) ? mid[p][1] : (mid[p][0] + mid[p][1]) / 2.0);
    const double pm = med * 2;
    if (l == 0) {
      a.phase_out[f] = pm;
      a.state[f].field14 = rs > (m / 4);
      a.state[f].nvalid = m;
    }
    // the next field starts from this field's median (its own median is
    // still 0 when its lines run, core.py:3366-3369)
    prev = pm;
  }
}

__global__ __launch_bounds__(kRefThreads) void k_burst_finish(RefineArgs a) {
  __shared__ double ll[kRefThreads], ll2[kRefThreads], tmp[kRefThreads];
  __shared__ double sadj[kRefThreads];
  __shared__ __align__(16) unsigned long long skey[kRefThreads];
  __shared__ int bad[kRefThreads], qual[kRefThreads];
  __shared__ int wcnt[kRefThreads / 64];
  __shared__ double mid[2];
  const int f = blockIdx.x;
  const ld_line_field fd = a.fields[f];
  const int n = fd.nlines;
  const int nb = kBurstLines < n ? kBurstLines : n;
  const int l = threadIdx.x, w = l >> 6, lane = l & 63;
  const double* g2 = a.ll2 + fd.lineloc_offset;
  if (l < n) {
    ll2[l] = g2[l];
    bad[l] = a.bad_in[fd.lineloc_offset + l] != 0;
  }
  __syncthreads();
  // Field.process: linebad |= deriv error of linelocs2; lines without a
  // burst result are bad (core.py:3398-3400)
  const BurstLine* bl = a.lines + (size_t)f * kBurstLines;
  int b = 0;
  if (l < n) {
    b = bad[l] || deriv_error(ll2, l, n);
    if (l >= 1 && l < kBurstLines && !bl[l].valid) b = 1;
  }
  // sample adjustments of the good lines (core.py:3402-3411)
  int u = 0;
  double sv = INFINITY;
  if (l < nb && !(isnan(ll2[l]) || b) && bl[l].valid) {
    const double lfreq = a.spl * get_linelen(ll2, l, fd.linecount, a.linelen);
    sv = bl[l].adj * lfreq * (1 / a.fsc_mhz);
    u = 1;
  }
  __syncthreads();
  if (l < n) bad[l] = b;
  sadj[l] = sv;
  const unsigned long long key = u ? rank_key(sv, l) : ~0ull;
  skey[l] = key;
  const unsigned long long ub = __ballot(u);
  if (lane == 0) wcnt[w] = __popcll(ub);
  __syncthreads();
  int m = 0;
#pragma unroll
  for (int i = 0; i < kRefThreads / 64; ++i) m += wcnt[i];
  if (u) {
    const int r = rank_in(skey, kBurstLines, key);
    if (r == (m - 1) / 2) mid[0] = sv;
    if (r == m / 2) mid[1] = sv;
  }
  __syncthreads();
  const double amed = m == 0 ? (double)NAN : ((m % 2) ? mid[1] : (mid[0] + mid[1]) / 2.0);
  // each line takes its own adjustment when within 2 samples of the median,
  // else the last such adjustment before it, else the median (core.py:3413-3421)
  qual[l] = u && (sv - amed >= -2) && (sv - amed <= 2);
  __syncthreads();
  double v = 0.0;
  if (l < n) {
    v = ll2[l];
    if (m > 0 && l < nb) {
      double add;
      if (qual[l]) {
        add = sadj[l];
      } else {
        int j = l - 1;
        while (j >= 0 && !qual[j]) --j;
        add = j >= 0 ? sadj[j] : amed;
      }
      v += add;
    }
  }
  if (l == 0) {
    if (m > 0) {
      const int f14 = a.state[f].field14;
      const int first = fd.first_field != 0;
      a.fpid_out[f] = first ? (f14 ? 1 : 3) : (f14 ? 4 : 2);
    } else {
      a.fpid_out[f] = 1;
    }
  }
  if (l < n) ll[l] = v;
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
         align256(sizeof(int) * ((size_t)nfields + 1)) + align256(sizeof(double) * (size_t)nfields) +
         align256(sizeof(BurstLine) * (size_t)nfields * kBurstLines) + 2 * align256(sizeof(double) * (size_t)nlinelocs);
}

static RefineArgs refine_args(ld_handle* h, const ld_line_params* p, int nfields, void* ws, int64_t nlinelocs) {
  RefineArgs a = {};
  char* w = (char*)ws;
  a.fields = (const ld_line_field*)w;
  w += align256(sizeof(ld_line_field) * (size_t)nfields);
  a.state = (FieldState*)w;
  w += align256(sizeof(FieldState) * (size_t)nfields);
  a.chain_first = (const int*)w;
  w += align256(sizeof(int) * ((size_t)nfields + 1));
  a.chain_seed = (const double*)w;
  w += align256(sizeof(double) * (size_t)nfields);
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

int ld_refine_burst_ntsc_chains(ld_handle* h, const float* demod, const float* demod_burst, int64_t demod_total,
                                const ld_line_field* fields, int nfields, const ld_line_params* params,
                                const double* linelocs2, const uint8_t* linebad_in, int64_t nlinelocs,
                                const double* chain_seeds, int nchains, double* linelocs_out, uint8_t* linebad_out,
                                double* phase_adjust_out, int32_t* field_phase_id, void* workspace, size_t ws_bytes,
                                void* stream) {
  if (!h || !demod || !demod_burst || !params || !linelocs2 || !linebad_in || !linelocs_out || !linebad_out ||
      !phase_adjust_out || !field_phase_id || !workspace || !chain_seeds) {
    set_error("ld_refine_burst_ntsc: null argument");
    return LD_EINVAL;
  }
  int rc = check_line_fields(fields, nfields, demod_total, nlinelocs, kBurstLines + 1, "ld_refine_burst_ntsc");
  if (rc != LD_OK) return rc;
  if (!(params->linelen > 0) || !(params->fsc_mhz > 0) || !(params->freq_mhz > 0)) {
    set_error("ld_refine_burst_ntsc: bad line parameters");
    return LD_EINVAL;
  }
  // chains: consecutive runs of fields with chain ids 0, 1, .., nchains-1
  std::vector<int> first(1, 0);
  for (int f = 0; f < nfields; ++f) {
    const int c = fields[f].chain, prevc = f ? fields[f - 1].chain : 0;
    if (c != prevc && c != prevc + 1) {
      set_error("ld_refine_burst_ntsc: field %d: chain %d after chain %d (chains must be consecutive runs "
                "numbered from 0)", f, c, prevc);
      return LD_EINVAL;
    }
    if (c == prevc + 1) first.push_back(f);
  }
  first.push_back(nfields);
  if ((int)first.size() - 1 != nchains || fields[0].chain != 0) {
    set_error("ld_refine_burst_ntsc: the fields hold %d chains, %d seeds given", (int)first.size() - 1, nchains);
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
  a.nchains = nchains;
  // pageable sources are staged before hipMemcpyAsync returns
  LD_HIP(hipMemcpyAsync((void*)a.fields, fields, sizeof(ld_line_field) * (size_t)nfields, hipMemcpyHostToDevice,
                        st));
  LD_HIP(hipMemcpyAsync((void*)a.chain_first, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice, st));
  LD_HIP(hipMemcpyAsync((void*)a.chain_seed, chain_seeds, sizeof(double) * (size_t)nchains, hipMemcpyHostToDevice,
                        st));
  hipLaunchKernelGGL(k_burst_scan, dim3((kBurstLines + kScanLines - 1) / kScanLines, nfields), dim3(64), 0, st, a,
                     demod, demod_burst);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_burst_chain, dim3(nchains), dim3(kRefThreads), 0, st, a);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_burst_finish, dim3(nfields), dim3(kRefThreads), 0, st, a);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

int ld_refine_burst_ntsc(ld_handle* h, const float* demod, const float* demod_burst, int64_t demod_total,
                         const ld_line_field* fields, int nfields, const ld_line_params* params,
                         const double* linelocs2, const uint8_t* linebad_in, int64_t nlinelocs,
                         double prev_phase_adjust, double* linelocs_out, uint8_t* linebad_out,
                         double* phase_adjust_out, int32_t* field_phase_id, void* workspace, size_t ws_bytes,
                         void* stream) {
  return ld_refine_burst_ntsc_chains(h, demod, demod_burst, demod_total, fields, nfields, params, linelocs2,
                                     linebad_in, nlinelocs, &prev_phase_adjust, 1, linelocs_out, linebad_out,
                                     phase_adjust_out, field_phase_id, workspace, ws_bytes, stream);
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
