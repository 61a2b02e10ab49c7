// GPU side of line location (Field.getpulses / skip_check, core.py:2122-2322;
// findpulses_numba_raw, utils.py:1101-1141):
//   k_edge_count / k_edge_scan / k_edge_write   findpulses over a whole
//       stitched buffer as an order-preserving stream compaction of the
//       threshold crossings: per-tile counts (and per-thread edge bits),
//       one exclusive scan, then each tile writes its edges at its offset
//   k_pulse_means   getpulses' vertical-sync levels: numpy's float32 mean of
//       the body of every long pulse, one wavefront per pulse, the samples
//       staged in LDS
//   k_range_means   the same mean over caller-listed ranges (black levels)
//   k_pulse_export  the edges found and their levels (not the whole
//       capacity) written straight into pinned host memory
//   k_field_facts   skip_check's line medians (radix select in LDS, one
//       workgroup per field line) and the final line location a field's
//       successor reads
// Edges and positions are int64; sample comparisons are made in double, as
// numba compares the float32 buffer with a float64 threshold.
#include <math.h>

#include "internal.hpp"
#include "npstat.hpp"

#pragma clang fp contract(off)

namespace ldgpu {

constexpr int kEdgeThreads = 256;
constexpr int kEdgePer = 16;                          // samples per thread
constexpr int kEdgeTile = kEdgeThreads * kEdgePer;    // samples per workgroup
constexpr int kScanThreads = 256;                     // small: the scan finds a CU beside other kernels
constexpr int kMeanMax = 5000;                        // findpulses' longest pulse
constexpr int kMeanStage = 1280;                      // pulses up to this long are staged in LDS
constexpr int kMeanPulses = 16;                       // pulses per wavefront of the fused mean kernel
constexpr long long kMeanWgMax = 8192;                // its workgroups per job
constexpr int kMeanWaves = 4;                         // pulses per workgroup
constexpr int kFactThreads = 256;
constexpr int kFactMax = 8192;                        // longest skip_check line slice

// Bit i of the result: sample base + i is "in pulse"; bit 16 (prev) for the
// sample before base.
__device__ __forceinline__ unsigned below_mask(const float* x, long long n, long long base, double high,
                                               unsigned& prev) {
  unsigned m = 0;
  if (base + kEdgePer <= n) {
    const float4* p = reinterpret_cast<const float4*>(x + base);
#pragma unroll
    for (int q = 0; q < kEdgePer / 4; ++q) {
      const float4 v = p[q];
      m |= ((double)v.x <= high ? 1u : 0u) << (4 * q);
      m |= ((double)v.y <= high ? 1u : 0u) << (4 * q + 1);
      m |= ((double)v.z <= high ? 1u : 0u) << (4 * q + 2);
      m |= ((double)v.w <= high ? 1u : 0u) << (4 * q + 3);
    }
  } else {
    for (int i = 0; i < kEdgePer; ++i)
      if (base + i < n && (double)x[base + i] <= high) m |= 1u << i;
  }
  prev = (base > 0 && base - 1 < n && (double)x[base - 1] <= high) ? 1u : 0u;
  return m;
}

// Edge bits among the thread's samples (valid samples only): a change of
// state against the previous sample; sample 0 counts as a start when it is
// in a pulse.
__device__ __forceinline__ unsigned edge_bits(unsigned m, unsigned prev, long long n, long long base) {
  unsigned e = m ^ ((m << 1) | prev);
  const long long left = n - base;
  if (left < kEdgePer) e &= left > 0 ? ((1u << left) - 1u) : 0u;
  return e & 0xffffu;
}

// Per tile: the edge count, and every thread's edge bits for k_edge_write
// (which then reads 2 bytes per 16 samples instead of the samples).
__global__ __launch_bounds__(kEdgeThreads) void k_edge_count(const float* __restrict__ x, long long n, double high,
                                                             long long* __restrict__ counts,
                                                             uint16_t* __restrict__ masks) {
  __shared__ int part[kEdgeThreads / 64];
  const int t = threadIdx.x;
  const long long base = (long long)blockIdx.x * kEdgeTile + (long long)t * kEdgePer;
  unsigned prev;
  const unsigned m = below_mask(x, n, base, high, prev);
  const unsigned e = edge_bits(m, prev, n, base);
  masks[(long long)blockIdx.x * kEdgeThreads + t] = (uint16_t)e;
  int c = __popc(e);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((t & 63) == 0) part[t >> 6] = c;
  __syncthreads();
  if (t == 0) {
    long long s = 0;
    for (int w = 0; w < kEdgeThreads / 64; ++w) s += part[w];
    counts[blockIdx.x] = s;
  }
}

// Exclusive scan of the per-tile counts in place; the total goes to *total.
// Each thread sums a run of consecutive counts (up to kScanPer of them held
// in registers, all loads in flight at once; longer runs four at a time and
// read again for the write), the wavefronts scan their runs with shuffles,
// and one wavefront scans the wavefront totals: two workgroup barriers.
constexpr int kScanPer = 32;
__device__ __forceinline__ void edge_scan_wg(long long* __restrict__ counts, long long ntiles,
                                             long long* __restrict__ total) {
  __shared__ long long wsum[kScanThreads / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long long per = (ntiles + kScanThreads - 1) / kScanThreads;
  const long long b = (long long)t * per;
  const long long e = b + per < ntiles ? b + per : ntiles;
  const bool held = per <= kScanPer;  // uniform
  long long c[kScanPer];
  long long s = 0;
  long long i = b;
  if (held) {
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) c[q] = b + q < e ? counts[b + q] : 0;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) s += c[q];
  } else {
    for (; i + 4 <= e; i += 4) {
      const long long c0 = counts[i], c1 = counts[i + 1], c2 = counts[i + 2], c3 = counts[i + 3];
      s += (c0 + c1) + (c2 + c3);
    }
    for (; i < e; ++i) s += counts[i];
  }
  long long inc = s;  // inclusive scan over the wavefront
  for (int o = 1; o < 64; o <<= 1) {
    const long long v = __shfl_up(inc, o);
    if (lane >= o) inc += v;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  if (w == 0) {
    long long v = lane < kScanThreads / 64 ? wsum[lane] : 0;
    for (int o = 1; o < kScanThreads / 64; o <<= 1) {
      const long long u = __shfl_up(v, o);
      if (lane >= o) v += u;
    }
    if (lane < kScanThreads / 64) wsum[lane] = v;  // inclusive wavefront offsets
  }
  __syncthreads();
  long long run = inc - s + (w > 0 ? wsum[w - 1] : 0);
  if (held) {
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) {
      if (b + q < e) counts[b + q] = run;
      run += c[q];
    }
  } else {
    for (i = b; i < e; ++i) {
      const long long ci = counts[i];
      counts[i] = run;
      run += ci;
    }
  }
  if (t == kScanThreads - 1) *total = wsum[kScanThreads / 64 - 1];
}
__global__ __launch_bounds__(kScanThreads) void k_edge_scan(long long* __restrict__ counts, long long ntiles,
                                                            long long* __restrict__ total) {
  edge_scan_wg(counts, ntiles, total);
}

__global__ __launch_bounds__(kEdgeThreads) void k_edge_write(const uint16_t* __restrict__ masks,
                                                             const long long* __restrict__ offsets,
                                                             long long* __restrict__ edges, long long cap) {
  __shared__ int part[kEdgeThreads];
  const int t = threadIdx.x;
  const long long base = (long long)blockIdx.x * kEdgeTile + (long long)t * kEdgePer;
  unsigned e = masks[(long long)blockIdx.x * kEdgeThreads + t];
  const int c = __popc(e);
  part[t] = c;
  __syncthreads();
  for (int o = 1; o < kEdgeThreads; o <<= 1) {
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long k = offsets[blockIdx.x] + part[t] - c;
  while (e) {
    const int i = __ffs(e) - 1;
    e &= e - 1;
    if (k < cap) edges[k] = base + i;
    ++k;
  }
}

// The same compaction from in-pulse bits, kBitGroups runs of 64 samples
// per thread (consecutive): a tile covers kBitTile samples.  Four runs per
// thread keep a wave's tiles few (about 2,000 per system for 1.3e8 samples),
// so the count and write launches are not bound by dispatching workgroups
// that have 2 KB to read each.  Word w holds samples 32 w .. 32 w + 31
// paired: bit j < 16 is sample 32 w + 2 j, bit 16 + j sample 32 w + 2 j + 1.
constexpr int kBitGroups = 4;
constexpr int kBitTile = kEdgeThreads * 64 * kBitGroups;

__device__ __forceinline__ uint32_t spread16(uint32_t x) {
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  return (x | (x << 1)) & 0x55555555u;
}
// sample-ordered bits of a paired word
__device__ __forceinline__ uint32_t unpair(uint32_t w) { return spread16(w & 0xffffu) | (spread16(w >> 16) << 1); }

__device__ __forceinline__ unsigned long long bit_edges(const uint32_t* __restrict__ bits, long long n,
                                                        long long base) {
  const long long w = base >> 5;
  const uint2 q = *reinterpret_cast<const uint2*>(bits + w);
  const unsigned long long m = (unsigned long long)unpair(q.x) | ((unsigned long long)unpair(q.y) << 32);
  // sample 32 w - 1 is the previous word's last odd sample: its bit 31
  const unsigned long long prev = base > 0 ? (bits[w - 1] >> 31) : 0ull;
  unsigned long long e = m ^ ((m << 1) | prev);
  const long long left = n - base;
  if (left < 64) e &= left > 0 ? ((1ull << left) - 1ull) : 0ull;
  return e;
}

__device__ __forceinline__ void bits_count_tile(const uint32_t* __restrict__ bits, long long n,
                                                long long* __restrict__ counts, long long tile) {
  __shared__ int part[kEdgeThreads / 64];
  const int t = threadIdx.x;
  const long long base = tile * kBitTile + (long long)t * (64 * kBitGroups);
  int c = 0;
#pragma unroll
  for (int g = 0; g < kBitGroups; ++g) {
    const long long bg = base + 64 * g;
    if (bg < n) c += __popcll(bit_edges(bits, n, bg));
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((t & 63) == 0) part[t >> 6] = c;
  __syncthreads();
  if (t == 0) {
    long long s = 0;
    for (int w = 0; w < kEdgeThreads / 64; ++w) s += part[w];
    counts[tile] = s;
  }
}

__device__ __forceinline__ void bits_write_tile(const uint32_t* __restrict__ bits, long long n,
                                                const long long* __restrict__ offsets, long long* __restrict__ edges,
                                                long long cap, long long tile) {
  __shared__ int part[kEdgeThreads];
  const int t = threadIdx.x;
  const long long base = tile * kBitTile + (long long)t * (64 * kBitGroups);
  unsigned long long e[kBitGroups];
  int c = 0;
#pragma unroll
  for (int g = 0; g < kBitGroups; ++g) {
    const long long bg = base + 64 * g;
    e[g] = bg < n ? bit_edges(bits, n, bg) : 0ull;
    c += __popcll(e[g]);
  }
  part[t] = c;
  __syncthreads();
  for (int o = 1; o < kEdgeThreads; o <<= 1) {
    const int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long k = offsets[tile] + part[t] - c;
#pragma unroll
  for (int g = 0; g < kBitGroups; ++g) {
    unsigned long long eg = e[g];
    while (eg) {
      const int i = __ffsll((long long)eg) - 1;
      eg &= eg - 1;
      if (k < cap) edges[k] = base + 64 * g + i;
      ++k;
    }
  }
}

__global__ __launch_bounds__(kEdgeThreads) void k_bits_count(const uint32_t* __restrict__ bits, long long n,
                                                             long long* __restrict__ counts) {
  bits_count_tile(bits, n, counts, blockIdx.x);
}

__global__ __launch_bounds__(kEdgeThreads) void k_bits_write(const uint32_t* __restrict__ bits, long long n,
                                                             const long long* __restrict__ offsets,
                                                             long long* __restrict__ edges, long long cap) {
  bits_write_tile(bits, n, offsets, edges, cap, blockIdx.x);
}

// numpy's float32 mean of an LDS-staged range of at most kMeanMax samples.
__device__ float lds_np_mean(const float* s, int m) {
  auto at = [s](int i) { return s[i]; };
  return np_pairwise_sum_any<6>(at, 0, m) / (float)m;
}

// numpy's pairwise float32 sum of s[0, m) (m <= kMeanMax) by one wavefront:
// the recursion's leaves (blocks of at most 128, at most 64 of them for
// m <= 5000) are listed in order, lane j sums leaf j with numpy's
// 8-accumulator block sum, and lane 0 adds the leaf sums up the recursion
// tree in numpy's order.
template <int D>
__device__ void np_leaves(int off, int n, int* loff, int* llen, int& cnt) {
  if constexpr (D == 0) {
    loff[cnt] = off;
    llen[cnt] = n;
    ++cnt;
  } else {
    if (n <= 128) {
      loff[cnt] = off;
      llen[cnt] = n;
      ++cnt;
      return;
    }
    int n2 = n / 2;
    n2 -= n2 % 8;
    np_leaves<D - 1>(off, n2, loff, llen, cnt);
    np_leaves<D - 1>(off + n2, n - n2, loff, llen, cnt);
  }
}
template <int D>
__device__ float np_tree(int n, const float* leaf, int& idx) {
  if constexpr (D == 0) {
    return leaf[idx++];
  } else {
    if (n <= 128) return leaf[idx++];
    int n2 = n / 2;
    n2 -= n2 % 8;
    const float left = np_tree<D - 1>(n2, leaf, idx);
    return left + np_tree<D - 1>(n - n2, leaf, idx);
  }
}

struct MeanScratch {
  float stage[kMeanStage];
  float leaf[64];
  int loff[64], llen[64];
};

// numpy's float32 mean of at(0 .. m-1) by the calling wavefront; the
// result is valid in lane 0.
template <class F>
__device__ float wave_np_mean_at(MeanScratch* sc, F at, int m, int lane) {
  if (lane == 0) {
    int cnt = 0;
    np_leaves<6>(0, m, sc->loff, sc->llen, cnt);
    for (int j = cnt; j < 64; ++j) sc->llen[j] = 0;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int n = sc->llen[lane];
  if (n > 0) sc->leaf[lane] = np_block_sum(at, sc->loff[lane], n);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float r = 0.0f;
  if (lane == 0) {
    int idx = 0;
    r = np_tree<6>(m, sc->leaf, idx) / (float)m;
  }
  return r;
}

// The mean of src[0, m) (m <= kMeanMax): staged in LDS when it fits (the
// leaf sums then read LDS), else the leaves read the samples in place.
__device__ float pulse_mean(MeanScratch* sc, const float* __restrict__ src, int m, int lane) {
  if (m <= kMeanStage) {
    for (int i = lane; i < m; i += 64) sc->stage[i] = src[i];
    const float* st = sc->stage;
    return wave_np_mean_at(sc, [st](int i) { return st[i]; }, m, lane);
  }
  return wave_np_mean_at(sc, [src](int i) { return src[i]; }, m, lane);
}

// 64 consecutive pulses per wavefront: lane i reads pulse k0 + i's edges and
// writes NaN unless the pulse is long; the long ones are then averaged one
// after the other by the whole wavefront.
__global__ __launch_bounds__(64 * kMeanWaves) void k_pulse_means(const float* __restrict__ x, long long n,
                                                                 const long long* __restrict__ edges,
                                                                 const long long* __restrict__ count,
                                                                 long long cap_pulses, double min_len, double trim,
                                                                 float* __restrict__ means) {
  __shared__ MeanScratch scratch[kMeanWaves];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  long long np = *count / 2;
  if (np > cap_pulses) np = cap_pulses;
  MeanScratch* sc = &scratch[w];
  const long long step = (long long)gridDim.x * kMeanWaves * 64;
  for (long long k0 = ((long long)blockIdx.x * kMeanWaves + w) * 64; k0 < np; k0 += step) {
    const long long k = k0 + lane;
    long long a = 0, b = 0;
    bool take = false;
    if (k < np) {
      const long long st = edges[2 * k], len = edges[2 * k + 1] - st;
      if ((double)len > min_len && len <= kMeanMax) {
        // int(start + trim) : int(start + len - trim), Python slice bounds
        np_slice((long long)((double)st + trim), (long long)((double)st + (double)len - trim), n, a, b);
        take = b - a > 0 && b - a <= kMeanMax;
      }
      if (!take) means[k] = __builtin_nanf("");
    }
    unsigned long long todo = __ballot(take);
    while (todo) {
      const int j = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const long long aj = __shfl(a, j), bj = __shfl(b, j);
      const float r = pulse_mean(sc, x + aj, (int)(bj - aj), lane);
      if (lane == 0) means[k0 + j] = r;
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// *hcount = *count; hedges[0, min(count, cap)) and hmeans[0, min(count, cap) / 2)
// from the device lists.  The host pointers are device addresses of pinned
// host memory; the host reads them after waiting for an event recorded
// behind this launch, whose completion releases the stores at system scope
// (no fence per workgroup: a system-scope fence writes back the L2).
__global__ __launch_bounds__(256) void k_pulse_export(const long long* __restrict__ edges,
                                                      const long long* __restrict__ count,
                                                      const float* __restrict__ means, long long cap,
                                                      long long* __restrict__ hedges, float* __restrict__ hmeans,
                                                      long long* __restrict__ hcount) {
  const long long total = *count;
  const long long n = total < cap ? total : cap;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long long i = t0; i < n; i += stride) hedges[i] = edges[i];
  for (long long i = t0; i < n / 2; i += stride) hmeans[i] = means[i];
  if (t0 == 0) *hcount = total;
}

// ------------------------------------------- all systems' searches at once
// The pulse search of several buffers (a batch's systems) in four launches
// instead of four or five per buffer: tile counts, one scan workgroup per
// buffer, edge compaction, then the sync levels together with the export of
// the list to pinned host memory.  Workgroup ranges per job come in the
// arguments.
constexpr int kMaxPulseJobs = 4;
struct PulseJobDev {
  const uint32_t* bits;
  const float* x;
  long long n;
  long long* edges;
  long long* count;
  float* means;
  long long cap;            // edges (even)
  double min_len, trim;
  long long* hedges;        // device addresses of the pinned host lists
  float* hmeans;
  long long* hcount;
  long long* tiles;         // per-tile counts (workspace)
  long long ntiles;
};
struct PulseSearchArgs {
  int njobs;
  PulseJobDev job[kMaxPulseJobs];
  long long tile0[kMaxPulseJobs + 1];   // first tile workgroup of each job
  long long mwg0[kMaxPulseJobs + 1];    // first means workgroup of each job
};

__device__ __forceinline__ int job_of(const long long* first, int njobs, long long b) {
  int j = 0;
  while (j + 1 < njobs && b >= first[j + 1]) ++j;
  return j;
}

__global__ __launch_bounds__(kEdgeThreads) void k_bits_count_multi(PulseSearchArgs a) {
  const int j = job_of(a.tile0, a.njobs, blockIdx.x);
  const PulseJobDev& d = a.job[j];
  bits_count_tile(d.bits, d.n, d.tiles, (long long)blockIdx.x - a.tile0[j]);
}

__global__ __launch_bounds__(kScanThreads) void k_edge_scan_multi(PulseSearchArgs a) {
  const PulseJobDev& d = a.job[blockIdx.x];
  edge_scan_wg(d.tiles, d.ntiles, d.count);
}

__global__ __launch_bounds__(kEdgeThreads) void k_bits_write_multi(PulseSearchArgs a) {
  const int j = job_of(a.tile0, a.njobs, blockIdx.x);
  const PulseJobDev& d = a.job[j];
  bits_write_tile(d.bits, d.n, d.tiles, d.edges, d.cap, (long long)blockIdx.x - a.tile0[j]);
}

// kMeanPulses pulses per one-wavefront workgroup and pass (a small LDS
// footprint and a short chain of means, so that the workgroups find room
// beside other kernels' and finish soon), as k_pulse_means, with each edge
// and level also written to the host lists: edges [0, min(count, cap)),
// levels of the pulses [0, min(count, cap) / 2), and the count.
__global__ __launch_bounds__(64) void k_means_export_multi(PulseSearchArgs a) {
  __shared__ MeanScratch scratch[1];
  const int j = job_of(a.mwg0, a.njobs, blockIdx.x);
  const PulseJobDev& d = a.job[j];
  const int lane = threadIdx.x;
  const long long total = *d.count;
  const long long ne = total < d.cap ? total : d.cap;
  const long long np = ne / 2;
  const long long wg = (long long)blockIdx.x - a.mwg0[j];
  if (wg == 0 && threadIdx.x == 0) *d.hcount = total;
  MeanScratch* sc = &scratch[0];
  const long long step = (a.mwg0[j + 1] - a.mwg0[j]) * kMeanPulses;
  for (long long k0 = wg * kMeanPulses; 2 * k0 < ne; k0 += step) {
    const long long k = k0 + lane;
    long long lo = 0, hi = 0;
    bool take = false;
    if (lane < kMeanPulses && 2 * k < ne) {
      const long long st = d.edges[2 * k];
      d.hedges[2 * k] = st;
      if (k < np) {
        const long long en = d.edges[2 * k + 1], len = en - st;
        d.hedges[2 * k + 1] = en;
        if ((double)len > d.min_len && len <= kMeanMax) {
          np_slice((long long)((double)st + d.trim), (long long)((double)st + (double)len - d.trim), d.n, lo, hi);
          take = hi - lo > 0 && hi - lo <= kMeanMax;
        }
        if (!take) {
          d.means[k] = __builtin_nanf("");
          d.hmeans[k] = __builtin_nanf("");
        }
      }
    }
    unsigned long long todo = __ballot(take);
    while (todo) {
      const int q = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const long long aq = __shfl(lo, q), bq = __shfl(hi, q);
      const float r = pulse_mean(sc, d.x + aq, (int)(bq - aq), lane);
      if (lane == 0) {
        d.means[k0 + q] = r;
        d.hmeans[k0 + q] = r;
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

__global__ __launch_bounds__(64) void k_range_means(const float* __restrict__ x, long long n,
                                                    const long long* __restrict__ begin,
                                                    const long long* __restrict__ end, int count,
                                                    float* __restrict__ out) {
  const int k = blockIdx.x;
  if (k >= count || threadIdx.x != 0) return;
  long long a, b;
  np_slice(begin[k], end[k], n, a, b);
  const long long m = b - a;
  float r = __builtin_nanf("");
  if (m > 0) {
    const float* p = x + a;
    auto at = [p](long long i) { return p[i]; };
    r = np_sum_f32(at, m) / (float)m;
  }
  out[k] = r;
}

struct FactArgs {
  const float* demod;
  const double* ll;
  ld_field_facts* facts;
  double line_samples;
  int outlinecount;
  int pal;
  int smax;  // samples the dynamic LDS holds: every slice's length
};

// Order-preserving unsigned key of a float (finite values).
__device__ __forceinline__ unsigned fkey(float v) {
  const unsigned u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// The k-th smallest (0-based) of the m keys in s[], by four 8-bit radix
// passes over LDS histograms; every thread of the workgroup returns it.
__device__ unsigned radix_select(const unsigned* s, int m, int k, unsigned* hist, unsigned* shared, int t) {
  unsigned prefix = 0, mask = 0;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[t] = 0;
    __syncthreads();
    for (int i = t; i < m; i += kFactThreads) {
      const unsigned v = s[i];
      if ((v & mask) == prefix) atomicAdd(&hist[(v >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (t < 64) {
      // each lane owns four bins; find the bin holding rank k
      const unsigned c0 = hist[4 * t], c1 = hist[4 * t + 1], c2 = hist[4 * t + 2], c3 = hist[4 * t + 3];
      const unsigned tot = c0 + c1 + c2 + c3;
      unsigned incl = tot;
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(incl, o);
        if (t >= o) incl += y;
      }
      const unsigned excl = incl - tot;
      if ((unsigned)k >= excl && (unsigned)k < incl) {
        unsigned r = (unsigned)k - excl, bin = 4 * t;
        if (r >= c0) { r -= c0; ++bin;
          if (r >= c1) { r -= c1; ++bin;
            if (r >= c2) { r -= c2; ++bin; } } }
        shared[0] = bin;
        shared[1] = r;
      }
    }
    __syncthreads();
    prefix |= shared[0] << shift;
    mask |= 255u << shift;
    k = (int)shared[1];
    __syncthreads();
  }
  return prefix;
}

// One workgroup per (field, j): the median of demod over skip_check's
// line slice j, as numba's np.median (exact order statistics by radix
// select); j == 0 also stores ll_linecount.
__global__ __launch_bounds__(kFactThreads) void k_field_facts(FactArgs a, const ld_line_field* __restrict__ fields) {
  extern __shared__ unsigned s[];  // a.smax keys (a slice of about one line, sized by the host)
  __shared__ unsigned hist[kFactThreads];
  __shared__ unsigned shared[4];
  const int f = blockIdx.x >> 3, j = blockIdx.x & 7;
  const ld_line_field fd = fields[f];
  const int t = threadIdx.x;
  const int first = fd.first_field != 0;
  const int lineoffset = a.pal ? (first ? 2 : 3) : 0;
  const int line = a.outlinecount + j + lineoffset;
  if (j == 0 && t == 0) {
    const int lc = a.pal ? (first ? 312 : 313) : (first ? 263 : 262);
    a.facts[f].ll_linecount = lc < fd.nlines ? a.ll[fd.lineloc_offset + lc] : __builtin_nan("");
    a.facts[f].known = 1;
    a.facts[f].reserved = 0;
  }
  double med = __builtin_nan("");
  long long lo = 0, hi = 0;
  if (line < fd.nlines) {
    const double b = a.ll[fd.lineloc_offset + line];
    // slice(int(b), int(b + length + 1)); a negative start would index from
    // the end in Python: counted as no data here
    if (b >= 0.0 && isfinite(b)) np_slice((long long)b, (long long)(b + a.line_samples + 1.0), fd.demod_len, lo, hi);
  }
  const long long m = hi - lo;
  if (m > 0 && m <= a.smax) {
    const float* src = a.demod + fd.demod_offset + lo;
    for (int i = t; i < m; i += kFactThreads) s[i] = fkey(src[i]);
    __syncthreads();
    const int h = (int)(m / 2);
    if (m & 1) {
      med = (double)fkey_inv(radix_select(s, (int)m, h, hist, shared, t));
    } else {
      // s[h-1] by selection; s[h] is the same key when it repeats past rank
      // h-1, else the smallest larger key
      const unsigned k1 = radix_select(s, (int)m, h - 1, hist, shared, t);
      unsigned le = 0, nxt = 0xffffffffu;
      for (int i = t; i < m; i += kFactThreads) {
        const unsigned v = s[i];
        le += v <= k1;
        if (v > k1 && v < nxt) nxt = v;
      }
      for (int o = 32; o > 0; o >>= 1) {
        le += __shfl_xor(le, o);
        nxt = min(nxt, (unsigned)__shfl_xor((int)nxt, o));
      }
      if ((t & 63) == 0) {
        hist[t >> 6] = le;
        hist[8 + (t >> 6)] = nxt;
      }
      __syncthreads();
      unsigned tle = 0, tn = 0xffffffffu;
      for (int w = 0; w < kFactThreads / 64; ++w) {
        tle += hist[w];
        tn = min(tn, hist[8 + w]);
      }
      const unsigned k2 = tle > (unsigned)h ? k1 : tn;
      med = (double)(fkey_inv(k1) + fkey_inv(k2)) / 2.0;
    }
  }
  if (t == 0) a.facts[f].medians[j] = med;
}

}  // namespace ldgpu

using namespace ldgpu;

extern "C" size_t ld_pulse_workspace_bytes(const ld_handle*, int64_t n) {
  const int64_t tiles = n > 0 ? (n + kEdgeTile - 1) / kEdgeTile : 1;
  return (size_t)(tiles + 1) * sizeof(long long) + (size_t)tiles * kEdgeThreads * sizeof(uint16_t);
}

extern "C" int ld_find_pulses(ld_handle* h, const float* x, int64_t n, double high, int64_t* edges, int64_t cap,
                              int64_t* count, void* workspace, size_t ws_bytes, void* stream) {
  if (!h || !x || n <= 0 || !count || cap < 0 || (cap > 0 && !edges) || !workspace) {
    set_error("ld_find_pulses: bad arguments");
    return LD_EINVAL;
  }
  if (((uintptr_t)x & 15) != 0) {
    set_error("ld_find_pulses: x must be 16-byte aligned");
    return LD_EINVAL;
  }
  if (ws_bytes < ld_pulse_workspace_bytes(h, n)) {
    set_error("ld_find_pulses: workspace %zu bytes < %zu needed", ws_bytes, ld_pulse_workspace_bytes(h, n));
    return LD_EINVAL;
  }
  const long long tiles = (n + kEdgeTile - 1) / kEdgeTile;
  if (tiles > 0x7fffffffLL) {
    set_error("ld_find_pulses: buffer too long");
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  long long* counts = (long long*)workspace;
  uint16_t* masks = (uint16_t*)(counts + tiles + 1);
  hipLaunchKernelGGL(k_edge_count, dim3((unsigned)tiles), dim3(kEdgeThreads), 0, st, x, (long long)n, high, counts,
                     masks);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_edge_scan, dim3(1), dim3(kScanThreads), 0, st, counts, tiles, (long long*)count);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_edge_write, dim3((unsigned)tiles), dim3(kEdgeThreads), 0, st, (const uint16_t*)masks,
                     (const long long*)counts, (long long*)edges, (long long)cap);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

extern "C" size_t ld_pulse_bits_workspace_bytes(const ld_handle*, int64_t n) {
  const int64_t tiles = n > 0 ? (n + kBitTile - 1) / kBitTile : 1;
  return (size_t)(tiles + 1) * sizeof(long long);
}

extern "C" int ld_find_pulses_bits(ld_handle* h, const uint32_t* bits, int64_t n, int64_t* edges, int64_t cap,
                                   int64_t* count, void* workspace, size_t ws_bytes, void* stream) {
  if (!h || !bits || n <= 0 || !count || cap < 0 || (cap > 0 && !edges) || !workspace) {
    set_error("ld_find_pulses_bits: bad arguments");
    return LD_EINVAL;
  }
  if (((uintptr_t)bits & 7) != 0) {
    set_error("ld_find_pulses_bits: bits must be 8-byte aligned");
    return LD_EINVAL;
  }
  if (ws_bytes < ld_pulse_bits_workspace_bytes(h, n)) {
    set_error("ld_find_pulses_bits: workspace %zu bytes < %zu needed", ws_bytes,
              ld_pulse_bits_workspace_bytes(h, n));
    return LD_EINVAL;
  }
  const long long tiles = (n + kBitTile - 1) / kBitTile;
  if (tiles > 0x7fffffffLL) {
    set_error("ld_find_pulses_bits: buffer too long");
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  long long* counts = (long long*)workspace;
  hipLaunchKernelGGL(k_bits_count, dim3((unsigned)tiles), dim3(kEdgeThreads), 0, st, bits, (long long)n, counts);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_edge_scan, dim3(1), dim3(kScanThreads), 0, st, counts, tiles, (long long*)count);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bits_write, dim3((unsigned)tiles), dim3(kEdgeThreads), 0, st, bits, (long long)n,
                     (const long long*)counts, (long long*)edges, (long long)cap);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

extern "C" int ld_pulse_means(ld_handle* h, const float* x, int64_t n, const int64_t* edges, const int64_t* count,
                              int64_t cap_pulses, double min_len, double trim, float* means, void* stream) {
  if (!h || !x || n <= 0 || !edges || !count || cap_pulses < 0 || (cap_pulses > 0 && !means)) {
    set_error("ld_pulse_means: bad arguments");
    return LD_EINVAL;
  }
  if (cap_pulses == 0) return LD_OK;
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  long long grid = (cap_pulses + 64 * kMeanWaves - 1) / (64 * kMeanWaves);
  if (grid > 1024) grid = 1024;
  hipLaunchKernelGGL(k_pulse_means, dim3((unsigned)grid), dim3(64 * kMeanWaves), 0, st, x, (long long)n,
                     (const long long*)edges, (const long long*)count, (long long)cap_pulses, min_len, trim, means);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

extern "C" int ld_pulse_export(ld_handle* h, const int64_t* edges, const int64_t* count, const float* means,
                               int64_t cap_edges, int64_t* host_edges, float* host_means, int64_t* host_count,
                               void* stream) {
  if (!h || !edges || !count || !means || cap_edges < 2 || !host_edges || !host_means || !host_count) {
    set_error("ld_pulse_export: bad arguments");
    return LD_EINVAL;
  }
  LD_HIP(hipSetDevice(h->device));
  // device addresses of the pinned host buffers (an error for memory the
  // GPU cannot reach, e.g. pageable memory)
  void *de = nullptr, *dm = nullptr, *dc = nullptr;
  LD_HIP(hipHostGetDevicePointer(&de, host_edges, 0));
  LD_HIP(hipHostGetDevicePointer(&dm, host_means, 0));
  LD_HIP(hipHostGetDevicePointer(&dc, host_count, 0));
  hipStream_t st = (hipStream_t)stream;
  long long grid = (cap_edges + 255) / 256;
  if (grid > 256) grid = 256;
  hipLaunchKernelGGL(k_pulse_export, dim3((unsigned)grid), dim3(256), 0, st, (const long long*)edges,
                     (const long long*)count, means, (long long)cap_edges, (long long*)de, (float*)dm,
                     (long long*)dc);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

extern "C" size_t ld_pulse_search_workspace_bytes(const ld_handle*, const ld_pulse_job* jobs, int njobs) {
  size_t b = 0;
  for (int i = 0; i < njobs; ++i) {
    const int64_t n = jobs ? jobs[i].n : 0;
    const int64_t tiles = n > 0 ? (n + kBitTile - 1) / kBitTile : 1;
    b += (size_t)(tiles + 1) * sizeof(long long);
  }
  return b > 0 ? b : sizeof(long long);
}

extern "C" int ld_pulse_search(ld_handle* h, const ld_pulse_job* jobs, int njobs, void* workspace, size_t ws_bytes,
                               void* stream) {
  if (!h || !jobs || njobs <= 0 || njobs > kMaxPulseJobs || !workspace) {
    set_error("ld_pulse_search: bad arguments (1 to %d jobs)", kMaxPulseJobs);
    return LD_EINVAL;
  }
  if (ws_bytes < ld_pulse_search_workspace_bytes(h, jobs, njobs)) {
    set_error("ld_pulse_search: workspace %zu bytes < %zu needed", ws_bytes,
              ld_pulse_search_workspace_bytes(h, jobs, njobs));
    return LD_EINVAL;
  }
  LD_HIP(hipSetDevice(h->device));
  PulseSearchArgs a{};
  a.njobs = njobs;
  long long* ws = (long long*)workspace;
  for (int i = 0; i < njobs; ++i) {
    const ld_pulse_job& q = jobs[i];
    if (!q.bits || q.n <= 0 || !q.x || !q.edges || !q.count || !q.means || q.cap < 2 || (q.cap & 1) ||
        !q.host_edges || !q.host_means || !q.host_count) {
      set_error("ld_pulse_search: job %d: bad arguments", i);
      return LD_EINVAL;
    }
    if (((uintptr_t)q.bits & 7) != 0) {
      set_error("ld_pulse_search: job %d: bits must be 8-byte aligned", i);
      return LD_EINVAL;
    }
    const long long tiles = (q.n + kBitTile - 1) / kBitTile;
    if (tiles > 0x3fffffffLL) {
      set_error("ld_pulse_search: job %d: buffer too long", i);
      return LD_EINVAL;
    }
    PulseJobDev& d = a.job[i];
    d.bits = q.bits;
    d.x = q.x;
    d.n = q.n;
    d.edges = (long long*)q.edges;
    d.count = (long long*)q.count;
    d.means = q.means;
    d.cap = q.cap;
    d.min_len = q.min_len;
    d.trim = q.trim;
    void *de = nullptr, *dm = nullptr, *dc = nullptr;
    LD_HIP(hipHostGetDevicePointer(&de, q.host_edges, 0));
    LD_HIP(hipHostGetDevicePointer(&dm, q.host_means, 0));
    LD_HIP(hipHostGetDevicePointer(&dc, q.host_count, 0));
    d.hedges = (long long*)de;
    d.hmeans = (float*)dm;
    d.hcount = (long long*)dc;
    d.tiles = ws;
    d.ntiles = tiles;
    ws += tiles + 1;
    a.tile0[i + 1] = a.tile0[i] + tiles;
    long long mwg = (q.cap / 2 + kMeanPulses - 1) / kMeanPulses;
    if (mwg > kMeanWgMax) mwg = kMeanWgMax;
    a.mwg0[i + 1] = a.mwg0[i] + mwg;
  }
  for (int i = njobs; i < kMaxPulseJobs; ++i) {
    a.tile0[i + 1] = a.tile0[njobs];
    a.mwg0[i + 1] = a.mwg0[njobs];
  }
  if (a.tile0[njobs] > 0x7fffffffLL) {
    set_error("ld_pulse_search: buffers too long");
    return LD_EINVAL;
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_bits_count_multi, dim3((unsigned)a.tile0[njobs]), dim3(kEdgeThreads), 0, st, a);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_edge_scan_multi, dim3(njobs), dim3(kScanThreads), 0, st, a);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_bits_write_multi, dim3((unsigned)a.tile0[njobs]), dim3(kEdgeThreads), 0, st, a);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_means_export_multi, dim3((unsigned)a.mwg0[njobs]), dim3(64), 0, st, a);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

extern "C" int ld_range_means(ld_handle* h, const float* x, int64_t n, const int64_t* begin, const int64_t* end,
                              int count, float* out, void* stream) {
  if (!h || !x || n <= 0 || count < 0 || (count > 0 && (!begin || !end || !out))) {
    set_error("ld_range_means: bad arguments");
    return LD_EINVAL;
  }
  if (count == 0) return LD_OK;
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  hipLaunchKernelGGL(k_range_means, dim3(count), dim3(64), 0, st, x, (long long)n, (const long long*)begin,
                     (const long long*)end, count, out);
  LD_HIP(hipGetLastError());
  return LD_OK;
}

extern "C" size_t ld_field_facts_workspace_bytes(const ld_handle*, int nfields) {
  return (size_t)(nfields > 0 ? nfields : 1) * sizeof(ld_line_field);
}

extern "C" int ld_field_facts_measure(ld_handle* h, const float* demod, int64_t demod_total,
                                      const ld_line_field* fields, int nfields, const double* linelocs,
                                      int64_t nlinelocs, int pal, int outlinecount, double line_samples,
                                      ld_field_facts* facts, void* workspace, size_t ws_bytes, void* stream) {
  if (!h || !demod || !fields || nfields <= 0 || !linelocs || !facts || outlinecount <= 0 ||
      !(line_samples > 0) || line_samples + 2 > kFactMax || !workspace) {
    set_error("ld_field_facts_measure: bad arguments");
    return LD_EINVAL;
  }
  if (ws_bytes < ld_field_facts_workspace_bytes(h, nfields)) {
    set_error("ld_field_facts_measure: workspace %zu bytes < %zu needed", ws_bytes,
              ld_field_facts_workspace_bytes(h, nfields));
    return LD_EINVAL;
  }
  for (int f = 0; f < nfields; ++f) {
    const ld_line_field& fd = fields[f];
    if (fd.nlines <= 0 || fd.lineloc_offset < 0 || fd.lineloc_offset + fd.nlines > nlinelocs) {
      set_error("ld_field_facts_measure: field %d: line range outside the arrays", f);
      return LD_EINVAL;
    }
    if (fd.demod_offset < 0 || fd.demod_len < 0 || fd.demod_offset + fd.demod_len > demod_total) {
      set_error("ld_field_facts_measure: field %d: demod span outside the buffer", f);
      return LD_ERANGE;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  // a pageable source is staged before hipMemcpyAsync returns
  ld_line_field* dfields = (ld_line_field*)workspace;
  LD_HIP(hipMemcpyAsync(dfields, fields, sizeof(ld_line_field) * nfields, hipMemcpyHostToDevice, st));
  // a slice int(b) .. int(b + line_samples + 1) holds at most
  // ceil(line_samples) + 2 samples: the LDS is sized for that (about 10 KB
  // at 40 MSPS instead of kFactMax's 32 KB, so that more workgroups fit)
  const int smax = (int)ceil(line_samples) + 3 < kFactMax ? (int)ceil(line_samples) + 3 : kFactMax;
  FactArgs a{demod, linelocs, facts, line_samples, outlinecount, pal ? 1 : 0, smax};
  hipLaunchKernelGGL(k_field_facts, dim3(nfields * 8), dim3(kFactThreads), sizeof(unsigned) * (size_t)smax, st, a,
                     (const ld_line_field*)dfields);
  LD_HIP(hipGetLastError());
  return LD_OK;
}
