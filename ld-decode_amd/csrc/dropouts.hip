// Dropout detection (Field.dropout_detect_demod, core.py:2830-2887), batched
// over fields: per field the spread of its rfhpf buffer, then the per-sample
// error test over its active lines, compacted to the error positions (the
// reference's errmap = np.nonzero(iserr)).
//   k_drop_partials  per chunk of the union of the field buffers: sum and sum
//                    of squares of rfhpf in double (overlapping buffers are
//                    read once)
//   k_drop_setup     per field: the rfhpf threshold float32(3 * std) and the
//                    sync-area table (line starts, running maximum of the
//                    sync-area ends)
//   k_drop_mask      per 4096-sample tile of a field buffer: the error bits
//                    (wave ballots, one 64-bit word per 64 samples) and
//                    their count; reads each channel once, coalesced
//   k_drop_scan      exclusive scan of the tile counts
//   k_drop_write     per tile with errors: positions from its bit words
// The grouping into dropout runs and their TBC coordinates (build_errlist,
// dropout_errlist_to_tbc) are serial and run on the host (ld_dropout_list).
#include <math.h>

#include <algorithm>
#include <climits>
#include <vector>

#include "internal.hpp"

#pragma clang fp contract(off)

namespace ldgpu {

constexpr int kDropThreads = 256;
constexpr int kDropPer = 16;
constexpr int kDropTile = kDropThreads * kDropPer;   // 4096 samples, 64 words
constexpr int kDropWords = kDropTile / 64;
constexpr long long kDropChunk = 16384;
constexpr int kDropScan = 1024;

struct DropTile {       // one 4096-sample tile of a field buffer
  int field;
  long long begin;      // first sample (buffer-relative)
};

struct DropField {      // device-side per-field plan
  int first_chunk, end_chunk;   // its chunks of the rfhpf union
  long long line_base;          // its rows of the sync-area table
};

struct DropArgs {
  const float* rfhpf;
  const float* demod;
  const float* demod_raw;
  const float* demod_05;
  const double* ll;
  const ld_dropout_field* fields;
  const DropField* plan;
  const long long* field_tile;   // first tile of every field, then the tile count
  int nfields;
  const float* thr;             // per field
  const long long* lstart;      // per field line: int(ll[l])
  const long long* lpmax;       // per field line: max sync-area end of lines 1..l
  unsigned long long* mask;     // per tile: kDropWords words
  double freq_hz_half;
  float vmin, vmax, vmin05, vmax05, sync_min, sync_min05;
  int video_rot;
  int pal;
};

// The vsync lines of get_vsync_lines (core.py:2816-2827).
__device__ __forceinline__ bool vsync_line(int l, bool first, bool pal) {
  if (l >= 1 && l < (first ? 10 : 9)) return true;
  return pal && l >= (first ? 311 : 310) && l < 318;
}

// Tile t: the field whose tiles include it and its first sample.
__device__ __forceinline__ DropTile drop_tile(const long long* __restrict__ field_tile, int nfields, long long t) {
  int lo = 0, hi = nfields - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (field_tile[mid] <= t) lo = mid; else hi = mid - 1;
  }
  DropTile d;
  d.field = lo;
  d.begin = (t - field_tile[lo]) * kDropTile;
  return d;
}

// The active range [int(ll[lineoffset + 1]), int(ll[lineoffset + linecount + 1]))
// clamped to the buffer.
__device__ __forceinline__ void drop_active(const double* ll, const ld_dropout_field& fd, long long& act0,
                                            long long& act1) {
  act0 = (long long)ll[fd.lineoffset + 1];
  act1 = (long long)ll[fd.lineoffset + fd.linecount + 1];
  if (act0 < 0) act0 = 0;
  if (act1 > fd.demod_len) act1 = fd.demod_len;
}

__global__ __launch_bounds__(kDropThreads) void k_drop_partials(const float* __restrict__ rfhpf,
                                                                const long long* __restrict__ chunk_begin,
                                                                const long long* __restrict__ chunk_end,
                                                                double2* __restrict__ partials) {
  __shared__ double2 red[kDropThreads / 64];
  const int c = blockIdx.x;
  const long long b = chunk_begin[c], e = chunk_end[c];
  double s0 = 0.0, q0 = 0.0, s1 = 0.0, q1 = 0.0;
  long long i = b + threadIdx.x;
  for (; i + kDropThreads < e; i += 2 * kDropThreads) {
    const double x = (double)rfhpf[i], y = (double)rfhpf[i + kDropThreads];
    s0 += x;
    q0 += x * x;
    s1 += y;
    q1 += y * y;
  }
  if (i < e) {
    const double x = (double)rfhpf[i];
    s0 += x;
    q0 += x * x;
  }
  double s = s0 + s1, q = q0 + q1;
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o);
    q += __shfl_xor(q, o);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_double2(s, q);
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t = red[0];
    for (int w = 1; w < kDropThreads / 64; ++w) {
      t.x += red[w].x;
      t.y += red[w].y;
    }
    partials[c] = t;
  }
}

// One block per field: the threshold, then the sync-area table.  Sample i
// lies in the union of the painted sync areas [int(ll[l]), end_l), l >= 1,
// exactly when the running maximum of end over lines 1..L exceeds i, L being
// the last line with int(ll[L]) <= i (line starts in order).
__global__ __launch_bounds__(kDropThreads) void k_drop_setup(DropArgs a, const double2* __restrict__ partials,
                                                             float* __restrict__ thr, long long* __restrict__ lstart,
                                                             long long* __restrict__ lpmax) {
  __shared__ long long scan[kDropThreads];
  __shared__ long long carry;
  const int f = blockIdx.x;
  const ld_dropout_field fd = a.fields[f];
  const DropField pl = a.plan[f];
  if (threadIdx.x == 0) {
    double s = 0.0, q = 0.0;
    for (int c = pl.first_chunk; c < pl.end_chunk; ++c) {
      s += partials[c].x;
      q += partials[c].y;
    }
    const double n = (double)fd.demod_len;
    double var = (q - s * (s / n)) / n;
    if (var < 0.0) var = 0.0;
    thr[f] = (float)(sqrt(var) * 3.0);
    carry = LLONG_MIN;
  }
  const double* ll = a.ll + fd.lineloc_offset;
  const long long hl = (long long)fd.hsync_max;
  const bool first = fd.first_field != 0, pal = a.pal != 0;
  __syncthreads();
  for (int l0 = 0; l0 < fd.nlines; l0 += kDropThreads) {
    const int l = l0 + threadIdx.x;
    long long st = 0, end = LLONG_MIN;
    if (l < fd.nlines) {
      st = (long long)ll[l];
      if (l >= 1) end = (vsync_line(l, first, pal) && l + 1 < fd.nlines) ? (long long)ll[l + 1] : st + hl;
    }
    scan[threadIdx.x] = end;
    __syncthreads();
    for (int o = 1; o < kDropThreads; o <<= 1) {
      const long long v = threadIdx.x >= (unsigned)o ? scan[threadIdx.x - o] : LLONG_MIN;
      __syncthreads();
      if (v > scan[threadIdx.x]) scan[threadIdx.x] = v;
      __syncthreads();
    }
    const long long m = scan[threadIdx.x] > carry ? scan[threadIdx.x] : carry;
    if (l < fd.nlines) {
      lstart[pl.line_base + l] = st;
      lpmax[pl.line_base + l] = m;
    }
    __syncthreads();
    if (threadIdx.x == kDropThreads - 1) carry = m;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kDropThreads) void k_drop_mask(DropArgs a, long long* __restrict__ counts) {
  __shared__ int part[kDropThreads / 64];
  const DropTile tile = drop_tile(a.field_tile, a.nfields, blockIdx.x);
  const ld_dropout_field fd = a.fields[tile.field];
  const DropField pl = a.plan[tile.field];
  const double* ll = a.ll + fd.lineloc_offset;
  unsigned long long* mw = a.mask + (size_t)blockIdx.x * kDropWords;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  long long act0, act1;
  drop_active(ll, fd, act0, act1);
  const long long b = tile.begin;
  if (b >= act1 || b + kDropTile <= act0) {     // uniform: no active sample
    if (threadIdx.x < kDropWords) mw[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) counts[blockIdx.x] = 0;
    return;
  }
  const long long* ls = a.lstart + pl.line_base;
  const long long* lp = a.lpmax + pl.line_base;
  // the last line starting at or before b (0 when none)
  int lo = 0, hi = fd.nlines - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ls[mid] <= b) lo = mid; else hi = mid - 1;
  }
  int l = lo;
  const float thr = a.thr[tile.field];
  const float* rf = a.rfhpf + fd.demod_offset;
  const float* dm = a.demod + fd.demod_offset;
  const float* dr = a.demod_raw + fd.demod_offset;
  const float* d5 = a.demod_05 + fd.demod_offset;
  const long long last = fd.demod_len - 1;
  int cnt = 0;
  constexpr int G = 4;   // samples per thread whose loads are issued together
  for (int k0 = 0; k0 < kDropPer; k0 += G) {
    float r[G], d[G], x[G], y[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      // loads clamped into the buffer (the active test below discards them)
      const long long i = b + (long long)(k0 + g) * kDropThreads + threadIdx.x;
      const long long ic = i < last ? i : last;
      const long long ir = ic - a.video_rot > 0 ? ic - a.video_rot : 0;
      r[g] = rf[ir];
      d[g] = dm[ic];
      x[g] = dr[ic];
      y[g] = d5[ic];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int k = k0 + g;
      const long long i = b + (long long)k * kDropThreads + threadIdx.x;
      bool err = false;
      if (i >= act0 && i < act1) {
        while (l + 1 < fd.nlines && ls[l + 1] <= i) ++l;
        const bool sync = l >= 1 && ls[l] <= i && lp[l] > i;
        if (i >= a.video_rot) err = (r[g] < -thr) | (r[g] > thr);
        err |= (double)x[g] > a.freq_hz_half;
        const float mn = sync ? a.sync_min : a.vmin, mn5 = sync ? a.sync_min05 : a.vmin05;
        err |= (d[g] < mn) | (d[g] > a.vmax);
        err |= (y[g] < mn5) | (y[g] > a.vmax05);
      }
      const unsigned long long bits = __ballot(err);
      if (lane == 0) {
        mw[k * (kDropThreads / 64) + w] = bits;
        cnt += __popcll(bits);
      }
    }
  }
  if (lane == 0) part[w] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int q = 0; q < kDropThreads / 64; ++q) t += part[q];
    counts[blockIdx.x] = t;
  }
}

// Exclusive scan of the tile counts in place; the total goes to *total.
__global__ __launch_bounds__(kDropScan) void k_drop_scan(long long* __restrict__ counts, long long ntiles,
                                                         long long* __restrict__ total) {
  __shared__ long long part[kDropScan];
  const int t = threadIdx.x;
  const long long per = (ntiles + kDropScan - 1) / kDropScan;
  const long long b = (long long)t * per;
  const long long e = b + per < ntiles ? b + per : ntiles;
  long long s = 0;
  for (long long i = b; i < e; ++i) s += counts[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < kDropScan; o <<= 1) {
    const long long v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  long long run = part[t] - s;
  for (long long i = b; i < e; ++i) {
    const long long c = counts[i];
    counts[i] = run;
    run += c;
  }
  if (t == kDropScan - 1) *total = part[t];
}

// One wave per tile: word j covers samples begin + 64 j .. + 63 in order.
__global__ __launch_bounds__(64) void k_drop_write(const long long* __restrict__ field_tile, int nfields,
                                                   const unsigned long long* __restrict__ mask,
                                                   const long long* __restrict__ offsets, long long* __restrict__ pos,
                                                   long long cap) {
  const long long o0 = offsets[blockIdx.x], o1 = offsets[blockIdx.x + 1];
  if (o0 == o1) return;
  const int lane = threadIdx.x;
  unsigned long long bits = mask[(size_t)blockIdx.x * kDropWords + lane];
  const int c = __popcll(bits);
  int incl = c;
  for (int o = 1; o < 64; o <<= 1) {
    const int v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  long long k = o0 + incl - c;
  const long long base = drop_tile(field_tile, nfields, blockIdx.x).begin + (long long)lane * 64;
  while (bits) {
    const int b = __builtin_ctzll(bits);
    bits &= bits - 1;
    if (k < cap) pos[k] = base + b;
    ++k;
  }
}

// field_start[f] = the scanned count of field f's first tile (f <= nfields).
__global__ void k_drop_starts(const long long* __restrict__ counts, const long long* __restrict__ field_tile,
                              int nfields, long long* __restrict__ field_start) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f <= nfields) field_start[f] = counts[field_tile[f]];
}

}  // namespace ldgpu

using namespace ldgpu;

namespace {

struct DropPlan {
  std::vector<long long> chunk_begin, chunk_end;   // absolute, sorted
  std::vector<DropField> fields;
  std::vector<long long> field_tile;     // first tile of every field (+ total)
  long long ntiles = 0;
  long long nlines = 0;
};

// The rfhpf union is cut at every field buffer's start and end, then into
// chunks of at most kDropChunk samples; each field sums a run of chunks.
DropPlan drop_plan(const ld_dropout_field* fields, int nfields) {
  DropPlan p;
  std::vector<std::pair<long long, long long>> iv;
  std::vector<long long> cuts;
  for (int f = 0; f < nfields; ++f) {
    iv.push_back({fields[f].demod_offset, fields[f].demod_offset + fields[f].demod_len});
    cuts.push_back(iv.back().first);
    cuts.push_back(iv.back().second);
  }
  std::sort(cuts.begin(), cuts.end());
  cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
  std::sort(iv.begin(), iv.end());
  size_t j = 0;
  long long reach = LLONG_MIN;   // furthest end of the intervals starting at or before the segment
  for (size_t c = 0; c + 1 < cuts.size(); ++c) {
    const long long s = cuts[c], e = cuts[c + 1];
    while (j < iv.size() && iv[j].first <= s) reach = std::max(reach, iv[j++].second);
    if (reach <= s) continue;      // not covered by any field buffer
    for (long long b = s; b < e; b += kDropChunk) {
      p.chunk_begin.push_back(b);
      p.chunk_end.push_back(std::min(e, b + kDropChunk));
    }
  }
  for (int f = 0; f < nfields; ++f) {
    const long long s = fields[f].demod_offset, e = s + fields[f].demod_len;
    DropField d;
    d.first_chunk = (int)(std::lower_bound(p.chunk_begin.begin(), p.chunk_begin.end(), s) - p.chunk_begin.begin());
    d.end_chunk = (int)(std::lower_bound(p.chunk_begin.begin(), p.chunk_begin.end(), e) - p.chunk_begin.begin());
    d.line_base = p.nlines;
    p.nlines += fields[f].nlines;
    p.fields.push_back(d);
    p.field_tile.push_back(p.ntiles);
    p.ntiles += (fields[f].demod_len + kDropTile - 1) / kDropTile;
  }
  p.field_tile.push_back(p.ntiles);
  return p;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct DropLayout {
  size_t fields, plan, chunk_begin, chunk_end, field_tile, partials, thr, lstart, lpmax, mask, counts, total;
};

DropLayout drop_layout(const DropPlan& p, int nfields) {
  DropLayout l;
  size_t o = 0;
  l.fields = o; o = align256(o + sizeof(ld_dropout_field) * nfields);
  l.plan = o; o = align256(o + sizeof(DropField) * nfields);
  l.chunk_begin = o; o = align256(o + sizeof(long long) * p.chunk_begin.size());
  l.chunk_end = o; o = align256(o + sizeof(long long) * p.chunk_end.size());
  l.field_tile = o; o = align256(o + sizeof(long long) * p.field_tile.size());
  l.partials = o; o = align256(o + sizeof(double2) * p.chunk_begin.size());
  l.thr = o; o = align256(o + sizeof(float) * nfields);
  l.lstart = o; o = align256(o + sizeof(long long) * p.nlines);
  l.lpmax = o; o = align256(o + sizeof(long long) * p.nlines);
  l.mask = o; o = align256(o + sizeof(unsigned long long) * kDropWords * p.ntiles);
  l.counts = o; o = align256(o + sizeof(long long) * (p.ntiles + 1));
  l.total = o;
  return l;
}

}  // namespace

extern "C" size_t ld_dropout_workspace_bytes(const ld_handle*, const ld_dropout_field* fields, int nfields) {
  if (!fields || nfields <= 0) return 0;
  const DropPlan p = drop_plan(fields, nfields);
  return drop_layout(p, nfields).total + 256;
}

extern "C" int ld_dropout_detect(ld_handle* h, const float* rfhpf, const float* demod, const float* demod_raw,
                                 const float* demod_05, int64_t total, const ld_dropout_field* fields, int nfields,
                                 const double* linelocs, int64_t nlinelocs, const ld_dropout_params* prm,
                                 int64_t* positions, int64_t cap, int64_t* field_start, void* workspace,
                                 size_t ws_bytes, void* stream) {
  if (!h || !rfhpf || !demod || !demod_raw || !demod_05 || !fields || nfields <= 0 || !linelocs || !prm ||
      cap < 0 || (cap > 0 && !positions) || !field_start || !workspace) {
    set_error("ld_dropout_detect: bad arguments");
    return LD_EINVAL;
  }
  for (int f = 0; f < nfields; ++f) {
    const ld_dropout_field& fd = fields[f];
    if (fd.demod_offset < 0 || fd.demod_len <= 0 || fd.demod_offset + fd.demod_len > total) {
      set_error("ld_dropout_detect: field %d: buffer outside the arrays", f);
      return LD_ERANGE;
    }
    if (fd.nlines < fd.lineoffset + fd.linecount + 2 || fd.lineloc_offset < 0 ||
        fd.lineloc_offset + fd.nlines > nlinelocs || fd.lineoffset < 0 || fd.linecount < 1) {
      set_error("ld_dropout_detect: field %d: line range", f);
      return LD_EINVAL;
    }
    if (!(fd.hsync_max >= 0.0 && fd.hsync_max < 1e9)) {
      set_error("ld_dropout_detect: field %d: hsync length %g", f, fd.hsync_max);
      return LD_EINVAL;
    }
  }
  if (prm->video_rot < 1) {
    set_error("ld_dropout_detect: video_rot %d (the reference's shift needs >= 1)", prm->video_rot);
    return LD_EINVAL;
  }
  const DropPlan p = drop_plan(fields, nfields);
  const DropLayout L = drop_layout(p, nfields);
  if (ws_bytes < L.total + 256) {
    set_error("ld_dropout_detect: workspace %zu bytes < %zu needed", ws_bytes, L.total + 256);
    return LD_EINVAL;
  }
  if (p.ntiles > 0x7fffffffLL || (int64_t)p.chunk_begin.size() > 0x7fffffffLL) {
    set_error("ld_dropout_detect: too many samples");
    return LD_EINVAL;
  }
  char* ws = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  // a pageable source is staged before hipMemcpyAsync returns
  LD_HIP(hipMemcpyAsync(ws + L.fields, fields, sizeof(ld_dropout_field) * nfields, hipMemcpyHostToDevice, st));
  LD_HIP(hipMemcpyAsync(ws + L.plan, p.fields.data(), sizeof(DropField) * nfields, hipMemcpyHostToDevice, st));
  LD_HIP(hipMemcpyAsync(ws + L.chunk_begin, p.chunk_begin.data(), sizeof(long long) * p.chunk_begin.size(),
                        hipMemcpyHostToDevice, st));
  LD_HIP(hipMemcpyAsync(ws + L.chunk_end, p.chunk_end.data(), sizeof(long long) * p.chunk_end.size(),
                        hipMemcpyHostToDevice, st));
  LD_HIP(hipMemcpyAsync(ws + L.field_tile, p.field_tile.data(), sizeof(long long) * p.field_tile.size(),
                        hipMemcpyHostToDevice, st));
  double2* partials = (double2*)(ws + L.partials);
  long long* counts = (long long*)(ws + L.counts);
  DropArgs a{};
  a.rfhpf = rfhpf;
  a.demod = demod;
  a.demod_raw = demod_raw;
  a.demod_05 = demod_05;
  a.ll = linelocs;
  a.fields = (const ld_dropout_field*)(ws + L.fields);
  a.plan = (const DropField*)(ws + L.plan);
  a.field_tile = (const long long*)(ws + L.field_tile);
  a.nfields = nfields;
  a.thr = (const float*)(ws + L.thr);
  a.lstart = (const long long*)(ws + L.lstart);
  a.lpmax = (const long long*)(ws + L.lpmax);
  a.mask = (unsigned long long*)(ws + L.mask);
  a.freq_hz_half = prm->freq_hz_half;
  a.vmin = prm->vmin;
  a.vmax = prm->vmax;
  a.vmin05 = prm->vmin05;
  a.vmax05 = prm->vmax05;
  a.sync_min = prm->sync_min;
  a.sync_min05 = prm->sync_min05;
  a.video_rot = prm->video_rot;
  a.pal = prm->pal;
  if (!p.chunk_begin.empty()) {
    hipLaunchKernelGGL(k_drop_partials, dim3((unsigned)p.chunk_begin.size()), dim3(kDropThreads), 0, st, rfhpf,
                       (const long long*)(ws + L.chunk_begin), (const long long*)(ws + L.chunk_end), partials);
    LD_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_drop_setup, dim3(nfields), dim3(kDropThreads), 0, st, a, (const double2*)partials,
                     (float*)(ws + L.thr), (long long*)(ws + L.lstart), (long long*)(ws + L.lpmax));
  LD_HIP(hipGetLastError());
  const unsigned ntiles = (unsigned)p.ntiles;
  hipLaunchKernelGGL(k_drop_mask, dim3(ntiles), dim3(kDropThreads), 0, st, a, counts);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_drop_scan, dim3(1), dim3(kDropScan), 0, st, counts, (long long)ntiles, counts + ntiles);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_drop_write, dim3(ntiles), dim3(64), 0, st, a.field_tile, nfields,
                     (const unsigned long long*)a.mask,
                     (const long long*)counts, (long long*)positions, (long long)cap);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_drop_starts, dim3((nfields + 64) / 64), dim3(64), 0, st, (const long long*)counts,
                     (const long long*)(ws + L.field_tile), nfields, (long long*)field_start);
  LD_HIP(hipGetLastError());
  return LD_OK;
}
