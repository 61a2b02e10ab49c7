// Dropout detection (Field.dropout_detect_demod, core.py:2830-2887), batched
// over fields: per field the spread of its rfhpf buffer, then the per-sample
// error test over its buffer compacted to the error positions (errmap).
//   k_drop_partials  per (field, 64 Ki-sample chunk): sum and sum of squares
//                    of rfhpf in double
//   k_drop_levels    per field: the rfhpf threshold float32(3 * std)
//   k_drop_count     per 4096-sample tile of a field buffer: error samples
//   k_drop_scan      exclusive scan of the tile counts
//   k_drop_write     per tile: the error positions, in order
// The grouping into dropout runs and their TBC coordinates (build_errlist,
// dropout_errlist_to_tbc) are serial and run on the host (ld_dropout_list).
#include <math.h>

#include <vector>

#include "internal.hpp"

#pragma clang fp contract(off)

namespace ldgpu {

constexpr int kDropThreads = 256;
constexpr int kDropPer = 16;
constexpr int kDropTile = kDropThreads * kDropPer;
constexpr int kDropChunk = 65536;
constexpr int kDropScan = 1024;

struct DropTile {       // one 4096-sample tile of a field buffer
  int field;
  int pad_;
  long long begin;      // first sample (buffer-relative)
};

struct DropArgs {
  const float* rfhpf;
  const float* demod;
  const float* demod_raw;
  const float* demod_05;
  const double* ll;
  const ld_dropout_field* fields;
  const DropTile* tiles;
  const float* thr;            // per field
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

// Whether sample i of field fd's buffer is an error sample.
__device__ __forceinline__ bool drop_err(const DropArgs& a, const ld_dropout_field& fd, float thr, long long i,
                                         long long act0, long long act1, int& l) {
  if (i < act0 || i >= act1) return false;
  const long long base = fd.demod_offset;
  const double* ll = a.ll + fd.lineloc_offset;
  bool err = false;
  if (i >= a.video_rot) {
    const float r = a.rfhpf[base + i - a.video_rot];
    err = (r < -thr) | (r > thr);
  }
  err |= (double)a.demod_raw[base + i] > a.freq_hz_half;
  // the sync-area minimum: the last line starting at or before i (lines in
  // order), over its hsync length or, on a vsync line, the whole line
  while (l + 1 < fd.nlines && (long long)ll[l + 1] <= i) ++l;
  bool sync = false;
  if (l >= 1 && (long long)ll[l] <= i) {
    const bool first = fd.first_field != 0;
    const long long end = (vsync_line(l, first, a.pal != 0) && l + 1 < fd.nlines)
                              ? (long long)ll[l + 1]
                              : (long long)ll[l] + (long long)fd.hsync_max;
    sync = i < end;
  }
  const float d = a.demod[base + i], d5 = a.demod_05[base + i];
  const float mn = sync ? a.sync_min : a.vmin, mn5 = sync ? a.sync_min05 : a.vmin05;
  err |= (d < mn) | (d > a.vmax);
  err |= (d5 < mn5) | (d5 > a.vmax05);
  return err;
}

// The active range [int(ll[lineoffset + 1]), int(ll[lineoffset + linecount + 1]))
// and the first line to search from for the tile starting at `begin`.
__device__ __forceinline__ void drop_range(const DropArgs& a, const ld_dropout_field& fd, long long begin,
                                           long long& act0, long long& act1, int& l) {
  const double* ll = a.ll + fd.lineloc_offset;
  act0 = (long long)ll[fd.lineoffset + 1];
  const int le = fd.lineoffset + fd.linecount + 1;
  act1 = le < fd.nlines ? (long long)ll[le] : fd.demod_len;
  if (act0 < 0) act0 = 0;
  if (act1 > fd.demod_len) act1 = fd.demod_len;
  // binary search: the last line l with int(ll[l]) <= begin (0 when none)
  int lo = 0, hi = fd.nlines - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((long long)ll[mid] <= begin) lo = mid; else hi = mid - 1;
  }
  l = lo;
}

__global__ __launch_bounds__(kDropThreads) void k_drop_partials(const float* __restrict__ rfhpf,
                                                                const ld_dropout_field* __restrict__ fields,
                                                                const int* __restrict__ chunk_field,
                                                                const long long* __restrict__ chunk_begin,
                                                                double2* __restrict__ partials) {
  __shared__ double2 red[kDropThreads / 64];
  const int c = blockIdx.x;
  const ld_dropout_field fd = fields[chunk_field[c]];
  const long long b = chunk_begin[c];
  const long long e = b + kDropChunk < fd.demod_len ? b + kDropChunk : fd.demod_len;
  double s = 0.0, q = 0.0;
  for (long long i = b + threadIdx.x; i < e; i += kDropThreads) {
    const double x = (double)rfhpf[fd.demod_offset + i];
    s += x;
    q += x * x;
  }
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

__global__ void k_drop_levels(const ld_dropout_field* __restrict__ fields, int nfields,
                              const int* __restrict__ first_chunk, const double2* __restrict__ partials,
                              float* __restrict__ thr) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nfields) return;
  double s = 0.0, q = 0.0;
  for (int c = first_chunk[f]; c < first_chunk[f + 1]; ++c) {
    s += partials[c].x;
    q += partials[c].y;
  }
  const double n = (double)fields[f].demod_len;
  const double m = s / n;
  double var = (q - s * m) / n;
  if (var < 0.0) var = 0.0;
  thr[f] = (float)(sqrt(var) * 3.0);
}

__global__ __launch_bounds__(kDropThreads) void k_drop_count(DropArgs a, long long* __restrict__ counts) {
  __shared__ int part[kDropThreads / 64];
  const DropTile tile = a.tiles[blockIdx.x];
  const ld_dropout_field fd = a.fields[tile.field];
  const float thr = a.thr[tile.field];
  const long long i0 = tile.begin + (long long)threadIdx.x * kDropPer;
  long long act0, act1;
  int l;
  drop_range(a, fd, i0, act0, act1, l);
  int c = 0;
  for (int k = 0; k < kDropPer; ++k) c += drop_err(a, fd, thr, i0 + k, act0, act1 < fd.demod_len ? act1 : fd.demod_len, l);
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int w = 0; w < kDropThreads / 64; ++w) t += part[w];
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

__global__ __launch_bounds__(kDropThreads) void k_drop_write(DropArgs a, const long long* __restrict__ offsets,
                                                             long long* __restrict__ pos, long long cap) {
  __shared__ int part[kDropThreads];
  const DropTile tile = a.tiles[blockIdx.x];
  const ld_dropout_field fd = a.fields[tile.field];
  const float thr = a.thr[tile.field];
  const long long i0 = tile.begin + (long long)threadIdx.x * kDropPer;
  long long act0, act1;
  int l;
  drop_range(a, fd, i0, act0, act1, l);
  unsigned bits = 0;
  for (int k = 0; k < kDropPer; ++k)
    if (drop_err(a, fd, thr, i0 + k, act0, act1 < fd.demod_len ? act1 : fd.demod_len, l)) bits |= 1u << k;
  const int c = __popc(bits);
  part[threadIdx.x] = c;
  __syncthreads();
  for (int o = 1; o < kDropThreads; o <<= 1) {
    const int v = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  long long k = offsets[blockIdx.x] + part[threadIdx.x] - c;
  while (bits) {
    const int i = __ffs(bits) - 1;
    bits &= bits - 1;
    if (k < cap) pos[k] = i0 + i;
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
  std::vector<int> chunk_field;
  std::vector<long long> chunk_begin;
  std::vector<int> first_chunk;
  std::vector<DropTile> tiles;
  std::vector<long long> field_tile;     // first tile of every field (+ total)
};

DropPlan drop_plan(const ld_dropout_field* fields, int nfields) {
  DropPlan p;
  for (int f = 0; f < nfields; ++f) {
    p.first_chunk.push_back((int)p.chunk_field.size());
    for (long long b = 0; b < fields[f].demod_len; b += kDropChunk) {
      p.chunk_field.push_back(f);
      p.chunk_begin.push_back(b);
    }
    p.field_tile.push_back((long long)p.tiles.size());
    for (long long b = 0; b < fields[f].demod_len; b += kDropTile) p.tiles.push_back({f, 0, b});
  }
  p.first_chunk.push_back((int)p.chunk_field.size());
  p.field_tile.push_back((long long)p.tiles.size());
  return p;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct DropLayout {
  size_t fields, chunk_field, chunk_begin, first_chunk, tiles, field_tile, partials, thr, counts, total;
};

DropLayout drop_layout(const DropPlan& p, int nfields) {
  DropLayout l;
  size_t o = 0;
  l.fields = o; o = align256(o + sizeof(ld_dropout_field) * nfields);
  l.chunk_field = o; o = align256(o + sizeof(int) * p.chunk_field.size());
  l.chunk_begin = o; o = align256(o + sizeof(long long) * p.chunk_begin.size());
  l.first_chunk = o; o = align256(o + sizeof(int) * p.first_chunk.size());
  l.tiles = o; o = align256(o + sizeof(DropTile) * p.tiles.size());
  l.field_tile = o; o = align256(o + sizeof(long long) * p.field_tile.size());
  l.partials = o; o = align256(o + sizeof(double2) * p.chunk_field.size());
  l.thr = o; o = align256(o + sizeof(float) * nfields);
  l.counts = o; o = align256(o + sizeof(long long) * (p.tiles.size() + 1));
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
  if ((int64_t)p.tiles.size() > 0x7fffffffLL) {
    set_error("ld_dropout_detect: too many samples");
    return LD_EINVAL;
  }
  char* ws = (char*)workspace;
  hipStream_t st = (hipStream_t)stream;
  LD_HIP(hipSetDevice(h->device));
  // a pageable source is staged before hipMemcpyAsync returns
  LD_HIP(hipMemcpyAsync(ws + L.fields, fields, sizeof(ld_dropout_field) * nfields, hipMemcpyHostToDevice, st));
  LD_HIP(hipMemcpyAsync(ws + L.chunk_field, p.chunk_field.data(), sizeof(int) * p.chunk_field.size(),
                        hipMemcpyHostToDevice, st));
  LD_HIP(hipMemcpyAsync(ws + L.chunk_begin, p.chunk_begin.data(), sizeof(long long) * p.chunk_begin.size(),
                        hipMemcpyHostToDevice, st));
  LD_HIP(hipMemcpyAsync(ws + L.first_chunk, p.first_chunk.data(), sizeof(int) * p.first_chunk.size(),
                        hipMemcpyHostToDevice, st));
  LD_HIP(hipMemcpyAsync(ws + L.tiles, p.tiles.data(), sizeof(DropTile) * p.tiles.size(), hipMemcpyHostToDevice,
                        st));
  LD_HIP(hipMemcpyAsync(ws + L.field_tile, p.field_tile.data(), sizeof(long long) * p.field_tile.size(),
                        hipMemcpyHostToDevice, st));
  const ld_dropout_field* dfields = (const ld_dropout_field*)(ws + L.fields);
  double2* partials = (double2*)(ws + L.partials);
  float* thr = (float*)(ws + L.thr);
  long long* counts = (long long*)(ws + L.counts);
  hipLaunchKernelGGL(k_drop_partials, dim3((unsigned)p.chunk_field.size()), dim3(kDropThreads), 0, st, rfhpf,
                     dfields, (const int*)(ws + L.chunk_field), (const long long*)(ws + L.chunk_begin), partials);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_drop_levels, dim3((nfields + 63) / 64), dim3(64), 0, st, dfields, nfields,
                     (const int*)(ws + L.first_chunk), (const double2*)partials, thr);
  LD_HIP(hipGetLastError());
  DropArgs a{};
  a.rfhpf = rfhpf;
  a.demod = demod;
  a.demod_raw = demod_raw;
  a.demod_05 = demod_05;
  a.ll = linelocs;
  a.fields = dfields;
  a.tiles = (const DropTile*)(ws + L.tiles);
  a.thr = thr;
  a.freq_hz_half = prm->freq_hz_half;
  a.vmin = prm->vmin;
  a.vmax = prm->vmax;
  a.vmin05 = prm->vmin05;
  a.vmax05 = prm->vmax05;
  a.sync_min = prm->sync_min;
  a.sync_min05 = prm->sync_min05;
  a.video_rot = prm->video_rot;
  a.pal = prm->pal;
  const unsigned ntiles = (unsigned)p.tiles.size();
  hipLaunchKernelGGL(k_drop_count, dim3(ntiles), dim3(kDropThreads), 0, st, a, counts);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_drop_scan, dim3(1), dim3(kDropScan), 0, st, counts, (long long)ntiles, counts + ntiles);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_drop_write, dim3(ntiles), dim3(kDropThreads), 0, st, a, (const long long*)counts,
                     (long long*)positions, (long long)cap);
  LD_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_drop_starts, dim3((nfields + 64) / 64), dim3(64), 0, st, (const long long*)counts,
                     (const long long*)(ws + L.field_tile), nfields, (long long*)field_start);
  LD_HIP(hipGetLastError());
  return LD_OK;
}
