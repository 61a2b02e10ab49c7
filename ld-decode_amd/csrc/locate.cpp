// Host side of line location: pulse classification, vertical-blanking
// analysis and the line-0 / line-start logic of Field.compute_linelocs
// (lddecode/core.py:1761-2481) over pulse lists the GPU found, and the
// field-to-field walk of LDdecode.decodefield (core.py:4018-4098).
//
// The arithmetic follows the reference's own types: pulse positions and
// lengths are integers, everything derived from them is double; getpulses'
// sync level is float32 (NumPy keeps float32 when a float32 scalar meets a
// Python float).  Python's negative indexing is reproduced where the
// reference can index with -1.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"

using ldgpu::set_error;

namespace {

enum Kind { kHsync = 0, kEq1 = 1, kVsync = 2, kEq2 = 3 };
constexpr int64_t kMaxSyncLen = 5000;      // findpulses_numba_raw max_synclen

struct Pulse {
  int64_t start, len;
};
struct VPulse {
  int kind;
  Pulse p;
  bool good;
};

inline bool inrange(double a, double lo, double hi) { return a >= lo && a <= hi; }
// round half to even, as np.round: adding and removing 1.5 * 2^52 rounds
// to an integer in the default rounding mode (|x| < 2^51; beyond that x is
// already an integer)
inline int64_t np_round(double x) {
  constexpr double kMagic = 6755399441055744.0;
  if (!(std::fabs(x) < 2251799813685248.0)) return (int64_t)x;
  return (int64_t)((x + kMagic) - kMagic);   // IEEE order kept: no fast-math
}

// The integers d with lo <= d / L <= hi in double arithmetic (the division
// correctly rounded, so d / L is non-decreasing in d): [first, last].
struct Span {
  int64_t first, last;
  bool has(int64_t d) const { return d >= first && d <= last; }
};
inline Span ratio_span(double L, double lo, double hi) {
  int64_t a = (int64_t)std::floor(lo * L) - 2, b = (int64_t)std::ceil(hi * L) + 2;
  while ((double)a / L < lo) ++a;
  while ((double)(a - 1) / L >= lo) --a;
  while ((double)b / L > hi) --b;
  while ((double)(b + 1) / L <= hi) ++b;
  return {a, b};
}

struct Geometry {
  const ld_locate_params& p;
  double linefreq;       // Field.get_linefreq() without a line: (freq / linelen) * linelen
  int outlinecount;      // frame_lines // 2 + 1
  int proclines;         // outlinecount + 10 (+3 PAL), lineoffset 0
  // pulse_qualitycheck's line-length ranges and computeLineLen's, as
  // integer sample spans (the tests divide integer gaps by linelen)
  Span q_half, q_full, q_mixed, line_ok;
  explicit Geometry(const ld_locate_params& q)
      : p(q),
        linefreq((q.freq / q.linelen) * q.linelen),
        outlinecount(q.frame_lines / 2 + 1),
        proclines(q.frame_lines / 2 + 1 + 10 + (q.pal ? 3 : 0)),
        q_half(ratio_span(q.linelen, 0.4, 0.6)),
        q_full(ratio_span(q.linelen, 0.9, 1.1)),
        q_mixed(ratio_span(q.linelen, 0.4, 1.1)),
        line_ok(ratio_span(q.linelen, 0.95, 1.05)) {}
  double us(double x) const { return x * linefreq; }
};

struct Timings {
  double hsync[2], eq[2], vsync[2];
};

// Per-call scratch, reused from field to field.
struct Scratch {
  std::vector<VPulse> valid, vb;
  std::vector<int64_t> loc, hist;
  std::vector<double> dist, fill;
  std::vector<uint8_t> have, err;
  std::vector<Pulse> pulses;
  std::vector<float> means;
};

// Field.get_timings (core.py:1761-1810).  The median of the hsync-like
// lengths (integers in [cmin, cmax]) comes from a histogram.
Timings get_timings(const Geometry& g, const std::vector<Pulse>& pulses, Scratch& sc) {
  const ld_locate_params& p = g.p;
  const double typical = g.us(p.hsync_us);
  const double cmin = g.us(p.hsync_us - 1.75), cmax = g.us(p.hsync_us + 2);
  const int64_t lo = (int64_t)std::ceil(cmin), hi = (int64_t)std::floor(cmax);
  double med = p.hsync_us;     // the reference falls back to microseconds here
  if (hi >= lo) {
    std::vector<int64_t>& h = sc.hist;
    h.assign((size_t)(hi - lo + 1), 0);
    int64_t cnt = 0;
    for (const Pulse& q : pulses)
      if (q.len >= lo && q.len <= hi) {
        ++h[q.len - lo];
        ++cnt;
      }
    if (cnt) {
      // the k-th smallest (0-based) for k = cnt / 2 and, for an even count, k - 1
      auto kth = [&](int64_t k) {
        int64_t c = 0;
        for (size_t i = 0; i < h.size(); ++i) {
          c += h[i];
          if (c > k) return (double)(lo + (int64_t)i);
        }
        return (double)hi;
      };
      med = (cnt % 2) ? kth(cnt / 2) : (kth(cnt / 2 - 1) + kth(cnt / 2)) / 2;
    }
  }
  Timings t;
  t.hsync[0] = med + g.us(-0.5);
  t.hsync[1] = med + g.us(0.5);
  const double off = med - typical;
  t.eq[0] = g.us(p.eq_us - 0.5) + off;
  t.eq[1] = g.us(p.eq_us + 0.5) + off;
  t.vsync[0] = g.us(p.vsync_us * 0.5) + off;
  t.vsync[1] = g.us(p.vsync_us + 1) + off;
  return t;
}

// Field.pulse_qualitycheck (core.py:1812-1823).
bool quality(const Geometry& g, int prev_kind, int64_t prev_start, int kind, int64_t start) {
  const int64_t d = start - prev_start;
  if (prev_kind > 0 && kind > 0) return g.q_half.has(d);        // (0.4, 0.6)
  if (prev_kind == 0 && kind == 0) return g.q_full.has(d);      // (0.9, 1.1)
  return g.q_mixed.has(d);                                      // (0.4, 1.1)
}

// Field.run_vblank_state_machine (core.py:1826-1911) over pulses[b, e).
bool vblank_state_machine(const Geometry& g, const Timings& t, const std::vector<Pulse>& pulses, size_t b,
                          size_t e, std::vector<VPulse>& valid) {
  valid.clear();
  double state_end = 0;
  double state_length = 0;     // 0: None (numPulses / 2 is never 0)
  const double half = g.p.num_pulses / 2.0;
  bool done = false;
  int last_kind = -1;          // valid[-1] kept in registers
  int64_t last_start = 0;
  for (size_t i = b; i < e; ++i) {
    const Pulse& q = pulses[i];
    const double n = (double)q.len;
    const int state = last_kind;
    int kind = -1;
    if (state == -1) {
      if (inrange(n, t.hsync[0], t.hsync[1])) kind = kHsync;
    } else if (state == kHsync) {
      if (inrange(n, t.hsync[0], t.hsync[1])) {
        kind = kHsync;
      } else if (inrange(n, t.eq[0], t.eq[1])) {
        kind = kEq1;
        state_length = half;
      } else if (inrange(n, t.vsync[0], t.vsync[1])) {
        kind = kVsync;
      }
    } else if (state == kEq1) {
      if (inrange(n, t.eq[0], t.eq[1])) {
        kind = kEq1;
      } else if (inrange(n, t.vsync[0], t.vsync[1])) {
        kind = kVsync;
        state_length = half;
      } else if (inrange(n, t.hsync[0], t.hsync[1])) {
        kind = kHsync;
      }
    } else if (state == kVsync) {
      if (inrange(n, t.eq[0], t.eq[1])) {
        kind = kEq2;
        state_length = half;
      } else if (inrange(n, t.vsync[0], t.vsync[1])) {
        kind = kVsync;
      } else if ((double)q.start > state_end && inrange(n, t.hsync[0], t.hsync[1])) {
        kind = kHsync;
      }
    } else if (state == kEq2) {
      if (inrange(n, t.eq[0], t.eq[1])) {
        kind = kEq2;
      } else if (inrange(n, t.hsync[0], t.hsync[1])) {
        kind = kHsync;
        done = true;
      }
    }
    if (kind >= 0 && kind != state) {
      if ((double)q.start < state_end) {
        kind = -1;
      } else if (state_length != 0) {
        state_end = (double)q.start + (state_length - 0.1) * g.p.linelen;
        state_length = 0;
      }
    }
    if (kind >= 0) {
      const bool good = last_kind < 0 ? false : quality(g, last_kind, last_start, kind, q.start);
      valid.push_back({kind, q, good});
      last_kind = kind;
      last_start = q.start;
    }
    if (done) return true;
  }
  return done;
}

// Field.refinepulses (core.py:1914-1949).
void refinepulses(const Geometry& g, const Timings& t, const std::vector<Pulse>& raw, Scratch& sc) {
  std::vector<VPulse>& valid = sc.valid;
  std::vector<VPulse>& vb = sc.vb;
  valid.clear();
  valid.reserve(raw.size() + 32);
  // the last entry's kind and start, kept in registers rather than read back
  int last_kind = -1;
  int64_t last_start = 0;
  size_t i = 0;
  while (i < raw.size()) {
    const Pulse& cur = raw[i];
    const double n = (double)cur.len;
    if (inrange(n, t.hsync[0], t.hsync[1])) {
      const bool good = last_kind < 0 ? false : quality(g, last_kind, last_start, kHsync, cur.start);
      valid.push_back({kHsync, cur, good});
      last_kind = kHsync;
      last_start = cur.start;
      ++i;
    } else if (i > 2 && inrange(n, t.eq[0], t.eq[1]) && last_kind == kHsync) {
      const bool done = vblank_state_machine(g, t, raw, i - 2, std::min(raw.size(), i + 24), vb);
      if (done) {
        for (size_t k = 2; k < vb.size(); ++k) valid.push_back(vb[k]);
        if (vb.size() > 2) {
          last_kind = vb.back().kind;
          last_start = vb.back().p.start;
        }
        i += vb.size() - 2;
      } else {
        ++i;
      }
    } else {
      ++i;
    }
  }
}

// Python's seq[i] for -len <= i < len.
template <class T>
const T& pyat(const std::vector<T>& v, int64_t i) {
  return v[i < 0 ? (int64_t)v.size() + i : i];
}

// Field.getBlankRange (core.py:1952-1977).  Returns false for (None, None).
bool blank_range(const std::vector<VPulse>& v, int64_t start, int64_t& firstblank, int64_t& lastblank) {
  const int64_t n = (int64_t)v.size();
  int64_t firstvsync = -1;
  for (int64_t i = std::max<int64_t>(start, 0); i < n; ++i)
    if (v[i].kind == kVsync) {
      firstvsync = i;
      break;
    }
  if (firstvsync < 0 || firstvsync < 10) return false;
  for (int64_t ns = firstvsync - 10; ns < firstvsync - 4; ++ns) {
    int64_t fb = -1;
    for (int64_t i = ns; i < n; ++i)
      if (v[i].kind > 0) {
        fb = i;
        break;
      }
    if (fb < 0) continue;
    int64_t hs = -1;
    for (int64_t i = fb; i < n; ++i)
      if (v[i].kind == 0) {
        hs = i;
        break;
      }
    if (hs < 0) continue;
    const int64_t lb = hs - 1;
    if (lb - fb > 12) {
      firstblank = fb;
      lastblank = lb;
      return true;
    }
  }
  return false;
}

struct VBlank {
  bool found = false;         // line0 is not None
  double line0 = 0;
  bool isfirst = false;
  bool blank = false;         // getBlankRange found a range (firstblank valid)
  int64_t firstblank = -1;
  int conf = -1;              // -1: None
  bool zeroed = false;        // the fallback that sets sync_confidence = 0 (core.py:2082)
  bool error = false;         // the reference raises (np.max of an empty array)
};

// Field.processVBlank (core.py:1988-2088) without a limit; limited() applies
// one afterwards (the limit only rejects a range that starts too late).
VBlank process_vblank(const Geometry& g, const std::vector<VPulse>& v, int64_t start) {
  const ld_locate_params& p = g.p;
  VBlank r;
  int64_t fb, lb;
  if (!blank_range(v, start, fb, lb)) return r;
  r.blank = true;
  r.firstblank = fb;
  const int64_t loc_presync = pyat(v, fb - 1).p.start;
  const int64_t m = (int64_t)v.size() - fb;        // entries of pt / pstart / plen
  const int np = p.num_pulses;
  const int order[3] = {kVsync, kEq1, kEq2};
  for (int oi = 0; oi < 3; ++oi) {
    const int kind = order[oi];
    int64_t grouploc = -1;
    for (int64_t j = 0; j < lb - fb; ++j) {
      const int64_t e1 = std::min<int64_t>(j + np, m);
      bool all = true;
      for (int64_t k = j; k < e1; ++k) all = all && v[fb + k].kind == kind;
      if (!all) continue;
      const int64_t e2 = std::min<int64_t>(j + np + 4, m);
      int64_t cnt = 0;
      for (int64_t k = j; k < e2; ++k) cnt += v[fb + k].kind == kind;
      if (cnt != np) break;
      const int64_t L = std::max<int64_t>(e1 - j, 0);
      if (L < 3) {      // np.max of an empty second difference raises in the reference
        r.error = true;
        return r;
      }
      int64_t gmax = INT64_MIN, lmax = INT64_MIN;
      for (int64_t k = j; k + 2 < e1; ++k) {
        const int64_t s0 = v[fb + k].p.start, s1 = v[fb + k + 1].p.start, s2 = v[fb + k + 2].p.start;
        gmax = std::max(gmax, (s2 - s1) - (s1 - s0));
      }
      for (int64_t k = j; k + 1 < e1; ++k) lmax = std::max(lmax, v[fb + k + 1].p.len - v[fb + k].p.len);
      if ((double)gmax < p.freq * 0.2 && (double)lmax < p.freq * 0.2) {
        grouploc = j;
        break;
      }
    }
    if (grouploc < 0) continue;
    const int64_t firstloc = v[fb + grouploc].p.start;
    const double distfroml1 = ((kind - 1) * np) * 0.5;
    const double dist = (double)(firstloc - loc_presync) / p.linelen;
    const int64_t hdist = np_round(dist * 2);
    bool isfirst = (((hdist % 2) + 2) % 2) == (p.first_field_h[1] != 1 ? 1 : 0);
    if (std::fmod(distfroml1 * 2, 2.0) != 0) isfirst = !isfirst;
    const double eqgap = p.first_field_h[isfirst ? 1 : 0];
    const double line0 = (double)firstloc - ((eqgap + distfroml1) * p.linelen);
    r.found = true;
    r.line0 = std::trunc(line0);
    r.isfirst = isfirst;
    r.conf = 100;
    return r;
  }
  if (pyat(v, fb - 1).good && v[fb].good && v[lb].good && v[lb + 1].good) {
    const double gap1 = (double)(v[fb].p.start - pyat(v, fb - 1).p.start);
    const double gap2 = (double)(v[lb + 1].p.start - v[lb].p.start);
    bool isfirst;
    if (p.pal && inrange(std::fabs(gap2 - gap1), 0, p.freq * 1)) {
      isfirst = inrange(gap1 / p.linelen, 0.45, 0.55);
    } else if (!p.pal && inrange(std::fabs(gap2 + gap1), p.linelen * 1.4, p.linelen * 1.6)) {
      isfirst = inrange(gap1 / p.linelen, 0.95, 1.05);
    } else {
      r.conf = 0;
      r.zeroed = true;
      return r;
    }
    r.found = true;
    r.line0 = (double)pyat(v, fb - 1).p.start;
    r.isfirst = isfirst;
    r.conf = 50;
    return r;
  }
  r.conf = 0;
  return r;
}

// processVBlank(validpulses, start, limit): a range starting past
// start + limit gives (None, None, None, None) before anything else.
VBlank limited(const VBlank& vb, int64_t start, int64_t limit) {
  if (limit >= 0 && vb.blank && vb.firstblank > start + limit) return VBlank();
  return vb;
}

// Field.computeLineLen (core.py:2091-2119).
double compute_line_len(const Geometry& g, const std::vector<VPulse>& v) {
  int64_t lr0 = -1, lr1 = -1;     // longrun
  int64_t c0 = -1, c1 = 0;        // currun (c0 < 0: None)
  for (int64_t i = 0; i < (int64_t)v.size(); ++i) {
    if (v[i].kind != 0) {
      if (c0 >= 0 && c1 > lr1) lr0 = c0, lr1 = c1;
      c0 = -1;
    } else if (c0 < 0) {
      c0 = i, c1 = 0;
    } else {
      ++c1;
    }
  }
  if (c0 >= 0 && c1 > lr1) lr0 = c0, lr1 = c1;
  double sum = 0;
  int64_t cnt = 0;
  for (int64_t i = lr0 + 1; i < lr0 + lr1; ++i) {
    const int64_t ll = v[i].p.start - v[i - 1].p.start;
    if (g.line_ok.has(ll)) {        // inrange(ll / linelen, 0.95, 1.05)
      sum += (double)ll;
      ++cnt;
    }
  }
  // np.mean of Python ints: a float64 pairwise sum; line lengths are exact
  // integers far below 2^53, so any summation order gives the same total
  return cnt ? sum / (double)cnt : g.p.linelen;
}

// Field.skip_check's classification (core.py:2139-2155).
int skip_score(const ld_locate_params& p, const double* med) {
  int score = 0, vsync_lines = 0;
  for (int j = 0; j < 8; ++j) {
    const double ire = (med[j] - p.ire0) / p.hz_ire;
    if (inrange(ire, p.vsync_ire - 10, p.vsync_ire / 2))
      ++vsync_lines;
    else if (inrange(ire, -5, 5))
      ++score;
    else
      --score;
  }
  if (vsync_lines >= 2) return 100;
  if (vsync_lines == 1 && score > 0) return 50;
  if (score > 0) return 25;
  return 0;
}

struct Line0 {
  bool found = false;
  double line0 = 0;
  bool has_next = false;
  double vblank_next = 0;
  bool isfirst = false;
  int conf = 100;
  bool error = false;
};

// Field.getLine0 (core.py:2157-2239) from processVBlank's unlimited results
// for the local vertical blanking and the next one (nx, computed from
// loc0.firstblank + 40 when loc0 was found).
Line0 get_line0(const Geometry& g, const VBlank& loc0, const VBlank& nx0, double meanlinelen,
                const ld_prev_field* prev, int64_t startloc) {
  const ld_locate_params& p = g.p;
  Line0 r;
  r.conf = 100;
  const int64_t limit = (prev && prev->valid && prev->skip_score >= 50) ? 100 : -1;
  const VBlank loc = limited(loc0, 0, limit);
  if (loc.error) {
    r.error = true;
    return r;
  }
  if (loc.zeroed) r.conf = 0;
  bool have_next = false, first_next = false;
  double l0_next = 0;
  int conf_next = -1;
  if (loc.found) {
    const VBlank& nx = nx0;
    if (nx.error) {
      r.error = true;
      return r;
    }
    if (nx.zeroed) r.conf = 0;
    conf_next = nx.conf;
    if (nx.found) {
      r.has_next = true;
      r.vblank_next = nx.line0;
      first_next = !nx.isfirst;
      const double fieldlen = meanlinelen * p.field_lines[first_next ? 0 : 1];
      l0_next = (double)np_round(nx.line0 - fieldlen);
      have_next = true;
      if (l0_next < 0) r.conf = 10;
    }
  }
  bool have_prev = false, first_prev = false;
  double l0_prev = 0;
  int conf_prev = 0;
  if (prev && prev->valid) {
    // prevfield.linelocs[linecount] - frameoffset (core.py:2199-2203)
    l0_prev = prev->ll_linecount - (double)(startloc - prev->startloc);
    first_prev = !prev->is_first_field;
    conf_prev = prev->sync_confidence;
    have_prev = true;
  }
  if (loc.found && have_next && have_prev) {
    double a[3] = {loc.line0, l0_next, l0_prev};
    std::sort(a, a + 3);
    r.found = true;
    r.line0 = a[1];
    r.isfirst = ((int)loc.isfirst + (int)first_prev + (int)first_next) >= 2;
    return r;
  }
  if (loc.found && loc.conf > 50) {
    r.found = true;
    r.line0 = loc.line0;
    r.isfirst = loc.isfirst;
    r.conf = std::min(r.conf, 90);
    return r;
  }
  if (have_prev) {
    r.found = true;
    r.line0 = l0_prev;
    r.isfirst = first_prev;
    r.conf = std::min(r.conf, std::max(conf_prev - 10, 10));
    return r;
  }
  if (have_next) {
    r.found = true;
    r.line0 = l0_next;
    r.isfirst = first_next;
    r.conf = conf_next;
    return r;
  }
  return r;
}

// getLine0's inputs kept in ld_field_loc, so that a walk can be checked
// against measured previous-field values without analysing the pulses again.
void store_vblank(const VBlank& vb, ld_field_loc* out, int i) {
  out->vb_line0[i] = vb.line0;
  out->vb_firstblank[i] = vb.firstblank;
  out->vb_conf[i] = vb.conf;
  out->vb_flags[i] = (vb.found ? LD_VB_FOUND : 0) | (vb.isfirst ? LD_VB_FIRST : 0) |
                     (vb.zeroed ? LD_VB_ZEROED : 0) | (vb.blank ? LD_VB_BLANK : 0) | (vb.error ? LD_VB_ERROR : 0);
}

VBlank load_vblank(const ld_field_loc& f, int i) {
  VBlank vb;
  vb.line0 = f.vb_line0[i];
  vb.firstblank = f.vb_firstblank[i];
  vb.conf = f.vb_conf[i];
  vb.found = f.vb_flags[i] & LD_VB_FOUND;
  vb.isfirst = f.vb_flags[i] & LD_VB_FIRST;
  vb.zeroed = f.vb_flags[i] & LD_VB_ZEROED;
  vb.blank = f.vb_flags[i] & LD_VB_BLANK;
  vb.error = f.vb_flags[i] & LD_VB_ERROR;
  return vb;
}

// Field.compute_linelocs (core.py:2325-2481) after getpulses.  Returns
// LD_OK / an LD_E* error; out->status says whether the field is valid.
// compute_linelocs from getLine0's result on: the line table, its gap fill
// and the next field's offset (core.py:2338-2481).  valid: refinepulses'
// list; last_start: rawpulses[-1].start.
int assign(const Geometry& g, const std::vector<VPulse>& valid, int64_t last_start, double meanlinelen,
           const Line0& l0, ld_field_loc* out, double* ll_out, uint8_t* bad_out, Scratch& sc) {
  const ld_locate_params& p = g.p;
  out->is_first_field = l0.isfirst ? 1 : 0;
  out->linecount = l0.isfirst ? 263 : 262;
  out->sync_confidence = l0.conf;
  out->vblank_next = l0.has_next ? (int64_t)l0.vblank_next : -1;
  out->line0loc = l0.found ? l0.line0 : NAN;
  out->status = LD_LOC_FAILED;
  out->nlines = 0;
  const int proclines = g.proclines;
  bool skipdetected = false;
  if (l0.has_next) skipdetected = (l0.vblank_next - l0.line0) / p.linelen < (out->linecount - 5);
  out->skipdetected = skipdetected;
  if (!l0.found) {
    out->nextfieldoffset = p.linelen * 200;
    return LD_OK;
  }
  const double line0 = l0.line0;
  const double lastline = ((double)last_start - line0) / meanlinelen;
  if (lastline < proclines) {
    out->nextfieldoffset = line0 - meanlinelen * 20;
    return LD_OK;
  }
  // line number -> (start, distance) for lines [0, proclines); others are
  // recorded by the reference too but never read
  std::vector<int64_t>& loc = sc.loc;
  std::vector<double>& dist = sc.dist;
  std::vector<uint8_t>& have = sc.have;
  loc.assign(proclines, -1);
  dist.assign(proclines, 0.0);
  have.assign(proclines, 0);
  for (const VPulse& vp : valid) {
    double lineloc = ((double)vp.p.start - line0) / meanlinelen;
    int64_t rl = np_round(lineloc);
    double d = std::fabs(lineloc - (double)rl);
    if (skipdetected) {
      const double le = out->linecount - ((l0.vblank_next - (double)vp.p.start) / meanlinelen);
      const int64_t rle = np_round(le);
      const double de = std::fabs(le - (double)rle);
      if (vp.kind == 0 && rl > 23 && de < d) {
        lineloc = le;
        rl = rle;
        d = de;
      }
    }
    const bool in = rl >= 0 && rl < proclines;
    if (d > p.hsync_tolerance || (in && have[rl] && d > dist[rl])) continue;
    if (rl > 0 && !vp.good) {
      if (vp.kind > 0 || (vp.kind == 0 && rl < 10)) continue;
    }
    if (in) {
      loc[rl] = vp.p.start;
      dist[rl] = d;
      have[rl] = 1;
    }
  }
  const int oc = g.outlinecount;
  std::vector<double>& fill = sc.fill;
  fill.resize(proclines);
  for (int l = 0; l < proclines; ++l) fill[l] = (double)loc[l];
  auto first_valid_from = [&](int l, int stop) {
    for (int i = l; i < stop; ++i)
      if (loc[i] > 0) return i;
    return -1;
  };
  if (fill[0] < 0) {
    const int nv = first_valid_from(0, std::min(oc + 1, proclines));
    if (nv < 0) {
      out->nextfieldoffset = line0 + (p.linelen * oc - 7);
      return LD_OK;
    }
    fill[0] = fill[nv] - nv * meanlinelen;
    if (fill[0] < p.linelen) {
      out->nextfieldoffset = line0 + (p.linelen * oc - 7);
      return LD_OK;
    }
  }
  std::vector<uint8_t>& err = sc.err;
  err.assign(proclines, 0);
  for (int l = 1; l < proclines; ++l) {
    if (fill[l] >= 0) continue;
    err[l] = 1;
    int pv = -1;
    for (int i = l; i >= 0; --i)
      if (loc[i] > 0) {
        pv = i;
        break;
      }
    const int nv = first_valid_from(l, std::min(oc + 1, proclines));
    if (pv < 0) {
      if (nv < 0) {
        set_error("ld_locate: line %d has no located line on either side (the reference raises)", l);
        return LD_EINVAL;
      }
      fill[l] = (double)loc[nv] - p.linelen * (nv - l);
    } else if (nv >= 0) {
      const double avg = (double)(loc[nv] - loc[pv]) / (nv - pv);
      fill[l] = (double)loc[pv] + avg * (l - pv);
    } else {
      fill[l] = (double)loc[pv] + p.linelen * (l - pv);
    }
  }
  for (int l = 0; l < proclines; ++l) {
    ll_out[l] = fill[l];
    bad_out[l] = err[l];
  }
  out->nlines = proclines;
  out->status = LD_LOC_OK;
  out->nextfieldoffset = l0.has_next ? l0.vblank_next - p.linelen * 8 : fill[oc - 7];
  return LD_OK;
}

// Field.compute_linelocs (core.py:2325-2481) after getpulses.  Returns
// LD_OK / an LD_E* error; out->status says whether the field is valid.
// When keep is non-NULL (room for keep_cap entries) refinepulses' list is
// copied there for ld_locate_recheck.
int locate(const Geometry& g, const std::vector<Pulse>& pulses, const ld_prev_field* prev, int64_t startloc,
           ld_field_loc* out, double* ll_out, uint8_t* bad_out, Scratch& sc, ld_vpulse* keep, int keep_cap) {
  const Timings t = get_timings(g, pulses, sc);
  out->hsync_max = t.hsync[1];
  refinepulses(g, t, pulses, sc);
  const std::vector<VPulse>& valid = sc.valid;
  const double meanlinelen = compute_line_len(g, valid);
  const VBlank loc0 = process_vblank(g, valid, 0);
  const VBlank nx0 = loc0.found ? process_vblank(g, valid, loc0.firstblank + 40) : VBlank();
  store_vblank(loc0, out, 0);
  store_vblank(nx0, out, 1);
  out->analyzed = 1;
  out->last_start = pulses.back().start;
  out->nvalid = -1;
  if (keep && (int64_t)valid.size() <= keep_cap) {
    for (size_t i = 0; i < valid.size(); ++i)
      keep[i] = {valid[i].p.start, (int32_t)valid[i].p.len, (int16_t)valid[i].kind, (int16_t)valid[i].good};
    out->nvalid = (int32_t)valid.size();
  }
  const Line0 l0 = get_line0(g, loc0, nx0, meanlinelen, prev, startloc);
  if (l0.error) {
    set_error("ld_locate: processVBlank would raise (np.max of an empty array) at buffer %lld",
              (long long)startloc);
    return LD_EINVAL;
  }
  out->npulses = (int32_t)pulses.size();
  out->meanlinelen = meanlinelen;
  return assign(g, valid, out->last_start, meanlinelen, l0, out, ll_out, bad_out, sc);
}

int check_params(const ld_locate_params* p, const char* who) {
  if (!p || !(p->freq > 0) || !(p->linelen > 0) || p->num_pulses <= 0 || p->frame_lines <= 0 ||
      !(p->hz_ire != 0) || p->block_stride <= 0 || p->blocklen <= 0 || p->readlen < 0) {
    set_error("%s: bad parameters", who);
    return LD_EINVAL;
  }
  return LD_OK;
}

// float32 arithmetic of getpulses' sync-level test (core.py:2288-2292):
// |hztoire(median) - vsync_ire| < 5 with NumPy keeping float32.
bool sync_level_ok(const ld_locate_params& p, std::vector<float>& means) {
  const size_t n = means.size(), m = n / 2;
  std::sort(means.begin(), means.end());
  float level = (n % 2) ? means[m] : (means[m - 1] + means[m]) / 2.0f;
  const float ire = (level - (float)p.ire0) / (float)p.hz_ire;
  return std::fabs(ire - (float)p.vsync_ire) < 5.0f;
}

// The state field f leaves for its successor (what getLine0 reads of it):
// measured facts when known, else the provisional estimate.
void next_prev(const ld_locate_params& p, const ld_field_loc& f, const double* ll, const ld_field_facts* fact,
               ld_prev_field& st) {
  const int lc = p.pal ? (f.is_first_field ? 312 : 313) : f.linecount;
  st.valid = 1;
  st.is_first_field = f.is_first_field;
  st.linecount = lc;
  st.sync_confidence = f.sync_confidence;
  st.startloc = f.startloc;
  if (fact && fact->known) {
    st.ll_linecount = fact->ll_linecount;
    st.skip_score = skip_score(p, fact->medians);
  } else {
    st.ll_linecount = ll[lc] + p.prov_delta;
    st.skip_score = 100;
  }
}

}  // namespace

extern "C" int ld_locate_recheck(const ld_locate_params* p, ld_field_loc* locs, double* linelocs1,
                                 uint8_t* linebad, int max_lines, const ld_vpulse* vpulses, int max_vpulses, int n,
                                 const ld_prev_field* prev, const ld_field_facts* facts, int* first_changed) {
  int rc = check_params(p, "ld_locate_recheck");
  if (rc != LD_OK) return rc;
  const Geometry g(*p);
  if (!locs || n < 0 || !linelocs1 || !linebad || max_lines < g.proclines || !first_changed ||
      (!vpulses && n > 0) || max_vpulses < 0) {
    set_error("ld_locate_recheck: bad arguments");
    return LD_EINVAL;
  }
  ld_prev_field st{};
  if (prev) st = *prev;
  const ld_prev_field* pv = (prev && prev->valid) ? &st : nullptr;
  *first_changed = -1;
  Scratch sc;
  std::vector<VPulse> valid;
  std::vector<double> ll(g.proclines);
  std::vector<uint8_t> bad(g.proclines);
  for (int k = 0; k < n; ++k) {
    ld_field_loc& f = locs[k];
    double* fll = linelocs1 + (int64_t)k * max_lines;
    uint8_t* fbad = linebad + (int64_t)k * max_lines;
    if (f.analyzed) {
      const Line0 l0 = get_line0(g, load_vblank(f, 0), load_vblank(f, 1), f.meanlinelen, pv, f.startloc);
      const bool found = !std::isnan(f.line0loc);
      const bool same_decision = !l0.error && l0.found == found && (!found || l0.line0 == f.line0loc) &&
                                 (int)l0.isfirst == f.is_first_field && l0.conf == f.sync_confidence &&
                                 (l0.has_next ? (int64_t)l0.vblank_next : -1) == f.vblank_next;
      if (!same_decision) {
        if (l0.error || f.nvalid < 0) {
          *first_changed = k;
          return LD_OK;
        }
        // the line table from the new line 0: same outputs -> keep the walk
        const ld_vpulse* kv = vpulses + (int64_t)k * max_vpulses;
        valid.resize(f.nvalid);
        for (int i = 0; i < f.nvalid; ++i) valid[i] = {kv[i].kind, {kv[i].start, (int64_t)kv[i].len}, kv[i].good != 0};
        ld_field_loc g2 = f;
        rc = assign(g, valid, f.last_start, f.meanlinelen, l0, &g2, ll.data(), bad.data(), sc);
        if (rc != LD_OK) return rc;
        bool same = g2.status == f.status && g2.is_first_field == f.is_first_field && g2.linecount == f.linecount &&
                    g2.sync_confidence == f.sync_confidence && g2.nextfieldoffset == f.nextfieldoffset;
        if (same && f.status == LD_LOC_OK)
          for (int l = 0; l < g.proclines && same; ++l) same = ll[l] == fll[l] && bad[l] == fbad[l];
        if (!same) {
          *first_changed = k;
          return LD_OK;
        }
        f.line0loc = g2.line0loc;
        f.vblank_next = g2.vblank_next;
        f.skipdetected = g2.skipdetected;
      }
    }
    if (f.status == LD_LOC_OK) {
      next_prev(*p, f, fll, facts ? &facts[k] : nullptr, st);
      pv = &st;
    } else {
      st.valid = 0;
      pv = nullptr;
    }
  }
  return LD_OK;
}

extern "C" int ld_skip_score(const ld_locate_params* p, const double* medians8) {
  if (check_params(p, "ld_skip_score") != LD_OK || !medians8) return LD_EINVAL;
  return skip_score(*p, medians8);
}

extern "C" int ld_locate_field(const ld_locate_params* p, const int64_t* starts, const int64_t* lens, int npulses,
                               const ld_prev_field* prev, int64_t startloc, int fields_written, ld_field_loc* out,
                               double* linelocs1, uint8_t* linebad, int max_lines) {
  int rc = check_params(p, "ld_locate_field");
  if (rc != LD_OK) return rc;
  const Geometry g(*p);
  if (!out || !linelocs1 || !linebad || max_lines < g.proclines || npulses < 0 ||
      (npulses > 0 && (!starts || !lens))) {
    set_error("ld_locate_field: bad arguments (max_lines %d, need %d)", max_lines, g.proclines);
    return LD_EINVAL;
  }
  std::memset(out, 0, sizeof(*out));
  out->startloc = startloc;
  out->vblank_next = -1;
  if (npulses == 0) {
    // compute_linelocs: no pulses -> skip a field (or a second before any
    // field was written); process() turns a missing offset into 200 lines
    out->status = LD_LOC_FAILED;
    out->line0loc = NAN;
    out->nextfieldoffset = fields_written ? p->linelen * 200 : (double)(int64_t)(p->freq * 1e6);
    return LD_OK;
  }
  Scratch sc;
  std::vector<Pulse> pl(npulses);
  for (int i = 0; i < npulses; ++i) pl[i] = {starts[i], lens[i]};
  return locate(g, pl, prev, startloc, out, linelocs1, linebad, sc, nullptr, 0);
}

extern "C" int ld_locate_fields(const ld_locate_params* p, const int64_t* edges, int64_t nedges, const float* means,
                                int64_t total, double start, const ld_prev_field* prev, const ld_field_facts* facts,
                                int max_fields, ld_field_loc* out, double* linelocs1, uint8_t* linebad,
                                int max_lines, ld_vpulse* vpulses, int max_vpulses, ld_prev_field* prev_out,
                                int* nfields_out) {
  int rc = check_params(p, "ld_locate_fields");
  if (rc != LD_OK) return rc;
  const Geometry g(*p);
  if (!edges || nedges < 0 || !means || total <= 0 || max_fields < 0 || !out || !linelocs1 || !linebad ||
      max_lines < g.proclines || !nfields_out || !(start >= 0)) {
    set_error("ld_locate_fields: bad arguments (max_lines %d, need %d)", max_lines, g.proclines);
    return LD_EINVAL;
  }
  const int64_t npairs = nedges / 2;
  const double minlength = g.us(10);
  const int64_t S = p->block_stride;
  const int64_t length = (p->readlen / p->blocklen + 2) * p->blocklen;
  ld_prev_field st{};
  if (prev) st = *prev;
  const ld_prev_field* pv = (prev && prev->valid) ? &st : nullptr;
  int written = prev ? prev->fields_written : 0;
  int nf = 0;
  Scratch sc;
  std::vector<Pulse>& pl = sc.pulses;
  std::vector<float>& vm = sc.means;
  while (nf < max_fields) {
    // decodefield's read (core.py:4030-4041) and DemodCache.read (:1358-1409)
    const int64_t readloc = std::max<int64_t>((int64_t)(start - (double)p->blockcut), 0);
    const int64_t begin = (readloc / p->blocklen) * p->blocklen;
    const int64_t b0 = begin / S, b1 = (begin + length) / S;
    const int64_t ws = b0 * S, we = (b1 + 1) * S;
    if (we > total) break;
    ld_field_loc& o = out[nf];
    double* ll = linelocs1 + (int64_t)nf * max_lines;
    uint8_t* bad = linebad + (int64_t)nf * max_lines;
    // the buffer's pulses: start after its first sample, end inside it
    pl.clear();
    vm.clear();
    int64_t lo = 0, hi = npairs;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (edges[2 * mid] <= ws) lo = mid + 1; else hi = mid;
    }
    for (int64_t k = lo; k < npairs && edges[2 * k + 1] < we; ++k) {
      const int64_t len = edges[2 * k + 1] - edges[2 * k];
      if (len <= 0 || (k > 0 && edges[2 * k] <= edges[2 * k - 1])) {
        set_error("ld_locate_fields: edges not increasing at %lld", (long long)(2 * k));
        return LD_EINVAL;
      }
      if (len > kMaxSyncLen) continue;
      pl.push_back({edges[2 * k] - ws, len});
      if ((double)len > minlength) vm.push_back(means[k]);
    }
    bool handled = false;
    std::memset(&o, 0, sizeof(o));
    if (pl.empty()) {
      if (!written) {
        o.status = LD_LOC_NEED_IRE0;
        handled = true;
      }
    } else if (vm.empty()) {
      // getpulses returns None: compute_linelocs skips a field, or a second
      o.status = LD_LOC_FAILED;
      o.line0loc = NAN;
      o.vblank_next = -1;
      o.nextfieldoffset = written ? p->linelen * 200 : (double)(int64_t)(p->freq * 1e6);
      o.npulses = (int32_t)pl.size();
      handled = true;
    } else if (!sync_level_ok(*p, vm)) {
      o.status = LD_LOC_NEED_LEVELS;
      handled = true;
    }
    if (!handled) {
      if (pl.empty()) {
        o.status = LD_LOC_FAILED;
        o.line0loc = NAN;
        o.vblank_next = -1;
        o.nextfieldoffset = p->linelen * 200;
      } else {
        rc = locate(g, pl, pv, ws, &o, ll, bad, sc, vpulses ? vpulses + (int64_t)nf * max_vpulses : nullptr,
                    max_vpulses);
        if (rc != LD_OK) return rc;
      }
    }
    o.start = start;
    o.readloc = readloc;
    o.startloc = ws;
    o.buflen = we - ws;
    ++nf;
    if (o.status == LD_LOC_NEED_LEVELS || o.status == LD_LOC_NEED_IRE0) break;
    if (o.status == LD_LOC_OK) {
      o.offset = o.nextfieldoffset - (double)(readloc - ws);
      next_prev(*p, o, ll, facts ? &facts[nf - 1] : nullptr, st);
      ++written;
      st.fields_written = written;
      pv = &st;
    } else {
      o.offset = o.nextfieldoffset;
      st.valid = 0;
      st.fields_written = written;
      pv = nullptr;
    }
    // readfield: fdoffset += offset (core.py:4187-4188), a float
    start = start + o.offset;
  }
  *nfields_out = nf;
  if (prev_out) *prev_out = st;
  return LD_OK;
}

// ------------------------------------------------------------- dropouts
extern "C" int ld_dropout_list(const int64_t* errmap, int64_t n, const double* ll, int nlines, int lineoffset,
                               int linecount, int outlinelen, double freq, int32_t* lines, int32_t* starts,
                               int32_t* ends, int cap, int* count) {
  if (!count || n < 0 || (n > 0 && !errmap) || !ll || lineoffset < 0 || linecount < 1 ||
      lineoffset + linecount + 1 > nlines || outlinelen <= 0 || cap < 0 || (cap > 0 && (!lines || !starts || !ends))) {
    set_error("ld_dropout_list: bad arguments");
    return LD_EINVAL;
  }
  *count = 0;
  if (n == 0 || !((double)errmap[n - 1] > ll[lineoffset])) return LD_OK;
  // build_errlist (core.py:2890-2908)
  int64_t fi = 0;
  while ((double)errmap[fi] < ll[lineoffset]) ++fi;
  const double firsterr = (double)errmap[fi];
  std::vector<std::pair<double, double>> el;
  double c0 = firsterr, c1 = firsterr;
  for (int64_t i = 0; i < n; ++i) {
    const double e = (double)errmap[i];
    if (e > c0 && e <= c1 + 20) {
      double pad = (e - c0) * 1.7;
      pad = std::min(pad, freq * 12);
      c1 = c0 + pad;
    } else if (e > firsterr) {
      el.push_back({c0 - 8, c1 + 4});
      c0 = c1 = e;
    }
  }
  el.push_back({c0, c1});
  // dropout_errlist_to_tbc (core.py:2911-2975)
  size_t idx = 0;
  std::pair<double, double> cur = el[idx++];
  bool have = true;
  while (idx < el.size() && cur.first < ll[lineoffset]) cur = el[idx++];
  int k = 0;
  auto emit = [&](int64_t line, int64_t s, int64_t e) {
    if (k < cap) {
      lines[k] = (int32_t)(line - 1);
      starts[k] = (int32_t)s;
      ends[k] = (int32_t)e;
    }
    ++k;
  };
  for (int line = lineoffset; line < linecount + lineoffset; ++line) {
    while (have && cur.first >= ll[line] && cur.first <= ll[line + 1]) {
      const double span = ll[line + 1] - ll[line];
      const int64_t sp = (int64_t)(((cur.first - ll[line]) / span) * outlinelen);
      const int64_t ep = np_round(((cur.second - ll[line]) / span) * outlinelen);
      const int64_t first_line = line + 1 - lineoffset;
      if (ep > outlinelen) {
        const int64_t num = ep / outlinelen;
        emit(first_line, sp, outlinelen);
        for (int64_t q = 0; q < num - 1; ++q) emit(first_line + q + 1, 0, outlinelen);
        emit(first_line + num, 0, ep % outlinelen);
      } else {
        emit(first_line, sp, ep);
      }
      if (idx < el.size()) cur = el[idx++]; else have = false;
    }
  }
  *count = k;
  return LD_OK;
}

// Many independent walks (one per capture of a batch) on a few host threads:
// each job is one ld_locate_fields call; job i's return code goes to
// jobs[i].status.  Returns the first failing job's code (its message in
// ld_last_error) or LD_OK.
// The walks of a batch on host threads, longest job (most pulses) first so
// that the last threads to finish do not start a long walk late.
struct WalkRun {
  ld_walk_job* jobs = nullptr;
  int njobs = 0;
  std::vector<int> order;
  std::vector<std::string> errs;
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  void work() {
    for (int i = next++; i < njobs; i = next++) {
      ld_walk_job& j = jobs[order[(size_t)i]];
      j.nfields = 0;
      j.status = ld_locate_fields(j.p, j.edges, j.nedges, j.means, j.total, j.start, j.prev, j.facts,
                                  j.max_fields, j.out, j.linelocs1, j.linebad, j.max_lines, j.vpulses,
                                  j.max_vpulses, j.prev_out, &j.nfields);
      if (j.status != LD_OK) errs[(size_t)order[(size_t)i]] = ld_last_error();
    }
  }
};

extern "C" int ld_locate_fields_many_start(ld_walk_job* jobs, int njobs, int nthreads, void** run) {
  if (njobs < 0 || (njobs > 0 && !jobs) || !run) {
    set_error("ld_locate_fields_many_start: bad arguments");
    return LD_EINVAL;
  }
  WalkRun* r = new WalkRun;
  r->jobs = jobs;
  r->njobs = njobs;
  r->errs.resize((size_t)njobs);
  r->order.resize((size_t)njobs);
  for (int i = 0; i < njobs; ++i) r->order[(size_t)i] = i;
  std::stable_sort(r->order.begin(), r->order.end(),
                   [jobs](int a, int b) { return jobs[a].nedges > jobs[b].nedges; });
  const int nt = std::max(1, std::min(nthreads, njobs));
  try {
    for (int t = 0; t < nt; ++t) r->pool.emplace_back([r] { r->work(); });
  } catch (...) {
    r->work();          // no more threads: the calling thread finishes the rest
  }
  *run = r;
  return LD_OK;
}

extern "C" int ld_locate_fields_many_wait(void* run) {
  if (!run) {
    set_error("ld_locate_fields_many_wait: no run");
    return LD_EINVAL;
  }
  WalkRun* r = static_cast<WalkRun*>(run);
  for (auto& t : r->pool) t.join();
  int rc = LD_OK;
  for (int i = 0; i < r->njobs; ++i)
    if (r->jobs[i].status != LD_OK) {
      set_error("ld_locate_fields_many: job %d: %s", i, r->errs[(size_t)i].c_str());
      rc = r->jobs[i].status;
      break;
    }
  delete r;
  return rc;
}

extern "C" int ld_locate_fields_many(ld_walk_job* jobs, int njobs, int nthreads) {
  void* run = nullptr;
  const int rc = ld_locate_fields_many_start(jobs, njobs, nthreads, &run);
  if (rc != LD_OK) return rc;
  return ld_locate_fields_many_wait(run);
}
