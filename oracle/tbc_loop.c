/* CPU restatement of the reference's field resampler inner loop
 * (scale_field, utils.py:151-183).  Test infrastructure: a second
 * implementation of oracle/tbc.py:scale_field (bit-identical to it, checked
 * by the CPU tests) that runs at compiled speed, used as the host-side TBC in
 * the benchmark's CPU baseline.  Not part of the product.
 *
 * Per output pixel q (pixel index i = start + q):
 *   coord = pixel_locs[i] (float32 when coord_f32, as utils.py:160 casts it)
 *   ci = int(coord), phase = frac * 65536, p0 = int(phase), alpha = phase - p0
 *   w[t] = lut[p0][t] + alpha * (lut[p0+1][t] - lut[p0][t])   (float32 result)
 *   out[q] = level[i] * sum_t float32(buf[ci - 7 + t] * w[t])  (float64 sum, tap order)
 * Taps outside [0, buflen) read 0; a phase outside the table (negative
 * coordinates) uses row 0 with alpha 0, as the GPU resampler does.
 */
#include <math.h>
#include <stdint.h>

/* built with -ffp-contract=off: every product and sum rounds on its own */


#define TAPS 16
#define PHASES 65536

int ora_scale_field(const float* buf, int64_t buflen, const double* pixel_locs, const double* level,
                    const float* lut, int64_t start, int64_t nout, int coord_f32, double* out) {
  for (int64_t q = 0; q < nout; ++q) {
    const int64_t i = start + q;
    const double loc = pixel_locs[i];
    int64_t ci, p0;
    float alpha;
    if (coord_f32) {
      const float c = (float)loc;
      ci = (int64_t)c;
      const float frac = c - (float)ci;
      const float ph = frac * (float)PHASES;
      p0 = (int64_t)ph;
      alpha = ph - (float)p0;
    } else {
      ci = (int64_t)floor(loc);
      const double frac = loc - (double)ci;
      const double ph = frac * (double)PHASES;
      p0 = (int64_t)ph;
      alpha = (float)(ph - (double)p0);
    }
    if (p0 < 0 || p0 >= PHASES) {
      p0 = 0;
      alpha = 0.0f;
    }
    const float* w0 = lut + p0 * TAPS;
    const float* w1 = w0 + TAPS;
    double res = 0.0;
    for (int t = 0; t < TAPS; ++t) {
      const float d = w1[t] - w0[t];
      /* float64 evaluation of w0 + alpha*d, rounded once to float32 */
      const float w = (float)((double)w0[t] + (double)alpha * (double)d);
      const int64_t k = ci - (TAPS / 2 - 1) + t;
      const float s = (k >= 0 && k < buflen) ? buf[k] : 0.0f;
      const float prod = s * w;
      res += (double)prod;
    }
    out[q] = level[i] * res;
  }
  return 0;
}
