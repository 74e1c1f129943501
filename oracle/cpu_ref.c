/*
 * C/OpenMP restatement of the reference GLASS CPU path -- TEST INFRASTRUCTURE.
 *
 * Used only as the CPU baseline of bench.py (and checked against the numpy
 * oracle in tests/test_cpu_ref.py); never linked into the product library.
 * Follows the same reference files as the numpy oracle:
 *   calibrate:   src/glass/cpu/calibration.py:51-76 (f32 ops, no contraction)
 *   warp:        src/glass/engine/warp.py:155-228, :99-118, :121-152
 *                (FP64 coordinates and Lanczos3 weights, FP64 sums)
 *   integrate:   src/glass/cpu/integration.py:36-118 +
 *                src/glass/engine/rejection.py:242-320 (per-pixel valid-count
 *                arithmetic; linear percentiles; FP64 sums)
 * Build: gcc -O3 -ffp-contract=off -fopenmp -shared -fPIC (oracle/cpu_ref.py).
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Worker threads of the OpenMP loops below (<= 0: the runtime default). */
void cpu_set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

static float decode_be(const uint8_t* p) { return (float)((((unsigned)p[0] << 8) | p[1]) ^ 0x8000u); }

/* raw: n frames of npix FITS u16 BE/BZERO samples; out: n*npix f32 */
void cpu_calibrate_u16be(const uint8_t* raw, int64_t n, int64_t npix, const float* bias, const float* dark,
                         const float* flat, float dark_scale, int dark_includes_bias, float flat_floor,
                         float pedestal, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < n; ++f) {
    for (int64_t i = 0; i < npix; ++i) {
      float v = decode_be(raw + 2 * (f * npix + i));
      if (dark) {
        v = v - dark[i] * dark_scale;
        if (!dark_includes_bias && bias) v = v - bias[i];
      } else if (bias) {
        v = v - bias[i];
      }
      if (flat) {
        float d = flat[i] < flat_floor ? flat_floor : flat[i];
        v = v / d;
      }
      if (pedestal != 0.0f) v = v + pedestal;
      out[f * npix + i] = v;
    }
  }
}

static double sinc(double x) {
  if (fabs(x) < 1.0e-8) return 1.0;
  const double pix = M_PI * x;
  return sin(pix) / pix;
}

static double lanczos3(double x) {
  x = fabs(x);
  if (x >= 3.0) return 0.0;
  return sinc(x) * sinc(x / 3.0);
}

/* One frame; inv = row-major FP64 inverse.  interp 1 bilinear, 3 lanczos3.
 * out gets fill outside the footprint; cov 0/1. */
void cpu_warp(const float* src, int64_t h, int64_t w, const double* inv, int interp, float fill, float* out,
              float* cov) {
  /* pixels in chunks, so a band of few rows still spreads over every thread */
#pragma omp parallel for schedule(dynamic, 2048)
  for (int64_t o = 0; o < h * w; ++o) {
    {
      const int64_t y = o / w, x = o - y * w;
      double den = inv[6] * (double)x + inv[7] * (double)y;
      den = den + inv[8];
      double sx = inv[0] * (double)x + inv[1] * (double)y;
      sx = sx + inv[2];
      double sy = inv[3] * (double)x + inv[4] * (double)y;
      sy = sy + inv[5];
      out[o] = fill;
      cov[o] = 0.0f;
      if (!(fabs(den) > 1.0e-12)) continue;
      sx /= den;
      sy /= den;
      if (interp == 1) {
        if (!(sx >= 0.0 && sx <= (double)(w - 1) && sy >= 0.0 && sy <= (double)(h - 1))) continue;
        const int64_t x0 = (int64_t)floor(sx), y0 = (int64_t)floor(sy);
        const int64_t x1 = x0 + 1 < w - 1 ? x0 + 1 : w - 1, y1 = y0 + 1 < h - 1 ? y0 + 1 : h - 1;
        const float tx = (float)(sx - (double)x0), ty = (float)(sy - (double)y0);
        const float top = src[y0 * w + x0] * (1.0f - tx) + src[y0 * w + x1] * tx;
        const float bot = src[y1 * w + x0] * (1.0f - tx) + src[y1 * w + x1] * tx;
        out[o] = top * (1.0f - ty) + bot * ty;
        cov[o] = 1.0f;
      } else {
        if (!(sx >= 2.0 && sx < (double)(w - 3) && sy >= 2.0 && sy < (double)(h - 3))) continue;
        const int64_t bx = (int64_t)floor(sx), by = (int64_t)floor(sy);
        double wxs[6];
        for (int k = 0; k < 6; ++k) wxs[k] = lanczos3(sx - (double)(bx - 2 + k));
        double acc = 0.0, ws = 0.0;
        for (int64_t yy = by - 2; yy <= by + 3; ++yy) {
          const double wy = lanczos3(sy - (double)yy);
          if (wy == 0.0) continue;
          for (int64_t xx = bx - 2; xx <= bx + 3; ++xx) {
            const double wx = wxs[xx - bx + 2];
            if (wx == 0.0) continue;
            const double wt = wx * wy;
            acc += (double)src[yy * w + xx] * wt;
            ws += wt;
          }
        }
        out[o] = fabs(ws) > 1.0e-12 ? (float)(acc / ws) : 0.0f;
        cov[o] = 1.0f;
      }
    }
  }
}

/* CPU-path Lanczos3 through an affine source map b (reference pixel ->
 * source pixel, s = (b0 x + b1 y) + b2, (b3 x + b4 y) + b5), every tap
 * summed (engine/warp.py:121-152 on finite sources), FP64 output: the finite
 * difference reference of the warp adjoint.  `fill` outside 2 <= s < dim-3. */
void cpu_warp_map64(const float* src, int64_t h, int64_t w, const double* b, double fill, double* out) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t y = 0; y < h; ++y) {
    for (int64_t x = 0; x < w; ++x) {
      double sx = b[0] * (double)x + b[1] * (double)y;
      sx = sx + b[2];
      double sy = b[3] * (double)x + b[4] * (double)y;
      sy = sy + b[5];
      if (!(sx >= 2.0 && sx < (double)(w - 3) && sy >= 2.0 && sy < (double)(h - 3))) {
        out[y * w + x] = fill;
        continue;
      }
      const int64_t bx = (int64_t)floor(sx), by = (int64_t)floor(sy);
      double wxs[6];
      for (int k = 0; k < 6; ++k) wxs[k] = lanczos3(sx - (double)(bx - 2 + k));
      double acc = 0.0, ws = 0.0;
      for (int j = 0; j < 6; ++j) {
        const double wy = lanczos3(sy - (double)(by - 2 + j));
        for (int k = 0; k < 6; ++k) {
          const double wt = wy * wxs[k];
          acc += (double)src[(by - 2 + j) * w + (bx - 2 + k)] * wt;
          ws += wt;
        }
      }
      out[y * w + x] = acc / ws;
    }
  }
}

static int cmp_float(const void* a, const void* b) {
  const float x = *(const float*)a, y = *(const float*)b;
  return (x > y) - (x < y);
}

static float percentile_valid(const float* s, int count, float fraction) {
  const int safe = count > 1 ? count : 1;
  const float position = (float)(safe - 1) * fraction;
  const int lower = (int)floorf(position);
  const int upper = lower + 1 < safe - 1 ? lower + 1 : safe - 1;
  const float t = position - (float)lower;
  return s[lower] * (1.0f - t) + s[upper] * t;
}

/* stack: f frames of npix (frame-major); weights[f]; outputs npix each. */
void cpu_integrate_winsorized(const float* stack, int64_t f, int64_t npix, const float* weights, float low_sigma,
                              float high_sigma, int min_samples, float max_frac, float* master, float* weight,
                              float* coverage, float* low_map, float* high_map) {
#pragma omp parallel
  {
    float* buf = (float*)malloc(sizeof(float) * (size_t)(f > 0 ? f : 1));
#pragma omp for schedule(static)
    for (int64_t p = 0; p < npix; ++p) {
      int n = 0;
      for (int64_t k = 0; k < f; ++k) {
        const float x = stack[k * npix + p];
        if (isfinite(x) && isfinite(weights[k]) && weights[k] > 0.0f) buf[n++] = x;
      }
      int n_low = 0, n_high = 0;
      float lo_t = 0.0f, hi_t = 0.0f;
      if (n > 0) {
        qsort(buf, (size_t)n, sizeof(float), cmp_float);
        const float c0 = percentile_valid(buf, n, 0.5f);
        const float q25 = percentile_valid(buf, n, 0.25f), q75 = percentile_valid(buf, n, 0.75f);
        float s0 = (q75 - q25) / 1.349f;
        if (!(s0 > 0.0f && isfinite(s0))) {
          double s = 0.0;
          for (int i = 0; i < n; ++i) s += (double)buf[i];
          const double m = (double)(float)(s / n);
          double ss = 0.0;
          for (int i = 0; i < n; ++i) ss += ((double)buf[i] - m) * ((double)buf[i] - m);
          s0 = (float)sqrt(ss / n);
        }
        const float lo = c0 - low_sigma * s0, hi = c0 + high_sigma * s0;
        double ws = 0.0;
        for (int i = 0; i < n; ++i) ws += (double)fminf(fmaxf(buf[i], lo), hi);
        const float wm = (float)(ws / n);
        double vs = 0.0;
        for (int i = 0; i < n; ++i) {
          const double d = (double)fminf(fmaxf(buf[i], lo), hi) - (double)wm;
          vs += d * d;
        }
        float sc = (float)sqrt(vs / n);
        if (!(sc > 0.0f && isfinite(sc))) sc = s0;
        lo_t = wm - low_sigma * sc;
        hi_t = wm + high_sigma * sc;
        for (int i = 0; i < n; ++i) {
          n_low += buf[i] < lo_t;
          n_high += buf[i] > hi_t;
        }
        const int rej = n_low + n_high;
        if (rej > 0) {
          const float frac = (float)((double)rej / (double)n);
          if (!(n - rej >= min_samples && frac <= max_frac)) n_low = n_high = 0;
        }
      }
      double sx = 0.0, sw = 0.0;
      int kept = 0;
      for (int64_t k = 0; k < f; ++k) {
        const float x = stack[k * npix + p];
        if (!(isfinite(x) && isfinite(weights[k]) && weights[k] > 0.0f)) continue;
        if (n_low + n_high > 0 && (x < lo_t || x > hi_t)) continue;
        sx += (double)x * (double)weights[k];
        sw += (double)weights[k];
        ++kept;
      }
      master[p] = sw > 0.0 ? (float)(sx / sw) : 0.0f;
      weight[p] = (float)sw;
      coverage[p] = (float)kept;
      low_map[p] = (float)n_low;
      high_map[p] = (float)n_high;
    }
    free(buf);
  }
}
