// Host-side (FP64) star measurement and star-list transform estimation.
//
// Star measurement follows src/glass/cpu/star_detect.py:45-118 on the window
// patches gathered by the GPU candidate pass (stars.hip); the transform search
// follows src/glass/cpu/registration.py:101-233, :306-358, :476-566 for the
// translation, similarity and affine models.  Everything here is scalar FP64
// built without contraction, with numpy's summation order where a sum feeds a
// result (pairwise_sum below), so star lists match the CPU path bit for bit.

#include "stars.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <set>
#include <stdexcept>
#include <utility>

#include "common.hpp"

namespace gh {

// numpy's pairwise summation of a contiguous float64 buffer (the add.reduce
// inner loop): below 8 elements a plain loop from 0.0; up to 128 eight
// interleaved accumulators combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7))
// plus the tail; above that the buffer splits at a multiple of 8 near the middle.
double pairwise_sum(const double* a, size_t n) {
  if (n < 8) {
    double res = 0.0;
    for (size_t i = 0; i < n; ++i) res += a[i];
    return res;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    size_t i = 8;
    for (; i < n - (n % 8); i += 8) {
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    }
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  size_t n2 = n / 2;
  n2 -= n2 % 8;
  return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
}

// The float32 flavour (float32 accumulators), as np.sum / np.mean of a
// float32 array use.
float pairwise_sum_f32(const float* a, size_t n) {
  if (n < 8) {
    float res = 0.0f;
    for (size_t i = 0; i < n; ++i) res += a[i];
    return res;
  }
  if (n <= 128) {
    float r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    size_t i = 8;
    for (; i < n - (n % 8); i += 8) {
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    }
    float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  size_t n2 = n / 2;
  n2 -= n2 % 8;
  return pairwise_sum_f32(a, n2) + pairwise_sum_f32(a + n2, n - n2);
}

float numpy_std_f32(const float* finite, size_t n) {
  // np.std of a float32 array: float32 mean, float32 squared deviations,
  // float32 mean of those, float32 sqrt (numpy/_core/_methods.py _var).
  const float mean = pairwise_sum_f32(finite, n) / static_cast<float>(n);
  std::vector<float> sq(n);
  for (size_t i = 0; i < n; ++i) {
    const float d = finite[i] - mean;
    sq[i] = d * d;
  }
  const float var = pairwise_sum_f32(sq.data(), n) / static_cast<float>(n);
  return std::sqrt(var);
}

bool measure_candidate(const CandidatePatch& c, int height, int width, int radius, double background, double noise,
                       int64_t origin_x, int64_t origin_y, Star& out) {
  const int y0 = std::max(0, c.y - radius), y1 = std::min(height, c.y + radius + 1);
  const int x0 = std::max(0, c.x - radius), x1 = std::min(width, c.x + radius + 1);
  const int pw = x1 - x0, ph = y1 - y0;
  const size_t n = static_cast<size_t>(pw) * static_cast<size_t>(ph);
  const int side = 2 * radius + 1;
  const float bg32 = static_cast<float>(background);
  std::vector<double> wts(n), sx(n), sy(n);
  bool any_finite = false;
  int npix = 0;
  for (int r = 0; r < ph; ++r) {
    for (int q = 0; q < pw; ++q) {
      // patch element (y0 + r, x0 + q) sits at window (y0 + r - (y - radius), x0 + q - (x - radius))
      const int wy = y0 + r - (c.y - radius), wx = x0 + q - (c.x - radius);
      const float v = c.values[static_cast<size_t>(wy) * side + wx];
      float w = 0.0f;
      if (std::isfinite(v)) {
        any_finite = true;
        const float d = v - bg32;
        w = d >= 0.0f ? d : 0.0f;  // np.maximum(d, 0.0); d is finite here
      }
      const size_t k = static_cast<size_t>(r) * pw + q;
      wts[k] = static_cast<double>(w);
      if (w > 0.0f) ++npix;
    }
  }
  if (!any_finite) return false;
  const double flux = pairwise_sum(wts.data(), n);
  if (!(flux > 0.0)) return false;
  for (int r = 0; r < ph; ++r) {
    for (int q = 0; q < pw; ++q) {
      const size_t k = static_cast<size_t>(r) * pw + q;
      sx[k] = static_cast<double>(x0 + q) * wts[k];
      sy[k] = static_cast<double>(y0 + r) * wts[k];
    }
  }
  const double cx = pairwise_sum(sx.data(), n) / flux;
  const double cy = pairwise_sum(sy.data(), n) / flux;
  std::vector<double> mxx_t(n), myy_t(n), mxy_t(n);
  for (int r = 0; r < ph; ++r) {
    for (int q = 0; q < pw; ++q) {
      const size_t k = static_cast<size_t>(r) * pw + q;
      const double dx = static_cast<double>(x0 + q) - cx;
      const double dy = static_cast<double>(y0 + r) - cy;
      mxx_t[k] = (dx * dx) * wts[k];
      myy_t[k] = (dy * dy) * wts[k];
      mxy_t[k] = (dx * dy) * wts[k];
    }
  }
  const double mxx = pairwise_sum(mxx_t.data(), n) / flux;
  const double myy = pairwise_sum(myy_t.data(), n) / flux;
  const double mxy = pairwise_sum(mxy_t.data(), n) / flux;
  const double trace = std::max(mxx + myy, 0.0);
  const double det_term = std::max((mxx - myy) * (mxx - myy) + 4.0 * mxy * mxy, 0.0);
  const double root = std::sqrt(det_term);
  const double major = std::max((trace + root) / 2.0, 0.0);
  const double minor = std::max((trace - root) / 2.0, 0.0);
  const double sigma = std::sqrt(std::max((major + minor) / 2.0, 0.0));
  out.has_fwhm = sigma > 0.0;
  out.fwhm_px = out.has_fwhm ? 2.354820045 * sigma : 0.0;
  out.has_eccentricity = major > 0.0;
  out.eccentricity = out.has_eccentricity ? std::sqrt(std::max(0.0, 1.0 - minor / major)) : 0.0;
  const float center = c.values[static_cast<size_t>(radius) * side + radius];
  out.peak = std::max(static_cast<double>(center) - background, 0.0);
  out.npix = npix;
  out.snr = flux / std::max(noise * std::sqrt(static_cast<double>(std::max(npix, 1))), 1.0e-6);
  out.sharpness = out.peak / std::max(flux, 1.0e-6);
  out.flux = flux;
  out.x = cx + static_cast<double>(origin_x);
  out.y = cy + static_cast<double>(origin_y);
  return true;
}

std::vector<Star> finish_star_list(std::vector<Star> stars, double min_separation_px, int64_t max_stars) {
  // Python's sorted(..., reverse=True) is stable: equal fluxes keep their order.
  auto by_flux = [](const Star& a, const Star& b) { return a.flux > b.flux; };
  if (min_separation_px > 0.0 && stars.size() > 1) {
    std::stable_sort(stars.begin(), stars.end(), by_flux);
    std::vector<Star> kept;
    const double limit = min_separation_px * min_separation_px;
    for (const Star& s : stars) {
      bool dup = false;
      for (const Star& k : kept) {
        const double dx = s.x - k.x, dy = s.y - k.y;
        if (dx * dx + dy * dy < limit) {
          dup = true;
          break;
        }
      }
      if (!dup) kept.push_back(s);
    }
    stars.swap(kept);
  }
  std::stable_sort(stars.begin(), stars.end(), by_flux);
  if (max_stars >= 0 && static_cast<int64_t>(stars.size()) > max_stars) stars.resize(static_cast<size_t>(max_stars));
  return stars;
}

// ------------------------------------------------------------ registration
namespace {

struct Mat3 {
  double m[3][3];
};

Mat3 translation(double dx, double dy) { return Mat3{{{1.0, 0.0, dx}, {0.0, 1.0, dy}, {0.0, 0.0, 1.0}}}; }

using Points = std::vector<std::pair<double, double>>;

Points apply_matrix(const Points& p, const Mat3& a) {
  Points out(p.size());
  for (size_t i = 0; i < p.size(); ++i) {
    const double x = p[i].first, y = p[i].second;
    const double u = (x * a.m[0][0] + y * a.m[0][1]) + a.m[0][2];
    const double v = (x * a.m[1][0] + y * a.m[1][1]) + a.m[1][2];
    const double w = (x * a.m[2][0] + y * a.m[2][1]) + a.m[2][2];
    if (std::fabs(w) > 1.0e-12) {
      out[i] = {u / w, v / w};
    } else {
      out[i] = {u, v};
    }
  }
  return out;
}

double median_of(std::vector<double> v) {
  // np.median of float64: mean of the middle pair for even counts.
  std::sort(v.begin(), v.end());
  const size_t n = v.size();
  if (n % 2) return v[n / 2];
  return (v[n / 2 - 1] + v[n / 2]) / 2.0;
}

Mat3 fit_translation(const Points& mov, const Points& ref) {
  std::vector<double> dx(mov.size()), dy(mov.size());
  for (size_t i = 0; i < mov.size(); ++i) {
    dx[i] = ref[i].first - mov[i].first;
    dy[i] = ref[i].second - mov[i].second;
  }
  return translation(median_of(dx), median_of(dy));
}

// Least squares for [x y 1] p = t via Householder QR (FP64).
void solve3(const Points& mov, const std::vector<double>& t, double p[3]) {
  const size_t n = mov.size();
  std::vector<double> a(n * 3), b(t);
  for (size_t i = 0; i < n; ++i) {
    a[i * 3 + 0] = mov[i].first;
    a[i * 3 + 1] = mov[i].second;
    a[i * 3 + 2] = 1.0;
  }
  double r[3][3] = {};
  for (int k = 0; k < 3; ++k) {
    double norm = 0.0;
    for (size_t i = k; i < n; ++i) norm += a[i * 3 + k] * a[i * 3 + k];
    norm = std::sqrt(norm);
    if (norm == 0.0) throw ValueError("affine registration matrix is rank deficient");
    const double alpha = a[k * 3 + k] > 0.0 ? -norm : norm;
    std::vector<double> v(n, 0.0);
    for (size_t i = k; i < n; ++i) v[i] = a[i * 3 + k];
    v[k] -= alpha;
    double vnorm = 0.0;
    for (size_t i = k; i < n; ++i) vnorm += v[i] * v[i];
    if (vnorm > 0.0) {
      for (int j = k; j < 3; ++j) {
        double s = 0.0;
        for (size_t i = k; i < n; ++i) s += v[i] * a[i * 3 + j];
        s = 2.0 * s / vnorm;
        for (size_t i = k; i < n; ++i) a[i * 3 + j] -= s * v[i];
      }
      double s = 0.0;
      for (size_t i = k; i < n; ++i) s += v[i] * b[i];
      s = 2.0 * s / vnorm;
      for (size_t i = k; i < n; ++i) b[i] -= s * v[i];
    }
    for (int j = k; j < 3; ++j) r[k][j] = a[k * 3 + j];
  }
  for (int k = 2; k >= 0; --k) {
    if (std::fabs(r[k][k]) <= 1.0e-12 * std::fabs(r[0][0])) throw ValueError("affine registration matrix is rank deficient");
    double s = b[k];
    for (int j = k + 1; j < 3; ++j) s -= r[k][j] * p[j];
    p[k] = s / r[k][k];
  }
}

Mat3 fit_affine(const Points& mov, const Points& ref) {
  if (mov.size() < 3) throw ValueError("affine registration requires at least three matched stars");
  std::vector<double> tx(mov.size()), ty(mov.size());
  for (size_t i = 0; i < mov.size(); ++i) {
    tx[i] = ref[i].first;
    ty[i] = ref[i].second;
  }
  double px[3], py[3];
  solve3(mov, tx, px);
  solve3(mov, ty, py);
  return Mat3{{{px[0], px[1], px[2]}, {py[0], py[1], py[2]}, {0.0, 0.0, 1.0}}};
}

Mat3 fit_similarity(const Points& mov, const Points& ref) {
  const size_t n = mov.size();
  if (n < 2) throw ValueError("similarity registration requires at least two matched stars");
  std::vector<double> mx(n), my(n), rx(n), ry(n);
  for (size_t i = 0; i < n; ++i) {
    mx[i] = mov[i].first;
    my[i] = mov[i].second;
    rx[i] = ref[i].first;
    ry[i] = ref[i].second;
  }
  // np.mean(points, axis=0) reduces along the outer axis: a plain running sum.
  auto running_mean = [n](const std::vector<double>& v) {
    double acc = v[0];
    for (size_t i = 1; i < n; ++i) acc += v[i];
    return acc / static_cast<double>(n);
  };
  const double mmx = running_mean(mx), mmy = running_mean(my);
  const double rmx = running_mean(rx), rmy = running_mean(ry);
  double var = 0.0, h00 = 0.0, h01 = 0.0, h10 = 0.0, h11 = 0.0;
  std::vector<double> sq(2 * n);
  for (size_t i = 0; i < n; ++i) {
    const double ax = mx[i] - mmx, ay = my[i] - mmy, bx = rx[i] - rmx, by = ry[i] - rmy;
    sq[2 * i] = ax * ax;
    sq[2 * i + 1] = ay * ay;
    h00 += ax * bx;
    h01 += ax * by;
    h10 += ay * bx;
    h11 += ay * by;
  }
  var = pairwise_sum(sq.data(), 2 * n);
  if (var <= 1.0e-12) throw ValueError("degenerate moving stars for similarity registration");
  // The proper rotation maximising tr(R H) (the SVD solution with the
  // reflection fix) has angle atan2(H01 - H10, H00 + H11); the scale uses the
  // sum of H's singular values, max(P, Q) below, as the reference does.
  const double theta = std::atan2(h01 - h10, h00 + h11);
  const double c = std::cos(theta), s = std::sin(theta);
  const double p = std::hypot(h00 + h11, h01 - h10), q = std::hypot(h00 - h11, h01 + h10);
  const double scale = std::max(p, q) / var;
  const double l00 = scale * c, l01 = -scale * s, l10 = scale * s, l11 = scale * c;
  const double t0 = rmx - (mmx * l00 + mmy * l01);
  const double t1 = rmy - (mmx * l10 + mmy * l11);
  return Mat3{{{l00, l01, t0}, {l10, l11, t1}, {0.0, 0.0, 1.0}}};
}

Mat3 fit_model(int model, const Points& mov, const Points& ref) {
  switch (model) {
    case kModelTranslation: return fit_translation(mov, ref);
    case kModelSimilarity: return fit_similarity(mov, ref);
    case kModelAffine: return fit_affine(mov, ref);
    default: throw ValueError("unsupported star registration model");
  }
}

struct Score {
  std::vector<std::pair<int64_t, int64_t>> pairs;
  double rms = INFINITY;
};

Score greedy_pairs(const Points& ref, const Points& moved, double tolerance) {
  Score sc;
  if (ref.empty() || moved.empty()) return sc;
  struct Cand {
    double d;
    int64_t i, j;
  };
  std::vector<Cand> cand;
  for (size_t i = 0; i < moved.size(); ++i) {
    for (size_t j = 0; j < ref.size(); ++j) {
      const double dx = moved[i].first - ref[j].first, dy = moved[i].second - ref[j].second;
      const double d = std::sqrt(dx * dx + dy * dy);
      if (d <= tolerance) cand.push_back({d, static_cast<int64_t>(i), static_cast<int64_t>(j)});
    }
  }
  if (cand.empty()) return sc;
  std::stable_sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) { return a.d < b.d; });
  std::vector<char> used_m(moved.size(), 0), used_r(ref.size(), 0);
  std::vector<double> sq;
  for (const Cand& c : cand) {
    if (used_m[c.i] || used_r[c.j]) continue;
    used_m[c.i] = used_r[c.j] = 1;
    sc.pairs.push_back({c.i, c.j});
    sq.push_back(c.d * c.d);
  }
  sc.rms = std::sqrt(pairwise_sum(sq.data(), sq.size()) / static_cast<double>(sq.size()));
  return sc;
}

std::vector<Mat3> translation_candidates(const Points& ref, const Points& mov, size_t brightest) {
  std::vector<Mat3> out;
  std::set<std::pair<long long, long long>> seen;
  const size_t nr = std::min(ref.size(), brightest), nm = std::min(mov.size(), brightest);
  for (size_t a = 0; a < nr; ++a) {
    for (size_t b = 0; b < nm; ++b) {
      const double dx = ref[a].first - mov[b].first, dy = ref[a].second - mov[b].second;
      // Python's round(): nearest, ties to even (the default FP rounding mode).
      const std::pair<long long, long long> key{static_cast<long long>(std::nearbyint(dx)),
                                                static_cast<long long>(std::nearbyint(dy))};
      if (seen.insert(key).second) out.push_back(translation(dx, dy));
    }
  }
  return out;
}

struct Triangle {
  double d0, d1;
  int i, j, k;
  double area;
};

std::vector<Triangle> triangle_descriptors(const Points& pts_all, size_t max_points, size_t max_triangles) {
  const size_t n = std::min(pts_all.size(), max_points);
  std::vector<Triangle> tris;
  auto dist = [&](int a, int b) {
    const double dx = pts_all[a].first - pts_all[b].first, dy = pts_all[a].second - pts_all[b].second;
    return std::sqrt(dx * dx + dy * dy);
  };
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = i + 1; j < n; ++j) {
      for (size_t k = j + 1; k < n; ++k) {
        const double ax = pts_all[j].first - pts_all[i].first, ay = pts_all[j].second - pts_all[i].second;
        const double bx = pts_all[k].first - pts_all[i].first, by = pts_all[k].second - pts_all[i].second;
        const double area = std::fabs(ax * by - ay * bx) * 0.5;
        if (area < 2.0) continue;
        double s[3] = {dist(i, j), dist(j, k), dist(k, i)};
        const double longest = std::max(s[0], std::max(s[1], s[2]));
        if (longest <= 1.0e-6) continue;
        std::sort(s, s + 3);
        tris.push_back({s[0] / longest, s[1] / longest, static_cast<int>(i), static_cast<int>(j), static_cast<int>(k),
                        area});
      }
    }
  }
  std::stable_sort(tris.begin(), tris.end(), [](const Triangle& a, const Triangle& b) { return a.area > b.area; });
  if (tris.size() > max_triangles) tris.resize(max_triangles);
  return tris;
}

std::vector<Mat3> asterism_candidates(const Points& ref, const Points& mov, int model, double tolerance,
                                      size_t max_triangles) {
  const auto rt = triangle_descriptors(ref, 30, max_triangles);
  const auto mt = triangle_descriptors(mov, 30, max_triangles);
  std::vector<Mat3> out;
  static const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (const Triangle& m : mt) {
    for (const Triangle& r : rt) {
      const double e0 = m.d0 - r.d0, e1 = m.d1 - r.d1;
      if (std::sqrt(e0 * e0 + e1 * e1) > tolerance) continue;
      const Points mtri = {mov[m.i], mov[m.j], mov[m.k]};
      const Points rraw = {ref[r.i], ref[r.j], ref[r.k]};
      for (const auto& pm : kPerm) {
        const Points rtri = {rraw[pm[0]], rraw[pm[1]], rraw[pm[2]]};
        try {
          out.push_back(fit_model(model, mtri, rtri));
        } catch (const ValueError&) {
        }
      }
      if (out.size() >= max_triangles) return out;
    }
  }
  return out;
}

}  // namespace

TransformResult estimate_star_transform(const double* ref_xy, int64_t n_ref, const double* mov_xy, int64_t n_mov,
                                        int model, int64_t min_inliers, double tolerance_px, int64_t max_stars) {
  if (model != kModelTranslation && model != kModelSimilarity && model != kModelAffine)
    throw ValueError("unsupported star registration model (translation, similarity or affine)");
  const size_t keep = static_cast<size_t>(std::max<int64_t>(0, max_stars));
  Points ref, mov;
  for (int64_t i = 0; i < n_ref && static_cast<size_t>(i) < keep; ++i) ref.push_back({ref_xy[2 * i], ref_xy[2 * i + 1]});
  for (int64_t i = 0; i < n_mov && static_cast<size_t>(i) < keep; ++i) mov.push_back({mov_xy[2 * i], mov_xy[2 * i + 1]});
  const size_t minimum = model == kModelAffine ? 3 : model == kModelSimilarity ? 2 : 1;
  TransformResult res;
  const Mat3 ident = translation(0.0, 0.0);
  std::copy(&ident.m[0][0], &ident.m[0][0] + 9, res.matrix);
  if (ref.size() < minimum || mov.size() < minimum) return res;
  std::vector<Mat3> cands = translation_candidates(ref, mov, 35);
  if (model == kModelSimilarity || model == kModelAffine) {
    std::vector<Mat3> ast = asterism_candidates(ref, mov, model, 0.025, 600);
    ast.insert(ast.end(), cands.begin(), cands.end());
    cands.swap(ast);
  }
  if (cands.empty()) cands.push_back(Mat3{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}});
  Mat3 best = cands[0];
  Score best_score;
  for (const Mat3& m : cands) {
    Score sc = greedy_pairs(ref, apply_matrix(mov, m), tolerance_px);
    if (sc.pairs.size() > best_score.pairs.size() ||
        (sc.pairs.size() == best_score.pairs.size() && sc.rms < best_score.rms)) {
      best = m;
      best_score = std::move(sc);
    }
  }
  for (int it = 0; it < 2; ++it) {
    if (best_score.pairs.size() < minimum) break;
    Points mf, rf;
    for (const auto& pr : best_score.pairs) {
      mf.push_back(mov[pr.first]);
      rf.push_back(ref[pr.second]);
    }
    Mat3 refined;
    try {
      refined = fit_model(model, mf, rf);
    } catch (const ValueError&) {
      break;
    }
    Score sc = greedy_pairs(ref, apply_matrix(mov, refined), tolerance_px);
    if (sc.pairs.size() > best_score.pairs.size() ||
        (sc.pairs.size() == best_score.pairs.size() && sc.rms <= best_score.rms)) {
      best = refined;
      best_score = std::move(sc);
    }
  }
  std::copy(&best.m[0][0], &best.m[0][0] + 9, res.matrix);
  res.pairs = best_score.pairs;
  res.rms = best_score.rms;
  res.ok = static_cast<int64_t>(best_score.pairs.size()) >= min_inliers;
  return res;
}

}  // namespace gh
