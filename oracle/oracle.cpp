// FP64 CPU oracle of the exact signed distance and its brick baker.
//
// TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared against and
// the timed CPU baseline. The product never links or calls it.
//
// A restatement, in C++, of the reference's double-precision evaluator
// `solid3::Exact` and the CPU rules of its brick cache. Every function cites the
// reference lines (paths under crates/) it follows. Differences, each stated
// where it occurs:
//  - the trim inside test answers by crossing parity against every loop piece
//    from the face's given reference point, instead of descending the cell tree
//    to a cell's reference (the tree is an acceleration of the same parity);
//  - a B-spline face's search starts from its span patches instead of the
//    quadtree over spans (same patches, same bounds);
//  - patches and spline edges are limited to degree 7 (the reference kernel's
//    MAX_DEGREE, crates/icam-core/src/solid3/gpu.rs:41-42); orc_open refuses more;
//  - closest_trimmed takes a critical point within DIST_EPS of a bound that no answer
//    has set, and a polished point within DIST_EPS of the point it started from
//    (surface3.rs:1276-1283 takes them only when strictly nearer): where the foot lies
//    on the shared border of two spans, the first span's Newton stops at the border,
//    not critical, and sets the bound, and the next span's critical point only ties it
//    in rounding, so the reference's search loses the face's answer there (a pocket
//    wall's point 0.165 mm nearer than the answer it gives).
#include "oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <queue>
#include <thread>
#include <vector>

namespace orc {

const double INF = std::numeric_limits<double>::infinity();
const double PI = 3.14159265358979323846;
const double TAU = 2.0 * PI;

// ------------------------------------------------------------------ work counts
// The primitives a run executes, counted per thread and added to the totals when a
// worker ends (oracle.h: orc_counts). Order as ORC_COUNT_* in oracle.h.
enum Count : int {
  C_PEVAL = 0,    // rational tensor Bezier point with first and second derivatives
  C_NEWTON2,      // a surface Newton step, its evaluation excluded
  C_CEVAL,        // rational Bezier curve point with first and second derivatives
  C_NEWTON1,      // a curve Newton step, its evaluation excluded
  C_FOOT,         // analytic feet: plane, cylinder, cone, sphere, torus (5)
  C_UV = C_FOOT + 5,  // analytic parameters, same kinds (5)
  C_CROSS = C_UV + 5, // a trim crossing's Newton iteration
  C_EDGE_LINE,    // a line edge's foot
  C_EDGE_ARC,     // an arc edge's foot
  C_VERT,         // a vertex (or pole, tip) distance
  NCOUNT
};
static_assert(int(NCOUNT) == int(ORC_COUNTS) && int(C_CROSS) == int(ORC_COUNT_CROSS) && int(C_VERT) == int(ORC_COUNT_VERT), "counts as oracle.h numbers them");
thread_local uint64_t t_count[NCOUNT];
std::atomic<uint64_t> g_count[NCOUNT];
inline void count(int k) { ++t_count[k]; }
void flush_counts() {
  for (int k = 0; k < NCOUNT; ++k)
    if (t_count[k]) g_count[k].fetch_add(t_count[k], std::memory_order_relaxed), t_count[k] = 0;
}

// ------------------------------------------------------------------ vectors
// crates/icam-core/src/curve3.rs:14-54
struct V3 {
  double x = 0, y = 0, z = 0;
  V3() = default;
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  static V3 of(const double* d) { return V3(d[0], d[1], d[2]); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) {
  double l = len(a);
  return l > 0 ? a * (1.0 / l) : a;
}
inline V3 vmin(V3 a, V3 b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }

struct Box {
  V3 lo{INF, INF, INF}, hi{-INF, -INF, -INF};
  void add(V3 p) { lo = vmin(lo, p), hi = vmax(hi, p); }
  void add(const Box& b) { lo = vmin(lo, b.lo), hi = vmax(hi, b.hi); }
};

// surface3.rs:553-555 box_dist
inline double box_dist(V3 p, const Box& b) {
  V3 d(std::max(std::max(b.lo.x - p.x, p.x - b.hi.x), 0.0), std::max(std::max(b.lo.y - p.y, p.y - b.hi.y), 0.0),
       std::max(std::max(b.lo.z - p.z, p.z - b.hi.z), 0.0));
  return len(d);
}

// curve3.rs:39-54 Frame
struct Frame {
  V3 o, x, y, z;
  V3 local(V3 p) const {
    V3 d = p - o;
    return {dot(d, x), dot(d, y), dot(d, z)};
  }
  V3 world(V3 l) const { return o + x * l.x + y * l.y + z * l.z; }
  V3 dir(V3 l) const { return x * l.x + y * l.y + z * l.z; }
};

// surface3.rs:36-53 axial coordinates
struct Axial {
  double theta, rho, h;
  V3 er, et;
};
Axial axial(const Frame& f, V3 p) {
  V3 l = f.local(p);
  double rho = std::sqrt(l.x * l.x + l.y * l.y);
  double th = rho > 0 ? std::atan2(l.y, l.x) : 0.0;
  double s = std::sin(th), c = std::cos(th);
  return {th, rho, l.z, f.dir(V3(c, s, 0)), f.dir(V3(-s, c, 0))};
}

// ------------------------------------------------------------------ analytic surfaces
// surface3.rs:85-149 Surf::foot / Surf::normal
struct Analytic {
  int kind = 0;
  Frame f;
  double c0 = 0, c1 = 0;  // r | r, a | major, minor
  // the exact rational form's parameter constants (trim2.rs:1807-1832)
  double psi0 = 0, span = TAU;
  int narc = 4;

  V3 foot(V3 p) const {
    if (kind >= ICAM_PLANE && kind <= ICAM_TORUS) count(C_FOOT + kind);
    switch (kind) {
      case ICAM_PLANE: return p - f.z * dot(p - f.o, f.z);
      case ICAM_CYLINDER: {
        Axial ax = axial(f, p);
        return f.o + ax.er * c0 + f.z * ax.h;
      }
      case ICAM_CONE: {
        Axial ax = axial(f, p);
        double sa = std::sin(c1), ca = std::cos(c1);
        double ha = -c0 / std::tan(c1);
        double s = ax.rho * sa + (ax.h - ha) * ca;
        if (s <= 0) return f.o + f.z * ha;
        return f.o + ax.er * (s * sa) + f.z * (ha + s * ca);
      }
      case ICAM_SPHERE: {
        V3 d = p - f.o;
        double l = len(d);
        return l > 0 ? f.o + d * (c0 / l) : f.o + f.x * c0;
      }
      case ICAM_TORUS: {
        Axial ax = axial(f, p);
        V3 c = f.o + ax.er * c0;
        V3 d = p - c;
        double l = len(d);
        V3 q = l > 0 ? c + d * (c1 / l) : c + ax.er * c1;
        if (dot(q - f.o, ax.er) >= 0 || std::fabs(c0) > c1) return q;
        double h = std::sqrt(std::max(c1 * c1 - c0 * c0, 0.0));
        return ax.h >= 0 ? f.o + f.z * h : f.o - f.z * h;
      }
    }
    return p;
  }
  V3 normal(V3 x) const {
    switch (kind) {
      case ICAM_PLANE: return f.z;
      case ICAM_CYLINDER: return axial(f, x).er;
      case ICAM_CONE: {
        Axial ax = axial(f, x);
        return ax.er * std::cos(c1) - f.z * std::sin(c1);
      }
      case ICAM_SPHERE: return unit(x - f.o);
      case ICAM_TORUS: {
        Axial ax = axial(f, x);
        return unit(x - (f.o + ax.er * c0));
      }
    }
    return f.z;
  }

  // trim2.rs:1789-1807 arc_param_over
  static double arc_param_over(double a, double a0, double span, int n) {
    double q = span / n;
    double x = a - a0;
    if (std::fabs(span - TAU) < 1e-12) {
      x = x - TAU * std::floor(x / TAU);
    } else {
      double y = x - TAU * std::round((x - 0.5 * span) / TAU);
      x = std::min(std::max(y, 0.0), span);
    }
    long k = std::min(std::max(long(std::floor(x / q)), 0L), long(n - 1));
    double alpha = x - q * k - 0.5 * q;
    double t = 0.5 * (std::tan(0.5 * alpha) / std::tan(0.25 * q) + 1.0);
    return k + std::min(std::max(t, 0.0), 1.0);
  }
  static double arc_param(double a, double a0, int n) { return arc_param_over(a, a0, 0.5 * PI * n, n); }

  // trim2.rs:1907-1934 Analytic::uv
  void uv(V3 q, double& u, double& v) const {
    if (kind >= ICAM_PLANE && kind <= ICAM_TORUS) count(C_UV + kind);
    switch (kind) {
      case ICAM_PLANE: {
        V3 l = f.local(q);
        u = l.x, v = l.y;
        return;
      }
      case ICAM_CYLINDER: {
        Axial ax = axial(f, q);
        u = arc_param(ax.theta, 0, 4), v = ax.h;
        return;
      }
      case ICAM_CONE: {
        Axial ax = axial(f, q);
        double sa = std::sin(c1), ca = std::cos(c1);
        double ha = -c0 / std::tan(c1);
        double s = std::max(ax.rho * sa + (ax.h - ha) * ca, 0.0);
        u = arc_param(ax.theta, 0, 4), v = ha + s * ca;
        return;
      }
      case ICAM_SPHERE: {
        Axial ax = axial(f, q);
        u = arc_param(ax.theta, 0, 4), v = arc_param(std::atan2(ax.h, ax.rho), -0.5 * PI, 2);
        return;
      }
      case ICAM_TORUS: {
        Axial ax = axial(f, q);
        u = arc_param(ax.theta, 0, 4), v = arc_param_over(std::atan2(ax.h, ax.rho - c0), psi0, span, narc);
        return;
      }
    }
    u = v = 0;
  }
};

// solid3.rs:1991-2005 singular_points / is_singular (cone apex, lemon/apple tips)
std::vector<std::pair<V3, V3>> singular_points(const Analytic& s, double sense) {
  std::vector<std::pair<V3, V3>> out;
  if (s.kind == ICAM_CONE) out.push_back({s.f.o + s.f.z * (-s.c0 / std::tan(s.c1)), s.f.z * (-std::sin(s.c1) * sense)});
  if (s.kind == ICAM_TORUS && std::fabs(s.c0) <= s.c1) {
    double h = std::sqrt(std::max(s.c1 * s.c1 - s.c0 * s.c0, 0.0));
    out.push_back({s.f.o + s.f.z * h, s.f.z * sense});
    out.push_back({s.f.o - s.f.z * h, s.f.z * -sense});
  }
  return out;
}
bool is_singular(const Analytic& s, V3 q) {
  for (auto& pt : singular_points(s, 1.0))
    if (len(pt.first - q) <= 1e-12 * (1.0 + len(pt.first))) return true;
  return false;
}

// ------------------------------------------------------------------ trim loops
// trim2.rs:104-136 Bernstein evaluation of a cubic coordinate
double b_eval3(const double c[4], double t) {
  if (c[0] == c[1] && c[1] == c[2] && c[2] == c[3]) return c[0];
  double b[4] = {c[0], c[1], c[2], c[3]};
  for (int l = 1; l <= 3; ++l)
    for (int i = 0; i <= 3 - l; ++i) b[i] = (1 - t) * b[i] + t * b[i + 1];
  return b[0];
}
double b_eval2(const double c[3], double t) {
  if (c[0] == c[1] && c[1] == c[2]) return c[0];
  double b[3] = {c[0], c[1], c[2]};
  for (int l = 1; l <= 2; ++l)
    for (int i = 0; i <= 2 - l; ++i) b[i] = (1 - t) * b[i] + t * b[i + 1];
  return b[0];
}

// trim2.rs:1150-1190 crossing: the parameter where a monotone cubic coordinate
// equals x, for x in its half-open range [min, max), by safeguarded Newton.
bool crossing(const double a[4], double x, double& out) {
  double p0 = a[0], p1 = a[3];
  if (!(std::min(p0, p1) <= x && x < std::max(p0, p1))) return false;
  double lo = 0, hi = 1;
  auto f = [&](double t) { return b_eval3(a, t) - x; };
  double flo = f(0.0);
  if (flo == 0.0) {
    out = 0.0;
    return true;
  }
  double t = std::min(std::max((x - p0) / (p1 - p0), 0.0), 1.0);
  const double d3[3] = {a[1] - a[0], a[2] - a[1], a[3] - a[2]};
  for (int it = 0; it < 60; ++it) {
    count(C_CROSS);
    double ft = f(t);
    if (ft == 0.0) {
      out = t;
      return true;
    }
    if ((ft < 0) == (flo < 0)) {
      lo = t;
      flo = ft;
    } else {
      hi = t;
    }
    double d = 3.0 * b_eval2(d3, t);
    double nt = d != 0 ? t - ft / d : 0.5 * (lo + hi);
    if (!(nt > lo && nt < hi)) nt = 0.5 * (lo + hi);
    if (std::fabs(nt - t) <= 1e-16) {
      out = nt;
      return true;
    }
    t = nt;
  }
  out = t;
  return true;
}

struct Trim {
  double lo[2], hi[2], period[2];
  std::vector<std::array<double, 8>> pieces;
  std::vector<std::array<double, 8>> inst;  // shifted copies
  double ref[2];
  bool ref_inside = true;

  // trim2.rs:1272-1296 the shifted copies reaching within a period of the domain
  void place() {
    double rlo[2] = {lo[0], lo[1]}, rhi[2] = {hi[0], hi[1]};
    for (int k = 0; k < 2; ++k)
      if (period[k] > 0) rlo[k] -= period[k], rhi[k] += period[k];
    inst.clear();
    for (auto& p : pieces) {
      double plo[2] = {INF, INF}, phi[2] = {-INF, -INF};
      for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 2; ++k) plo[k] = std::min(plo[k], p[2 * i + k]), phi[k] = std::max(phi[k], p[2 * i + k]);
      long r0[2], r1[2];
      for (int k = 0; k < 2; ++k) {
        if (period[k] > 0) {
          r0[k] = long(std::ceil((rlo[k] - phi[k]) / period[k]));
          r1[k] = long(std::floor((rhi[k] - plo[k]) / period[k]));
        } else {
          r0[k] = r1[k] = 0;
        }
      }
      for (long a = r0[0]; a <= r1[0]; ++a)
        for (long c = r0[1]; c <= r1[1]; ++c) {
          std::array<double, 8> q = p;
          for (int i = 0; i < 4; ++i) q[2 * i] += a * period[0], q[2 * i + 1] += c * period[1];
          inst.push_back(q);
        }
    }
  }

  // trim2.rs:1205-1260 parity of the two-leg path a -> (a.u, b.v) -> b, moved up by
  // an infinitesimal e and right by e^2
  bool parity(const double a[2], const double b[2]) const {
    unsigned n = 0;
    bool up = a[1] < b[1], right = a[0] < b[0];
    for (auto& c : inst) {
      double cu[4] = {c[0], c[2], c[4], c[6]}, cv[4] = {c[1], c[3], c[5], c[7]};
      double t;
      if (a[1] != b[1] && crossing(cu, a[0], t)) {
        double y = b_eval3(cv, t);
        if (up ? (y > a[1] && y <= b[1]) : (y > b[1] && y <= a[1])) ++n;
      }
      if (a[0] != b[0] && crossing(cv, b[1], t)) {
        double x = b_eval3(cu, t);
        auto slope_pos = [&]() {
          const double du3[3] = {cu[1] - cu[0], cu[2] - cu[1], cu[3] - cu[2]};
          const double dv3[3] = {cv[1] - cv[0], cv[2] - cv[1], cv[3] - cv[2]};
          double du = b_eval2(du3, t), dv = b_eval2(dv3, t);
          if (dv == 0) dv = cv[3] - cv[0];
          return du * dv > 0;
        };
        bool counted;
        if (x == b[0])
          counted = (!slope_pos()) == right;
        else if (x == a[0])
          counted = slope_pos() == right;
        else if (right)
          counted = x > a[0] && x < b[0];
        else
          counted = x > b[0] && x < a[0];
        if (counted) ++n;
      }
    }
    return n % 2 == 1;
  }

  void reduce(double& u, double& v) const {
    double q[2] = {u, v};
    for (int k = 0; k < 2; ++k) {
      if (period[k] > 0)
        q[k] = lo[k] + (q[k] - lo[k] - period[k] * std::floor((q[k] - lo[k]) / period[k]));
      else
        q[k] = std::min(std::max(q[k], lo[k]), hi[k]);
    }
    u = q[0], v = q[1];
  }

  // trim2.rs:1433-1458 inside: the reference's answer flipped by the path's parity
  bool inside(double u, double v) const {
    if (pieces.empty()) return true;
    reduce(u, v);
    double q[2] = {u, v};
    return ref_inside ^ parity(ref, q);
  }

  // trim2.rs:1115-1132 meets_box for a monotone piece
  static bool meets_box(const std::array<double, 8>& c, const double lo[2], const double hi[2]) {
    double b0[2] = {INF, INF}, b1[2] = {-INF, -INF};
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 2; ++k) b0[k] = std::min(b0[k], c[2 * i + k]), b1[k] = std::max(b1[k], c[2 * i + k]);
    for (int k = 0; k < 2; ++k)
      if (!(b0[k] <= hi[k] && lo[k] <= b1[k])) return false;
    double cu[4] = {c[0], c[2], c[4], c[6]}, cv[4] = {c[1], c[3], c[5], c[7]};
    double u0 = cu[0], u3 = cu[3], ta, tb;
    auto param_at = [&](double x) {
      if (x == cu[0]) return 0.0;
      if (x == cu[3]) return 1.0;
      double t;
      if (crossing(cu, x, t)) return t;
      return std::fabs(x - cu[0]) < std::fabs(x - cu[3]) ? 0.0 : 1.0;
    };
    if (u0 == u3) {
      ta = 0, tb = 1;
    } else {
      double a = std::max(lo[0], std::min(u0, u3)), b = std::min(hi[0], std::max(u0, u3));
      ta = param_at(a), tb = param_at(b);
    }
    double va = b_eval3(cv, ta), vb = b_eval3(cv, tb);
    return std::min(va, vb) <= hi[1] && std::max(va, vb) >= lo[1];
  }

  // trim2.rs:1570-1610 box_class: 1 wholly in, 0 wholly out, -1 straddles
  int box_class(const double blo[2], const double bhi[2]) const {
    if (pieces.empty()) return 1;
    double g = 1e-12 * (std::fabs(hi[0] - lo[0]) + std::fabs(hi[1] - lo[1]));
    double glo[2] = {blo[0] - g, blo[1] - g}, ghi[2] = {bhi[0] + g, bhi[1] + g};
    for (auto& c : inst)
      if (meets_box(c, glo, ghi)) return -1;
    return inside(0.5 * (blo[0] + bhi[0]), 0.5 * (blo[1] + bhi[1])) ? 1 : 0;
  }
};

// ------------------------------------------------------------------ rational Bezier patches
const int MAXP = 8;  // control points a direction: degree 7 at most

// surface3.rs:246-253 SEval
struct SEval {
  V3 p, ps, pt, pss, pst, ptt;
};

// Bernstein basis of degree n at t with first and second derivatives
// (surface3.rs:208-238 bernstein_small).
void bernstein(int n, double t, double b[MAXP], double d1[MAXP], double d2[MAXP]) {
  double cur[MAXP + 1] = {1.0}, l1[MAXP + 1] = {0}, l2[MAXP + 1] = {0};
  if (n >= 1 && n - 1 == 0) std::copy(cur, cur + MAXP + 1, l1);
  if (n >= 2 && n - 2 == 0) std::copy(cur, cur + MAXP + 1, l2);
  for (int k = 1; k <= n; ++k) {
    for (int i = k; i >= 0; --i) {
      double left = i > 0 ? cur[i - 1] : 0.0, right = i < k ? cur[i] : 0.0;
      cur[i] = (1.0 - t) * right + t * left;
    }
    if (n >= 1 && k == n - 1) std::copy(cur, cur + MAXP + 1, l1);
    if (n >= 2 && k == n - 2) std::copy(cur, cur + MAXP + 1, l2);
  }
  auto get = [](const double* lv, int ln, int i) { return (i < 0 || i >= ln) ? 0.0 : lv[i]; };
  for (int i = 0; i <= n; ++i) {
    b[i] = cur[i];
    d1[i] = n >= 1 ? n * (get(l1, n, i - 1) - get(l1, n, i)) : 0.0;
    d2[i] = n >= 2 ? double(n * (n - 1)) * (get(l2, n - 1, i - 2) - 2.0 * get(l2, n - 1, i - 1) + get(l2, n - 1, i)) : 0.0;
  }
}

// surface3.rs:254-263 Patch: pw[i][j] homogeneous, i along u
struct Patch {
  int m = 0, n = 0;
  double pw[MAXP][MAXP][4];
  double u0 = 0, u1 = 1, v0 = 0, v1 = 1;

  V3 ctrl(int i, int j) const { return V3(pw[i][j][0] / pw[i][j][3], pw[i][j][1] / pw[i][j][3], pw[i][j][2] / pw[i][j][3]); }
  Box bbox() const {
    Box b;
    for (int i = 0; i <= m; ++i)
      for (int j = 0; j <= n; ++j) b.add(ctrl(i, j));
    return b;
  }
  // surface3.rs:322-351 eval_small
  SEval eval(double s, double t) const {
    count(C_PEVAL);
    double bu[MAXP], bu1[MAXP], bu2[MAXP], bv[MAXP], bv1[MAXP], bv2[MAXP];
    bernstein(m, s, bu, bu1, bu2);
    bernstein(n, t, bv, bv1, bv2);
    double a[6][4] = {{0}};
    for (int i = 0; i <= m; ++i)
      for (int j = 0; j <= n; ++j) {
        const double* h = pw[i][j];
        double w[6] = {bu[i] * bv[j], bu1[i] * bv[j], bu[i] * bv1[j], bu2[i] * bv[j], bu1[i] * bv1[j], bu[i] * bv2[j]};
        for (int k = 0; k < 6; ++k)
          for (int c = 0; c < 4; ++c) a[k][c] += w[k] * h[c];
      }
    auto v3 = [&](int k) { return V3(a[k][0], a[k][1], a[k][2]); };
    double w = a[0][3];
    SEval e;
    e.p = v3(0) * (1.0 / w);
    e.ps = (v3(1) - e.p * a[1][3]) * (1.0 / w);
    e.pt = (v3(2) - e.p * a[2][3]) * (1.0 / w);
    e.pss = (v3(3) - e.ps * (2.0 * a[1][3]) - e.p * a[3][3]) * (1.0 / w);
    e.pst = (v3(4) - e.ps * a[2][3] - e.pt * a[1][3] - e.p * a[4][3]) * (1.0 / w);
    e.ptt = (v3(5) - e.pt * (2.0 * a[2][3]) - e.p * a[5][3]) * (1.0 / w);
    return e;
  }
  // surface3.rs:265-283, 352-380 de Casteljau splits
  // the parts' headers (every control point they use is written by the split)
  void head(Patch& o) const { o.m = m, o.n = n, o.u0 = u0, o.u1 = u1, o.v0 = v0, o.v1 = v1; }
  void split_u(double s, Patch& l, Patch& r) const {
    head(l), head(r);
    for (int j = 0; j <= n; ++j) {
      double b[MAXP][4];
      for (int i = 0; i <= m; ++i)
        for (int c = 0; c < 4; ++c) b[i][c] = pw[i][j][c];
      for (int c = 0; c < 4; ++c) l.pw[0][j][c] = b[0][c], r.pw[m][j][c] = b[m][c];
      for (int lev = 1; lev <= m; ++lev) {
        for (int i = 0; i <= m - lev; ++i)
          for (int c = 0; c < 4; ++c) b[i][c] = (1.0 - s) * b[i][c] + s * b[i + 1][c];
        for (int c = 0; c < 4; ++c) l.pw[lev][j][c] = b[0][c], r.pw[m - lev][j][c] = b[m - lev][c];
      }
    }
    double um = u0 + s * (u1 - u0);
    l.u1 = um, r.u0 = um;
  }
  void split_v(double t, Patch& l, Patch& r) const {
    head(l), head(r);
    for (int i = 0; i <= m; ++i) {
      double b[MAXP][4];
      for (int j = 0; j <= n; ++j)
        for (int c = 0; c < 4; ++c) b[j][c] = pw[i][j][c];
      for (int c = 0; c < 4; ++c) l.pw[i][0][c] = b[0][c], r.pw[i][n][c] = b[n][c];
      for (int lev = 1; lev <= n; ++lev) {
        for (int j = 0; j <= n - lev; ++j)
          for (int c = 0; c < 4; ++c) b[j][c] = (1.0 - t) * b[j][c] + t * b[j + 1][c];
        for (int c = 0; c < 4; ++c) l.pw[i][lev][c] = b[0][c], r.pw[i][n - lev][c] = b[n - lev][c];
      }
    }
    double vm = v0 + t * (v1 - v0);
    l.v1 = vm, r.v0 = vm;
  }
  // surface3.rs:979-985 quarters: [a0, a1, b0, b1]
  void quarters(Patch q[4]) const {
    Patch a, b;
    split_u(0.5, a, b);
    a.split_v(0.5, q[0], q[1]);
    b.split_v(0.5, q[2], q[3]);
  }

  // surface3.rs:401-452 newton_iters: projected Newton on |S - p|^2, a parameter on
  // its bound with its step pointing out held there
  void newton(V3 p, double& s, double& t, int iters, double s0 = 0, double s1 = 1, double t0 = 0, double t1 = 1) const {
    bool hs = false, ht = false;
    for (int it = 0; it < iters; ++it) {
      SEval e = eval(s, t);
      count(C_NEWTON2);
      V3 r = e.p - p;
      double g0 = dot(e.ps, r), g1 = dot(e.pt, r);
      double jss = dot(e.ps, e.ps), jst = dot(e.ps, e.pt), jtt = dot(e.pt, e.pt);
      double h0 = jss + dot(e.pss, r), h1 = jst + dot(e.pst, r), h2 = jtt + dot(e.ptt, r);
      double ds, dt;
      if (hs && ht) break;
      if (hs) {
        ds = 0;
        dt = -g1 / (h2 > 0 ? h2 : std::max(jtt, 1e-300));
      } else if (ht) {
        ds = -g0 / (h0 > 0 ? h0 : std::max(jss, 1e-300));
        dt = 0;
      } else {
        if (!(h0 > 0 && h0 * h2 - h1 * h1 > 0)) h0 = jss, h1 = jst, h2 = jtt;
        double det = h0 * h2 - h1 * h1;
        if (std::fabs(det) < 1e-300) break;
        ds = -(h2 * g0 - h1 * g1) / det;
        dt = -(h0 * g1 - h1 * g0) / det;
      }
      auto at_bound = [](double x, double d, double a, double b) { return (x <= a && d < 0) || (x >= b && d > 0); };
      if (!hs && at_bound(s, ds, s0, s1)) {
        hs = true;
        continue;
      }
      if (!ht && at_bound(t, dt, t0, t1)) {
        ht = true;
        continue;
      }
      double ns = std::min(std::max(s + ds, s0), s1), nt = std::min(std::max(t + dt, t0), t1);
      double moved = std::fabs(ns - s) + std::fabs(nt - t);
      s = ns, t = nt;
      if (moved < 1e-13) break;
    }
  }
};

// surface3.rs:564-570 critical
bool critical(V3 p, const SEval& e) {
  V3 r = e.p - p;
  auto big = [](V3 x) { return std::max(std::fabs(x.x), std::max(std::fabs(x.y), std::fabs(x.z))); };
  double round = 1024.0 * std::numeric_limits<double>::epsilon() * (big(e.p) + big(p) + len(r));
  auto ok = [&](V3 d) { return std::fabs(dot(r, d)) <= (round + 1e-12 * len(d)) * len(d); };
  return ok(e.ps) && ok(e.pt);
}

// surface3.rs:593-690 Bern2: a tensor Bernstein polynomial, coefficient (i, j) at i (n + 1) + j
struct Bern2 {
  int m, n;
  std::vector<double> c;
};
std::vector<double> binomials(int n) {
  std::vector<double> b(n + 1, 1.0);
  for (int k = 1; k <= n; ++k) b[k] = b[k - 1] * (n + 1 - k) / k;
  return b;
}
Bern2 bmul(const Bern2& a, const Bern2& o) {
  int m = a.m + o.m, n = a.n + o.n;
  auto ba = binomials(a.m), bb = binomials(a.n), bo = binomials(o.m), bp = binomials(o.n), bm = binomials(m), bn = binomials(n);
  std::vector<double> c((m + 1) * (n + 1), 0.0);
  for (int i = 0; i <= a.m; ++i)
    for (int j = 0; j <= a.n; ++j) {
      double x = a.c[i * (a.n + 1) + j] * ba[i] * bb[j];
      if (x == 0) continue;
      for (int k = 0; k <= o.m; ++k)
        for (int l = 0; l <= o.n; ++l) c[(i + k) * (n + 1) + j + l] += x * o.c[k * (o.n + 1) + l] * bo[k] * bp[l];
    }
  for (int k = 0; k <= m; ++k)
    for (int l = 0; l <= n; ++l) c[k * (n + 1) + l] /= bm[k] * bn[l];
  return {m, n, c};
}
Bern2 d_u(const Bern2& a) {
  if (a.m == 0) return {0, a.n, std::vector<double>(a.n + 1, 0.0)};
  int n1 = a.n + 1;
  std::vector<double> c;
  for (int i = 0; i < a.m; ++i)
    for (int j = 0; j < n1; ++j) c.push_back(a.m * (a.c[(i + 1) * n1 + j] - a.c[i * n1 + j]));
  return {a.m - 1, a.n, c};
}
Bern2 d_v(const Bern2& a) {
  if (a.n == 0) return {a.m, 0, std::vector<double>(a.m + 1, 0.0)};
  int n1 = a.n + 1;
  std::vector<double> c;
  for (int i = 0; i <= a.m; ++i)
    for (int j = 0; j < a.n; ++j) c.push_back(a.n * (a.c[i * n1 + j + 1] - a.c[i * n1 + j]));
  return {a.m, a.n - 1, c};
}
Bern2 elevate_u(const Bern2& a) {
  int m = a.m + 1, n1 = a.n + 1;
  std::vector<double> c((m + 1) * n1, 0.0);
  for (int i = 0; i <= m; ++i) {
    double al = double(i) / m;
    for (int j = 0; j < n1; ++j) {
      double lo = i > 0 ? a.c[(i - 1) * n1 + j] : 0.0, hi = i < m ? a.c[i * n1 + j] : 0.0;
      c[i * n1 + j] = al * lo + (1.0 - al) * hi;
    }
  }
  return {m, a.n, c};
}
Bern2 elevate_v(const Bern2& a) {
  int n = a.n + 1, n0 = a.n + 1;
  std::vector<double> c((a.m + 1) * (n + 1), 0.0);
  for (int i = 0; i <= a.m; ++i)
    for (int j = 0; j <= n; ++j) {
      double al = double(j) / n;
      double lo = j > 0 ? a.c[i * n0 + j - 1] : 0.0, hi = j < n ? a.c[i * n0 + j] : 0.0;
      c[i * (n + 1) + j] = al * lo + (1.0 - al) * hi;
    }
  return {a.m, n, c};
}
Bern2 axpy(const Bern2& x, double a, const Bern2& o, double b) {
  Bern2 r = x;
  for (size_t k = 0; k < r.c.size(); ++k) r.c[k] = a * x.c[k] + b * o.c[k];
  return r;
}
double max_abs(const Bern2& a) {
  double m = 0;
  for (double x : a.c) m = std::max(m, std::fabs(x));
  return m;
}

// surface3.rs:694-731 no_critical_point: the squared distance has no critical point
// on the closed patch when the combination of its gradient along the middle's
// gradient has Bernstein coefficients clear of zero
bool no_critical_point(const Patch& pa, V3 p) {
  int m = pa.m, n = pa.n;
  auto mk = [&](int which) {
    Bern2 b{m, n, std::vector<double>((m + 1) * (n + 1))};
    for (int i = 0; i <= m; ++i)
      for (int j = 0; j <= n; ++j) {
        const double* h = pa.pw[i][j];
        double x = which == 3 ? h[3] : h[which] - (which == 0 ? p.x : which == 1 ? p.y : p.z) * h[3];
        b.c[i * (n + 1) + j] = x;
      }
    return b;
  };
  Bern2 w = mk(3);
  for (double x : w.c)
    if (!(x > 0)) return false;
  SEval e = pa.eval(0.5, 0.5);
  V3 r = e.p - p;
  double a = dot(e.ps, r), b = dot(e.pt, r);
  double l = std::hypot(a, b);
  if (!(l > 0)) return false;
  a /= l, b /= l;
  Bern2 q[3] = {mk(0), mk(1), mk(2)};
  Bern2 qu[3] = {d_u(q[0]), d_u(q[1]), d_u(q[2])}, qv[3] = {d_v(q[0]), d_v(q[1]), d_v(q[2])};
  auto dot3 = [](const Bern2* x, const Bern2* y) {
    Bern2 s = bmul(x[0], y[0]);
    for (int k = 1; k < 3; ++k) s = axpy(s, 1.0, bmul(x[k], y[k]), 1.0);
    return s;
  };
  Bern2 qq = dot3(q, q);
  Bern2 gs = axpy(bmul(w, dot3(q, qu)), 1.0, bmul(d_u(w), qq), -1.0);
  Bern2 gt = axpy(bmul(w, dot3(q, qv)), 1.0, bmul(d_v(w), qq), -1.0);
  Bern2 f = axpy(elevate_u(gs), a, elevate_v(gt), b);
  auto qmax = [](const Bern2* x) { return std::max(max_abs(x[0]), std::max(max_abs(x[1]), max_abs(x[2]))); };
  double q0 = qmax(q), q1 = qmax(qu), q2 = qmax(qv);
  double size = max_abs(w) * 3.0 * q0 * (std::fabs(a) * q1 + std::fabs(b) * q2) +
                (std::fabs(a) * max_abs(d_u(w)) + std::fabs(b) * max_abs(d_v(w))) * 3.0 * q0 * q0;
  for (double x : f.c)
    if (!(x > 1e-11 * size)) return false;
  return true;
}

// surface3.rs:973-975 slab_bound
double slab_bound(const Patch& pa, V3 p, V3 n) {
  double m = INF;
  for (int i = 0; i <= pa.m; ++i)
    for (int j = 0; j <= pa.n; ++j) m = std::min(m, dot(p - pa.ctrl(i, j), n));
  return m;
}

const double DIST_EPS = 1e-9;  // surface3.rs:555
const int SEARCH_NEWTON = 12;  // surface3.rs:552

struct SurfHit {
  V3 q;
  double dist;
  double u, v;
  V3 n;
};

// Bezier decomposition of a clamped B-spline by knot insertion (The NURBS Book
// algorithm 5.1; curve3.rs:232-290), flat points of dimension N.
std::vector<std::pair<std::vector<double>, std::pair<double, double>>> decompose(int p, std::vector<double> kn, std::vector<double> pts, int N) {
  auto count = [&](double u) {
    int c = 0;
    for (double k : kn) c += (k == u);
    return c;
  };
  std::vector<double> interior;
  for (size_t i = p + 1; i + p + 1 < kn.size(); ++i)
    if (kn[i] > kn[p] && kn[i] < kn[kn.size() - 1 - p] && (interior.empty() || interior.back() != kn[i])) interior.push_back(kn[i]);
  for (double u : interior) {
    while (count(u) < p) {
      int k = 0;
      for (int i = int(kn.size()) - 2; i >= 0; --i)
        if (kn[i] <= u && kn[i + 1] > u) {
          k = i;
          break;
        }
      int s = 0;
      for (int i = k; i >= 0 && kn[i] == u; --i) ++s;
      int np = int(pts.size()) / N;
      std::vector<double> q;
      for (int i = 0; i <= k - p; ++i)
        for (int c = 0; c < N; ++c) q.push_back(pts[i * N + c]);
      for (int i = k - p + 1; i <= k - s; ++i) {
        double a = (u - kn[i]) / (kn[i + p] - kn[i]);
        for (int c = 0; c < N; ++c) q.push_back(a * pts[i * N + c] + (1 - a) * pts[(i - 1) * N + c]);
      }
      for (int i = k - s; i < np; ++i)
        for (int c = 0; c < N; ++c) q.push_back(pts[i * N + c]);
      kn.insert(kn.begin() + k + 1, u);
      pts = q;
    }
  }
  // spans: distinct consecutive knots
  std::vector<double> brk;
  for (size_t i = p; i + p < kn.size(); ++i)
    if (brk.empty() || brk.back() != kn[i]) brk.push_back(kn[i]);
  std::vector<std::pair<std::vector<double>, std::pair<double, double>>> out;
  int np = int(pts.size()) / N;
  int sp = 0;
  for (int s = 0; s + p < np; s += p, ++sp)
    out.push_back({std::vector<double>(pts.begin() + s * N, pts.begin() + (s + p + 1) * N), {brk[sp], brk[sp + 1]}});
  return out;
}

// surface3.rs:482-520 NurbsSurf, kept as its span patches
struct Nurbs {
  int du, dv, nsu, nsv;
  std::vector<Patch> spans;  // span (a, b) at a * nsv + b
  std::vector<double> su0, sv0;  // span starts, plus the end
  std::vector<double> knots_u, knots_v;
  Box box;
  // surface3.rs:579-591, 1378-1388 Creases: the parameter values of the sides the surface
  // collapses on (a patch touching one is never cleared by the critical-point test)
  std::vector<double> crease_u, crease_v;
  bool touches(const Patch& pa) const {
    auto on = [](double x, const std::vector<double>& lines, double a, double b) {
      for (double c : lines)
        if (std::fabs(x - c) <= 1e-12 * std::max(std::fabs(b - a), std::fabs(c))) return true;
      return false;
    };
    return on(pa.u0, crease_u, pa.u0, pa.u1) || on(pa.u1, crease_u, pa.u0, pa.u1) || on(pa.v0, crease_v, pa.v0, pa.v1) || on(pa.v1, crease_v, pa.v0, pa.v1);
  }

  void build(const icam_face_desc& f) {
    du = f.du, dv = f.dv;
    std::vector<double> ku(f.ku, f.ku + f.nu + f.du + 1), kv(f.kv, f.kv + f.nv + f.dv + 1);
    // along u for each column j: the rows of Bezier points
    std::vector<std::vector<std::vector<double>>> cols;  // [j][span a] flat (du+1)*4
    std::vector<std::pair<double, double>> ur, vr;
    for (uint32_t j = 0; j < f.nv; ++j) {
      std::vector<double> pts;
      for (uint32_t i = 0; i < f.nu; ++i) pts.insert(pts.end(), f.pw + (i * f.nv + j) * 4, f.pw + (i * f.nv + j) * 4 + 4);
      auto d = decompose(du, ku, pts, 4);
      std::vector<std::vector<double>> segs;
      ur.clear();
      for (auto& s : d) segs.push_back(s.first), ur.push_back(s.second);
      cols.push_back(segs);
    }
    nsu = int(ur.size());
    // for each u span a and Bezier row i: decompose along v
    std::vector<std::vector<std::vector<std::vector<double>>>> rows(nsu);  // [a][i][span b]
    for (int a = 0; a < nsu; ++a) {
      rows[a].resize(du + 1);
      for (int i = 0; i <= du; ++i) {
        std::vector<double> pts;
        for (uint32_t j = 0; j < f.nv; ++j) pts.insert(pts.end(), cols[j][a].begin() + i * 4, cols[j][a].begin() + i * 4 + 4);
        auto d = decompose(dv, kv, pts, 4);
        vr.clear();
        for (auto& s : d) rows[a][i].push_back(s.first), vr.push_back(s.second);
      }
    }
    nsv = int(vr.size());
    spans.resize(nsu * nsv);
    for (int a = 0; a < nsu; ++a)
      for (int b = 0; b < nsv; ++b) {
        Patch& pa = spans[a * nsv + b];
        pa.m = du, pa.n = dv;
        for (int i = 0; i <= du; ++i)
          for (int j = 0; j <= dv; ++j)
            for (int c = 0; c < 4; ++c) pa.pw[i][j][c] = rows[a][i][b][j * 4 + c];
        pa.u0 = ur[a].first, pa.u1 = ur[a].second, pa.v0 = vr[b].first, pa.v1 = vr[b].second;
        box.add(pa.bbox());
      }
    knots_u = ku, knots_v = kv;
    for (auto& r : ur) su0.push_back(r.first);
    su0.push_back(ur.back().second);
    for (auto& r : vr) sv0.push_back(r.first);
    sv0.push_back(vr.back().second);
  }
  int span_of(const std::vector<double>& s0, double x) const {
    int n = int(s0.size()) - 1;
    for (int i = n - 1; i >= 0; --i)
      if (s0[i] <= x) return i;
    return 0;
  }
  SEval eval(double u, double v) const {
    int a = span_of(su0, u), b = span_of(sv0, v);
    const Patch& pa = spans[a * nsv + b];
    double s = (u - pa.u0) / (pa.u1 - pa.u0), t = (v - pa.v0) / (pa.v1 - pa.v0);
    SEval e = pa.eval(std::min(std::max(s, 0.0), 1.0), std::min(std::max(t, 0.0), 1.0));
    double su = 1.0 / (pa.u1 - pa.u0), sv = 1.0 / (pa.v1 - pa.v0);
    e.ps = e.ps * su, e.pt = e.pt * sv, e.pss = e.pss * (su * su), e.pst = e.pst * (su * sv), e.ptt = e.ptt * (sv * sv);
    return e;
  }

  // surface3.rs:1323-1365 descend: Newton carried across spans to a critical point
  bool polish(V3 p, double u, double v, SurfHit& out) const {
    int a = span_of(su0, u), b = span_of(sv0, v);
    const Patch* pa = &spans[a * nsv + b];
    double s = (u - pa->u0) / (pa->u1 - pa->u0), t = (v - pa->v0) / (pa->v1 - pa->v0);
    for (int hop = 0; hop < 16; ++hop) {
      double ns = std::min(std::max(s, 0.0), 1.0), nt = std::min(std::max(t, 0.0), 1.0);
      pa->newton(p, ns, nt, 40);
      SEval e = pa->eval(ns, nt);
      if (critical(p, e)) {
        out = {e.p, len(e.p - p), pa->u0 + ns * (pa->u1 - pa->u0), pa->v0 + nt * (pa->v1 - pa->v0), unit(cross(e.ps, e.pt))};
        return true;
      }
      V3 r = e.p - p;
      double gs = dot(r, e.ps), gt = dot(r, e.pt);
      bool moved = false;
      int na = a, nb = b;
      if (ns >= 1.0 && gs < 0 && a + 1 < nsu) {
        na = a + 1, s = 0.0, moved = true;
      } else if (ns <= 0.0 && gs > 0 && a > 0) {
        na = a - 1, s = 1.0, moved = true;
      } else {
        s = ns;
      }
      if (nt >= 1.0 && gt < 0 && b + 1 < nsv) {
        nb = b + 1, t = 0.0, moved = true;
      } else if (nt <= 0.0 && gt > 0 && b > 0) {
        nb = b - 1, t = 1.0, moved = true;
      } else {
        t = nt;
      }
      if (!moved) return false;
      a = na, b = nb;
      pa = &spans[a * nsv + b];
    }
    return false;
  }

  struct Item {
    double lb;
    int idx;  // into a pool
    int depth;
    bool inside;
    bool operator<(const Item& o) const { return lb > o.lb; }
  };

  // surface3.rs:1214-1298 closest_trimmed: branch and bound over patches; only
  // critical points inside the loops answer; returns (hit?, complete)
  bool closest_trimmed(V3 p, double cutoff, const Trim* trim, SurfHit& best, bool& found, int& pops_out) const {
    found = false;
    double best_d = cutoff;
    // the thread's pool, kept between calls (a patch of degree 7 takes 2 KB)
    thread_local std::vector<Patch> pool;
    pool.clear();
    std::priority_queue<Item> heap;
    for (int k = 0; k < nsu * nsv; ++k) {
      double lb = box_dist(p, spans[k].bbox());
      if (lb < best_d - DIST_EPS) {
        pool.push_back(spans[k]);
        heap.push({lb, int(pool.size() - 1), 0, false});
      }
    }
    int pops = 0;
    auto answer = [&](const SurfHit& h) {
      if (!found || h.dist < best.dist) best = h, found = true;
    };
    while (!heap.empty()) {
      Item it = heap.top();
      heap.pop();
      if (it.lb >= best_d - DIST_EPS || best_d <= DIST_EPS) break;
      if (++pops > 20000) {
        pops_out = pops;
        return false;
      }
      const Patch& pa = pool[it.idx];  // read before the pool grows
      int state = it.inside ? 1 : (trim ? trim->box_class(std::array<double, 2>{pa.u0, pa.v0}.data(), std::array<double, 2>{pa.u1, pa.v1}.data()) : 1);
      if (state == 0) continue;
      bool now_in = state == 1;
      double s = 0.5, t = 0.5;
      pa.newton(p, s, t, SEARCH_NEWTON);
      SEval e = pa.eval(s, t);
      double d = len(e.p - p);
      V3 n = unit(cross(e.ps, e.pt));
      double u = pa.u0 + s * (pa.u1 - pa.u0), v = pa.v0 + t * (pa.v1 - pa.v0);
      // (a critical point within DIST_EPS of a bound no answer set yet, and a polished
      // one within DIST_EPS of its start, are taken: where the foot lies on the shared
      // border of two spans, the first span's Newton stops at the border, not critical,
      // and sets the bound; the next span's critical point then ties it in rounding)
      if (d < best_d + (found ? 0.0 : DIST_EPS) && (now_in || !trim || trim->inside(u, v))) {
        best_d = std::min(best_d, d);
        if (critical(p, e)) {
          answer({e.p, d, u, v, n});
        } else {
          SurfHit h;
          if (polish(p, u, v, h) && h.dist <= d + DIST_EPS && (!trim || trim->inside(h.u, h.v))) {
            best_d = std::min(best_d, h.dist);
            answer(h);
          }
        }
      }
      double slab = slab_bound(pa, p, d > 1e-12 ? (p - e.p) * (1.0 / d) : n);
      if (slab >= best_d - DIST_EPS || it.depth >= 40) continue;
      if (!touches(pa) && no_critical_point(pa, p)) continue;
      Patch q[4];
      pa.quarters(q);
      for (int k = 0; k < 4; ++k) {
        double lbq = std::max(box_dist(p, q[k].bbox()), slab);
        if (lbq < best_d - DIST_EPS) {
          pool.push_back(q[k]);
          heap.push({lbq, int(pool.size() - 1), it.depth + 1, now_in});
        }
      }
    }
    pops_out = pops;
    return true;
  }

  // surface3.rs:1444-1500 closest (untrimmed)
  bool closest(V3 p, double cutoff, SurfHit& best) const {
    bool found = false;
    double best_d = cutoff;
    thread_local std::vector<Patch> pool;
    pool.clear();
    std::priority_queue<Item> heap;
    for (int k = 0; k < nsu * nsv; ++k) {
      double lb = box_dist(p, spans[k].bbox());
      if (lb < best_d - DIST_EPS) {
        pool.push_back(spans[k]);
        heap.push({lb, int(pool.size() - 1), 0, false});
      }
    }
    int pops = 0;
    while (!heap.empty()) {
      Item it = heap.top();
      heap.pop();
      if (it.lb >= best_d - DIST_EPS || best_d <= DIST_EPS) break;
      if (++pops > 20000) break;
      const Patch& pa = pool[it.idx];  // read before the pool grows
      double s = 0.5, t = 0.5;
      pa.newton(p, s, t, SEARCH_NEWTON);
      SEval e = pa.eval(s, t);
      double d = len(e.p - p);
      V3 n = unit(cross(e.ps, e.pt));
      if (d < best_d) {
        best_d = d;
        best = {e.p, d, pa.u0 + s * (pa.u1 - pa.u0), pa.v0 + t * (pa.v1 - pa.v0), n};
        found = true;
      }
      double slab = slab_bound(pa, p, d > 1e-12 ? (p - e.p) * (1.0 / d) : n);
      if (slab >= best_d - DIST_EPS || it.depth >= 40) continue;
      Patch q[4];
      pa.quarters(q);
      for (int k = 0; k < 4; ++k) {
        double lbq = std::max(box_dist(p, q[k].bbox()), slab);
        if (lbq < best_d - DIST_EPS) {
          pool.push_back(q[k]);
          heap.push({lbq, int(pool.size() - 1), it.depth + 1, false});
        }
      }
    }
    return found;
  }
};

// ------------------------------------------------------------------ curves
// curve3.rs:64-213 Bez3: a rational Bezier piece, homogeneous, local s in [0, 1]
struct Bez3 {
  int deg = 0;
  double pw[MAXP][4];
  double t0 = 0, t1 = 1;
  V3 pt(int i) const { return V3(pw[i][0] / pw[i][3], pw[i][1] / pw[i][3], pw[i][2] / pw[i][3]); }
  V3 start() const { return pt(0); }
  V3 end() const { return pt(deg); }
  struct E {
    V3 p, d1, d2;
  };
  // curve3.rs:129-156
  E eval(double s) const {
    count(C_CEVAL);
    int n = deg;
    double b[MAXP][4];
    std::memcpy(b, pw, sizeof(b[0]) * (deg + 1));
    double d1h[4] = {0}, d2h[4] = {0};
    for (int lev = 1; lev <= n; ++lev) {
      if (n >= 2 && lev == n - 1)
        for (int c = 0; c < 4; ++c) d2h[c] = double(n * (n - 1)) * (b[2][c] - 2.0 * b[1][c] + b[0][c]);
      if (lev == n)
        for (int c = 0; c < 4; ++c) d1h[c] = n * (b[1][c] - b[0][c]);
      for (int i = 0; i <= n - lev; ++i)
        for (int c = 0; c < 4; ++c) b[i][c] = (1.0 - s) * b[i][c] + s * b[i + 1][c];
    }
    double w = b[0][3];
    E e;
    e.p = V3(b[0][0] / w, b[0][1] / w, b[0][2] / w);
    e.d1 = (V3(d1h[0], d1h[1], d1h[2]) - e.p * d1h[3]) * (1.0 / w);
    e.d2 = (V3(d2h[0], d2h[1], d2h[2]) - e.d1 * (2.0 * d1h[3]) - e.p * d2h[3]) * (1.0 / w);
    return e;
  }
  // curve3.rs:157-174
  void split(double s, Bez3& l, Bez3& r) const {
    l = *this, r = *this;
    double b[MAXP][4];
    std::memcpy(b, pw, sizeof(b[0]) * (deg + 1));
    for (int c = 0; c < 4; ++c) l.pw[0][c] = b[0][c], r.pw[deg][c] = b[deg][c];
    for (int lev = 1; lev <= deg; ++lev) {
      for (int i = 0; i <= deg - lev; ++i)
        for (int c = 0; c < 4; ++c) b[i][c] = (1.0 - s) * b[i][c] + s * b[i + 1][c];
      for (int c = 0; c < 4; ++c) l.pw[lev][c] = b[0][c], r.pw[deg - lev][c] = b[deg - lev][c];
    }
    double tm = t0 + s * (t1 - t0);
    l.t1 = tm, r.t0 = tm;
  }
  // curve3.rs:180-186 turning of the control polygon
  double turning() const {
    std::vector<V3> legs;
    for (int i = 0; i < deg; ++i) {
      V3 d = pt(i + 1) - pt(i);
      if (len(d) > 1e-14) legs.push_back(d);
    }
    double t = 0;
    for (size_t i = 0; i + 1 < legs.size(); ++i) t += std::atan2(len(cross(legs[i], legs[i + 1])), dot(legs[i], legs[i + 1]));
    return t;
  }
  static double seg_dist(V3 p, V3 a, V3 b) {
    V3 ab = b - a;
    double l2 = dot(ab, ab);
    double t = l2 > 0 ? std::min(std::max(dot(p - a, ab) / l2, 0.0), 1.0) : 0.0;
    return len(p - (a + ab * t));
  }
  // curve3.rs:187-191
  double thickness() const {
    double m = 0;
    for (int i = 0; i <= deg; ++i) m = std::max(m, seg_dist(pt(i), start(), end()));
    return m;
  }
  // curve3.rs:192-213 newton
  double newton(V3 p, double s) const {
    for (int it = 0; it < 40; ++it) {
      E e = eval(s);
      count(C_NEWTON1);
      V3 r = e.p - p;
      double f = dot(e.d1, r);
      double fp = dot(e.d2, r) + dot(e.d1, e.d1);
      if (fp <= 0) fp = dot(e.d1, e.d1);
      if (std::fabs(fp) < 1e-300) break;
      double ns = std::min(std::max(s - f / fp, 0.0), 1.0);
      if (std::fabs(ns - s) < 1e-15) return ns;
      s = ns;
    }
    return s;
  }
  Box box() const {
    Box b;
    for (int i = 0; i <= deg; ++i) b.add(pt(i));
    return b;
  }
};

enum Near { INTERIOR, START, END };
struct Foot {
  V3 q;
  double dist;
  double t;
  Near near;
};

// curve3.rs:338-476 Spline3: flat pieces (turning <= 0.35 rad) and their B&B
struct Spline {
  std::vector<Bez3> pieces;
  std::vector<double> thick;
  std::vector<Box> boxes;
  void build(const std::vector<Bez3>& in) {
    std::vector<Bez3> stack(in.rbegin(), in.rend());
    while (!stack.empty()) {
      Bez3 z = stack.back();
      stack.pop_back();
      if (z.turning() <= 0.35 || std::fabs(z.t1 - z.t0) < 1e-12) {
        pieces.push_back(z);
      } else {
        Bez3 l, r;
        z.split(0.5, l, r);
        stack.push_back(r);
        stack.push_back(l);
      }
    }
    for (auto& z : pieces) thick.push_back(z.thickness()), boxes.push_back(z.box());
  }
  // curve3.rs:408-476 closest
  bool closest(V3 p, double cutoff, Foot& best) const {
    bool found = false;
    double best_d = cutoff;
    struct It {
      double lb;
      int i;
      Bez3 z;
      int depth;
    };
    std::vector<It> stack;
    for (size_t i = 0; i < pieces.size(); ++i) {
      double lb = std::max(box_dist(p, boxes[i]), Bez3::seg_dist(p, pieces[i].start(), pieces[i].end()) - thick[i]);
      if (lb < best_d) stack.push_back({lb, int(i), pieces[i], 0});
    }
    auto sort = [&]() { std::stable_sort(stack.begin(), stack.end(), [](const It& a, const It& b) { return a.lb > b.lb; }); };
    sort();
    int popped = 0;
    while (!stack.empty()) {
      It it = stack.back();
      stack.pop_back();
      if (it.lb >= best_d - 1e-12) continue;
      if (++popped > 4096) break;
      const Bez3& z = it.z;
      V3 a = z.start(), b = z.end(), ab = b - a;
      double l2 = dot(ab, ab);
      double f = l2 > 0 ? std::min(std::max(dot(p - a, ab) / l2, 0.0), 1.0) : 0.5;
      for (double s0 : {f, 0.0, 1.0, -1.0, 2.0}) {
        double s = s0 < 0 ? 0.0 : s0 > 1 ? 1.0 : z.newton(p, s0);
        Bez3::E e = z.eval(s);
        double d = len(p - e.p);
        if (d < best_d) {
          best_d = d;
          bool last = it.i == int(pieces.size()) - 1 && s >= 1.0 && z.t1 == pieces[it.i].t1;
          bool first = it.i == 0 && s <= 0.0 && z.t0 == pieces[it.i].t0;
          best = {e.p, d, z.t0 + s * (z.t1 - z.t0), first ? START : last ? END : INTERIOR};
          found = true;
        }
      }
      if (it.depth < 40 && z.thickness() > 1e-12) {
        Bez3 l, r;
        z.split(0.5, l, r);
        for (Bez3* c : {&l, &r}) {
          double lbc = Bez3::seg_dist(p, c->start(), c->end()) - c->thickness();
          if (lbc < best_d - 1e-12) stack.push_back({lbc, it.i, *c, it.depth + 1});
        }
        sort();
      }
    }
    return found;
  }
};

// curve3.rs:56-60 swept
double swept(double t0, double t, double sweep) {
  double d = sweep >= 0 ? t - t0 : t0 - t;
  return d - TAU * std::floor(d / TAU);
}

// curve3.rs:489-610 Edge3
struct Edge {
  int kind;
  V3 a, b;
  Frame f;
  double r, r2 = 0, t0, sweep;
  Spline spline;
  int faces[2];
  const double* uvmap = nullptr;  // per piece, per face (u0 v0 u1 v1)
  std::vector<double> uvm;
  std::vector<Bez3> raw;  // the pieces as given (for the uv map)
  Box box;

  V3 at(double s) const {
    double t = t0 + s * sweep;
    return f.world(V3(r * std::cos(t), r * std::sin(t), 0));
  }
  // curve3.rs:525-562 (line, arc) and :605-610 closest_below
  bool closest_below(V3 p, double cutoff, Foot& out) const {
    if (kind == ICAM_SPLINE) return spline.closest(p, cutoff, out) && out.dist < cutoff;
    if (kind == ICAM_ELLIPSE) {
      // curve3.rs:557-590 in the ellipse's plane: the critical points of |E(t) - p|^2 (at
      // most four) by Newton from 64 seeds over the arc, and the two ends
      const V3 l = f.local(p);
      const double a = r, b = r2;
      auto e = [&](double t) { return V3(a * std::cos(t), b * std::sin(t), 0); };
      auto dist = [&](double t) { return len(e(t) - l); };
      double bd = dist(t0), bt = t0;
      Near near = START;
      const double te = t0 + sweep;
      if (dist(te) < bd) bd = dist(te), bt = te, near = END;
      for (int k = 0; k < 64; ++k) {
        double t = t0 + sweep * (k + 0.5) / 64.0;
        for (int it = 0; it < 30; ++it) {
          const double sn = std::sin(t), cs = std::cos(t);
          const double g = (b * b - a * a) * sn * cs + a * l.x * sn - b * l.y * cs;
          const double gd = (b * b - a * a) * (cs * cs - sn * sn) + a * l.x * cs + b * l.y * sn;
          if (std::fabs(gd) < 1e-300) break;
          const double nt = t - g / gd;
          if (std::fabs(nt - t) < 1e-15) {
            t = nt;
            break;
          }
          t = nt;
        }
        if (swept(t0, t, sweep) <= std::fabs(sweep) && dist(t) < bd) bd = dist(t), bt = t, near = INTERIOR;
      }
      out = {f.world(e(bt)), bd, bt, near};
      return out.dist < cutoff;
    }
    count(kind == ICAM_LINE ? C_EDGE_LINE : C_EDGE_ARC);
    if (kind == ICAM_LINE) {
      V3 ab = b - a;
      double l2 = dot(ab, ab);
      double t = l2 > 0 ? std::min(std::max(dot(p - a, ab) / l2, 0.0), 1.0) : 0.0;
      V3 q = a + ab * t;
      out = {q, len(p - q), t, t <= 0 ? START : t >= 1 ? END : INTERIOR};
    } else {
      V3 l = f.local(p);
      bool full = std::fabs(std::fabs(sweep) - TAU) < 1e-12;
      double t = std::atan2(l.y, l.x);
      if ((l.x != 0 || l.y != 0) && (full || swept(t0, t, sweep) <= std::fabs(sweep))) {
        V3 q = f.world(V3(r * std::cos(t), r * std::sin(t), 0));
        out = {q, len(p - q), t, INTERIOR};
      } else {
        V3 qa = at(0.0), qb = at(1.0);
        if (len(p - qa) <= len(p - qb))
          out = {qa, len(p - qa), t0, START};
        else
          out = {qb, len(p - qb), t0 + sweep, END};
      }
    }
    return out.dist < cutoff;
  }
};

// ------------------------------------------------------------------ the evaluator
struct Face {
  int kind;
  double sense;
  Analytic an;
  std::unique_ptr<Nurbs> ns;
  Trim trim;
  std::vector<std::pair<V3, V3>> tips;
  Box box;
};

// solid3.rs:96-236 BoxTree: median split over boxes sorted by centre x, leaves of
// at most 4 boxes, best-first traversal
struct BoxTree {
  std::vector<Box> boxes;
  std::vector<int> ids;
  struct Node {
    Box b;
    int a, e;
  };
  std::vector<Node> nodes;
  void build(const std::vector<Box>& in) {
    std::vector<int> order(in.size());
    for (size_t i = 0; i < in.size(); ++i) order[i] = int(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return in[a].lo.x + in[a].hi.x < in[b].lo.x + in[b].hi.x; });
    for (int i : order) boxes.push_back(in[i]), ids.push_back(i);
    if (!boxes.empty()) build_at(0, 0, int(boxes.size()));
  }
  Box build_at(int k, int a, int b) {
    if (int(nodes.size()) <= k) nodes.resize(k + 1);
    Box bx;
    if (b - a <= 4) {
      for (int i = a; i < b; ++i) bx.add(boxes[i]);
    } else {
      int m = (a + b) / 2;
      Box l = build_at(2 * k + 1, a, m), r = build_at(2 * k + 2, m, b);
      bx.add(l), bx.add(r);
    }
    nodes[k] = {bx, a, b};
    return bx;
  }
  template <class F>
  void nearest_first(V3 p, double bound, F&& f) const {
    if (nodes.empty()) return;
    struct It {
      double lb;
      int k;
      bool leaf;
      bool operator<(const It& o) const { return lb > o.lb; }
    };
    std::priority_queue<It> heap;
    heap.push({box_dist(p, nodes[0].b), 0, false});
    while (!heap.empty()) {
      It it = heap.top();
      heap.pop();
      if (it.lb >= bound) break;
      if (it.leaf) {
        bound = f(ids[it.k]);
        continue;
      }
      const Node& nd = nodes[it.k];
      if (nd.e - nd.a <= 4) {
        for (int i = nd.a; i < nd.e; ++i) {
          double d = box_dist(p, boxes[i]);
          if (d < bound) heap.push({d, i, true});
        }
      } else {
        for (int c : {2 * it.k + 1, 2 * it.k + 2}) {
          double d = box_dist(p, nodes[c].b);
          if (d < bound) heap.push({d, c, false});
        }
      }
    }
  }
};

struct Sd {
  double d;
  V3 q;
  uint32_t feat;
  bool unsure;
  bool cusp;
};

// surface3.rs:1392-1411 collapsed_sides: the sides of a B-spline surface's domain along
// which it collapses to a point (a pole): per side, the direction held (0: u), its
// value and the point
struct Collapsed {
  int k;
  double value;
  V3 x;
};
std::vector<Collapsed> collapsed_sides(const Nurbs& ns) {
  const double scale = std::max(len(ns.box.hi - ns.box.lo), 1e-9);
  const double r[2][2] = {{ns.su0.front(), ns.su0.back()}, {ns.sv0.front(), ns.sv0.back()}};
  std::vector<Collapsed> out;
  for (int k = 0; k < 2; ++k)
    for (double side : {r[k][0], r[k][1]}) {
      const double* other = r[1 - k];
      auto at = [&](double t) {
        const double w = other[0] + (other[1] - other[0]) * t;
        return k == 0 ? ns.eval(side, w).p : ns.eval(w, side).p;
      };
      const V3 p0 = at(0.0);
      bool all = true;
      for (int i = 1; i <= 16; ++i) all &= len(at(i / 16.0) - p0) <= 1e-12 * scale;
      if (all) out.push_back({k, side, p0});
    }
  return out;
}

// solid3.rs:1955-1986 fan_normal: the pseudo-normal at a pole, the surface there being
// the cone of its limiting tangents S_k(w) with limiting normals S_k x S_kw (u held) or
// S_kw x S_v (v held), each weighted by the angle its tangent turns through, over the
// part of the side the trim covers; in the surface's own orientation
V3 fan_normal(const Nurbs& ns, const Trim& trim, int k, double value) {
  const double r[2][2] = {{ns.su0.front(), ns.su0.back()}, {ns.sv0.front(), ns.sv0.back()}};
  const double* range = r[k];
  const double* other = r[1 - k];
  const double h = value == range[0] ? 1.0 : -1.0;
  const double step = 1e-9 * (range[1] - range[0]) * h;
  const int n = 512;
  V3 out(0, 0, 0), pt(0, 0, 0), pn(0, 0, 0);
  bool have = false, pc = false;
  for (int i = 0; i <= n; ++i) {
    const double w = other[0] + (other[1] - other[0]) * i / n;
    const double u = k == 0 ? value : w, v = k == 0 ? w : value;
    const SEval e = ns.eval(u, v);
    V3 tan = k == 0 ? e.ps : e.pt, nrm = k == 0 ? cross(e.ps, e.pst) : cross(e.pst, e.pt);
    tan = tan * h, nrm = nrm * h;
    const bool covered = k == 0 ? trim.inside(value + step, w) : trim.inside(w, value + step);
    if (have && pc && covered && len(tan) > 0 && len(pt) > 0) {
      const V3 a = pt * (1.0 / len(pt)), b = tan * (1.0 / len(tan));
      const double ang = std::atan2(len(cross(a, b)), dot(a, b));
      const V3 mid = pn * (1.0 / len(pn)) + nrm * (1.0 / len(nrm));
      if (len(mid) > 0) out = out + mid * (ang / len(mid));
    }
    pt = tan, pn = nrm, pc = covered, have = true;
  }
  return out;
}

// surface3.rs:734-760 curve_no_critical_point: |C - p|^2 has no critical point on a
// rational Bezier curve (positive weights) when w Q.Q' - w' Q.Q, Q = P - p w, has
// Bernstein coefficients of one strict sign past a rounding allowance
bool curve_no_critical_point(const Bez3& z, V3 p) {
  const int m = z.deg;
  auto mk = [&](int c) {
    Bern2 b{m, 0, std::vector<double>(m + 1)};
    for (int i = 0; i <= m; ++i) b.c[i] = c == 3 ? z.pw[i][3] : z.pw[i][c] - (c == 0 ? p.x : c == 1 ? p.y : p.z) * z.pw[i][3];
    return b;
  };
  const Bern2 w = mk(3);
  for (double x : w.c)
    if (!(x > 0)) return false;
  const Bern2 q[3] = {mk(0), mk(1), mk(2)};
  const Bern2 qu[3] = {d_u(q[0]), d_u(q[1]), d_u(q[2])};
  auto dotb = [](const Bern2* x, const Bern2* y) {
    Bern2 s = bmul(x[0], y[0]);
    for (int k = 1; k < 3; ++k) s = axpy(s, 1.0, bmul(x[k], y[k]), 1.0);
    return s;
  };
  const Bern2 wu = d_u(w);
  const Bern2 f = axpy(bmul(w, dotb(q, qu)), 1.0, bmul(wu, dotb(q, q)), -1.0);
  double q0 = 0, q1 = 0;
  for (int k = 0; k < 3; ++k) q0 = std::max(q0, max_abs(q[k])), q1 = std::max(q1, max_abs(qu[k]));
  const double size = max_abs(w) * 3.0 * q0 * q1 + max_abs(wu) * 3.0 * q0 * q0;
  double lo = INF, hi = -INF;
  for (double x : f.c) lo = std::min(lo, x), hi = std::max(hi, x);
  return lo > 1e-11 * size || hi < -1e-11 * size;
}

// surface3.rs:1413-1443 c0_lines: the interior knot lines of multiplicity at least the
// degree, where the surface is only C0 and may fold: per line the direction held (0: u),
// its value, and one piece per span of the other direction (the near row or column of
// the patch after the line, parameters along the line) with the spans on either side
struct C0Piece {
  Bez3 z;
  int before, after;
};
struct C0Line {
  int face = 0, k = 0;
  double value = 0;
  std::vector<C0Piece> pieces;
};
std::vector<C0Line> c0_lines(const Nurbs& ns, int face) {
  std::vector<C0Line> out;
  for (int k = 0; k < 2; ++k) {
    const std::vector<double>& s0 = k == 0 ? ns.su0 : ns.sv0;
    const std::vector<double>& kn = k == 0 ? ns.knots_u : ns.knots_v;
    const int deg = k == 0 ? ns.du : ns.dv, nspan = k == 0 ? ns.nsu : ns.nsv, others = k == 0 ? ns.nsv : ns.nsu;
    for (int a = 1; a < nspan; ++a) {
      const double x = s0[a];
      if (!(x > s0.front() && x < s0.back()) || std::count(kn.begin(), kn.end(), x) < deg) continue;
      C0Line l;
      l.face = face, l.k = k, l.value = x;
      for (int b = 0; b < others; ++b) {
        C0Piece pc;
        pc.before = k == 0 ? (a - 1) * ns.nsv + b : b * ns.nsv + a - 1;
        pc.after = k == 0 ? a * ns.nsv + b : b * ns.nsv + a;
        const Patch& af = ns.spans[pc.after];
        pc.z.deg = k == 0 ? af.n : af.m;
        for (int i = 0; i <= pc.z.deg; ++i)
          for (int c = 0; c < 4; ++c) pc.z.pw[i][c] = k == 0 ? af.pw[0][i][c] : af.pw[i][0][c];
        pc.z.t0 = k == 0 ? af.v0 : af.u0, pc.z.t1 = k == 0 ? af.v1 : af.u1;
        l.pieces.push_back(pc);
      }
      out.push_back(l);
    }
  }
  return out;
}

struct Exact {
  std::vector<Face> faces;
  std::vector<Edge> edges;
  std::vector<std::pair<V3, V3>> verts;
  // solid3.rs:1553 poles: a B-spline face's collapsed side away from the vertices and
  // inside its loops, answering as a point with its fan normal times the face's sense
  struct Pole {
    int face;
    V3 x, n;
  };
  std::vector<Pole> poles;
  // solid3.rs:1549-1568 the interior C0 knot lines of B-spline faces, searched like edges
  std::vector<C0Line> lines;
  double tol;
  BoxTree tree;
  Box part;

  void open(const icam_part_desc* d) {
    tol = d->tol;
    part.lo = V3::of(d->lo), part.hi = V3::of(d->hi);
    faces.resize(d->n_faces);
    for (uint32_t i = 0; i < d->n_faces; ++i) {
      const icam_face_desc& fd = d->faces[i];
      Face& f = faces[i];
      f.kind = int(fd.kind);
      f.sense = fd.sense;
      if (fd.kind == ICAM_NURBS) {
        f.ns.reset(new Nurbs());
        f.ns->build(fd);
        for (const Collapsed& c : collapsed_sides(*f.ns)) (c.k == 0 ? f.ns->crease_u : f.ns->crease_v).push_back(c.value);
      } else {
        Analytic& a = f.an;
        a.kind = int(fd.kind);
        a.f.o = V3::of(fd.o), a.f.x = V3::of(fd.x), a.f.z = V3::of(fd.z);
        a.f.y = cross(a.f.z, a.f.x);
        a.c0 = fd.c[0], a.c1 = fd.c[1];
        if (a.kind == ICAM_TORUS) {
          // trim2.rs:1807-1818 the meridian arcs a torus's form covers
          if (a.c0 >= a.c1) {
            a.psi0 = 0, a.span = TAU;
          } else {
            double pp = std::acos(std::min(std::max(-a.c0 / a.c1, -1.0), 1.0));
            a.psi0 = -pp, a.span = 2 * pp;
          }
          a.narc = int(std::max(std::ceil(a.span / (0.5 * PI) - 1e-9), 1.0));
        }
      }
      Trim& t = f.trim;
      for (int k = 0; k < 2; ++k) t.lo[k] = fd.lo[k], t.hi[k] = fd.hi[k], t.period[k] = fd.period[k], t.ref[k] = fd.ref[k];
      t.ref_inside = fd.ref_inside != 0;
      for (uint32_t k = 0; k < fd.n_pieces; ++k) {
        std::array<double, 8> pc;
        std::copy(fd.pieces + 8 * k, fd.pieces + 8 * k + 8, pc.begin());
        t.pieces.push_back(pc);
      }
      t.place();
    }
    edges.resize(d->n_edges);
    for (uint32_t i = 0; i < d->n_edges; ++i) {
      const icam_edge_desc& ed = d->edges[i];
      Edge& e = edges[i];
      e.kind = int(ed.kind);
      e.faces[0] = int(ed.faces[0]), e.faces[1] = int(ed.faces[1]);
      if (ed.kind == ICAM_LINE) {
        e.a = V3::of(ed.a), e.b = V3::of(ed.b);
        e.box.add(e.a), e.box.add(e.b);
      } else if (ed.kind == ICAM_ELLIPSE) {
        e.f.o = V3::of(ed.o), e.f.x = V3::of(ed.x), e.f.z = V3::of(ed.z);
        e.f.y = cross(e.f.z, e.f.x);
        e.r = ed.r, e.r2 = ed.r2, e.t0 = ed.t0, e.sweep = ed.sweep;
        // curve3.rs:698-704 the box of the whole ellipse
        V3 axes[3] = {V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
        for (V3 ax : axes) {
          double ex = std::hypot(e.r * dot(e.f.x, ax), e.r2 * dot(e.f.y, ax));
          e.box.add(e.f.o + ax * ex), e.box.add(e.f.o - ax * ex);
        }
      } else if (ed.kind == ICAM_ARC) {
        e.f.o = V3::of(ed.o), e.f.x = V3::of(ed.x), e.f.z = V3::of(ed.z);
        e.f.y = cross(e.f.z, e.f.x);
        e.r = ed.r, e.t0 = ed.t0, e.sweep = ed.sweep;
        // solid3.rs:1217-1223 the box of its circle
        V3 axes[3] = {V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
        for (V3 ax : axes) {
          double ex = std::sqrt(std::max(1.0 - dot(e.f.z, ax) * dot(e.f.z, ax), 0.0)) * e.r;
          e.box.add(e.f.o + ax * ex), e.box.add(e.f.o - ax * ex);
        }
      } else {
        int n = int(ed.degree);
        double tpos = 0;
        for (uint32_t k = 0; k < ed.n_pieces; ++k) {
          Bez3 z;
          z.deg = n;
          for (int i2 = 0; i2 <= n; ++i2)
            for (int c = 0; c < 4; ++c) z.pw[i2][c] = ed.pw[(k * (n + 1) + i2) * 4 + c];
          z.t0 = tpos, z.t1 = tpos + 1.0;
          tpos += 1.0;
          e.raw.push_back(z);
          e.box.add(z.box());
        }
        e.spline.build(e.raw);
      }
      if (ed.uvmap) e.uvm.assign(ed.uvmap, ed.uvmap + 8 * std::max<uint32_t>(ed.n_pieces, 1));
    }
    verts.resize(d->n_verts);
    for (uint32_t i = 0; i < d->n_verts; ++i) verts[i] = {V3::of(d->verts[i].p), V3::of(d->verts[i].pn)};
    // solid3.rs:1211-1282 face boxes
    for (uint32_t i = 0; i < faces.size(); ++i) {
      Face& f = faces[i];
      // solid3.rs:1211-1282 face_box: the face's edges and its vertex loops' points
      const icam_face_desc& fdesc = d->faces[i];
      if (f.kind == ICAM_NURBS) {
        f.box = f.ns->box;
      } else if (f.kind == ICAM_PLANE) {
        for (auto& e : edges)
          if (e.faces[0] == int(i) || e.faces[1] == int(i)) f.box.add(e.box);
        for (uint32_t k = 0; k < fdesc.n_points; ++k) f.box.add(V3::of(fdesc.points + 3 * k));
      } else if (f.kind == ICAM_CYLINDER || f.kind == ICAM_CONE) {
        double h0 = INF, h1 = -INF;
        for (uint32_t k = 0; k < fdesc.n_points; ++k) {
          double h = dot(V3::of(fdesc.points + 3 * k) - f.an.f.o, f.an.f.z);
          h0 = std::min(h0, h), h1 = std::max(h1, h);
        }
        for (auto& e : edges)
          if (e.faces[0] == int(i) || e.faces[1] == int(i))
            for (int k = 0; k < 8; ++k) {
              V3 q((k & 1) ? e.box.hi.x : e.box.lo.x, (k & 2) ? e.box.hi.y : e.box.lo.y, (k & 4) ? e.box.hi.z : e.box.lo.z);
              double h = dot(q - f.an.f.o, f.an.f.z);
              h0 = std::min(h0, h), h1 = std::max(h1, h);
            }
        for (double h : {h0, h1}) {
          double rad = f.kind == ICAM_CONE ? std::fabs(f.an.c0 + h * std::tan(f.an.c1)) : f.an.c0;
          V3 axes[3] = {V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
          V3 c = f.an.f.o + f.an.f.z * h;
          for (V3 ax : axes) {
            double ex = std::sqrt(std::max(1.0 - dot(f.an.f.z, ax) * dot(f.an.f.z, ax), 0.0)) * rad;
            f.box.add(c + ax * ex), f.box.add(c - ax * ex);
          }
        }
      } else if (f.kind == ICAM_SPHERE) {
        V3 r(f.an.c0, f.an.c0, f.an.c0);
        f.box.add(f.an.f.o - r), f.box.add(f.an.f.o + r);
      } else {
        for (double h : {-f.an.c1, f.an.c1}) {
          V3 axes[3] = {V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
          V3 c = f.an.f.o + f.an.f.z * h;
          for (V3 ax : axes) {
            double ex = std::sqrt(std::max(1.0 - dot(f.an.f.z, ax) * dot(f.an.f.z, ax), 0.0)) * (f.an.c0 + f.an.c1);
            f.box.add(c + ax * ex), f.box.add(c - ax * ex);
          }
        }
      }
      // solid3.rs:1628-1642 tips: singular points inside the face, away from vertices
      if (f.kind != ICAM_NURBS) {
        for (auto& sp : singular_points(f.an, f.sense)) {
          bool near_vertex = false;
          for (auto& v : verts) near_vertex |= len(v.first - sp.first) <= tol;
          double u, v;
          f.an.uv(sp.first, u, v);
          if (!near_vertex && f.trim.inside(u, v)) f.tips.push_back(sp);
        }
      }
    }
    // solid3.rs:1609-1625 the C0 lines and the poles
    for (size_t fi = 0; fi < faces.size(); ++fi) {
      const Face& f = faces[fi];
      if (f.kind != ICAM_NURBS) continue;
      for (C0Line& l : c0_lines(*f.ns, int(fi))) lines.push_back(l);
      for (const Collapsed& c : collapsed_sides(*f.ns)) {
        bool near_vertex = false;
        for (auto& v : verts) near_vertex |= len(v.first - c.x) <= tol;
        if (near_vertex) continue;
        const double o0 = c.k == 0 ? f.ns->sv0.front() : f.ns->su0.front(), o1 = c.k == 0 ? f.ns->sv0.back() : f.ns->su0.back();
        const double mid = 0.5 * (o0 + o1);
        if (!(c.k == 0 ? f.trim.inside(c.value, mid) : f.trim.inside(mid, c.value))) continue;
        poles.push_back({int(fi), c.x, fan_normal(*f.ns, f.trim, c.k, c.value) * f.sense});
      }
    }
    std::vector<Box> boxes;
    for (auto& f : faces) boxes.push_back(f.box);
    for (auto& e : edges) boxes.push_back(e.box);
    for (auto& v : verts) {
      Box b;
      b.add(v.first);
      boxes.push_back(b);
    }
    // solid3.rs:1642-1650 a line's box holds its pieces' control points
    for (auto& l : lines) {
      Box b;
      for (auto& pc : l.pieces) b.add(pc.z.box());
      boxes.push_back(b);
    }
    for (auto& pl : poles) {
      Box b;
      b.add(pl.x);
      boxes.push_back(b);
    }
    tree.build(boxes);
  }

  // solid3.rs:1891-1897 normal_at: a face's outward normal at a point on it
  V3 normal_at(int fi, V3 x) const {
    const Face& f = faces[fi];
    if (f.kind == ICAM_NURBS) {
      SurfHit h;
      if (f.ns->closest(x, 4.0 * tol + 1e-9, h) || f.ns->closest(x, INF, h)) return h.n * f.sense;
      return V3(0, 0, 1) * f.sense;
    }
    return f.an.normal(x) * f.sense;
  }

  // solid3.rs:1902-1940 face: its nearest critical point inside it below cutoff
  bool face(int fi, V3 p, double cutoff, double& d, V3& q, V3& n, bool& capped) const {
    const Face& f = faces[fi];
    if (f.kind == ICAM_NURBS) {
      SurfHit h;
      bool found;
      int pops;
      bool complete = f.ns->closest_trimmed(p, cutoff, f.trim.pieces.empty() ? nullptr : &f.trim, h, found, pops);
      capped |= !complete;
      if (!found) return false;
      if (!(h.dist < cutoff && len(h.n) > 0.5)) return false;
      d = h.dist, q = h.q, n = h.n * f.sense;
      return true;
    }
    q = f.an.foot(p);
    d = len(p - q);
    if (!(d < cutoff)) return false;
    for (auto& tp : f.tips)
      if (len(tp.first - q) <= 1e-12 * (1.0 + len(tp.first))) {
        d = len(p - tp.first), q = tp.first, n = tp.second;
        return true;
      }
    if (is_singular(f.an, q)) return false;
    double u, v;
    f.an.uv(q, u, v);
    if (!f.trim.inside(u, v)) return false;
    n = f.an.normal(q) * f.sense;
    return true;
  }

  // solid3.rs:1818-1886 c0_line: the nearest point of a C0 line inside its face below
  // cutoff where the distance along the line is critical (where the line meets the
  // face's boundary the edges answer), best first over halves of its pieces, and the
  // two sides' outward normals summed there
  bool c0_line(const C0Line& l, V3 p, double cutoff, double& dist, V3& q, V3& n, bool& capped) const {
    const Face& f = faces[l.face];
    const Trim* t = f.trim.pieces.empty() ? nullptr : &f.trim;
    auto at = [&](double w, double uv[2]) { uv[0] = l.k == 0 ? l.value : w, uv[1] = l.k == 0 ? w : l.value; };
    struct Item {
      double lb;
      int i;
      Bez3 z;
      int depth;
    };
    std::vector<Item> heap;
    for (size_t i = 0; i < l.pieces.size(); ++i) {
      const double lb = box_dist(p, l.pieces[i].z.box());
      if (lb < cutoff) heap.push_back({lb, int(i), l.pieces[i].z, 0});
    }
    double bound = cutoff, bw = 0;
    int bi = -1, pops = 0;
    auto big = [](V3 x) { return std::max(std::fabs(x.x), std::max(std::fabs(x.y), std::fabs(x.z))); };
    while (!heap.empty()) {
      size_t k = 0;
      for (size_t j = 1; j < heap.size(); ++j)
        if (heap[j].lb < heap[k].lb) k = j;
      const Item it = heap[k];
      heap[k] = heap.back();
      heap.pop_back();
      if (it.lb >= bound - DIST_EPS) break;
      if (++pops > 4096) {
        capped = true;
        break;
      }
      double a[2], b[2];
      at(it.z.t0, a), at(it.z.t1, b);
      const double lo[2] = {std::min(a[0], b[0]), std::min(a[1], b[1])}, hi[2] = {std::max(a[0], b[0]), std::max(a[1], b[1])};
      const int state = t ? t->box_class(lo, hi) : 1;
      if (state == 0) continue;
      const double s = it.z.newton(p, 0.5);
      const Bez3::E e = it.z.eval(s);
      const double d = len(e.p - p), w = it.z.t0 + s * (it.z.t1 - it.z.t0);
      double uv[2];
      at(w, uv);
      if (d < bound && (state == 1 || t->inside(uv[0], uv[1]))) {
        bound = d;
        const V3 r = e.p - p;
        const double round = 1024.0 * std::numeric_limits<double>::epsilon() * (big(e.p) + big(p) + d);
        if (std::fabs(dot(r, e.d1)) <= (round + 1e-12 * len(e.d1)) * len(e.d1)) dist = d, q = e.p, bi = it.i, bw = w;
      }
      if (it.depth >= 40 || curve_no_critical_point(it.z, p)) continue;
      Bez3 c[2];
      it.z.split(0.5, c[0], c[1]);
      for (const Bez3& ch : c) {
        const double lbc = box_dist(p, ch.box());
        if (lbc < bound - DIST_EPS) heap.push_back({lbc, it.i, ch, it.depth + 1});
      }
    }
    if (bi < 0) return false;
    // the patch before the line at its far side, the patch after at its near side
    const C0Piece& pc = l.pieces[bi];
    auto side_normal = [&](const Patch& pa, double side) {
      const double s = l.k == 0 ? side : (bw - pa.u0) / (pa.u1 - pa.u0), tt = l.k == 0 ? (bw - pa.v0) / (pa.v1 - pa.v0) : side;
      const SEval e = pa.eval(s, tt);
      return unit(cross(e.ps, e.pt));
    };
    n = (side_normal(f.ns->spans[pc.before], 1.0) + side_normal(f.ns->spans[pc.after], 0.0)) * f.sense;
    return true;
  }

  // solid3.rs:1750-1816 sdf_within
  Sd sdf_within(V3 p, double bound) const {
    int nf = int(faces.size()), ne = int(edges.size());
    Sd best{bound, p, 2u << 28, false, false};
    bool found = false, cusp = false, capped = false;
    auto sgn = [](double x) { return x < 0 ? -1.0 : 1.0; };
    tree.nearest_first(p, bound, [&](int id) {
      double cur = std::fabs(best.d), before = best.d;
      if (id < nf) {
        double d;
        V3 q, n;
        if (face(id, p, cur, d, q, n, capped)) best = {d * sgn(dot(p - q, n)), q, uint32_t(id), false, false};
      } else if (id < nf + ne) {
        int ei = id - nf;
        const Edge& e = edges[ei];
        Foot c;
        if (e.closest_below(p, cur, c) && c.near == INTERIOR) {
          V3 pn = normal_at(e.faces[0], c.q) + normal_at(e.faces[1], c.q);
          best = {c.dist * sgn(dot(p - c.q, pn)), c.q, (1u << 28) | uint32_t(ei), false, false};
          cusp = len(pn) < 1e-9;
        }
      } else if (id < nf + ne + int(verts.size())) {
        int vi = id - nf - ne;
        count(C_VERT);
        double d = len(p - verts[vi].first);
        if (d < cur) best = {d * sgn(dot(p - verts[vi].first, verts[vi].second)), verts[vi].first, (2u << 28) | uint32_t(vi), false, false};
      } else if (id < nf + ne + int(verts.size() + lines.size())) {
        // solid3.rs:1790-1794 a C0 line: its face's answer
        const C0Line& l = lines[id - nf - ne - int(verts.size())];
        double d;
        V3 q, n;
        if (c0_line(l, p, cur, d, q, n, capped)) best = {d * sgn(dot(p - q, n)), q, uint32_t(l.face), false, false};
      } else {
        // solid3.rs:1794-1799 a pole: its face's answer
        const Pole& pl = poles[id - nf - ne - int(verts.size() + lines.size())];
        count(C_VERT);
        double d = len(p - pl.x);
        if (d < cur) best = {d * sgn(dot(p - pl.x, pl.n)), pl.x, uint32_t(pl.face), false, false};
      }
      if (best.d != before) {
        found = true;
        if ((best.feat >> 28) != 1) cusp = false;
      }
      return std::fabs(best.d);
    });
    if (!found && std::isfinite(bound)) return sdf_within(p, INF);
    best.cusp = cusp;
    best.unsure = capped || cusp;
    return best;
  }
};

// ------------------------------------------------------------------ bricks

// partcache.rs:962-970 packed: the class (0 words, 1 AIR, 2 SOLID)
int packed_class(const int16_t q[729]) {
  bool all_pos = true, all_nonpos = true;
  for (int n = 0; n < 729; ++n) all_pos &= q[n] > 0, all_nonpos &= q[n] <= 0;
  return all_pos ? 1 : all_nonpos ? 2 : 0;
}
void pack_words(const int16_t q[729], uint32_t* w) {
  for (int k = 0; k < ICAM_WORDS; ++k)
    w[k] = uint32_t(uint16_t(q[2 * k])) | (2 * k + 1 < 729 ? uint32_t(uint16_t(q[2 * k + 1])) << 16 : 0u);
}
int16_t quantize(double f, double clamp) {
  double x = std::min(std::max(f / clamp, -1.0), 1.0) * 32767.0;
  return int16_t(std::round(x));  // f64::round: half away from zero
}

// partcache.rs:979-1058 exact_brick: nodes read coarse to fine, each within the
// bound the brick's centre gives, and every node within |d| - clamp of a reading
// past the clamp settled as the clamp with its sign
int exact_brick(const Exact& ex, V3 p, double h, int16_t q[729]) {
  double clamp = h * std::sqrt(3.0), half = 4.0 * h;
  V3 c = p + V3(half, half, half);
  double d0 = std::fabs(ex.sdf_within(c, INF).d);
  auto at = [&](int n) { return p + V3(h * (n % 9), h * ((n / 9) % 9), h * (n / 81)); };
  double v[729];
  bool have[729] = {false};
  for (int step : {8, 4, 2, 1})
    for (int k = 0; k < 9; k += step)
      for (int j = 0; j < 9; j += step)
        for (int i = 0; i < 9; i += step) {
          int n = i + 9 * j + 81 * k;
          if (have[n]) continue;
          V3 x = at(n);
          double f = ex.sdf_within(x, d0 + len(x - c) + 1e-9).d;
          v[n] = f, have[n] = true;
          double r = std::fabs(f) - clamp;
          if (r > 0) {
            long cells = long(std::floor(r / h));
            for (long kk = std::max(k - cells, 0L); kk <= std::min(k + cells, 8L); ++kk)
              for (long jj = std::max(j - cells, 0L); jj <= std::min(j + cells, 8L); ++jj)
                for (long ii = std::max(i - cells, 0L); ii <= std::min(i + cells, 8L); ++ii) {
                  int m = int(ii + 9 * jj + 81 * kk);
                  if (!have[m] && len(at(m) - x) < r) v[m] = std::copysign(clamp, f), have[m] = true;
                }
          }
        }
  for (int n = 0; n < 729; ++n) q[n] = quantize(have[n] ? v[n] : clamp, clamp);
  return packed_class(q);
}

template <class F>
void parallel_for(long n, int threads, F&& f) {
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  if (threads == 1 || n < 2) {
    for (long i = 0; i < n; ++i) f(i);
    flush_counts();
    return;
  }
  std::atomic<long> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&]() {
      for (;;) {
        long i = next.fetch_add(1);
        if (i >= n) break;
        f(i);
      }
      flush_counts();
    });
  for (auto& t : pool) t.join();
}

// partcache.rs:226-331 decide: coarse-to-fine boxes of bricks, each read at its
// centre (the CPU path of box_bounds, :359-372), AIR when d - r > clamp, SOLID when
// d + r < -clamp, split while undecided; boxes of one brick left are read
void decide(const Exact& ex, const icam_level* lv, std::vector<uint32_t>& map, std::vector<std::pair<long, double>>& read, int threads) {
  double h = lv->h, side = 8.0 * h, clamp = h * std::sqrt(3.0);
  V3 lo = V3::of(lv->lo);
  long nb[3] = {lv->nb[0], lv->nb[1], lv->nb[2]};
  auto first = [&](long a[3]) { return lo + V3(side * a[0], side * a[1], side * a[2]); };
  map.assign(nb[0] * nb[1] * nb[2], ICAM_AIR);
  long size = 1;
  while (size < std::max(nb[0], std::max(nb[1], nb[2]))) size *= 2;
  struct Bx {
    long a[3];
  };
  std::vector<Bx> boxes;
  for (long z = 0; z < nb[2]; z += size)
    for (long y = 0; y < nb[1]; y += size)
      for (long x = 0; x < nb[0]; x += size) boxes.push_back({{x, y, z}});
  while (!boxes.empty()) {
    std::vector<double> dc(boxes.size());
    std::vector<double> rr(boxes.size());
    std::vector<char> unsure(boxes.size());
    parallel_for(long(boxes.size()), threads, [&](long i) {
      long b[3] = {boxes[i].a[0], boxes[i].a[1], boxes[i].a[2]};
      long e[3] = {b[0] + size, b[1] + size, b[2] + size};
      V3 p = first(b) - V3(clamp, clamp, clamp), q = first(e) + V3(clamp, clamp, clamp);
      V3 c = (p + q) * 0.5;
      rr[i] = 0.5 * len(q - p);
      Sd s = ex.sdf_within(c, INF);
      dc[i] = s.d, unsure[i] = s.unsure;
    });
    std::vector<Bx> next;
    for (size_t i = 0; i < boxes.size(); ++i) {
      double lo_f = unsure[i] ? -INF : dc[i] - rr[i], hi_f = unsure[i] ? INF : dc[i] + rr[i];
      long* a = boxes[i].a;
      if (lo_f > clamp || hi_f < -clamp) {
        uint32_t v = lo_f > clamp ? ICAM_AIR : ICAM_SOLID;
        for (long z = std::max(a[2], 0L); z < std::min(a[2] + size, nb[2]); ++z)
          for (long y = std::max(a[1], 0L); y < std::min(a[1] + size, nb[1]); ++y)
            for (long x = std::max(a[0], 0L); x < std::min(a[0] + size, nb[0]); ++x) map[x + nb[0] * (y + nb[1] * z)] = v;
      } else if (size == 1) {
        read.push_back({a[0] + nb[0] * (a[1] + nb[1] * a[2]), unsure[i] ? std::nan("") : dc[i]});
      } else {
        long hf = size / 2;
        for (int qd = 0; qd < 8; ++qd) {
          long k[3] = {a[0] + hf * (qd & 1), a[1] + hf * ((qd >> 1) & 1), a[2] + hf * ((qd >> 2) & 1)};
          if (k[0] < nb[0] && k[1] < nb[1] && k[2] < nb[2]) next.push_back({{k[0], k[1], k[2]}});
        }
      }
    }
    boxes = next;
    size /= 2;
  }
  std::sort(read.begin(), read.end());
}

}  // namespace orc

using namespace orc;

struct orc_exact {
  Exact ex;
};

extern "C" {

orc_exact* orc_open(const icam_part_desc* part) {
  for (uint32_t f = 0; f < part->n_faces; ++f)
    if (part->faces[f].kind == ICAM_NURBS && (part->faces[f].du >= uint32_t(MAXP) || part->faces[f].dv >= uint32_t(MAXP))) return nullptr;
  for (uint32_t e = 0; e < part->n_edges; ++e)
    if (part->edges[e].kind == ICAM_SPLINE && part->edges[e].degree >= uint32_t(MAXP)) return nullptr;
  orc_exact* o = new orc_exact();
  o->ex.open(part);
  return o;
}
void orc_close(orc_exact* o) { delete o; }

void orc_sdf(orc_exact* o, const double* xyz, size_t n, const double* bound, double* d, uint32_t* feat, uint32_t* flags, int threads) {
  parallel_for(long(n), threads, [&](long i) {
    Sd s = o->ex.sdf_within(V3::of(xyz + 3 * i), bound ? bound[i] : INF);
    d[i] = s.d;
    if (feat) feat[i] = s.feat;
    if (flags) flags[i] = (s.unsure ? 1u : 0u) | (s.cusp ? 2u : 0u);
  });
}

long orc_decide(orc_exact* o, const icam_level* lv, int64_t* out, size_t cap, int threads) {
  std::vector<uint32_t> map;
  std::vector<std::pair<long, double>> read;
  decide(o->ex, lv, map, read, threads);
  for (size_t i = 0; i < read.size() && i < cap; ++i) out[i] = read[i].first;
  return long(read.size());
}

int orc_brick(orc_exact* o, const double first_node[3], double h, int16_t* q) { return exact_brick(o->ex, V3::of(first_node), h, q); }

long orc_bake(orc_exact* o, const icam_level* lv, uint32_t* map_out, uint32_t* atlas, size_t atlas_cap_words, long* bricks_read, int threads) {
  std::vector<uint32_t> map;
  std::vector<std::pair<long, double>> read;
  decide(o->ex, lv, map, read, threads);
  if (bricks_read) *bricks_read = long(read.size());
  std::vector<int16_t> q(read.size() * 729);
  std::vector<int> cls(read.size());
  double side = 8.0 * lv->h;
  long nb0 = lv->nb[0], nb1 = lv->nb[1];
  parallel_for(long(read.size()), threads, [&](long i) {
    long b = read[i].first;
    long x = b % nb0, y = (b / nb0) % nb1, z = b / (nb0 * nb1);
    V3 p = V3::of(lv->lo) + V3(side * x, side * y, side * z);
    cls[i] = exact_brick(o->ex, p, lv->h, &q[729 * i]);
  });
  long slots = 0;
  bool fits = true;
  for (size_t i = 0; i < read.size(); ++i) {
    long b = read[i].first;
    if (cls[i] == 1) {
      map[b] = ICAM_AIR;
    } else if (cls[i] == 2) {
      map[b] = ICAM_SOLID;
    } else {
      if (size_t(slots + 1) * ICAM_WORDS <= atlas_cap_words)
        pack_words(&q[729 * i], atlas + slots * ICAM_WORDS);
      else
        fits = false;
      map[b] = uint32_t(slots++);
    }
  }
  std::copy(map.begin(), map.end(), map_out);
  return fits ? slots : -1;
}

int orc_counts(uint64_t* out, int cap, int reset) {
  for (int k = 0; k < NCOUNT; ++k) {
    const uint64_t v = reset ? g_count[k].exchange(0) : g_count[k].load();
    if (k < cap) out[k] = v;
  }
  return NCOUNT;
}

int orc_face_inside(orc_exact* o, uint32_t face, double u, double v) { return o->ex.faces[face].trim.inside(u, v) ? 1 : 0; }
void orc_face_uv(orc_exact* o, uint32_t face, const double q[3], double uv[2]) { o->ex.faces[face].an.uv(V3::of(q), uv[0], uv[1]); }
void orc_face_foot(orc_exact* o, uint32_t face, const double p[3], double q[3]) {
  V3 r = o->ex.faces[face].an.foot(V3::of(p));
  q[0] = r.x, q[1] = r.y, q[2] = r.z;
}
int orc_nurbs_closest(orc_exact* o, uint32_t face, const double p[3], double cutoff, double* dist, double q[3]) {
  SurfHit h;
  if (!o->ex.faces[face].ns || !o->ex.faces[face].ns->closest(V3::of(p), cutoff, h)) return 0;
  *dist = h.dist;
  q[0] = h.q.x, q[1] = h.q.y, q[2] = h.q.z;
  return 1;
}
void orc_nurbs_eval(orc_exact* o, uint32_t face, double u, double v, double q[3]) {
  SEval e = o->ex.faces[face].ns->eval(u, v);
  q[0] = e.p.x, q[1] = e.p.y, q[2] = e.p.z;
}
int orc_edge_closest(orc_exact* o, uint32_t edge, const double p[3], double* dist, double q[3]) {
  Foot f;
  o->ex.edges[edge].closest_below(V3::of(p), INF, f);
  *dist = f.dist;
  q[0] = f.q.x, q[1] = f.q.y, q[2] = f.q.z;
  return f.near == INTERIOR ? 1 : 0;
}

}  // extern "C"
