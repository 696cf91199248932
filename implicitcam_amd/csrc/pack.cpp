// Host packer: the part as record tables for the kernels (layout.hpp).
//
// What it builds, each after the reference's pack (crates/icam-core/src/solid3/gpu.rs):
//  - analytic faces with their exact rational forms' parameter constants
//    (trim2.rs:1807-1832) and their trim cell trees (trim2.rs:1262-1347);
//  - B-spline faces cut into knot-span Bezier patches (surface3.rs:454-467),
//    raised to bicubic (gpu.rs:149-171), each with a tree of sub-patches split
//    breadth first while its radius of convexity is short of its size and the
//    need (gpu.rs:110-112, 292-350), radii from interval bounds on the
//    derivatives (surface3.rs:819-926);
//  - spline edges cut into cubic pieces, each with a binary tree of parts under
//    the same rule (gpu.rs:501-577), radii |C'|^2 / |C''| (surface3.rs:931-970);
//  - line and arc edges, vertices with their pseudo-normals;
//  - the flat primitive list with float32 boxes rounded outward.
#include "pack.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstring>
#include <limits>

namespace icam {
namespace {

const double INF = std::numeric_limits<double>::infinity();
const double PI = 3.14159265358979323846;
const double TAU = 2 * PI;

struct V {
  double x = 0, y = 0, z = 0;
};
inline V mk(double a, double b, double c) { return V{a, b, c}; }
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len(V a) { return std::sqrt(dot(a, a)); }
inline V of(const double* d) { return {d[0], d[1], d[2]}; }

struct Box {
  V lo{INF, INF, INF}, hi{-INF, -INF, -INF};
  void add(V p) {
    lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
    hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
  }
  void add(const Box& b) { add(b.lo), add(b.hi); }
  double diag() const { return len(hi - lo); }
};

// ------------------------------------------------------------------ trim cell trees
// trim2.rs:104-136, 1150-1190: Bernstein evaluation and the half-open crossing.
double beval3(const double* c, double t) {
  if (c[0] == c[1] && c[1] == c[2] && c[2] == c[3]) return c[0];
  double b[4] = {c[0], c[1], c[2], c[3]};
  for (int l = 1; l <= 3; ++l)
    for (int i = 0; i <= 3 - l; ++i) b[i] = (1 - t) * b[i] + t * b[i + 1];
  return b[0];
}
double beval2(const double* c, double t) {
  if (c[0] == c[1] && c[1] == c[2]) return c[0];
  double b[3] = {c[0], c[1], c[2]};
  for (int l = 1; l <= 2; ++l)
    for (int i = 0; i <= 2 - l; ++i) b[i] = (1 - t) * b[i] + t * b[i + 1];
  return b[0];
}
bool cross_at(const double a[4], double x, double& out) {
  double p0 = a[0], p1 = a[3];
  if (!(std::min(p0, p1) <= x && x < std::max(p0, p1))) return false;
  double lo = 0, hi = 1, flo = beval3(a, 0) - x;
  if (flo == 0) return out = 0, true;
  double t = std::min(std::max((x - p0) / (p1 - p0), 0.0), 1.0);
  double d3[3] = {a[1] - a[0], a[2] - a[1], a[3] - a[2]};
  for (int it = 0; it < 60; ++it) {
    double ft = beval3(a, t) - x;
    if (ft == 0) return out = t, true;
    if ((ft < 0) == (flo < 0))
      lo = t, flo = ft;
    else
      hi = t;
    double d = 3 * beval2(d3, t);
    double nt = d != 0 ? t - ft / d : 0.5 * (lo + hi);
    if (!(nt > lo && nt < hi)) nt = 0.5 * (lo + hi);
    if (std::fabs(nt - t) <= 1e-16) return out = nt, true;
    t = nt;
  }
  return out = t, true;
}

struct TrimBuild {
  double lo[2], hi[2], per[2];
  std::vector<std::array<double, 8>> inst;  // shifted copies
  std::vector<uint32_t> inst_piece;
  std::vector<std::array<double, 2>> inst_shift;
  struct Cell {
    double lo[2], hi[2], r[2];
    uint32_t inside, kind, a, b;
  };
  std::vector<Cell> cells;
  std::vector<uint32_t> among;

  // trim2.rs:1205-1260 parity of the two-leg path with the symbolic perturbation
  bool parity(const double a[2], const double b[2], const std::vector<uint32_t>& ids) const {
    unsigned n = 0;
    bool up = a[1] < b[1], right = a[0] < b[0];
    for (uint32_t i : ids) {
      const auto& c = inst[i];
      double cu[4] = {c[0], c[2], c[4], c[6]}, cv[4] = {c[1], c[3], c[5], c[7]}, t;
      if (a[1] != b[1] && cross_at(cu, a[0], t)) {
        double y = beval3(cv, t);
        if (up ? (y > a[1] && y <= b[1]) : (y > b[1] && y <= a[1])) ++n;
      }
      if (a[0] != b[0] && cross_at(cv, b[1], t)) {
        double x = beval3(cu, t);
        auto slope_pos = [&]() {
          double du3[3] = {cu[1] - cu[0], cu[2] - cu[1], cu[3] - cu[2]}, dv3[3] = {cv[1] - cv[0], cv[2] - cv[1], cv[3] - cv[2]};
          double du = beval2(du3, t), dv = beval2(dv3, t);
          if (dv == 0) dv = cv[3] - cv[0];
          return du * dv > 0;
        };
        bool counted = x == b[0] ? (!slope_pos()) == right : x == a[0] ? slope_pos() == right : right ? (x > a[0] && x < b[0]) : (x > b[0] && x < a[0]);
        if (counted) ++n;
      }
    }
    return n % 2 == 1;
  }

  static void bbox(const std::array<double, 8>& c, double blo[2], double bhi[2]) {
    blo[0] = blo[1] = INF, bhi[0] = bhi[1] = -INF;
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 2; ++k) blo[k] = std::min(blo[k], c[2 * i + k]), bhi[k] = std::max(bhi[k], c[2 * i + k]);
  }

  // trim2.rs:1262-1347 place and build_cell, with the root answered from the given
  // reference point instead of an offset reference
  void build(const icam_face_desc& f) {
    for (int k = 0; k < 2; ++k) lo[k] = f.lo[k], hi[k] = f.hi[k], per[k] = f.period[k];
    double rlo[2] = {lo[0], lo[1]}, rhi[2] = {hi[0], hi[1]};
    for (int k = 0; k < 2; ++k)
      if (per[k] > 0) rlo[k] -= per[k], rhi[k] += per[k];
    for (uint32_t p = 0; p < f.n_pieces; ++p) {
      std::array<double, 8> c;
      std::copy(f.pieces + 8 * p, f.pieces + 8 * p + 8, c.begin());
      double plo[2], phi[2];
      bbox(c, plo, phi);
      long r0[2], r1[2];
      for (int k = 0; k < 2; ++k) {
        r0[k] = per[k] > 0 ? long(std::ceil((rlo[k] - phi[k]) / per[k])) : 0;
        r1[k] = per[k] > 0 ? long(std::floor((rhi[k] - plo[k]) / per[k])) : 0;
      }
      for (long a = r0[0]; a <= r1[0]; ++a)
        for (long b = r0[1]; b <= r1[1]; ++b) {
          auto q = c;
          for (int i = 0; i < 4; ++i) q[2 * i] += a * per[0], q[2 * i + 1] += b * per[1];
          inst.push_back(q);
          inst_piece.push_back(p);
          inst_shift.push_back({a * per[0], b * per[1]});
        }
    }
    const double REF[2] = {0.5137, 0.4871};  // trim2.rs:101
    std::vector<uint32_t> all(inst.size());
    for (size_t i = 0; i < all.size(); ++i) all[i] = uint32_t(i);
    Cell root;
    for (int k = 0; k < 2; ++k) root.lo[k] = lo[k], root.hi[k] = hi[k], root.r[k] = lo[k] + (hi[k] - lo[k]) * REF[k];
    double ref[2] = {f.ref[0], f.ref[1]};
    root.inside = (f.n_pieces == 0) ? 1u : uint32_t((f.ref_inside != 0) ^ parity(ref, root.r, all));
    root.kind = 0, root.a = root.b = 0;
    cells.push_back(root);
    if (f.n_pieces) build_cell(0, all, 0, REF);
  }

  void build_cell(uint32_t id, const std::vector<uint32_t>& cand, int depth, const double REF[2]) {
    Cell c = cells[id];
    std::vector<uint32_t> touching;
    for (uint32_t i : cand) {
      double blo[2], bhi[2];
      bbox(inst[i], blo, bhi);
      if (c.lo[0] <= bhi[0] && blo[0] <= c.hi[0] && c.lo[1] <= bhi[1] && blo[1] <= c.hi[1]) touching.push_back(i);
    }
    if (touching.empty()) return;
    if (depth >= 10 || touching.size() <= 4) {
      cells[id].kind = 1, cells[id].a = uint32_t(among.size()), cells[id].b = uint32_t(touching.size());
      among.insert(among.end(), touching.begin(), touching.end());
      return;
    }
    double m[2] = {0.5 * (c.lo[0] + c.hi[0]), 0.5 * (c.lo[1] + c.hi[1])};
    uint32_t first = uint32_t(cells.size());
    cells[id].kind = 2, cells[id].a = first;
    for (int q = 0; q < 4; ++q) {
      Cell k;
      k.lo[0] = (q & 1) ? m[0] : c.lo[0], k.hi[0] = (q & 1) ? c.hi[0] : m[0];
      k.lo[1] = (q & 2) ? m[1] : c.lo[1], k.hi[1] = (q & 2) ? c.hi[1] : m[1];
      for (int a = 0; a < 2; ++a) k.r[a] = k.lo[a] + (k.hi[a] - k.lo[a]) * REF[a];
      k.inside = uint32_t((c.inside != 0) ^ parity(c.r, k.r, touching));
      k.kind = 0, k.a = k.b = 0;
      cells.push_back(k);
    }
    for (int q = 0; q < 4; ++q) build_cell(first + q, touching, depth + 1, REF);
  }

  void reduce(double& u, double& v) const {
    double q[2] = {u, v};
    for (int k = 0; k < 2; ++k)
      q[k] = per[k] > 0 ? lo[k] + (q[k] - lo[k] - per[k] * std::floor((q[k] - lo[k]) / per[k])) : std::min(std::max(q[k], lo[k]), hi[k]);
    u = q[0], v = q[1];
  }
  bool inside(double u, double v) const {
    if (inst.empty() && cells.size() == 1 && cells[0].kind == 0) return cells[0].inside != 0;
    reduce(u, v);
    uint32_t id = 0;
    for (;;) {
      const Cell& c = cells[id];
      if (c.kind == 0) return c.inside != 0;
      if (c.kind == 1) {
        std::vector<uint32_t> ids(among.begin() + c.a, among.begin() + c.a + c.b);
        double q[2] = {u, v};
        return (c.inside != 0) ^ parity(c.r, q, ids);
      }
      double m[2] = {0.5 * (c.lo[0] + c.hi[0]), 0.5 * (c.lo[1] + c.hi[1])};
      id = c.a + uint32_t(u >= m[0]) + 2 * uint32_t(v >= m[1]);
    }
  }
  // trim2.rs:1115-1132 meets_box
  bool meets_box(const std::array<double, 8>& c, const double blo[2], const double bhi[2]) const {
    double b0[2], b1[2];
    bbox(c, b0, b1);
    for (int k = 0; k < 2; ++k)
      if (!(b0[k] <= bhi[k] && blo[k] <= b1[k])) return false;
    double cu[4] = {c[0], c[2], c[4], c[6]}, cv[4] = {c[1], c[3], c[5], c[7]};
    auto param_at = [&](double x) {
      if (x == cu[0]) return 0.0;
      if (x == cu[3]) return 1.0;
      double t;
      if (cross_at(cu, x, t)) return t;
      return std::fabs(x - cu[0]) < std::fabs(x - cu[3]) ? 0.0 : 1.0;
    };
    double ta = 0, tb = 1;
    if (cu[0] != cu[3]) {
      double a = std::max(blo[0], std::min(cu[0], cu[3])), b = std::min(bhi[0], std::max(cu[0], cu[3]));
      ta = param_at(a), tb = param_at(b);
    }
    double va = beval3(cv, ta), vb = beval3(cv, tb);
    return std::min(va, vb) <= bhi[1] && std::max(va, vb) >= blo[1];
  }
  // trim2.rs:1570-1610 box_class, answered from the loops directly: 1 in, 0 out, -1 straddles
  int box_class(const double blo[2], const double bhi[2]) const {
    if (inst.empty()) return cells[0].inside ? 1 : 0;
    double g = 1e-12 * (std::fabs(hi[0] - lo[0]) + std::fabs(hi[1] - lo[1]));
    double glo[2] = {blo[0] - g, blo[1] - g}, ghi[2] = {bhi[0] + g, bhi[1] + g};
    for (auto& c : inst)
      if (meets_box(c, glo, ghi)) return -1;
    return inside(0.5 * (blo[0] + bhi[0]), 0.5 * (blo[1] + bhi[1])) ? 1 : 0;
  }
};

// ------------------------------------------------------------------ Bezier helpers
using H = std::array<double, 4>;

// Bezier decomposition of a clamped B-spline by knot insertion (The NURBS Book
// algorithm 5.1), flat points of dimension N; returns pieces and their knot ranges.
void decompose(int p, std::vector<double> kn, std::vector<double> pts, int N, std::vector<std::vector<double>>& pieces,
               std::vector<std::pair<double, double>>& ranges) {
  auto count = [&](double u) {
    int c = 0;
    for (double k : kn) c += (k == u);
    return c;
  };
  std::vector<double> interior;
  for (size_t i = p + 1; i + p + 1 < kn.size(); ++i)
    if (kn[i] > kn[p] && kn[i] < kn[kn.size() - 1 - p] && (interior.empty() || interior.back() != kn[i])) interior.push_back(kn[i]);
  for (double u : interior)
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
  std::vector<double> brk;
  for (size_t i = p; i + p < kn.size(); ++i)
    if (brk.empty() || brk.back() != kn[i]) brk.push_back(kn[i]);
  int np = int(pts.size()) / N, sp = 0;
  for (int s = 0; s + p < np; s += p, ++sp) {
    pieces.emplace_back(pts.begin() + s * N, pts.begin() + (s + p + 1) * N);
    ranges.push_back({brk[sp], brk[sp + 1]});
  }
}

// gpu.rs:128-140 a Bezier curve raised one degree
std::vector<H> raise(const std::vector<H>& p) {
  size_t n = p.size();
  std::vector<H> r(n + 1);
  for (size_t i = 0; i <= n; ++i) {
    double a = double(i) / n;
    for (int k = 0; k < 4; ++k) r[i][k] = a * (i > 0 ? p[i - 1][k] : 0.0) + (1 - a) * (i < n ? p[i][k] : 0.0);
  }
  return r;
}

struct Patch4 {  // bicubic, homogeneous, i along u
  H p[4][4];
};
struct Piece4 {
  H p[4];
};

void split_curve(const H* in, double s, H* l, H* r) {  // 4 points
  H b[4];
  for (int i = 0; i < 4; ++i) b[i] = in[i];
  l[0] = b[0], r[3] = b[3];
  for (int lev = 1; lev <= 3; ++lev) {
    for (int i = 0; i <= 3 - lev; ++i)
      for (int k = 0; k < 4; ++k) b[i][k] = (1 - s) * b[i][k] + s * b[i + 1][k];
    l[lev] = b[0], r[3 - lev] = b[3 - lev];
  }
}
void quarter(const Patch4& pa, Patch4 q[4]) {  // q index = (s half) + 2 (t half)
  Patch4 a, b;  // u halves
  for (int j = 0; j < 4; ++j) {
    H col[4], l[4], r[4];
    for (int i = 0; i < 4; ++i) col[i] = pa.p[i][j];
    split_curve(col, 0.5, l, r);
    for (int i = 0; i < 4; ++i) a.p[i][j] = l[i], b.p[i][j] = r[i];
  }
  const Patch4* halves[2] = {&a, &b};
  for (int hs = 0; hs < 2; ++hs)
    for (int i = 0; i < 4; ++i) {
      H l[4], r[4];
      split_curve(halves[hs]->p[i], 0.5, l, r);
      for (int j = 0; j < 4; ++j) q[hs + 0].p[i][j] = l[j], q[hs + 2].p[i][j] = r[j];
    }
}

V point(const H& h) { return {h[0] / h[3], h[1] / h[3], h[2] / h[3]}; }

// Intervals (surface3.rs:773-805)
struct Iv {
  double a, b;
};
Iv iv_of(std::initializer_list<double> xs) {
  Iv r{INF, -INF};
  for (double x : xs) r.a = std::min(r.a, x), r.b = std::max(r.b, x);
  return r;
}
Iv add(Iv x, Iv y) { return {x.a + y.a, x.b + y.b}; }
Iv sub(Iv x, Iv y) { return {x.a - y.b, x.b - y.a}; }
Iv mul(Iv x, Iv y) {
  double p[4] = {x.a * y.a, x.a * y.b, x.b * y.a, x.b * y.b};
  return {*std::min_element(p, p + 4), *std::max_element(p, p + 4)};
}
Iv scale(Iv x, double k) { return k >= 0 ? Iv{x.a * k, x.b * k} : Iv{x.b * k, x.a * k}; }
Iv divpos(Iv x, Iv w) { return mul(x, Iv{1.0 / w.b, 1.0 / w.a}); }
double mig(Iv x) { return (x.a <= 0 && x.b >= 0) ? 0.0 : std::min(std::fabs(x.a), std::fabs(x.b)); }
double mag(Iv x) { return std::max(std::fabs(x.a), std::fabs(x.b)); }

// Bernstein evaluation of a bicubic with derivatives (point and first derivatives only)
void eval_patch(const Patch4& pa, double s, double t, V& p, V& ps, V& pt) {
  auto basis = [](double x, double b[4], double d[4]) {
    double y = 1 - x;
    b[0] = y * y * y, b[1] = 3 * x * y * y, b[2] = 3 * x * x * y, b[3] = x * x * x;
    d[0] = -3 * y * y, d[1] = 3 * y * y - 6 * x * y, d[2] = 6 * x * y - 3 * x * x, d[3] = 3 * x * x;
  };
  double bu[4], du[4], bv[4], dv[4];
  basis(s, bu, du), basis(t, bv, dv);
  double a[3][4] = {{0}};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        a[0][k] += bu[i] * bv[j] * pa.p[i][j][k];
        a[1][k] += du[i] * bv[j] * pa.p[i][j][k];
        a[2][k] += bu[i] * dv[j] * pa.p[i][j][k];
      }
  double w = a[0][3];
  p = mk(a[0][0], a[0][1], a[0][2]) * (1 / w);
  ps = (mk(a[1][0], a[1][1], a[1][2]) - p * a[1][3]) * (1 / w);
  pt = (mk(a[2][0], a[2][1], a[2][2]) - p * a[2][3]) * (1 / w);
}

// The same with the mixed second derivative (the fan of normals at a pole).
void eval_patch2(const Patch4& pa, double s, double t, V& p, V& ps, V& pt, V& pst) {
  auto basis = [](double x, double b[4], double d[4]) {
    double y = 1 - x;
    b[0] = y * y * y, b[1] = 3 * x * y * y, b[2] = 3 * x * x * y, b[3] = x * x * x;
    d[0] = -3 * y * y, d[1] = 3 * y * y - 6 * x * y, d[2] = 6 * x * y - 3 * x * x, d[3] = 3 * x * x;
  };
  double bu[4], du[4], bv[4], dv[4];
  basis(s, bu, du), basis(t, bv, dv);
  double a[4][4] = {{0}};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) {
        a[0][k] += bu[i] * bv[j] * pa.p[i][j][k];
        a[1][k] += du[i] * bv[j] * pa.p[i][j][k];
        a[2][k] += bu[i] * dv[j] * pa.p[i][j][k];
        a[3][k] += du[i] * dv[j] * pa.p[i][j][k];
      }
  const double w = a[0][3];
  p = mk(a[0][0], a[0][1], a[0][2]) * (1 / w);
  ps = (mk(a[1][0], a[1][1], a[1][2]) - p * a[1][3]) * (1 / w);
  pt = (mk(a[2][0], a[2][1], a[2][2]) - p * a[2][3]) * (1 / w);
  pst = (mk(a[3][0], a[3][1], a[3][2]) - ps * a[2][3] - pt * a[1][3] - p * a[3][3]) * (1 / w);
}

// What a leaf keeps of its bounds for the per-point convexity test (gpu.wgsl:1581-1614
// reads the same quantities): lambda_min(G) from below, then the intervals of S_s, S_t,
// S_ss, S_st and S_tt (x lo, x hi, y lo, y hi, z lo, z hi each), in the parameters
// scaled by the middle's tangent lengths. Unusable bounds are left wide open.
constexpr int PATCH_CERT = 31;
// The same for a curve leaf: |C'|^2 from below, then the intervals of C' and C''.
constexpr int PIECE_CERT = 13;
void open_cert(double* out, int n) {
  out[0] = 0.0;
  for (int k = 1; k < n; k += 2) out[k] = -1e30, out[k + 1] = 1e30;
}
void put_iv(double* out, const Iv* iv) {
  for (int k = 0; k < 3; ++k) {
    // widened by their own rounding (the hulls and quotients are summed in double)
    double pad = 1e-13 * (std::fabs(iv[k].a) + std::fabs(iv[k].b)) + 1e-300;
    out[2 * k] = iv[k].a - pad, out[2 * k + 1] = iv[k].b + pad;
  }
}

// ---- bounds over a rational Bezier patch or piece, from exact Bernstein forms
//
// A rational bicubic S = P / w (P the homogeneous x, y, z part, w the weights; polynomials
// of bidegree (3, 3) in Bernstein form) has derivatives whose denominators are powers of w:
//   S_s  = N_s / w^2,                       N_s = P_s w - P w_s        bidegree (5, 6)
//   S_t  = N_t / w^2,                       N_t = P_t w - P w_t        (6, 5)
//   S_ss = (N_s,s w - 2 N_s w_s) / w^3                                 (7, 9)
//   S_st = (N_s,t w - 2 N_s w_t) / w^3                                 (8, 8)
//   S_tt = (N_t,t w - 2 N_t w_t) / w^3                                 (9, 7)
// and its Gram matrix G = [S_s.S_s, S_s.S_t; S_t.S_s, S_t.S_t] has entries (N_s.N_s,
// N_s.N_t, N_t.N_t) / w^4, numerators of bidegree (10, 12), (11, 11), (12, 10). Products
// and derivatives of Bernstein forms are exact operations on their coefficients, and a
// Bernstein form lies within the hull of its coefficients over the unit square, so each
// quantity lies in the quotient of its numerator's hull by its (positive) denominator's
// hull. The forms are exact in the coefficients of P - c w (c the control points' middle,
// which leaves every derivative unchanged and keeps the coefficients small).
//
// The forms are kept scaled, coefficient (i, j) times binom(m, i) binom(n, j): a product
// is then a plain convolution of the scaled coefficients, and a derivative along s has
// scaled coefficients (i + 1) a_(i+1) j - (m - i) a_ij. Each form carries one bound E on
// the rounding error of every scaled coefficient: a product's is (k + 2) u |A|_1 |B|_max
// + |A|_1 E_B + |B|_1 E_A + k E_A E_B (k terms a coefficient at most, u = 2^-53), and a
// hull widens each coefficient by E over its binomials. The least eigenvalue of G is
// bounded below from the entries' enclosures: it grows with either diagonal entry and
// shrinks as |G_12| grows, so it is at least the least eigenvalue of [a_lo, b_mag; b_mag,
// c_lo], less the rounding of that formula. A piece (a rational cubic curve) is the same
// with one parameter.
constexpr double U53 = 1.1102230246251565e-16;
constexpr int BMAX = 13;  // coefficients a side, at most (bidegree 12)
struct BForm {
  int m = 0, n = 0;        // bidegree
  double c[BMAX * BMAX];   // scaled coefficients, (i, j) at i (n + 1) + j
  double E = 0.0;          // rounding error bound of every scaled coefficient
};
const double* binom_row(int n) {
  static const struct Table {
    double t[BMAX][BMAX];
    Table() {
      for (int a = 0; a < BMAX; ++a) {
        t[a][0] = 1.0;
        for (int b = 1; b < BMAX; ++b) t[a][b] = b > a ? 0.0 : t[a - 1][b - 1] + (b <= a - 1 ? t[a - 1][b] : 0.0);
      }
    }
  } T;
  return T.t[n];
}
double l1(const BForm& A) {
  double s = 0;
  for (int q = 0; q < (A.m + 1) * (A.n + 1); ++q) s += std::fabs(A.c[q]);
  return s;
}
double linf(const BForm& A) {
  double s = 0;
  for (int q = 0; q < (A.m + 1) * (A.n + 1); ++q) s = std::max(s, std::fabs(A.c[q]));
  return s;
}
// the derivative along the first (s) or second (t) parameter
BForm bf_d(const BForm& A, bool along_s) {
  BForm R;
  R.m = along_s ? A.m - 1 : A.m, R.n = along_s ? A.n : A.n - 1;
  double big = 0;
  for (int i = 0; i <= R.m; ++i)
    for (int j = 0; j <= R.n; ++j) {
      double x, y;
      if (along_s)
        x = (i + 1) * A.c[(i + 1) * (A.n + 1) + j], y = (A.m - i) * A.c[i * (A.n + 1) + j];
      else
        x = (j + 1) * A.c[i * (A.n + 1) + j + 1], y = (A.n - j) * A.c[i * (A.n + 1) + j];
      R.c[i * (R.n + 1) + j] = x - y;
      big = std::max(big, std::fabs(x) + std::fabs(y));
    }
  R.E = double(along_s ? A.m + 1 : A.n + 1) * A.E + 3.0 * U53 * big;
  return R;
}
// the product
BForm bf_mul(const BForm& A, const BForm& B) {
  BForm R;
  R.m = A.m + B.m, R.n = A.n + B.n;
  const int N = (R.m + 1) * (R.n + 1), rn = R.n + 1;
  for (int q = 0; q < N; ++q) R.c[q] = 0.0;
  for (int i1 = 0; i1 <= A.m; ++i1)
    for (int j1 = 0; j1 <= A.n; ++j1) {
      const double a = A.c[i1 * (A.n + 1) + j1];
      double* r = R.c + i1 * rn + j1;
      const double* b = B.c;
      for (int i2 = 0; i2 <= B.m; ++i2, b += B.n + 1)
        for (int j2 = 0; j2 <= B.n; ++j2) r[i2 * rn + j2] += a * b[j2];
    }
  const double k = double((std::min(A.m, B.m) + 1) * (std::min(A.n, B.n) + 1));
  const double la = l1(A), lb = l1(B);
  R.E = ((k + 2.0) * U53 * la * linf(B) + la * B.E + lb * A.E + k * A.E * B.E) * (1.0 + 1e-10);
  return R;
}
// A - k B (the same bidegree), and A + B
BForm bf_sub(const BForm& A, const BForm& B, double k = 1.0) {
  BForm R;
  R.m = A.m, R.n = A.n;
  double big = 0;
  for (int q = 0; q < (R.m + 1) * (R.n + 1); ++q) {
    const double kb = k * B.c[q];
    R.c[q] = A.c[q] - kb;
    big = std::max(big, std::fabs(A.c[q]) + std::fabs(kb));
  }
  R.E = A.E + std::fabs(k) * B.E + 2.0 * U53 * big;
  return R;
}
BForm bf_add(const BForm& A, const BForm& B) { return bf_sub(A, B, -1.0); }
// the hull of the coefficients (unscaled), widened by their errors
Iv bf_hull(const BForm& A) {
  Iv r{INF, -INF};
  const double* bm = binom_row(A.m);
  const double* bn = binom_row(A.n);
  for (int i = 0; i <= A.m; ++i)
    for (int j = 0; j <= A.n; ++j) {
      const double s = bm[i] * bn[j], v = A.c[i * (A.n + 1) + j] / s, e = A.E / s + 2.0 * U53 * std::fabs(v);
      r.a = std::min(r.a, v - e), r.b = std::max(r.b, v + e);
    }
  return r;
}
// x / w for w a positive interval, widened by the rounding of the reciprocals and products
Iv div_hull(Iv x, Iv w) {
  Iv r = divpos(x, w);
  const double pad = 4.0 * U53 * std::max(std::fabs(r.a), std::fabs(r.b));
  return {r.a - pad, r.b + pad};
}

// The scaled forms of a patch: component k of P - c w (k < 3), or w (k = 3).
BForm patch_form(const Patch4& pw, V c, int k) {
  BForm R;
  R.m = R.n = 3;
  const double* b3 = binom_row(3);
  const double ck = k == 0 ? c.x : k == 1 ? c.y : c.z;
  double big = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const H& h = pw.p[i][j];
      const double sc = b3[i] * b3[j], cw = k == 3 ? 0.0 : ck * h[3], v = k == 3 ? h[3] : h[k] - cw;
      R.c[i * 4 + j] = v * sc;
      big = std::max(big, (std::fabs(k == 3 ? h[3] : h[k]) + std::fabs(cw)) * sc);
    }
  R.E = (k == 3 ? 1.0 : 3.0) * U53 * big;
  return R;
}

// Enclosures over a patch: lambda_min(G) from below and S_s, S_t, S_ss, S_st, S_tt, all
// in parameters scaled by the tangents' largest lengths (any scaling D of the parameters
// turns G into D G D and the Hessian of the squared distance into D H D, which keeps
// every convexity test made with them); false when a weight is not positive. G's
// entries are first enclosed from the tangents' enclosures (products of intervals); when
// the radius that gives falls short of `enough`, also from G's own numerators (exact,
// which keeps a tangent turning over the patch from driving a component's interval
// through zero), and the larger bound is kept.
struct PatchBounds {
  double lambda;
  Iv d[5][3];
};
double lambda2(double alo, double bm, double clo) {
  const double lam = 0.5 * ((alo + clo) - std::sqrt((alo - clo) * (alo - clo) + 4.0 * bm * bm));
  return lam - 16.0 * U53 * (alo + clo + 2.0 * bm);
}
double frob(const Iv d[5][3]) {
  auto mag2 = [](const Iv* v) { return mag(v[0]) * mag(v[0]) + mag(v[1]) * mag(v[1]) + mag(v[2]) * mag(v[2]); };
  return std::sqrt(mag2(d[2]) + 2 * mag2(d[3]) + mag2(d[4])) * (1 + 1e-12);
}
bool patch_bounds(const Patch4& pw, PatchBounds& out, double enough) {
  V c{0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) c = c + point(pw.p[i][j]);
  c = c * (1.0 / 16);
  const BForm w = patch_form(pw, c, 3);
  if (!(bf_hull(w).a > 0)) return false;
  const BForm ws = bf_d(w, true), wt = bf_d(w, false);
  const BForm w2 = bf_mul(w, w), w3 = bf_mul(w2, w);
  const Iv h2 = bf_hull(w2), h3 = bf_hull(w3);
  if (!(h2.a > 0 && h3.a > 0)) return false;
  BForm Ns[3], Nt[3];
  Iv S[5][3];
  for (int k = 0; k < 3; ++k) {
    const BForm q = patch_form(pw, c, k);
    Ns[k] = bf_sub(bf_mul(bf_d(q, true), w), bf_mul(q, ws));
    Nt[k] = bf_sub(bf_mul(bf_d(q, false), w), bf_mul(q, wt));
    S[0][k] = div_hull(bf_hull(Ns[k]), h2);
    S[1][k] = div_hull(bf_hull(Nt[k]), h2);
    S[2][k] = div_hull(bf_hull(bf_sub(bf_mul(bf_d(Ns[k], true), w), bf_mul(Ns[k], ws), 2.0)), h3);
    S[3][k] = div_hull(bf_hull(bf_sub(bf_mul(bf_d(Ns[k], false), w), bf_mul(Ns[k], wt), 2.0)), h3);
    S[4][k] = div_hull(bf_hull(bf_sub(bf_mul(bf_d(Nt[k], false), w), bf_mul(Nt[k], wt), 2.0)), h3);
  }
  // G's entries from the tangents' enclosures: |S_s|^2 in [sum mig^2, sum mag^2], S_s.S_t
  // in the sum of the interval products
  double alo = 0, ahi = 0, clo = 0, chi = 0;
  Iv b{0, 0};
  for (int k = 0; k < 3; ++k) {
    alo += mig(S[0][k]) * mig(S[0][k]), ahi += mag(S[0][k]) * mag(S[0][k]);
    clo += mig(S[1][k]) * mig(S[1][k]), chi += mag(S[1][k]) * mag(S[1][k]);
    b = add(b, mul(S[0][k], S[1][k]));
  }
  alo *= 1 - 8 * U53, ahi *= 1 + 8 * U53, clo *= 1 - 8 * U53, chi *= 1 + 8 * U53;
  const double bmag = mag(b) * (1 + 8 * U53);
  // the scaling: the tangents' largest lengths
  const double sa = ahi > 0 ? std::sqrt(ahi) : 1.0, sb = chi > 0 ? std::sqrt(chi) : 1.0;
  const double k_s[5] = {1 / sa, 1 / sb, 1 / (sa * sa), 1 / (sa * sb), 1 / (sb * sb)};
  for (int q = 0; q < 5; ++q)
    for (int k = 0; k < 3; ++k) out.d[q][k] = scale(S[q][k], k_s[q]);
  double lam = lambda2(alo / (sa * sa), bmag / (sa * sb), clo / (sb * sb));
  const double hh = frob(out.d);
  if (!(hh > 0 && 0.99 * lam / hh >= enough)) {
    // G's entries from their own numerators
    const BForm w4 = bf_mul(w2, w2);
    const Iv h4 = bf_hull(w4);
    if (h4.a > 0) {
      BForm ga = bf_mul(Ns[0], Ns[0]), gb = bf_mul(Ns[0], Nt[0]), gc = bf_mul(Nt[0], Nt[0]);
      for (int k = 1; k < 3; ++k) ga = bf_add(ga, bf_mul(Ns[k], Ns[k])), gb = bf_add(gb, bf_mul(Ns[k], Nt[k])), gc = bf_add(gc, bf_mul(Nt[k], Nt[k]));
      const Iv ea = div_hull(bf_hull(ga), h4), eb = div_hull(bf_hull(gb), h4), ec = div_hull(bf_hull(gc), h4);
      lam = std::max(lam, lambda2(std::max(ea.a, 0.0) / (sa * sa), mag(eb) / (sa * sb), std::max(ec.a, 0.0) / (sb * sb)));
    }
  }
  out.lambda = lam;
  return true;
}

// surface3.rs:819-830 the radius of convexity, 0.99 lambda_min(G) / sqrt(|S_ss|^2 + 2 |S_st|^2
// + |S_tt|^2) with both bounded over the patch (patch_bounds; the tighter bound on G is
// sought when the first gives less than `enough`); `cert` (when not NULL) receives lambda
// and the derivatives' enclosures (PATCH_CERT words)
double patch_radius(const Patch4& pw, double* cert = nullptr, double enough = INF) {
  if (cert) open_cert(cert, PATCH_CERT);
  PatchBounds pb;
  if (!patch_bounds(pw, pb, enough)) return 0.0;
  if (cert) {
    cert[0] = pb.lambda > 0 ? pb.lambda : 0.0;
    for (int q = 0; q < 5; ++q) put_iv(cert + 1 + 6 * q, pb.d[q]);
  }
  if (!(pb.lambda > 0)) return 0.0;
  const double hh = frob(pb.d);
  return hh > 0 ? 0.99 * pb.lambda / hh : INF;  // RHO_MARGIN, surface3.rs:807
}

// A piece's scaled forms: component k of P - c w (k < 3), or w (k = 3), one parameter.
BForm piece_form(const H* pw, V c, int k) {
  BForm R;
  R.m = 3, R.n = 0;
  const double* b3 = binom_row(3);
  const double ck = k == 0 ? c.x : k == 1 ? c.y : c.z;
  double big = 0;
  for (int i = 0; i < 4; ++i) {
    const double cw = k == 3 ? 0.0 : ck * pw[i][3], v = k == 3 ? pw[i][3] : pw[i][k] - cw;
    R.c[i] = v * b3[i];
    big = std::max(big, (std::fabs(k == 3 ? pw[i][3] : pw[i][k]) + std::fabs(cw)) * b3[i]);
  }
  R.E = (k == 3 ? 1.0 : 3.0) * U53 * big;
  return R;
}

// surface3.rs:931-970 a cubic piece's radius |C'|^2_min / |C''|_max: C' = N / w^2 with
// N = P' w - P w' (degree 5), C'' = (N' w - 2 N w') / w^3 (degree 7), |C'|^2 = N.N / w^4
// (degree 10), each enclosed as the patch's are; `cert` (when not NULL) receives |C'|^2
// from below and the enclosures of C' and C'' (PIECE_CERT words)
double piece_radius(const H* pw, double* cert = nullptr) {
  if (cert) open_cert(cert, PIECE_CERT);
  V c{0, 0, 0};
  for (int i = 0; i < 4; ++i) c = c + point(pw[i]);
  c = c * 0.25;
  const BForm w = piece_form(pw, c, 3);
  if (!(bf_hull(w).a > 0)) return 0.0;
  const BForm w1 = bf_d(w, true), w2 = bf_mul(w, w), w3 = bf_mul(w2, w), w4 = bf_mul(w2, w2);
  const Iv h2 = bf_hull(w2), h3 = bf_hull(w3), h4 = bf_hull(w4);
  if (!(h2.a > 0 && h3.a > 0 && h4.a > 0)) return 0.0;
  Iv d1[3], d2[3];
  BForm len2;
  for (int k = 0; k < 3; ++k) {
    const BForm q = piece_form(pw, c, k);
    const BForm N = bf_sub(bf_mul(bf_d(q, true), w), bf_mul(q, w1));
    d1[k] = div_hull(bf_hull(N), h2);
    d2[k] = div_hull(bf_hull(bf_sub(bf_mul(bf_d(N, true), w), bf_mul(N, w1), 2.0)), h3);
    const BForm nn = bf_mul(N, N);
    len2 = k == 0 ? nn : bf_add(len2, nn);
  }
  const double l = std::max(div_hull(bf_hull(len2), h4).a, 0.0);
  if (cert) cert[0] = l, put_iv(cert + 1, d1), put_iv(cert + 7, d2);
  if (!(l > 0)) return 0.0;
  const double h = std::sqrt(mag(d2[0]) * mag(d2[0]) + mag(d2[1]) * mag(d2[1]) + mag(d2[2]) * mag(d2[2])) * (1 + 1e-12);
  return h > 0 ? 0.99 * l / h : INF;
}

// trim2.rs:1789-1807 arc_param_over: an angle's parameter in the exact rational form of
// n equal arcs covering span from a0 (each arc's half-angle tangent, rescaled to [k, k+1])
double arc_param_over(double a, double a0, double span, int n) {
  const double q = span / n;
  double x = a - a0;
  if (std::fabs(span - TAU) < 1e-12) {
    x = x - TAU * std::floor(x / TAU);
  } else {
    const double y = x - TAU * std::floor((x - 0.5 * span) / TAU + 0.5);
    x = std::min(std::max(y, 0.0), span);
  }
  const double k = std::min(std::max(std::floor(x / q), 0.0), double(n - 1));
  const double alpha = x - q * k - 0.5 * q;
  const double t = 0.5 * (std::tan(0.5 * alpha) / std::tan(0.25 * q) + 1.0);
  return k + std::min(std::max(t, 0.0), 1.0);
}

// gpu.rs:110-112 the split rule
bool split_rule(double rho, double diag, double need, int depth) { return depth < 8 && rho < 2.0 * diag + need && diag > 0.25 * need; }

float down(double x) {
  float f = float(x);
  if (double(f) > x) f = std::nextafter(f, -INFINITY);
  return f;
}
float up(double x) {
  float f = float(x);
  if (double(f) < x) f = std::nextafter(f, INFINITY);
  return f;
}

}  // namespace

std::vector<float> to_f32(const std::vector<double>& f) {
  std::vector<float> r(f.size());
  for (size_t i = 0; i < f.size(); ++i) r[i] = float(f[i]);
  return r;
}

namespace {
// f(0) .. f(n - 1) on up to 16 host threads (inline for small n); f must only write
// what its index owns.
template <class F> void parallel_for(size_t n, F&& f) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 8 || hw == 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
  };
  std::vector<std::thread> ts;
  for (unsigned t = 1; t < hw; ++t) ts.emplace_back(work);
  work();
  for (auto& t : ts) t.join();
}
}  // namespace

namespace {
// The primitives' hierarchy (Layout::bvh_*): the primitives sorted by the Morton code of
// their box centres (10 bits an axis over the primitives' extent), then levels of nodes
// each bounding 64 of the level below, until a level has at most 64. A node holds a
// boundary point when one of its primitives does: faces, edges, vertices and edge pieces
// do; a sub-patch when it lies wholly inside its face's loops; a C0 line's piece never.
void build_hierarchy(Packed& P) {
  Layout& L = P.L;
  const size_t n = P.pcode.size();
  float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], P.pbox[6 * i + k]), hi[k] = std::max(hi[k], P.pbox[6 * i + 3 + k]);
  std::vector<std::pair<uint32_t, uint32_t>> key(n);
  for (size_t i = 0; i < n; ++i) {
    uint32_t code = 0;
    for (int k = 0; k < 3; ++k) {
      const double c = 0.5 * (double(P.pbox[6 * i + k]) + P.pbox[6 * i + 3 + k]), side = double(hi[k]) - lo[k];
      const double f = side > 0 ? (c - lo[k]) / side * 1024.0 : 0.0;
      const uint32_t q = f <= 0 ? 0u : f >= 1023.0 ? 1023u : uint32_t(f);
      for (int b = 0; b < 10; ++b) code |= ((q >> b) & 1u) << (3 * b + k);
    }
    key[i] = {code, uint32_t(i)};
  }
  std::sort(key.begin(), key.end());
  P.bvh_prim.resize(n);
  for (size_t i = 0; i < n; ++i) P.bvh_prim[i] = key[i].second;
  auto holds = [&](uint32_t code) {
    const uint32_t kind = code >> 28, id = code & 0x0fffffffu;
    if (kind == P_LEAF) return P.U[L.leaf_u + LEAF_U * id + 2] != 0u;
    if (kind == P_CLEAF) return P.U[L.edge_u + EDGE_U * P.U[L.cleaf_u + CLEAF_U * id + 1]] != EDGE_CREASE;
    return true;
  };
  P.bvh_box.clear(), P.bvh_holds.clear();
  for (size_t i = 0; i < n; ++i) {
    const uint32_t q = P.bvh_prim[i];
    P.bvh_box.insert(P.bvh_box.end(), P.pbox.begin() + 6 * q, P.pbox.begin() + 6 * q + 6);
    P.bvh_holds.push_back(holds(P.pcode[q]) ? 1 : 0);
  }
  for (int l = 0; l < 8; ++l) L.bvh_off[l] = 0, L.bvh_cnt[l] = 0;
  L.bvh_off[0] = 0, L.bvh_cnt[0] = uint32_t(n);
  int l = 0;
  while (L.bvh_cnt[l] > 64 && l < 7) {
    const uint32_t below = L.bvh_off[l], cnt = L.bvh_cnt[l], m = (cnt + 63) / 64;
    L.bvh_off[l + 1] = uint32_t(P.bvh_holds.size()), L.bvh_cnt[l + 1] = m;
    for (uint32_t j = 0; j < m; ++j) {
      float b[6] = {3e38f, 3e38f, 3e38f, -3e38f, -3e38f, -3e38f};
      uint8_t h = 0;
      for (uint32_t c = 64 * j; c < std::min(cnt, 64 * j + 64); ++c) {
        const float* x = &P.bvh_box[6 * size_t(below + c)];
        for (int k = 0; k < 3; ++k) b[k] = std::min(b[k], x[k]), b[3 + k] = std::max(b[3 + k], x[3 + k]);
        h |= P.bvh_holds[below + c];
      }
      P.bvh_box.insert(P.bvh_box.end(), b, b + 6);
      P.bvh_holds.push_back(h);
    }
    ++l;
  }
  L.bvh_top = uint32_t(l);
}
}  // namespace

bool pack_part(const icam_part_desc* part, double need, Packed& out) {
  Packed& P = out;
  Layout& L = P.L;
  V o = mk(0.5 * (part->lo[0] + part->hi[0]), 0.5 * (part->lo[1] + part->hi[1]), 0.5 * (part->lo[2] + part->hi[2]));
  P.origin[0] = o.x, P.origin[1] = o.y, P.origin[2] = o.z;
  auto rel = [&](const double* w) { return of(w) - o; };
  auto relh = [&](const double* h) { return H{h[0] - o.x * h[3], h[1] - o.y * h[3], h[2] - o.z * h[3], h[3]}; };

  // sections are built into separate vectors, then concatenated
  std::vector<double> face_f, edge_f, vert_f, leaf_f, span_f, cleaf_f, cpiece_f, cell_f, inst_f, tpiece_f, uvmap_f;
  std::vector<double> pole_f;  // B-spline poles, point records placed after the tips
  // B-spline faces' interior C0 knot lines: the face, the direction held (0: u), the
  // knot, and per piece its four points and the span before the line
  struct Crease {
    uint32_t face, k;
    double value;
    std::vector<std::vector<H>> pieces;
    std::vector<uint32_t> before;
  };
  std::vector<Crease> creases;
  std::vector<uint32_t> face_u, edge_u, leaf_u, span_u, cleaf_u, cpiece_u, cell_u, inst_u, among_u;
  std::vector<Box> pbox;
  std::vector<uint32_t> pcode;
  std::vector<Box> face_box(part->n_faces);
  std::vector<Box> edge_box(part->n_edges);

  // a curve piece's tree of parts, breadth first; its leaves are primitives
  auto curve_leaves = [&](const std::vector<H>& pts, uint32_t cp, uint32_t e) {
    struct Part_ {
      H p[4];
      double t0, t1;
      int depth;
    };
    std::vector<Part_> todo;
    Part_ root;
    for (int i = 0; i < 4; ++i) root.p[i] = pts[i];
    root.t0 = 0, root.t1 = 1, root.depth = 0;
    todo.push_back(root);
    while (!todo.empty()) {
      Part_ pp = todo.back();
      todo.pop_back();
      Box b;
      for (int i = 0; i < 4; ++i) b.add(point(pp.p[i]));
      double cert[PIECE_CERT];
      const double rho = piece_radius(pp.p, cert);
      ++P.nodes;
      if (split_rule(rho, b.diag(), need, pp.depth)) {
        Part_ l, r;
        split_curve(pp.p, 0.5, l.p, r.p);
        double tm = 0.5 * (pp.t0 + pp.t1);
        l.t0 = pp.t0, l.t1 = tm, r.t0 = tm, r.t1 = pp.t1;
        l.depth = r.depth = pp.depth + 1;
        todo.push_back(r), todo.push_back(l);
        continue;
      }
      uint32_t id = uint32_t(cleaf_f.size() / CLEAF_F);
      cleaf_f.insert(cleaf_f.end(), {rho, pp.t0, pp.t1, 0.0});
      cleaf_f.insert(cleaf_f.end(), cert, cert + PIECE_CERT);
      cleaf_u.insert(cleaf_u.end(), {cp, e});
      pbox.push_back(b);
      pcode.push_back((P_CLEAF << 28) | id);
    }
  };

  // Host features, which the device does not read and a point within reach of which is
  // left to the caller (the reference's Code::Cpu, crates/icam-core/src/solid3/gpu.rs:
  // 271, 287-289, 714-716): a B-spline face past bicubic, a spline edge past cubic, an
  // ellipse edge on a B-spline face (its pieces carry no parameter map), and every edge
  // of a host face (an edge's sign takes its faces' normals). Each is one primitive with
  // its box (a B-spline's from its control points: positive weights keep the surface in
  // their hull).
  if (part->n_faces >= HOST_EDGE || part->n_edges >= HOST_EDGE) {
    P.error = "too many faces or edges";
    return false;
  }
  auto host_face = [&](uint32_t f) {
    return f < part->n_faces && part->faces[f].kind == ICAM_NURBS && (part->faces[f].du > 3 || part->faces[f].dv > 3);
  };
  std::vector<uint8_t> host_edge(part->n_edges, 0);

  // ---- edges first (face boxes use them); their records and pieces
  for (uint32_t e = 0; e < part->n_edges; ++e) {
    const icam_edge_desc& ed = part->edges[e];
    const bool host = (ed.kind == ICAM_SPLINE && ed.degree > 3) || (ed.kind == ICAM_ELLIPSE && ed.uvmap) || host_face(ed.faces[0]) ||
                      host_face(ed.faces[1]);
    host_edge[e] = host ? 1 : 0;
    double rec[EDGE_F] = {0};
    uint32_t ru[EDGE_U] = {0};
    ru[0] = ed.kind, ru[1] = ed.faces[0], ru[2] = ed.faces[1], ru[3] = 0xffffffffu;
    if (ed.uvmap) {
      ru[3] = uint32_t(uvmap_f.size());
      uint32_t np = std::max<uint32_t>(ed.n_pieces, 1);
      uvmap_f.insert(uvmap_f.end(), ed.uvmap, ed.uvmap + 8 * np);
    }
    Box bx;
    if (ed.kind == ICAM_LINE) {
      V a = rel(ed.a), b = rel(ed.b);
      rec[0] = a.x, rec[1] = a.y, rec[2] = a.z, rec[3] = b.x, rec[4] = b.y, rec[5] = b.z;
      bx.add(a), bx.add(b);
    } else if (ed.kind == ICAM_ARC) {
      V c = rel(ed.o), x = of(ed.x), z = of(ed.z), y = cross(z, x);
      double r[13] = {c.x, c.y, c.z, x.x, x.y, x.z, y.x, y.y, y.z, ed.r, ed.t0, ed.sweep, std::fabs(std::fabs(ed.sweep) - TAU) < 1e-12 ? 1.0 : 0.0};
      std::copy(r, r + 13, rec);
      V axes[3] = {mk(1, 0, 0), mk(0, 1, 0), mk(0, 0, 1)};
      for (V ax : axes) {
        double ex = std::sqrt(std::max(1 - dot(z, ax) * dot(z, ax), 0.0)) * ed.r;
        bx.add(c + ax * ex), bx.add(c - ax * ex);
      }
    } else if (ed.kind == ICAM_ELLIPSE) {
      // curve3.rs:986-1008 bez_pieces: k arcs of at most a quarter turn, each the rational
      // quadratic through its ends with its middle point pushed out by 1 / cos(half its
      // turn) and weighted cos(half its turn)
      V c = rel(ed.o), x = of(ed.x), z = of(ed.z), y = cross(z, x);
      const int k = std::max(1, int(std::ceil(std::fabs(ed.sweep) / (0.5 * PI))));
      const double dt = ed.sweep / k, phi = 0.5 * dt, w = std::cos(phi);
      auto at = [&](double t, double sc) { return c + x * (ed.r * std::cos(t) * sc) + y * (ed.r2 * std::sin(t) * sc); };
      ru[4] = uint32_t(cpiece_f.size() / CPIECE_F), ru[5] = host ? 0u : uint32_t(k);
      for (int i = 0; i < k; ++i) {
        const double ta = ed.t0 + dt * i;
        const V a = at(ta, 1.0), m = at(ta + phi, 1.0 / w), b = at(ta + dt, 1.0);
        std::vector<H> pts = {H{a.x, a.y, a.z, 1.0}, H{m.x * w, m.y * w, m.z * w, w}, H{b.x, b.y, b.z, 1.0}};
        pts = raise(pts);
        for (auto& h : pts) bx.add(point(h));
        if (host) continue;
        const uint32_t cp = uint32_t(cpiece_f.size() / CPIECE_F);
        for (auto& h : pts) cpiece_f.insert(cpiece_f.end(), h.begin(), h.end());
        cpiece_u.insert(cpiece_u.end(), {e, uint32_t(i), uint32_t(k), 0});
        curve_leaves(pts, cp, e);
      }
    } else {
      if (ed.degree < 1) {
        P.error = "spline edge of degree " + std::to_string(ed.degree);
        return false;
      }
      ru[4] = uint32_t(cpiece_f.size() / CPIECE_F), ru[5] = host ? 0u : ed.n_pieces;
      for (uint32_t k = 0; k < ed.n_pieces; ++k) {
        std::vector<H> pts;
        for (uint32_t i = 0; i <= ed.degree; ++i) pts.push_back(relh(ed.pw + (k * (ed.degree + 1) + i) * 4));
        for (auto& h : pts) {
          if (!(h[3] > 0)) {
            P.error = "spline edge with a weight not positive";
            return false;
          }
          bx.add(point(h));
        }
        if (host) continue;
        while (pts.size() < 4) pts = raise(pts);
        uint32_t cp = uint32_t(cpiece_f.size() / CPIECE_F);
        for (auto& h : pts) cpiece_f.insert(cpiece_f.end(), h.begin(), h.end());
        cpiece_u.insert(cpiece_u.end(), {e, k, ed.n_pieces, 0});
        curve_leaves(pts, cp, e);
      }
    }
    edge_f.insert(edge_f.end(), rec, rec + EDGE_F);
    edge_u.insert(edge_u.end(), ru, ru + EDGE_U);
    edge_box[e] = bx;
  }

  // ---- faces
  for (uint32_t fi = 0; fi < part->n_faces; ++fi) {
    const icam_face_desc& fd = part->faces[fi];
    // trim cells
    TrimBuild tb;
    tb.build(fd);
    uint32_t cell0 = uint32_t(cell_f.size() / CELL_F), inst0 = uint32_t(inst_f.size() / INST_F), piece0 = uint32_t(tpiece_f.size() / TPIECE_F),
             among0 = uint32_t(among_u.size());
    for (uint32_t p = 0; p < fd.n_pieces; ++p) tpiece_f.insert(tpiece_f.end(), fd.pieces + 8 * p, fd.pieces + 8 * p + 8);
    for (size_t i = 0; i < tb.inst.size(); ++i) {
      inst_f.insert(inst_f.end(), {tb.inst_shift[i][0], tb.inst_shift[i][1]});
      inst_u.push_back(piece0 + tb.inst_piece[i]);
    }
    for (auto& c : tb.cells) {
      cell_f.insert(cell_f.end(), {c.lo[0], c.lo[1], c.hi[0], c.hi[1], c.r[0], c.r[1]});
      uint32_t a = c.kind == 1 ? among0 + c.a : c.kind == 2 ? cell0 + c.a : 0;
      cell_u.insert(cell_u.end(), {c.inside, c.kind, a, c.b});
    }
    for (uint32_t x : tb.among) among_u.push_back(inst0 + x);
    bool whole = fd.n_pieces == 0;

    double rec[FACE_F] = {0};
    uint32_t ru[FACE_U] = {0};
    ru[0] = fd.kind, ru[1] = cell0, ru[2] = whole ? 1u : 0u;
    rec[14] = fd.sense;
    rec[17] = fd.lo[0], rec[18] = fd.lo[1], rec[19] = fd.hi[0], rec[20] = fd.hi[1], rec[21] = fd.period[0], rec[22] = fd.period[1];
    Box fb;
    if (host_face(fi)) {
      for (uint32_t k = 0; k < fd.nu * fd.nv; ++k) {
        const H h = relh(fd.pw + 4 * k);
        if (!(h[3] > 0)) {
          P.error = "B-spline face with a weight not positive";
          return false;
        }
        fb.add(point(h));
      }
      pbox.push_back(fb);
      pcode.push_back((P_HOST << 28) | fi);
    } else if (fd.kind == ICAM_NURBS) {
      // spans: along u per column, then along v per Bezier row (surface3.rs:1150-1165)
      std::vector<double> ku(fd.ku, fd.ku + fd.nu + fd.du + 1), kv(fd.kv, fd.kv + fd.nv + fd.dv + 1);
      std::vector<std::vector<std::vector<double>>> cols;
      std::vector<std::pair<double, double>> ur, vr;
      for (uint32_t j = 0; j < fd.nv; ++j) {
        std::vector<double> pts;
        for (uint32_t i = 0; i < fd.nu; ++i) pts.insert(pts.end(), fd.pw + (i * fd.nv + j) * 4, fd.pw + (i * fd.nv + j) * 4 + 4);
        std::vector<std::vector<double>> pcs;
        ur.clear();
        decompose(int(fd.du), ku, pts, 4, pcs, ur);
        cols.push_back(pcs);
      }
      int nsu = int(ur.size()), nsv = 0;
      uint32_t span0 = uint32_t(span_f.size() / SPAN_F);
      std::vector<std::vector<std::vector<std::vector<double>>>> rows(nsu);
      for (int a = 0; a < nsu; ++a) {
        rows[a].resize(fd.du + 1);
        for (uint32_t i = 0; i <= fd.du; ++i) {
          std::vector<double> pts;
          for (uint32_t j = 0; j < fd.nv; ++j) pts.insert(pts.end(), cols[j][a].begin() + i * 4, cols[j][a].begin() + i * 4 + 4);
          vr.clear();
          decompose(int(fd.dv), kv, pts, 4, rows[a][i], vr);
        }
        nsv = int(vr.size());
      }
      // the spans' patches, raised to bicubic (gpu.rs:149-171)
      std::vector<Patch4> fp(size_t(nsu) * nsv);
      for (int a = 0; a < nsu; ++a)
        for (int b = 0; b < nsv; ++b) {
          std::vector<std::vector<H>> pw(fd.du + 1);
          for (uint32_t i = 0; i <= fd.du; ++i)
            for (uint32_t j = 0; j <= fd.dv; ++j) pw[i].push_back(relh(&rows[a][i][b][j * 4]));
          for (auto& r : pw)
            while (r.size() < 4) r = raise(r);
          while (pw.size() < 4) {
            std::vector<std::vector<H>> np(pw.size() + 1, std::vector<H>(4));
            for (int j = 0; j < 4; ++j) {
              std::vector<H> col;
              for (auto& r : pw) col.push_back(r[j]);
              col = raise(col);
              for (size_t i = 0; i < col.size(); ++i) np[i][j] = col[i];
            }
            pw = np;
          }
          Patch4& pa = fp[size_t(a) * nsv + b];
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) pa.p[i][j] = pw[i][j];
        }
      // solid3.rs:1609-1625 poles: a side of the domain along which the surface
      // collapses to a point (surface3.rs:1392-1411), away from the vertices and
      // inside the loops, answers as a point record (after the tips) with the fan of
      // the surface's normals around it times the face's sense (solid3.rs:1955-1986)
      struct FacePole {
        int k;
        double value;
        V x;
      };
      std::vector<FacePole> fpoles;
      {
        auto feval = [&](double u, double v, V& p, V& ps, V& pt, V& pst) {
          int a = 0, b = 0;
          for (int k = 1; k < nsu; ++k)
            if (ur[k].first <= u) a = k;
          for (int k = 1; k < nsv; ++k)
            if (vr[k].first <= v) b = k;
          const double su = ur[a].second - ur[a].first, sv = vr[b].second - vr[b].first;
          const double s = std::min(std::max((u - ur[a].first) / su, 0.0), 1.0), t = std::min(std::max((v - vr[b].first) / sv, 0.0), 1.0);
          eval_patch2(fp[size_t(a) * nsv + b], s, t, p, ps, pt, pst);
          ps = ps * (1 / su), pt = pt * (1 / sv), pst = pst * (1 / (su * sv));
        };
        auto inside = [&](double u, double v) { return whole || tb.inside(u, v); };
        Box cb;
        for (auto& pa : fp)
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) cb.add(point(pa.p[i][j]));
        const double scale = std::max(cb.diag(), 1e-9);
        const double range[2][2] = {{ur.front().first, ur.back().second}, {vr.front().first, vr.back().second}};
        for (int k = 0; k < 2; ++k)
          for (double side : {range[k][0], range[k][1]}) {
            const double* other = range[1 - k];
            auto at = [&](double t) {
              V p, ps, pt, pst;
              const double w = other[0] + (other[1] - other[0]) * t;
              feval(k == 0 ? side : w, k == 0 ? w : side, p, ps, pt, pst);
              return p;
            };
            const V p0 = at(0.0);
            bool all = true;
            for (int i = 1; i <= 16; ++i) all &= len(at(i / 16.0) - p0) <= 1e-12 * scale;
            if (!all) continue;
            bool near_vertex = false;
            for (uint32_t vi = 0; vi < part->n_verts; ++vi) near_vertex |= len(of(part->verts[vi].p) - (p0 + o)) <= part->tol;
            const double mid = 0.5 * (other[0] + other[1]);
            if (near_vertex || !(k == 0 ? inside(side, mid) : inside(mid, side))) continue;
            // the fan: limiting tangents S_k(w) and normals S_k x S_kw (u held) or
            // S_kw x S_v (v held), oriented into the domain, each weighted by the angle
            // its tangent turns through where the loops cover the side
            const double h = side == range[k][0] ? 1.0 : -1.0, step = 1e-9 * (range[k][1] - range[k][0]) * h;
            V out = mk(0, 0, 0), ptan = mk(0, 0, 0), pnrm = mk(0, 0, 0);
            bool have = false, pcov = false;
            for (int i = 0; i <= 512; ++i) {
              const double w = other[0] + (other[1] - other[0]) * i / 512.0;
              V p, ps, pt, pst;
              feval(k == 0 ? side : w, k == 0 ? w : side, p, ps, pt, pst);
              V tan = (k == 0 ? ps : pt) * h, nrm = (k == 0 ? cross(ps, pst) : cross(pst, pt)) * h;
              const bool cov = k == 0 ? inside(side + step, w) : inside(w, side + step);
              if (have && pcov && cov && len(tan) > 0 && len(ptan) > 0) {
                const V ua = ptan * (1 / len(ptan)), ub = tan * (1 / len(tan));
                const double ang = std::atan2(len(cross(ua, ub)), dot(ua, ub));
                const V m = pnrm * (1 / len(pnrm)) + nrm * (1 / len(nrm));
                if (len(m) > 0) out = out + m * (ang / len(m));
              }
              ptan = tan, pnrm = nrm, pcov = cov, have = true;
            }
            const V pn = out * fd.sense;
            pole_f.insert(pole_f.end(), {p0.x, p0.y, p0.z, pn.x, pn.y, pn.z, double(fi + 1), 0});
            fpoles.push_back({k, side, p0});
          }
      }
      // surface3.rs:1413-1443 c0_lines: an interior knot of multiplicity at least the
      // degree leaves the surface only C0 along its line, where it may fold; the line's
      // pieces are the near row (u held) or column (v held) of each span after it
      for (int k = 0; k < 2; ++k) {
        const std::vector<std::pair<double, double>>& rr = k == 0 ? ur : vr;
        const std::vector<double>& kn = k == 0 ? ku : kv;
        const int deg = int(k == 0 ? fd.du : fd.dv), nspan = k == 0 ? nsu : nsv, others = k == 0 ? nsv : nsu;
        for (int a = 1; a < nspan; ++a) {
          const double x = rr[a].first;
          if (int(std::count(kn.begin(), kn.end(), x)) < deg) continue;
          Crease c{fi, uint32_t(k), x, {}, {}};
          for (int b = 0; b < others; ++b) {
            const int sa = k == 0 ? a * nsv + b : b * nsv + a, sb = k == 0 ? (a - 1) * nsv + b : b * nsv + a - 1;
            std::vector<H> pts;
            for (int i = 0; i < 4; ++i) pts.push_back(k == 0 ? fp[sa].p[0][i] : fp[sa].p[i][0]);
            c.pieces.push_back(pts);
            c.before.push_back(span0 + uint32_t(sb));
          }
          creases.push_back(c);
        }
      }
      // each span's tree of sub-patches, breadth first (gpu.rs:292-350), built on
      // several threads, then appended in span order
      struct SpanLeaf {
        double f[LEAF_F];
        uint32_t u[LEAF_U];
        Box b;
      };
      struct SpanTree {
        std::vector<SpanLeaf> leaves;
        size_t nodes = 0;
      };
      std::vector<SpanTree> trees(size_t(nsu) * nsv);
      auto span_tree = [&](size_t idx) {
        const int a = int(idx / size_t(nsv)), b = int(idx % size_t(nsv));
        const Patch4& pa = fp[idx];
        const uint32_t sid = span0 + uint32_t(idx);
        SpanTree& out = trees[idx];
        struct Sub {
          Patch4 p;
          double s0, s1, t0, t1;
          int depth;
          int inside;
        };
        std::vector<Sub> queue;
        double klo[2] = {ur[a].first, vr[b].first}, khi[2] = {ur[a].second, vr[b].second};
        queue.push_back({pa, 0, 1, 0, 1, 0, whole ? 1 : tb.box_class(klo, khi)});
        for (size_t qi = 0; qi < queue.size(); ++qi) {
          Sub sb = queue[qi];
          if (sb.inside == 0) continue;  // wholly outside the face's loops
          Box bb;
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) bb.add(point(sb.p.p[i][j]));
          double cert[PATCH_CERT + 1] = {0};
          const double rho = patch_radius(sb.p, cert, 2.0 * bb.diag() + need);
          ++out.nodes;
          if (split_rule(rho, bb.diag(), need, sb.depth)) {
            Patch4 q[4];
            quarter(sb.p, q);
            double sm = 0.5 * (sb.s0 + sb.s1), tm = 0.5 * (sb.t0 + sb.t1);
            for (int k = 0; k < 4; ++k) {
              Sub c;
              c.p = q[k];
              c.s0 = (k & 1) ? sm : sb.s0, c.s1 = (k & 1) ? sb.s1 : sm;
              c.t0 = (k & 2) ? tm : sb.t0, c.t1 = (k & 2) ? sb.t1 : tm;
              c.depth = sb.depth + 1;
              if (sb.inside == 1) {
                c.inside = 1;
              } else {
                double du_ = ur[a].second - ur[a].first, dv_ = vr[b].second - vr[b].first;
                double clo[2] = {ur[a].first + c.s0 * du_, vr[b].first + c.t0 * dv_}, chi[2] = {ur[a].first + c.s1 * du_, vr[b].first + c.t1 * dv_};
                c.inside = tb.box_class(clo, chi);
              }
              queue.push_back(c);
            }
            continue;
          }
          // the slab between two planes normal to the sub-patch's middle normal that
          // holds its control points (the patch lies in their hull)
          V p0, ps0, pt0;
          eval_patch(sb.p, 0.5, 0.5, p0, ps0, pt0);
          V nn = cross(ps0, pt0);
          double nl = len(nn);
          double slo = -INF, shi = INF;
          if (nl > 0) {
            nn = nn * (1 / nl);
            slo = INF, shi = -INF;
            for (int i = 0; i < 4; ++i)
              for (int j = 0; j < 4; ++j) {
                double x = dot(point(sb.p.p[i][j]), nn);
                slo = std::min(slo, x), shi = std::max(shi, x);
              }
            double pad = 1e-9 * (std::fabs(slo) + std::fabs(shi) + bb.diag());
            slo -= pad, shi += pad;
          } else {
            nn = mk(0, 0, 1);
          }
          SpanLeaf lf;
          const double head[12] = {rho, sb.s0, sb.s1, sb.t0, sb.t1, nn.x, nn.y, nn.z, slo, shi, 0, 0};
          std::copy(head, head + 12, lf.f);
          std::copy(cert, cert + PATCH_CERT + 1, lf.f + 12);
          // a pole whose side the sub-patch touches, and that side in the span's
          // parameters (1 s = 0, 2 s = 1, 3 t = 0, 4 t = 1)
          double pole[4] = {0, 0, 0, 0};
          for (const FacePole& fpo : fpoles) {
            int side = 0;
            if (fpo.k == 0 && fpo.value == ur[a].first && sb.s0 == 0.0) side = 1;
            if (fpo.k == 0 && fpo.value == ur[a].second && sb.s1 == 1.0) side = 2;
            if (fpo.k == 1 && fpo.value == vr[b].first && sb.t0 == 0.0) side = 3;
            if (fpo.k == 1 && fpo.value == vr[b].second && sb.t1 == 1.0) side = 4;
            if (side) pole[0] = fpo.x.x, pole[1] = fpo.x.y, pole[2] = fpo.x.z, pole[3] = side;
          }
          std::copy(pole, pole + 4, lf.f + 12 + PATCH_CERT + 1);
          lf.u[0] = sid, lf.u[1] = fi, lf.u[2] = sb.inside == 1 ? 1u : 0u, lf.u[3] = 0;
          lf.b = bb;
          out.leaves.push_back(lf);
        }
      };
      parallel_for(trees.size(), span_tree);
      for (int a = 0; a < nsu; ++a)
        for (int b = 0; b < nsv; ++b) {
          const Patch4& pa = fp[size_t(a) * nsv + b];
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) span_f.insert(span_f.end(), pa.p[i][j].begin(), pa.p[i][j].end());
          span_f.insert(span_f.end(), {ur[a].first, ur[a].second, vr[b].first, vr[b].second, 0, 0, 0, 0});
          span_u.insert(span_u.end(), {fi, whole ? 0u : 1u, cell0, 0});
          const SpanTree& tr = trees[size_t(a) * nsv + b];
          P.nodes += tr.nodes;
          for (const SpanLeaf& lf : tr.leaves) {
            const uint32_t id = uint32_t(leaf_f.size() / LEAF_F);
            leaf_f.insert(leaf_f.end(), lf.f, lf.f + LEAF_F);
            leaf_u.insert(leaf_u.end(), lf.u, lf.u + LEAF_U);
            pbox.push_back(lf.b);
            pcode.push_back((P_LEAF << 28) | id);
            fb.add(lf.b);
          }
        }
      ru[3] = span0, ru[4] = uint32_t(nsu), ru[5] = uint32_t(nsv);
    } else {
      V c = rel(fd.o), x = of(fd.x), z = of(fd.z), y = cross(z, x);
      double r[12] = {c.x, c.y, c.z, x.x, x.y, x.z, y.x, y.y, y.z, z.x, z.y, z.z};
      std::copy(r, r + 12, rec);
      rec[12] = fd.c[0], rec[13] = fd.c[1];
      double psi0 = 0, span = TAU;
      int narc = 4;
      if (fd.kind == ICAM_TORUS && fd.c[0] < fd.c[1]) {
        double pp = std::acos(std::min(std::max(-fd.c[0] / fd.c[1], -1.0), 1.0));
        psi0 = -pp, span = 2 * pp;
        narc = int(std::max(std::ceil(span / (0.5 * PI) - 1e-9), 1.0));
      }
      rec[15] = psi0, rec[16] = span, ru[3] = uint32_t(narc);
      // solid3.rs:1211-1282 face_box
      V axes[3] = {mk(1, 0, 0), mk(0, 1, 0), mk(0, 0, 1)};
      auto circle = [&](V cc, V zz, double rad) {
        for (V ax : axes) {
          double ex = std::sqrt(std::max(1 - dot(zz, ax) * dot(zz, ax), 0.0)) * rad;
          fb.add(cc + ax * ex), fb.add(cc - ax * ex);
        }
      };
      // the face's edges and its vertex loops' points (a cone's apex)
      if (fd.kind == ICAM_PLANE) {
        for (uint32_t e = 0; e < part->n_edges; ++e)
          if (part->edges[e].faces[0] == fi || part->edges[e].faces[1] == fi) fb.add(edge_box[e]);
        for (uint32_t k = 0; k < fd.n_points; ++k) fb.add(rel(fd.points + 3 * k));
      } else if (fd.kind == ICAM_CYLINDER || fd.kind == ICAM_CONE) {
        double h0 = INF, h1 = -INF;
        for (uint32_t k = 0; k < fd.n_points; ++k) {
          double h = dot(rel(fd.points + 3 * k) - c, z);
          h0 = std::min(h0, h), h1 = std::max(h1, h);
        }
        for (uint32_t e = 0; e < part->n_edges; ++e)
          if (part->edges[e].faces[0] == fi || part->edges[e].faces[1] == fi)
            for (int k = 0; k < 8; ++k) {
              const Box& eb = edge_box[e];
              V q = mk((k & 1) ? eb.hi.x : eb.lo.x, (k & 2) ? eb.hi.y : eb.lo.y, (k & 4) ? eb.hi.z : eb.lo.z);
              double h = dot(q - c, z);
              h0 = std::min(h0, h), h1 = std::max(h1, h);
            }
        for (double h : {h0, h1}) circle(c + z * h, z, fd.kind == ICAM_CONE ? std::fabs(fd.c[0] + h * std::tan(fd.c[1])) : fd.c[0]);
      } else if (fd.kind == ICAM_SPHERE) {
        fb.add(c - mk(fd.c[0], fd.c[0], fd.c[0])), fb.add(c + mk(fd.c[0], fd.c[0], fd.c[0]));
      } else {
        for (double h : {-fd.c[1], fd.c[1]}) circle(c + z * h, z, fd.c[0] + fd.c[1]);
      }
      pbox.push_back(fb);
      pcode.push_back((P_FACE << 28) | fi);
    }
    face_box[fi] = fb;
    face_f.insert(face_f.end(), rec, rec + FACE_F);
    face_u.insert(face_u.end(), ru, ru + FACE_U);
  }

  // ---- edge and vertex primitives (after the faces: their order in the list
  // follows the reference's tree ids, faces, edges, vertices)
  for (uint32_t e = 0; e < part->n_edges; ++e)
    if (host_edge[e]) {
      pbox.push_back(edge_box[e]);
      pcode.push_back((P_HOST << 28) | HOST_EDGE | e);
    } else if (part->edges[e].kind == ICAM_LINE || part->edges[e].kind == ICAM_ARC) {
      pbox.push_back(edge_box[e]);
      pcode.push_back((P_EDGE << 28) | e);
    }
  for (uint32_t v = 0; v < part->n_verts; ++v) {
    V p = rel(part->verts[v].p), pn = of(part->verts[v].pn);
    vert_f.insert(vert_f.end(), {p.x, p.y, p.z, pn.x, pn.y, pn.z, 0, 0});
    Box b;
    b.add(p);
    pbox.push_back(b);
    pcode.push_back((P_VERT << 28) | v);
  }
  // solid3.rs:1628-1642 tips: a cone's apex inside its face and away from the
  // vertices answers as a point with the pseudo-normal z (-sin a) sense
  // (solid3.rs:1991-2000); it is packed after the vertices as one more point record,
  // word 6 holding its face + 1 (the feature it reports)
  // A spindle torus's tips (|major| <= minor: lemon or apple) likewise, on the axis at
  // +-h, h = sqrt(minor^2 - major^2), pseudo-normals +-z sense.
  uint32_t n_tips = 0;
  auto tip = [&](uint32_t fi, V at, V pn, double u, double v) {
    const icam_face_desc& fd = part->faces[fi];
    bool near_vertex = false;
    for (uint32_t k = 0; k < part->n_verts; ++k) near_vertex |= len(of(part->verts[k].p) - at) <= part->tol;
    TrimBuild tb;
    tb.build(fd);
    if (near_vertex || !tb.inside(u, v)) return;
    V p = at - o;
    vert_f.insert(vert_f.end(), {p.x, p.y, p.z, pn.x, pn.y, pn.z, double(fi + 1), 0});
    Box b;
    b.add(p);
    pbox.push_back(b);
    pcode.push_back((P_VERT << 28) | (part->n_verts + n_tips));
    ++n_tips;
  };
  for (uint32_t fi = 0; fi < part->n_faces; ++fi) {
    const icam_face_desc& fd = part->faces[fi];
    V z = of(fd.z);
    if (fd.kind == ICAM_CONE) {
      // the apex's parameters: angle 0 about the axis (arc parameter 0), height ha
      const double ha = -fd.c[0] / std::tan(fd.c[1]);
      tip(fi, of(fd.o) + z * ha, z * (-std::sin(fd.c[1]) * fd.sense), 0.0, ha);
    } else if (fd.kind == ICAM_TORUS && std::fabs(fd.c[0]) <= fd.c[1]) {
      // the tips' parameters: angle 0, tube angle atan2(+-h, -major) (arc parameter)
      const double h = std::sqrt(std::max(fd.c[1] * fd.c[1] - fd.c[0] * fd.c[0], 0.0));
      const double pp = std::acos(std::min(std::max(-fd.c[0] / fd.c[1], -1.0), 1.0));
      const int narc = int(std::max(std::ceil(2 * pp / (0.5 * PI) - 1e-9), 1.0));
      for (int sgn = 1; sgn >= -1; sgn -= 2)
        tip(fi, of(fd.o) + z * (sgn * h), z * (sgn * fd.sense), 0.0, arc_param_over(std::atan2(sgn * h, -fd.c[0]), -pp, 2 * pp, narc));
    }
  }

  // solid3.rs:1642-1650 the C0 lines, after the tips as the reference's tree holds
  // them after the vertices: an edge record each after the part's edges, their pieces,
  // and their trees of curve leaves
  for (const Crease& c : creases) {
    const uint32_t e = uint32_t(edge_u.size() / EDGE_U);
    double rec[EDGE_F] = {0};
    uint32_t ru[EDGE_U] = {0};
    rec[0] = c.value;
    ru[0] = EDGE_CREASE, ru[1] = ru[2] = c.face, ru[3] = 0xffffffffu;
    ru[4] = uint32_t(cpiece_f.size() / CPIECE_F), ru[5] = uint32_t(c.pieces.size()), ru[6] = c.k;
    for (size_t k = 0; k < c.pieces.size(); ++k) {
      const uint32_t cp = uint32_t(cpiece_f.size() / CPIECE_F);
      for (const H& h : c.pieces[k]) cpiece_f.insert(cpiece_f.end(), h.begin(), h.end());
      cpiece_u.insert(cpiece_u.end(), {e, uint32_t(k), uint32_t(c.pieces.size()), c.before[k]});
      curve_leaves(c.pieces[k], cp, e);
    }
    edge_f.insert(edge_f.end(), rec, rec + EDGE_F);
    edge_u.insert(edge_u.end(), ru, ru + EDGE_U);
  }

  // the poles, after the tips: point records reporting their face
  for (size_t k = 0; k + VERT_F <= pole_f.size(); k += VERT_F) {
    vert_f.insert(vert_f.end(), pole_f.begin() + k, pole_f.begin() + k + VERT_F);
    Box b;
    b.add(mk(pole_f[k], pole_f[k + 1], pole_f[k + 2]));
    pbox.push_back(b);
    pcode.push_back((P_VERT << 28) | (part->n_verts + n_tips));
    ++n_tips;
  }

  // ---- concatenate
  auto put = [](std::vector<double>& dst, const std::vector<double>& src) {
    uint32_t at = uint32_t(dst.size());
    dst.insert(dst.end(), src.begin(), src.end());
    return at;
  };
  auto putu = [](std::vector<uint32_t>& dst, const std::vector<uint32_t>& src) {
    uint32_t at = uint32_t(dst.size());
    dst.insert(dst.end(), src.begin(), src.end());
    return at;
  };
  P.F.clear(), P.U.clear();
  L.face_f = put(P.F, face_f), L.edge_f = put(P.F, edge_f), L.vert_f = put(P.F, vert_f), L.leaf_f = put(P.F, leaf_f);
  L.span_f = put(P.F, span_f), L.cleaf_f = put(P.F, cleaf_f), L.cpiece_f = put(P.F, cpiece_f), L.cell_f = put(P.F, cell_f);
  L.inst_f = put(P.F, inst_f), L.tpiece_f = put(P.F, tpiece_f), L.uvmap_f = put(P.F, uvmap_f);
  P.F.push_back(0);
  L.face_u = putu(P.U, face_u), L.edge_u = putu(P.U, edge_u), L.leaf_u = putu(P.U, leaf_u), L.span_u = putu(P.U, span_u);
  L.cleaf_u = putu(P.U, cleaf_u), L.cpiece_u = putu(P.U, cpiece_u), L.cell_u = putu(P.U, cell_u), L.inst_u = putu(P.U, inst_u);
  L.among_u = putu(P.U, among_u);
  P.U.push_back(0);
  // uvmap offsets are relative to the uvmap section
  L.n_faces = part->n_faces, L.n_edges = part->n_edges + uint32_t(creases.size()), L.n_verts = part->n_verts + n_tips;
  L.n_spans = uint32_t(span_u.size() / SPAN_U), L.n_leaves = uint32_t(leaf_u.size() / LEAF_U), L.n_cleaves = uint32_t(cleaf_u.size() / CLEAF_U);
  L.n_cpieces = uint32_t(cpiece_u.size() / CPIECE_U);
  L.n_prims = uint32_t(pcode.size());
  P.pbox.resize(6 * pbox.size());
  for (size_t i = 0; i < pbox.size(); ++i) {
    P.pbox[6 * i + 0] = down(pbox[i].lo.x), P.pbox[6 * i + 1] = down(pbox[i].lo.y), P.pbox[6 * i + 2] = down(pbox[i].lo.z);
    P.pbox[6 * i + 3] = up(pbox[i].hi.x), P.pbox[6 * i + 4] = up(pbox[i].hi.y), P.pbox[6 * i + 5] = up(pbox[i].hi.z);
  }
  P.pcode = pcode;
  V plo = rel(part->lo), phi = rel(part->hi);
  L.lo[0] = down(plo.x), L.lo[1] = down(plo.y), L.lo[2] = down(plo.z);
  L.hi[0] = up(phi.x), L.hi[1] = up(phi.y), L.hi[2] = up(phi.z);
  L.tol = float(part->tol);
  L.diag = len(phi - plo);
  double ext = std::max(0.5 * L.diag, 1.0);
  L.eps32 = float(8.0 * 1.1920929e-7 * ext);  // partgpu.rs:139
  L.eps64 = 8.0 * 2.220446049250313e-16 * ext;
  L.max_rounds = 4096;
  build_hierarchy(P);
  return true;
}

}  // namespace icam
