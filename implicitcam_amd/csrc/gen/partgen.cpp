// Synthetic parts for the exact signed-distance evaluator, written directly in
// the evaluator's data model (include/icam_part.h): faces on exact surfaces with
// their trim loops given as monotone polynomial cubics in the face's parameter
// domain, edges with the faces that share them, vertices with angle-weighted
// pseudo-normals.
//
// The reference builds such a solid from a STEP file (Solid3::from_step,
// crates/icam-core/src/solid3.rs:597-851) and fits and certifies the trim loops
// (trim2::build, crates/icam-core/src/trim2.rs:853-1090). Here every loop is
// known in closed form: iso-parameter lines are exact cubics, a NURBS boundary
// curve seen in a side plane is an exact polynomial cubic per knot span (unit
// boundary weights, control points at the Greville abscissae, so the plane's
// first coordinate is linear in the knot parameter), and a circle in a plane is
// approximated by 16 cubic pieces a quadrant (radial deviation below 1e-10 of
// the radius).
//
// Parts:
//   "can"   solid cylinder r 3, z 0..5 (the reference's `can`, solid3.rs:2013-2040)
//   "bump"  10 x 10 box with a biquadratic B-spline top (`bump_box`, solid3.rs:2094-2143)
//   "c1"    box 100 x 60 x 40 minus a through hole r 10 along z with a torus
//           fillet r 3 on the hole's top edge
//   "c2"    120 x 120 block, top a rational bicubic NURBS, 16 x 16 net, clamped
//           uniform knots, heights from seeded 3-octave value noise scaled to 5 mm,
//           interior weights uniform in [0.8, 1.2]
//   "c2flat" the same with every control point at the top's base height: its top
//           is the plane z = 20 and the part the box [-60, 60]^2 x [-20, 20]
//   "oblique:<k>" a cylinder on a disc cut at the top by an oblique plane: an ellipse
//           edge (radius, height and angle from the seed and k)
//   "crease:<k>" the c2 block on a 17 x 17 net with a knot of multiplicity three at
//           0.5 along u (k even) or v (k odd): the top folds along that knot line, a
//           ridge of height 12 (a valley for k = 2, 3) fading to zero at the boundary
//   "quintic:<k>" an 80 x 80 block with a rational degree-(5, 5) top whose outer ring
//           of control points is flat (so its boundary curves are straight): top
//           edges as lines (k = 0) or as the boundary rows' degree-5 pieces (k = 1)
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "../../../include/icam_part.h"

namespace {

struct V {
  double x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double len(V a) { return std::sqrt(dot(a, a)); }
inline V unit(V a) { return a * (1.0 / len(a)); }

// xorshift64 as the reference's checks seed it (crates/icam-check/src/exact_check.rs:46-53).
struct Rng {
  uint64_t s;
  double next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return double(s >> 11) / double(uint64_t(1) << 53);
  }
  uint64_t next_u64() {
    next();
    return s;
  }
};

struct Owned {
  icam_part_desc desc;  // first member: the handle handed out is this
  std::vector<icam_face_desc> faces;
  std::vector<icam_edge_desc> edges;
  std::vector<icam_vert_desc> verts;
  std::deque<std::vector<double>> arrays;
  // what the part was made from, for closed forms (icam_gen_describe)
  std::vector<double> info;

  const double* keep(std::vector<double> v) {
    arrays.push_back(std::move(v));
    return arrays.back().data();
  }
  void finish(V lo, V hi, double tol) {
    std::memset(&desc, 0, sizeof(desc));
    desc.n_faces = uint32_t(faces.size());
    desc.n_edges = uint32_t(edges.size());
    desc.n_verts = uint32_t(verts.size());
    desc.faces = faces.data();
    desc.edges = edges.data();
    desc.verts = verts.data();
    desc.lo[0] = lo.x, desc.lo[1] = lo.y, desc.lo[2] = lo.z;
    desc.hi[0] = hi.x, desc.hi[1] = hi.y, desc.hi[2] = hi.z;
    desc.tol = tol;
  }
};

void put3(double* d, V v) { d[0] = v.x, d[1] = v.y, d[2] = v.z; }

// ------------------------------------------------------------- trim pieces

using Piece = std::array<double, 8>;

Piece line_piece(double u0, double v0, double u1, double v1) {
  return {u0, v0, u0 + (u1 - u0) / 3.0, v0 + (v1 - v0) / 3.0, u0 + 2.0 * (u1 - u0) / 3.0, v0 + 2.0 * (v1 - v0) / 3.0, u1, v1};
}

// The straight line from (u0, v0) to (u1, v1) as k cubic pieces end to end.
void add_line(std::vector<Piece>& out, double u0, double v0, double u1, double v1, int k) {
  for (int i = 0; i < k; ++i)
    out.push_back(line_piece(u0 + (u1 - u0) * i / k, v0 + (v1 - v0) * i / k, u0 + (u1 - u0) * (i + 1) / k, v0 + (v1 - v0) * (i + 1) / k));
}

void add_polyline(std::vector<Piece>& out, const std::vector<std::array<double, 2>>& pts, bool closed) {
  size_t n = pts.size();
  for (size_t i = 0; i + 1 < n + (closed ? 1 : 0); ++i) {
    auto a = pts[i], b = pts[(i + 1) % n];
    out.push_back(line_piece(a[0], a[1], b[0], b[1]));
  }
}

// A circle (cu, cv, r) in a plane's parameters, 16 cubic pieces a quadrant, each
// monotone in both parameters (the quadrant ends are piece ends).
void add_circle(std::vector<Piece>& out, double cu, double cv, double r) {
  const int n = 64;
  for (int i = 0; i < n; ++i) {
    double a = 2.0 * M_PI * i / n, b = 2.0 * M_PI * (i + 1) / n;
    double k = 4.0 / 3.0 * std::tan((b - a) / 4.0);
    double p0u = cu + r * std::cos(a), p0v = cv + r * std::sin(a);
    double p3u = cu + r * std::cos(b), p3v = cv + r * std::sin(b);
    double p1u = p0u - k * r * std::sin(a), p1v = p0v + k * r * std::cos(a);
    double p2u = p3u + k * r * std::sin(b), p2v = p3v - k * r * std::cos(b);
    out.push_back({p0u, p0v, p1u, p1v, p2u, p2v, p3u, p3v});
  }
}

// de Casteljau split of a 2-D cubic at t.
void split2(const Piece& c, double t, Piece& l, Piece& r) {
  double b[4][2];
  for (int i = 0; i < 4; ++i) b[i][0] = c[2 * i], b[i][1] = c[2 * i + 1];
  double L[4][2], R[4][2];
  L[0][0] = b[0][0], L[0][1] = b[0][1];
  R[3][0] = b[3][0], R[3][1] = b[3][1];
  for (int lev = 1; lev <= 3; ++lev) {
    for (int i = 0; i <= 3 - lev; ++i)
      for (int k = 0; k < 2; ++k) b[i][k] = (1 - t) * b[i][k] + t * b[i + 1][k];
    L[lev][0] = b[0][0], L[lev][1] = b[0][1];
    R[3 - lev][0] = b[3 - lev][0], R[3 - lev][1] = b[3 - lev][1];
  }
  for (int i = 0; i < 4; ++i) l[2 * i] = L[i][0], l[2 * i + 1] = L[i][1], r[2 * i] = R[i][0], r[2 * i + 1] = R[i][1];
}

// Roots in (0, 1) of the derivative of one coordinate of a cubic.
void deriv_roots(const Piece& c, int k, std::vector<double>& ts) {
  double p0 = c[k], p1 = c[2 + k], p2 = c[4 + k], p3 = c[6 + k];
  // derivative / 3 = (p1-p0)(1-t)^2 + 2(p2-p1)t(1-t) + (p3-p2)t^2 = A t^2 + B t + C
  double d0 = p1 - p0, d1 = p2 - p1, d2 = p3 - p2;
  double A = d0 - 2 * d1 + d2, B = 2 * (d1 - d0), C = d0;
  auto keep = [&](double t) {
    if (t > 1e-12 && t < 1 - 1e-12) ts.push_back(t);
  };
  if (std::fabs(A) < 1e-14 * (std::fabs(B) + std::fabs(C) + 1e-300)) {
    if (B != 0) keep(-C / B);
    return;
  }
  double disc = B * B - 4 * A * C;
  if (disc < 0) return;
  double s = std::sqrt(disc);
  double q = -0.5 * (B + (B >= 0 ? s : -s));
  if (q != 0) keep(C / q);
  keep(q / A);
}

// A cubic cut into pieces monotone in both coordinates.
void add_monotone(std::vector<Piece>& out, const Piece& c) {
  std::vector<double> ts;
  deriv_roots(c, 0, ts);
  deriv_roots(c, 1, ts);
  std::sort(ts.begin(), ts.end());
  Piece rest = c;
  double done = 0.0;
  for (double t : ts) {
    if (t - done < 1e-9) continue;
    double local = (t - done) / (1.0 - done);
    Piece l, r;
    split2(rest, local, l, r);
    out.push_back(l);
    rest = r;
    done = t;
  }
  out.push_back(rest);
}

// ------------------------------------------------------------- B-spline helpers

// Bezier pieces of a clamped B-spline of degree p (points of dimension N, flat),
// by inserting every interior knot until it has multiplicity p.
std::vector<std::vector<double>> decompose(int p, std::vector<double> kn, std::vector<double> pts, int N) {
  auto count = [&](double u) {
    int c = 0;
    for (double k : kn) c += (k == u);
    return c;
  };
  std::vector<double> interior;
  for (size_t i = p + 1; i + p + 1 < kn.size(); ++i)
    if (interior.empty() || interior.back() != kn[i]) interior.push_back(kn[i]);
  for (double u : interior) {
    while (count(u) < p) {
      // The NURBS Book algorithm 5.1, one insertion.
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
  std::vector<std::vector<double>> out;
  int np = int(pts.size()) / N;
  for (int s = 0; s + p < np; s += p) out.emplace_back(pts.begin() + s * N, pts.begin() + (s + p + 1) * N);
  return out;
}

// Degree elevation of a 2-D Bezier piece to a cubic (degree 1, 2 or 3).
Piece to_cubic2(const std::vector<double>& c, int deg) {
  if (deg == 3) return {c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]};
  if (deg == 1) return line_piece(c[0], c[1], c[2], c[3]);
  // quadratic
  return {c[0], c[1], c[0] / 3 + 2 * c[2] / 3, c[1] / 3 + 2 * c[3] / 3, 2 * c[2] / 3 + c[4] / 3, 2 * c[3] / 3 + c[5] / 3, c[4], c[5]};
}

// ------------------------------------------------------------- builders

struct Builder {
  Owned* o;

  uint32_t plane(V org, V z, V x, double sense, double lo0, double lo1, double hi0, double hi1, const std::vector<Piece>& pcs, double r0, double r1,
                 bool rin) {
    icam_face_desc f;
    std::memset(&f, 0, sizeof(f));
    f.kind = ICAM_PLANE;
    f.sense = sense;
    put3(f.o, org), put3(f.x, x), put3(f.z, z);
    set_trim(f, lo0, lo1, hi0, hi1, 0, 0, pcs, r0, r1, rin);
    o->faces.push_back(f);
    return uint32_t(o->faces.size() - 1);
  }
  uint32_t analytic(uint32_t kind, V org, V z, V x, double c0, double c1, double sense, double lo0, double lo1, double hi0, double hi1, double per0,
                    double per1, const std::vector<Piece>& pcs, double r0, double r1, bool rin) {
    icam_face_desc f;
    std::memset(&f, 0, sizeof(f));
    f.kind = kind;
    f.sense = sense;
    put3(f.o, org), put3(f.x, x), put3(f.z, z);
    f.c[0] = c0, f.c[1] = c1;
    set_trim(f, lo0, lo1, hi0, hi1, per0, per1, pcs, r0, r1, rin);
    o->faces.push_back(f);
    return uint32_t(o->faces.size() - 1);
  }
  void set_trim(icam_face_desc& f, double lo0, double lo1, double hi0, double hi1, double per0, double per1, const std::vector<Piece>& pcs, double r0,
                double r1, bool rin) {
    f.lo[0] = lo0, f.lo[1] = lo1, f.hi[0] = hi0, f.hi[1] = hi1;
    f.period[0] = per0, f.period[1] = per1;
    std::vector<double> flat;
    for (auto& p : pcs) flat.insert(flat.end(), p.begin(), p.end());
    f.n_pieces = uint32_t(pcs.size());
    f.pieces = pcs.empty() ? nullptr : o->keep(flat);
    f.ref[0] = r0, f.ref[1] = r1;
    f.ref_inside = rin ? 1u : 0u;
  }
  void line(V a, V b, uint32_t f0, uint32_t f1, const double* uvmap = nullptr) {
    icam_edge_desc e;
    std::memset(&e, 0, sizeof(e));
    e.kind = ICAM_LINE;
    put3(e.a, a), put3(e.b, b);
    e.n_pieces = 1;
    e.faces[0] = f0, e.faces[1] = f1;
    e.uvmap = uvmap;
    o->edges.push_back(e);
  }
  void circle(V org, V z, V x, double r, uint32_t f0, uint32_t f1) {
    icam_edge_desc e;
    std::memset(&e, 0, sizeof(e));
    e.kind = ICAM_ARC;
    put3(e.o, org), put3(e.x, x), put3(e.z, z);
    e.r = r, e.t0 = 0.0, e.sweep = 2.0 * M_PI;
    e.n_pieces = 1;
    e.faces[0] = f0, e.faces[1] = f1;
    o->edges.push_back(e);
  }
  void spline(uint32_t degree, const std::vector<std::vector<double>>& pieces, uint32_t f0, uint32_t f1, const std::vector<double>& uvmap) {
    icam_edge_desc e;
    std::memset(&e, 0, sizeof(e));
    e.kind = ICAM_SPLINE;
    e.degree = degree;
    e.n_pieces = uint32_t(pieces.size());
    std::vector<double> flat;
    for (auto& p : pieces) flat.insert(flat.end(), p.begin(), p.end());
    e.pw = o->keep(flat);
    e.faces[0] = f0, e.faces[1] = f1;
    e.uvmap = uvmap.empty() ? nullptr : o->keep(uvmap);
    o->edges.push_back(e);
  }
  void vert(V p, V pn) {
    icam_vert_desc v;
    put3(v.p, p), put3(v.pn, pn);
    o->verts.push_back(v);
  }
};

const double NaN = std::nan("");

// ------------------------------------------------------------- can

Owned* make_can() {
  Owned* o = new Owned();
  Builder b{o};
  V X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1};
  (void)Y;
  std::vector<Piece> disc;
  add_circle(disc, 0, 0, 3.0);
  // bottom: the plane's normal +z, outward -z
  uint32_t bot = b.plane({0, 0, 0}, Z, X, -1.0, -4, -4, 4, 4, disc, -3.5, 0.0123, false);
  uint32_t top = b.plane({0, 0, 5}, Z, X, 1.0, -4, -4, 4, 4, disc, -3.5, 0.0123, false);
  std::vector<Piece> band = {line_piece(0, 0, 4, 0), line_piece(0, 5, 4, 5)};
  uint32_t cyl = b.analytic(ICAM_CYLINDER, {0, 0, 0}, Z, X, 3.0, 0, 1.0, 0, -1, 4, 6, 4, 0, band, 2.0, 2.5, true);
  b.circle({0, 0, 0}, Z, X, 3.0, bot, cyl);
  b.circle({0, 0, 5}, Z, X, 3.0, top, cyl);
  // the seam vertices of the two circles (the fixture's vertices 1 and 2), each
  // between a disc and the cylinder at angle pi in both
  b.vert({3, 0, 0}, V{1, 0, -1} * M_PI);
  b.vert({3, 0, 5}, V{1, 0, 1} * M_PI);
  o->finish({-3, -3, 0}, {3, 3, 5}, 1e-9);
  return o;
}

// ------------------------------------------------------------- bump box

Owned* make_bump() {
  Owned* o = new Owned();
  Builder b{o};
  const double zc[3][3] = {{5, 5, 5}, {7, 9, 5}, {5, 6, 5}};  // zc[i][j], i along x
  // the net, i along u (x), j along v (y)
  std::vector<double> pw;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) pw.insert(pw.end(), {5.0 * i, 5.0 * j, zc[i][j], 1.0});
  icam_face_desc top;
  std::memset(&top, 0, sizeof(top));
  top.kind = ICAM_NURBS;
  top.sense = 1.0;
  top.du = top.dv = 2;
  top.nu = top.nv = 3;
  top.pw = o->keep(pw);
  top.ku = o->keep({0, 0, 0, 1, 1, 1});
  top.kv = o->keep({0, 0, 0, 1, 1, 1});
  b.set_trim(top, 0, 0, 1, 1, 0, 0, {}, 0.5, 0.5, true);
  o->faces.push_back(top);
  uint32_t ft = 0;
  V X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1};
  V b00{0, 0, 0}, b10{10, 0, 0}, b11{10, 10, 0}, b01{0, 10, 0};
  V t00{0, 0, 5}, t10{10, 0, 5}, t11{10, 10, 5}, t01{0, 10, 5};
  // bottom: normal -z, frame x = X, y = (-z) x X = -Y: params (x, -y)
  uint32_t fb = b.plane(b00, Z * -1.0, X, 1.0, -1, -11, 11, 1, [] {
    std::vector<Piece> p;
    add_polyline(p, {{0, 0}, {10, 0}, {10, -10}, {0, -10}}, true);
    return p;
  }(), -0.5, 0.0123, false);
  // y = 0 side: normal -y, frame x = X, y axis = (-Y) x X = +Z: params (x, z)
  std::vector<Piece> s0;
  add_polyline(s0, {{0, 5}, {0, 0}, {10, 0}, {10, 5}}, false);
  for (auto& pc : std::vector<Piece>{to_cubic2({0, 5, 5, 7, 10, 5}, 2)}) add_monotone(s0, pc);
  uint32_t fy0 = b.plane(b00, Y * -1.0, X, 1.0, -1, -1, 11, 11, s0, 5.0, -0.5, false);
  // x = 10 side: normal +x, frame x = Y, y axis = X x Y = Z: params (y, z)
  std::vector<Piece> s1;
  add_polyline(s1, {{0, 5}, {0, 0}, {10, 0}, {10, 5}}, false);
  add_monotone(s1, to_cubic2({0, 5, 5, 6, 10, 5}, 2));
  uint32_t fx1 = b.plane(b10, X, Y, 1.0, -1, -1, 11, 11, s1, 5.0, -0.5, false);
  // y = 10 side: normal +y, frame x = X, y axis = Y x X = -Z: params (x, -z)
  std::vector<Piece> s2;
  add_polyline(s2, {{0, 0}, {10, 0}, {10, -5}, {0, -5}}, true);
  uint32_t fy1 = b.plane(b01, Y, X, 1.0, -1, -11, 11, 1, s2, 5.0, 0.5, false);
  // x = 0 side: normal -x, frame x = Y, y axis = (-X) x Y = -Z: params (y, -z)
  std::vector<Piece> s3;
  add_polyline(s3, {{0, 0}, {10, 0}, {10, -5}, {0, -5}}, true);
  uint32_t fx0 = b.plane(b00, X * -1.0, Y, 1.0, -1, -11, 11, 1, s3, 5.0, 0.5, false);
  // edges: bottom square, vertical, top
  b.line(b00, b10, fb, fy0);
  b.line(b10, b11, fb, fx1);
  b.line(b11, b01, fb, fy1);
  b.line(b01, b00, fb, fx0);
  b.line(b00, t00, fy0, fx0);
  b.line(b10, t10, fy0, fx1);
  b.line(b11, t11, fx1, fy1);
  b.line(b01, t01, fy1, fx0);
  auto nan4 = std::vector<double>{NaN, NaN, NaN, NaN};
  auto cat = [](std::vector<double> a, const std::vector<double>& c) {
    a.insert(a.end(), c.begin(), c.end());
    return a;
  };
  // top boundary y = 0 (v = 0, u 0 -> 1) and x = 10 (u = 1, v 0 -> 1): quadratic B-splines
  b.spline(2, {{0, 0, 5, 1, 5, 0, 7, 1, 10, 0, 5, 1}}, ft, fy0, cat({0, 0, 1, 0}, nan4));
  b.spline(2, {{10, 0, 5, 1, 10, 5, 6, 1, 10, 10, 5, 1}}, ft, fx1, cat({1, 0, 1, 1}, nan4));
  // straight top boundaries y = 10 (v = 1, u 1 -> 0) and x = 0 (u = 0, v 1 -> 0)
  b.line(t11, t01, ft, fy1, o->keep(cat({1, 1, 0, 1}, nan4)));
  b.line(t01, t00, ft, fx0, o->keep(cat({0, 1, 0, 0}, nan4)));
  // vertices: bottom corners (three right angles), top corners from the faces' normals
  double q = M_PI / 2;
  b.vert(b00, V{-1, -1, -1} * q);
  b.vert(b10, V{1, -1, -1} * q);
  b.vert(b11, V{1, 1, -1} * q);
  b.vert(b01, V{-1, 1, -1} * q);
  // top corners: the surface's tangents at its corners (quadratic: 2 (P1 - P0))
  auto corner = [&](V p, V su, V sv, V side_a_n, V side_b_n, V down) {
    // the face angles: top between the two boundary tangents, each side between
    // its boundary tangent and the vertical edge
    auto ang = [](V a, V c) { return std::atan2(len(cross(a, c)), dot(a, c)); };
    V n = unit(cross(su, sv));
    if (n.z < 0) n = n * -1.0;
    V pn = n * ang(su, sv) + side_a_n * ang(su, down) + side_b_n * ang(sv, down);
    b.vert(p, pn);
  };
  // corner t00: tangents toward the surface's inside: +Su along y = 0, +Sv along x = 0
  corner(t00, V{5, 0, 2}, V{0, 5, 0}, Y * -1.0, X * -1.0, Z * -1.0);
  corner(t10, V{-5, 0, 2}, V{0, 5, 1}, Y * -1.0, X, Z * -1.0);
  corner(t11, V{-5, 0, 0}, V{0, -5, 1}, Y, X, Z * -1.0);
  corner(t01, V{5, 0, 0}, V{0, -5, 0}, Y, X * -1.0, Z * -1.0);
  o->finish({0, 0, 0}, {10, 10, 9}, 1e-9);
  return o;
}

// ------------------------------------------------------------- C1

Owned* make_c1() {
  Owned* o = new Owned();
  Builder b{o};
  V X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1};
  const double ax = 50, ay = 30, az = 20, rh = 10, rf = 3;
  const double R = rh + rf, zc = az - rf;  // fillet tube centre circle
  auto rect = [](double u0, double v0, double u1, double v1) {
    std::vector<Piece> p;
    add_polyline(p, {{u0, v0}, {u1, v0}, {u1, v1}, {u0, v1}}, true);
    return p;
  };
  // F0 top z = 20 (params x, y), minus the fillet's disc r 13
  std::vector<Piece> top = rect(-ax, -ay, ax, ay);
  add_circle(top, 0, 0, R);
  uint32_t f_top = b.plane({0, 0, az}, Z, X, 1.0, -ax - 1, -ay - 1, ax + 1, ay + 1, top, -ax - 0.5, 0.0123, false);
  // F1 bottom z = -20, normal -z (params x, -y), minus the hole's disc r 10
  std::vector<Piece> bot = rect(-ax, -ay, ax, ay);
  add_circle(bot, 0, 0, rh);
  uint32_t f_bot = b.plane({0, 0, -az}, Z * -1.0, X, 1.0, -ax - 1, -ay - 1, ax + 1, ay + 1, bot, -ax - 0.5, 0.0123, false);
  // F2 x = 50: frame x = Y, y axis = X x Y = Z: params (y, z)
  uint32_t f_xp = b.plane({ax, 0, 0}, X, Y, 1.0, -ay - 1, -az - 1, ay + 1, az + 1, rect(-ay, -az, ay, az), -ay - 0.5, 0.0123, false);
  // F3 x = -50: frame x = Y, y axis = (-X) x Y = -Z: params (y, -z)
  uint32_t f_xm = b.plane({-ax, 0, 0}, X * -1.0, Y, 1.0, -ay - 1, -az - 1, ay + 1, az + 1, rect(-ay, -az, ay, az), -ay - 0.5, 0.0123, false);
  // F4 y = 30: frame x = X, y axis = Y x X = -Z: params (x, -z)
  uint32_t f_yp = b.plane({0, ay, 0}, Y, X, 1.0, -ax - 1, -az - 1, ax + 1, az + 1, rect(-ax, -az, ax, az), -ax - 0.5, 0.0123, false);
  // F5 y = -30: frame x = X, y axis = (-Y) x X = Z: params (x, z)
  uint32_t f_ym = b.plane({0, -ay, 0}, Y * -1.0, X, 1.0, -ax - 1, -az - 1, ax + 1, az + 1, rect(-ax, -az, ax, az), -ax - 0.5, 0.0123, false);
  // F6 the hole: cylinder r 10 about z, material outside it (sense -1), z from -20 to 17
  std::vector<Piece> wall = {line_piece(0, -az, 4, -az), line_piece(0, zc, 4, zc)};
  uint32_t f_cyl = b.analytic(ICAM_CYLINDER, {0, 0, 0}, Z, X, rh, 0, -1.0, 0, -az - 1, 4, zc + 1, 4, 0, wall, 2.0, 0.0123, true);
  // F7 the fillet: torus major 13 minor 3 about z at z 17, the quarter of its tube from
  // its top (psi = 90 degrees, arc parameter 1) to its inner equator (180, 2)
  std::vector<Piece> fil = {line_piece(0, 1, 4, 1), line_piece(0, 2, 4, 2)};
  uint32_t f_tor = b.analytic(ICAM_TORUS, {0, 0, zc}, Z, X, R, rf, 1.0, 0, 0, 4, 4, 4, 4, fil, 2.0, 1.5, true);
  // edges
  V c[2][2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) c[i][j][k] = V{i ? ax : -ax, j ? ay : -ay, k ? az : -az};
  for (int k = 0; k < 2; ++k) {
    uint32_t f = k ? f_top : f_bot;
    b.line(c[0][0][k], c[1][0][k], f, f_ym);
    b.line(c[1][0][k], c[1][1][k], f, f_xp);
    b.line(c[1][1][k], c[0][1][k], f, f_yp);
    b.line(c[0][1][k], c[0][0][k], f, f_xm);
  }
  b.line(c[0][0][0], c[0][0][1], f_xm, f_ym);
  b.line(c[1][0][0], c[1][0][1], f_xp, f_ym);
  b.line(c[1][1][0], c[1][1][1], f_xp, f_yp);
  b.line(c[0][1][0], c[0][1][1], f_xm, f_yp);
  b.circle({0, 0, az}, Z, X, R, f_top, f_tor);
  b.circle({0, 0, zc}, Z, X, rh, f_cyl, f_tor);
  b.circle({0, 0, -az}, Z, X, rh, f_cyl, f_bot);
  // the circles' seam vertices (at angle 0): the fillet meets the top and the hole
  // tangentially, the hole meets the bottom at a right angle
  b.vert({R, 0, az}, V{0, 0, 2 * M_PI});
  b.vert({rh, 0, zc}, V{-2 * M_PI, 0, 0});
  b.vert({rh, 0, -az}, V{-1, 0, -1} * M_PI);
  double q = M_PI / 2;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) b.vert(c[i][j][k], V{i ? 1.0 : -1.0, j ? 1.0 : -1.0, k ? 1.0 : -1.0} * q);
  (void)Y;
  o->finish({-ax, -ay, -az}, {ax, ay, az}, 1e-9);
  return o;
}

// ------------------------------------------------------------- C2

double hash01(uint64_t seed, int64_t ix, int64_t iy, int oct) {
  uint64_t h = seed ^ (uint64_t(ix) * 0x9E3779B97F4A7C15ull) ^ (uint64_t(iy) * 0xC2B2AE3D27D4EB4Full) ^ (uint64_t(oct + 1) * 0x165667B19E3779F9ull);
  Rng r{h ? h : 1};
  r.next();
  r.next();
  return r.next();
}

// Seeded value noise in [-1, 1]: lattice values smoothly interpolated, octaves of
// periods 80, 40 and 20 mm with amplitudes 1, 1/2, 1/4, normalized by their sum.
double value_noise(uint64_t seed, double x, double y) {
  double sum = 0, amp = 1, norm = 0, period = 80;
  for (int oct = 0; oct < 3; ++oct) {
    double fx = x / period, fy = y / period;
    int64_t ix = int64_t(std::floor(fx)), iy = int64_t(std::floor(fy));
    double tx = fx - ix, ty = fy - iy;
    auto s = [](double t) { return t * t * (3 - 2 * t); };
    double sx = s(tx), sy = s(ty);
    double v00 = 2 * hash01(seed, ix, iy, oct) - 1, v10 = 2 * hash01(seed, ix + 1, iy, oct) - 1;
    double v01 = 2 * hash01(seed, ix, iy + 1, oct) - 1, v11 = 2 * hash01(seed, ix + 1, iy + 1, oct) - 1;
    double v = (v00 * (1 - sx) + v10 * sx) * (1 - sy) + (v01 * (1 - sx) + v11 * sx) * sy;
    sum += amp * v;
    norm += amp;
    amp *= 0.5;
    period *= 0.5;
  }
  return sum / norm;
}

// crease >= 0: a net of 17 x 17 whose knots along direction crease & 1 (0: u) hold 0.5
// three times (the degree), so the top is only C0 along that knot line, and a tent of
// height 12 (a ridge; a valley when crease & 2) across it, faded to zero at the two
// boundary rows it meets: the top folds along the line and stays smooth at its edges.
Owned* make_c2(uint64_t seed, bool flat, int n = 16, double half = 60, double amp = 5, int crease = -1) {
  Owned* o = new Owned();
  Builder b{o};
  const int p = 3;
  if (crease >= 0) n = 17;
  const double zb = -half / 3, ztop = half / 3;
  auto knots = [&](bool folded) {
    std::vector<double> kn;
    for (int i = 0; i <= p; ++i) kn.push_back(0);
    if (folded) {
      for (int i = 1; i < 12; ++i) {
        if (i == 6)
          kn.insert(kn.end(), {0.5, 0.5, 0.5});
        else
          kn.push_back(i / 12.0);
      }
    } else {
      const int spans = n - p;
      for (int i = 1; i < spans; ++i) kn.push_back(double(i) / spans);
    }
    for (int i = 0; i <= p; ++i) kn.push_back(1);
    return kn;
  };
  const std::vector<double> ku = knots(crease >= 0 && (crease & 1) == 0), kv = knots(crease >= 0 && (crease & 1) == 1);
  // the pieces' knot intervals: the distinct knots
  auto breaks = [](const std::vector<double>& kn) {
    std::vector<double> r;
    for (double x : kn)
      if (r.empty() || x > r.back()) r.push_back(x);
    return r;
  };
  const std::vector<double> bu = breaks(ku), bv = breaks(kv);
  // Greville abscissae: the net's x (and y) are then linear in u (and v)
  auto knot_sum = [&](const std::vector<double>& kn, int i) { return kn[i + 1] + kn[i + 2] + kn[i + 3]; };
  std::vector<double> gx(n), gy(n);
  for (int i = 0; i < n; ++i) gx[i] = -half + 2 * half * knot_sum(ku, i) / 3.0, gy[i] = -half + 2 * half * knot_sum(kv, i) / 3.0;
  Rng rng{seed ? seed : 1};
  std::vector<double> pw(n * n * 4);
  std::vector<double> zz(n * n), ww(n * n);
  // the noise scaled so its largest height over the net is the amplitude
  std::vector<double> raw(n * n);
  double peak = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) raw[i * n + j] = value_noise(seed, gx[i], gy[j]), peak = std::max(peak, std::fabs(raw[i * n + j]));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double z = flat ? ztop : ztop + amp * raw[i * n + j] / peak;
      if (crease >= 0) {
        // the tent across the line times a bump along it that is zero on the boundary
        const int a = (crease & 1) == 0 ? i : j, c = (crease & 1) == 0 ? j : i;
        const double g = ((crease & 1) == 0 ? knot_sum(ku, a) : knot_sum(kv, a)) / 3.0;
        const double tent = 1.0 - 2.0 * std::fabs(g - 0.5), bump = (c == 0 || c == n - 1) ? 0.0 : 1.0;
        z += ((crease & 2) ? -12.0 : 12.0) * tent * bump;
      }
      double w = (i == 0 || j == 0 || i == n - 1 || j == n - 1) ? 1.0 : 0.8 + 0.4 * rng.next();
      zz[i * n + j] = z, ww[i * n + j] = w;
      double* q = &pw[(i * n + j) * 4];
      q[0] = gx[i] * w, q[1] = gy[j] * w, q[2] = z * w, q[3] = w;
    }
  icam_face_desc top;
  std::memset(&top, 0, sizeof(top));
  top.kind = ICAM_NURBS;
  top.sense = 1.0;
  top.du = top.dv = p;
  top.nu = top.nv = n;
  top.pw = o->keep(pw);
  top.ku = o->keep(ku);
  top.kv = o->keep(kv);
  b.set_trim(top, 0, 0, 1, 1, 0, 0, {}, 0.5, 0.5, true);
  o->faces.push_back(top);
  const uint32_t ft = 0;
  // the four boundary curves as 3-D cubic Bezier pieces (unit weights), each with
  // its knot intervals: side 0 v = 0 (y = -60), 1 u = 1 (x = 60), 2 v = 1 (y = 60),
  // 3 u = 0 (x = -60), each in the direction of increasing knot parameter
  auto boundary = [&](int side) {
    std::vector<double> pts;
    for (int k = 0; k < n; ++k) {
      int i = 0, j = 0;
      if (side == 0) i = k, j = 0;
      if (side == 1) i = n - 1, j = k;
      if (side == 2) i = k, j = n - 1;
      if (side == 3) i = 0, j = k;
      pts.insert(pts.end(), {gx[i], gy[j], zz[i * n + j], 1.0});
    }
    return decompose(p, side % 2 == 0 ? ku : kv, pts, 4);
  };
  V X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1};
  double top_max = ztop + amp + (crease >= 0 ? 13 : 1);
  uint32_t side_face[4];
  for (int side = 0; side < 4; ++side) {
    auto pcs = boundary(side);
    // the side plane's frame: outward normal, x axis, and which world axis is its
    // first parameter; the second parameter is the height z
    V nrm, xa;
    if (side == 0) nrm = Y * -1.0, xa = X;  // y axis = (-Y) x X = Z
    if (side == 1) nrm = X, xa = Y;         // X x Y = Z
    if (side == 2) nrm = Y, xa = X * -1.0;  // Y x (-X) = Z
    if (side == 3) nrm = X * -1.0, xa = Y * -1.0;  // (-X) x (-Y) = Z
    V org = nrm * half;
    auto param = [&](V w) { return std::array<double, 2>{dot(w - org, xa), w.z}; };
    std::vector<Piece> loop;
    for (auto& pc : pcs) {
      std::vector<double> c2;
      for (int k = 0; k <= p; ++k) {
        auto q = param(V{pc[4 * k], pc[4 * k + 1], pc[4 * k + 2]});
        c2.push_back(q[0]), c2.push_back(q[1]);
      }
      add_monotone(loop, to_cubic2(c2, p));
    }
    V a{pcs.front()[0], pcs.front()[1], pcs.front()[2]};
    V e{pcs.back()[4 * p], pcs.back()[4 * p + 1], pcs.back()[4 * p + 2]};
    auto pa = param(a), pe = param(e);
    auto pab = param(V{a.x, a.y, zb}), peb = param(V{e.x, e.y, zb});
    add_polyline(loop, {{pe[0], pe[1]}, {peb[0], peb[1]}, {pab[0], pab[1]}, {pa[0], pa[1]}}, false);
    side_face[side] = b.plane(org, nrm, xa, 1.0, -half - 1, zb - 1, half + 1, top_max, loop, 0.0123, zb - 0.5, false);
  }
  // bottom: normal -z, frame x = X, y axis = -Y: params (x, -y)
  std::vector<Piece> bot;
  add_polyline(bot, {{-half, -half}, {half, -half}, {half, half}, {-half, half}}, true);
  uint32_t fb = b.plane({0, 0, zb}, Z * -1.0, X, 1.0, -half - 1, -half - 1, half + 1, half + 1, bot, -half - 0.5, 0.0123, false);
  // top spline edges with their pieces' knot intervals on the NURBS face
  for (int side = 0; side < 4; ++side) {
    auto pcs = boundary(side);
    std::vector<double> uv;
    const std::vector<double>& br = side % 2 == 0 ? bu : bv;
    for (int k = 0; k < int(pcs.size()); ++k) {
      double a = br[k], c = br[k + 1];
      if (side == 0) uv.insert(uv.end(), {a, 0, c, 0});
      if (side == 1) uv.insert(uv.end(), {1, a, 1, c});
      if (side == 2) uv.insert(uv.end(), {a, 1, c, 1});
      if (side == 3) uv.insert(uv.end(), {0, a, 0, c});
      uv.insert(uv.end(), {NaN, NaN, NaN, NaN});
    }
    b.spline(p, pcs, ft, side_face[side], uv);
  }
  // corners
  V bc[4] = {{-half, -half, zb}, {half, -half, zb}, {half, half, zb}, {-half, half, zb}};
  auto zat = [&](int i, int j) { return zz[i * n + j]; };
  V tc[4] = {{-half, -half, zat(0, 0)}, {half, -half, zat(n - 1, 0)}, {half, half, zat(n - 1, n - 1)}, {-half, half, zat(0, n - 1)}};
  // side faces around: corner k sits between side (k+3)%4 and side k
  for (int k = 0; k < 4; ++k) {
    int s0 = (k + 3) % 4, s1 = k;
    b.line(bc[k], bc[(k + 1) % 4], fb, side_face[k]);
    b.line(bc[k], tc[k], side_face[s0], side_face[s1]);
  }
  // vertices: bottom corners, and top corners from the three faces' normals and angles
  double q = M_PI / 2;
  for (int k = 0; k < 4; ++k) {
    V d = bc[k] * (1.0 / half);
    b.vert(bc[k], V{d.x > 0 ? 1.0 : -1.0, d.y > 0 ? 1.0 : -1.0, -1.0} * q);
  }
  auto ang = [](V a, V c) { return std::atan2(len(cross(a, c)), dot(a, c)); };
  for (int k = 0; k < 4; ++k) {
    int ci = (k == 1 || k == 2) ? n - 1 : 0, cj = (k >= 2) ? n - 1 : 0;
    int di = ci == 0 ? 1 : -1, dj = cj == 0 ? 1 : -1;
    // tangents into the surface along its two boundaries (unit weights there)
    auto P = [&](int i, int j) { return V{gx[i], gy[j], zat(i, j)}; };
    V tu = P(ci + di, cj) - P(ci, cj), tv = P(ci, cj + dj) - P(ci, cj);
    V n0 = unit(cross(tu, tv));
    if (n0.z < 0) n0 = n0 * -1.0;
    V down{0, 0, -1};
    V nu_side = (cj == 0) ? Y * -1.0 : Y, nv_side = (ci == 0) ? X * -1.0 : X;
    V pn = n0 * ang(tu, tv) + nu_side * ang(tu, down) + nv_side * ang(tv, down);
    b.vert(tc[k], pn);
  }
  double zmax = -1e300;
  for (double z : zz) zmax = std::max(zmax, z);
  o->finish({-half, -half, zb}, {half, half, zmax}, 1e-9);
  return o;
}

// ------------------------------------------------------------- quintic top

// An 80 x 80 block whose top is a rational B-spline of degree (5, 5) on a 10 x 10 net
// (clamped uniform knots, five spans a direction, control points at the Greville
// abscissae), interior heights from the seeded noise scaled to 4 mm and interior
// weights in [0.8, 1.2]. The net's outer ring lies at the top's base height with unit
// weights, so each boundary curve is the straight segment between two top corners
// (a B-spline reproduces a linear function from its Greville abscissae). The four top
// edges are lines (spline_edges false) or the boundary rows' degree-5 Bezier pieces
// (true); the sides and the bottom are planes.
Owned* make_quintic(uint64_t seed, bool spline_edges) {
  Owned* o = new Owned();
  Builder b{o};
  const int p = 5, n = 10;
  const double half = 40, zb = -half / 3, ztop = half / 3, amp = 4;
  std::vector<double> kn;
  for (int i = 0; i <= p; ++i) kn.push_back(0);
  for (int i = 1; i < n - p; ++i) kn.push_back(double(i) / (n - p));
  for (int i = 0; i <= p; ++i) kn.push_back(1);
  std::vector<double> br;  // the distinct knots
  for (double x : kn)
    if (br.empty() || x > br.back()) br.push_back(x);
  std::vector<double> g(n);
  for (int i = 0; i < n; ++i) {
    double sum = 0;
    for (int k = 1; k <= p; ++k) sum += kn[i + k];
    g[i] = -half + 2 * half * sum / p;
  }
  Rng rng{seed ? seed : 1};
  std::vector<double> raw(n * n), zz(n * n), pw(n * n * 4);
  double peak = 1e-300;
  for (int i = 1; i < n - 1; ++i)
    for (int j = 1; j < n - 1; ++j) raw[i * n + j] = value_noise(seed, g[i], g[j]), peak = std::max(peak, std::fabs(raw[i * n + j]));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const bool ring = i == 0 || j == 0 || i == n - 1 || j == n - 1;
      const double z = ring ? ztop : ztop + amp * raw[i * n + j] / peak;
      const double w = ring ? 1.0 : 0.8 + 0.4 * rng.next();
      zz[i * n + j] = z;
      double* q = &pw[(i * n + j) * 4];
      q[0] = g[i] * w, q[1] = g[j] * w, q[2] = z * w, q[3] = w;
    }
  icam_face_desc top;
  std::memset(&top, 0, sizeof(top));
  top.kind = ICAM_NURBS;
  top.sense = 1.0;
  top.du = top.dv = p;
  top.nu = top.nv = n;
  top.pw = o->keep(pw);
  top.ku = o->keep(kn);
  top.kv = o->keep(kn);
  b.set_trim(top, 0, 0, 1, 1, 0, 0, {}, 0.5, 0.5, true);
  o->faces.push_back(top);
  const uint32_t ft = 0;
  V X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1};
  // side k: 0 y = -half, 1 x = +half, 2 y = +half, 3 x = -half; the plane's
  // parameters (along its x axis, z), each side a rectangle
  uint32_t side_face[4];
  for (int k = 0; k < 4; ++k) {
    V nrm, xa;
    if (k == 0) nrm = Y * -1.0, xa = X;
    if (k == 1) nrm = X, xa = Y;
    if (k == 2) nrm = Y, xa = X * -1.0;
    if (k == 3) nrm = X * -1.0, xa = Y * -1.0;
    std::vector<Piece> loop;
    add_polyline(loop, {{-half, zb}, {half, zb}, {half, ztop}, {-half, ztop}}, true);
    side_face[k] = b.plane(nrm * half, nrm, xa, 1.0, -half - 1, zb - 1, half + 1, ztop + 1, loop, 0.0123, zb - 0.5, false);
  }
  std::vector<Piece> bot;
  add_polyline(bot, {{-half, -half}, {half, -half}, {half, half}, {-half, half}}, true);
  const uint32_t fb = b.plane({0, 0, zb}, Z * -1.0, X, 1.0, -half - 1, -half - 1, half + 1, half + 1, bot, -half - 0.5, 0.0123, false);
  V bc[4] = {{-half, -half, zb}, {half, -half, zb}, {half, half, zb}, {-half, half, zb}};
  V tc[4] = {{-half, -half, ztop}, {half, -half, ztop}, {half, half, ztop}, {-half, half, ztop}};
  // top edge k from tc[k] to tc[k + 1], and where it runs on the top: side 0 v = 0
  // (u up), 1 u = 1 (v up), 2 v = 1 (u down), 3 u = 0 (v down)
  const double ua[4][4] = {{0, 0, 1, 0}, {1, 0, 1, 1}, {1, 1, 0, 1}, {0, 1, 0, 0}};
  for (int k = 0; k < 4; ++k) {
    if (!spline_edges) {
      const double uv[8] = {ua[k][0], ua[k][1], ua[k][2], ua[k][3], NaN, NaN, NaN, NaN};
      b.line(tc[k], tc[(k + 1) % 4], ft, side_face[k], o->keep(std::vector<double>(uv, uv + 8)));
      continue;
    }
    // the boundary row in the edge's direction, cut into its Bezier pieces
    std::vector<double> pts;
    for (int m = 0; m < n; ++m) {
      const int t = (k >= 2) ? n - 1 - m : m;
      const int i = k == 0 ? t : k == 1 ? n - 1 : k == 2 ? t : 0;
      const int j = k == 0 ? 0 : k == 1 ? t : k == 2 ? n - 1 : t;
      pts.insert(pts.end(), {g[i], g[j], zz[i * n + j], 1.0});
    }
    // (the clamped uniform knots are symmetric: a reversed row has the same)
    auto pcs = decompose(p, kn, pts, 4);
    std::vector<double> uv;
    for (int c = 0; c < int(pcs.size()); ++c) {
      const double a = double(c) / pcs.size(), e = double(c + 1) / pcs.size();
      auto lerp = [&](double t, int w) { return ua[k][w] + t * (ua[k][w + 2] - ua[k][w]); };
      uv.insert(uv.end(), {lerp(a, 0), lerp(a, 1), lerp(e, 0), lerp(e, 1), NaN, NaN, NaN, NaN});
    }
    b.spline(p, pcs, ft, side_face[k], uv);
  }
  for (int k = 0; k < 4; ++k) {
    b.line(bc[k], bc[(k + 1) % 4], fb, side_face[k]);
    b.line(bc[k], tc[k], side_face[(k + 3) % 4], side_face[k]);
  }
  // vertices: every corner meets three faces at right angles
  const double q = M_PI / 2;
  for (int k = 0; k < 4; ++k) {
    const double sx = (k == 1 || k == 2) ? 1.0 : -1.0, sy = k >= 2 ? 1.0 : -1.0;
    b.vert(bc[k], V{sx, sy, -1.0} * q);
    b.vert(tc[k], V{sx, sy, 1.0} * q);
  }
  double zmax = ztop;
  for (double z : zz) zmax = std::max(zmax, z);
  o->info = {5.0, double(p), double(n)};
  o->info.insert(o->info.end(), kn.begin(), kn.end());
  o->info.insert(o->info.end(), pw.begin(), pw.end());
  o->finish({-half, -half, zb}, {half, half, zmax}, 1e-9);
  return o;
}

// ------------------------------------------------------------- solids of revolution

// The parameter of the exact rational arc form along an angle (the trim domain's
// coordinate on spheres and tori, trim2.rs:1789-1807): n quarter-or-less arcs over
// [a0, a0 + span], each mapped to [k, k + 1] by the half-angle tangent.
double arc_param_over(double a, double a0, double span, int n) {
  const double TAU = 2.0 * M_PI;
  double q = span / n, x = a - a0;
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

// One element of a profile in the (rho, z) half-plane: a segment, or a circular arc
// about (rc, zc) of radius rad from angle psi0 to psi1 (angles from the +rho
// direction towards +z). A profile is a closed loop, counterclockwise with rho as the
// first coordinate, so the region it bounds is on its left. `face`: the element is
// a face of the solid (true), or not (an axis segment, or a hole's end cap lying in
// one of the plate's planes: then `ext` names that plane's face, or is NONE).
struct Prof {
  bool arc = false;
  double r0 = 0, z0 = 0, r1 = 0, z1 = 0;
  double rc = 0, zc = 0, rad = 0, psi0 = 0, psi1 = 0;
  bool face = true;
  uint32_t ext = 0xffffffffu;
  double ext_nz = 0;  // an end cap's plane: its outward normal's z (+-1)
};
Prof seg(double r0, double z0, double r1, double z1) {
  Prof p;
  p.r0 = r0, p.z0 = z0, p.r1 = r1, p.z1 = z1;
  p.face = !(r0 == 0 && r1 == 0);
  return p;
}
Prof arc(double rc, double zc, double rad, double psi0, double psi1) {
  Prof p;
  p.arc = true;
  p.rc = rc, p.zc = zc, p.rad = rad, p.psi0 = psi0, p.psi1 = psi1;
  p.r0 = rc + rad * std::cos(psi0), p.z0 = zc + rad * std::sin(psi0);
  p.r1 = rc + rad * std::cos(psi1), p.z1 = zc + rad * std::sin(psi1);
  if (std::fabs(p.r0) < 1e-12) p.r0 = 0;
  if (std::fabs(p.r1) < 1e-12) p.r1 = 0;
  return p;
}

// The faces and circle edges of a profile revolved about the vertical axis through
// (cx, cy). hole: the region the profile bounds is empty (a hole in a plate), so the
// solid's outward normal points into it. The faces each element makes:
//   horizontal segment: a plane (disc or annulus), frame z the outward normal;
//   vertical segment: a cylinder;  slanted segment: a cone, its frame's z along the
//   axis in the direction its radius grows (semi-angle in (0, 90) degrees);
//   arc about a point of the axis: a sphere;  arc about a point off it: a ring torus.
// Analytic trim domains as in the part model: (quarter-arc parameter of the angle
// about the axis, height | latitude arc parameter | tube arc parameter).
void revolve(Builder& b, double cx, double cy, std::vector<Prof> prof, bool hole, int split = 1) {
  // a trim line u = 0 .. 4 at v, as `split` pieces
  auto tline = [&](std::vector<Piece>& pcs, double v) { add_line(pcs, 0, v, 4, v, split); };
  V X{1, 0, 0}, Z{0, 0, 1};
  const size_t n = prof.size();
  std::vector<uint32_t> fid(n, 0xffffffffu);
  const double flip = hole ? -1.0 : 1.0;
  for (size_t i = 0; i < n; ++i) {
    Prof& e = prof[i];
    if (!e.face) {
      fid[i] = e.ext;
      continue;
    }
    V o{cx, cy, 0};
    if (e.arc) {
      const bool ccw = e.psi1 > e.psi0;
      double sense = (ccw ? 1.0 : -1.0) * flip;
      if (e.rc == 0) {
        // sphere: latitude arc parameter v in [0, 2]
        double v0 = arc_param_over(e.psi0, -0.5 * M_PI, M_PI, 2), v1 = arc_param_over(e.psi1, -0.5 * M_PI, M_PI, 2);
        double lo = std::min(v0, v1), hi = std::max(v0, v1);
        std::vector<Piece> pcs;
        if (lo > 1e-12) tline(pcs, lo);
        if (hi < 2 - 1e-12) tline(pcs, hi);
        fid[i] = b.analytic(ICAM_SPHERE, V{cx, cy, e.zc}, Z, X, e.rad, 0, sense, 0, -1, 4, 3, 4, 0, pcs, 2.0, 0.5 * (lo + hi), true);
      } else if (std::fabs(e.rc) > e.rad) {
        // ring torus: tube arc parameter over the whole meridian circle, periodic
        double pm = 0.5 * (e.psi0 + e.psi1);
        double v0 = arc_param_over(e.psi0, 0, 2 * M_PI, 4), v1 = arc_param_over(e.psi1, 0, 2 * M_PI, 4);
        double vm = arc_param_over(pm, 0, 2 * M_PI, 4);
        std::vector<Piece> pcs;
        tline(pcs, v0), tline(pcs, v1);
        fid[i] = b.analytic(ICAM_TORUS, V{cx, cy, e.zc}, Z, X, e.rc, e.rad, sense, 0, 0, 4, 4, 4, 4, pcs, 2.0, vm, true);
      } else {
        // spindle torus (the arc crosses the axis: a lemon when its centre lies across
        // the axis, an apple when on this side): the tube angle over the sheet on the
        // axis's own side, [-pp, pp] with pp = acos(-major / minor), in narc quarter arcs
        // (trim2.rs:1807-1818), not periodic; no trim line where the arc meets the axis
        const double pp = std::acos(std::min(std::max(-e.rc / e.rad, -1.0), 1.0));
        const int narc = int(std::max(std::ceil(2 * pp / (0.5 * M_PI) - 1e-9), 1.0));
        double v0 = arc_param_over(e.psi0, -pp, 2 * pp, narc), v1 = arc_param_over(e.psi1, -pp, 2 * pp, narc);
        double vm = arc_param_over(0.5 * (e.psi0 + e.psi1), -pp, 2 * pp, narc);
        std::vector<Piece> pcs;
        if (e.r0 > 0) tline(pcs, v0);
        if (e.r1 > 0) tline(pcs, v1);
        fid[i] = b.analytic(ICAM_TORUS, V{cx, cy, e.zc}, Z, X, e.rc, e.rad, sense, 0, 0, 4, double(narc), 4, 0, pcs, 2.0, vm, true);
        // the tip on the axis: a vertex loop of the face, as STEP writes it, and a vertex
        // whose pseudo-normal is the face's normal there weighted by the full turn
        for (int end = 0; end < 2; ++end) {
          if ((end ? e.r1 : e.r0) > 0) continue;
          const double a = end ? e.psi1 : e.psi0, s = e.psi1 > e.psi0 ? 1.0 : -1.0;
          V tip{cx, cy, end ? e.z1 : e.z0};
          icam_face_desc& f = b.o->faces.back();
          f.n_points = 1;
          f.points = b.o->keep({tip.x, tip.y, tip.z});
          b.vert(tip, Z * (s * std::sin(a) * flip * 2.0 * M_PI));
        }
      }
      continue;
    }
    const double dr = e.r1 - e.r0, dz = e.z1 - e.z0;
    // outward normal of the region (its right side) in (rho, z)
    const double on_r = dz, on_z = -dr;
    if (dz == 0) {
      // plane at z0: frame z the solid's outward normal
      V zn = Z * ((on_z > 0 ? 1.0 : -1.0) * flip);
      double ra = std::min(e.r0, e.r1), rb = std::max(e.r0, e.r1);
      std::vector<Piece> pcs;
      add_circle(pcs, 0, 0, rb);
      if (ra > 0) add_circle(pcs, 0, 0, ra);
      fid[i] = b.plane(V{cx, cy, e.z0}, zn, X, 1.0, -rb - 1, -rb - 1, rb + 1, rb + 1, pcs, 0.5 * (ra + rb), 0.0123, true);
    } else if (dr == 0) {
      double lo = std::min(e.z0, e.z1), hi = std::max(e.z0, e.z1);
      std::vector<Piece> pcs;
      tline(pcs, lo), tline(pcs, hi);
      double sense = (on_r > 0 ? 1.0 : -1.0) * flip;
      fid[i] = b.analytic(ICAM_CYLINDER, o, Z, X, e.r0, 0, sense, 0, lo - 1, 4, hi + 1, 4, 0, pcs, 2.0, 0.5 * (lo + hi), true);
    } else {
      // cone from its narrow end A along zdir
      const bool a_is_0 = e.r0 < e.r1;
      const double ra = a_is_0 ? e.r0 : e.r1, za = a_is_0 ? e.z0 : e.z1, zb = a_is_0 ? e.z1 : e.z0;
      const double zdir = zb > za ? 1.0 : -1.0, len_h = std::fabs(zb - za);
      const double ang = std::atan(std::fabs(dr) / std::fabs(dz));
      // the surface's own normal in (rho, z): (cos a, -zdir sin a)
      const double own_r = std::cos(ang), own_z = -zdir * std::sin(ang);
      double sense = ((own_r * on_r + own_z * on_z) > 0 ? 1.0 : -1.0) * flip;
      std::vector<Piece> pcs;
      if (ra > 0) tline(pcs, 0);
      tline(pcs, len_h);
      fid[i] = b.analytic(ICAM_CONE, V{cx, cy, za}, Z * zdir, X, ra, ang, sense, 0, -1, 4, len_h + 1, 4, 0, pcs, 2.0, 0.5 * len_h, true);
      if (ra == 0) {
        // the apex: a vertex loop of the face, as STEP writes it, and a vertex whose
        // pseudo-normal is the cone's axis weighted by the full turn (solid3.rs:838-846)
        V apex{cx, cy, za};
        icam_face_desc& f = b.o->faces.back();
        f.n_points = 1;
        f.points = b.o->keep({apex.x, apex.y, apex.z});
        b.vert(apex, Z * (zdir * -std::sin(ang) * sense * 2.0 * M_PI));
      }
    }
  }
  // the solid's outward normal (rho, z) of element e at its start (at_end false) or end
  auto normal = [&](const Prof& e, bool at_end) {
    double nr, nz;
    if (!e.face) return std::array<double, 2>{0.0, e.ext_nz};
    if (e.arc) {
      double a = at_end ? e.psi1 : e.psi0, s = e.psi1 > e.psi0 ? 1.0 : -1.0;
      nr = s * std::cos(a), nz = s * std::sin(a);
    } else {
      double dr = e.r1 - e.r0, dz = e.z1 - e.z0, l = std::hypot(dr, dz);
      nr = dz / l, nz = -dr / l;
    }
    return std::array<double, 2>{nr * flip, nz * flip};
  };
  // circle edges where consecutive elements with faces meet off the axis, each with
  // the seam vertex STEP gives a closed edge (at angle 0), its pseudo-normal the
  // faces' normals weighted by their angle there, pi each (solid3.rs:797-842)
  for (size_t i = 0; i < n; ++i) {
    const Prof &e = prof[i], &f = prof[(i + 1) % n];
    uint32_t a = fid[i], c = fid[(i + 1) % n];
    if (a == 0xffffffffu || c == 0xffffffffu || (!e.face && !f.face)) continue;
    if (e.r1 <= 0) continue;
    b.circle(V{cx, cy, e.z1}, Z, X, e.r1, a, c);
    auto na = normal(e, true), nc = normal(f, false);
    b.vert(V{cx + e.r1, cy, e.z1}, V{na[0] + nc[0], 0, na[1] + nc[1]} * M_PI);
  }
}

// A solid of revolution about the z axis from its profile; the box from the profile's
// extremes.
void describe(std::vector<double>& info, const std::vector<Prof>& prof) {
  info.push_back(double(prof.size()));
  for (auto& e : prof) info.insert(info.end(), {e.arc ? 1.0 : 0.0, e.r0, e.z0, e.r1, e.z1, e.rc, e.zc, e.rad, e.psi0, e.psi1, e.face ? 1.0 : 0.0});
}

Owned* make_revolved(const std::vector<Prof>& prof) {
  Owned* o = new Owned();
  Builder b{o};
  revolve(b, 0, 0, prof, false);
  o->info = {1.0};
  describe(o->info, prof);
  double rmax = 0, zlo = 1e300, zhi = -1e300;
  for (auto& e : prof) {
    rmax = std::max({rmax, e.r0, e.r1});
    zlo = std::min({zlo, e.z0, e.z1}), zhi = std::max({zhi, e.z0, e.z1});
    if (e.arc) {
      // extremes the arc passes through
      for (int k = -4; k <= 4; ++k) {
        double a = 0.5 * M_PI * k;
        double lo = std::min(e.psi0, e.psi1), hi = std::max(e.psi0, e.psi1);
        if (a > lo && a < hi) {
          rmax = std::max(rmax, e.rc + e.rad * std::cos(a));
          zlo = std::min(zlo, e.zc + e.rad * std::sin(a)), zhi = std::max(zhi, e.zc + e.rad * std::sin(a));
        }
      }
    }
  }
  o->finish({-rmax, -rmax, zlo}, {rmax, rmax, zhi}, 1e-9);
  return o;
}

// A cylinder of radius R and height L on a disc, capped by a hemisphere written as a
// rational biquadratic B-spline face (the nine-point circle along u, a quarter circle
// along v, weights 1 and 1/sqrt 2): its v = 1 side collapses to the pole on the axis,
// which no vertex marks, so the pole answers as a point with the fan of the surface's
// normals around it (solid3.rs:1609-1625, 1955-1986). The equator is a spline edge of
// four quadratic pieces mapped onto the dome's u. The closed form is the capsule
// profile's (a quarter circle about a point of the axis).
Owned* make_dome(double R, double L) {
  Owned* o = new Owned();
  Builder b{o};
  Prof cap = arc(0, L, R, 0, 0.5 * M_PI);
  cap.face = false;  // made below as the B-spline face
  revolve(b, 0, 0, {seg(0, 0, R, 0), seg(R, 0, R, L), cap, seg(0, L + R, 0, 0)}, false);
  const uint32_t f_cyl = 1;  // the profile's faces in order: the disc, the cylinder
  const double s = std::sqrt(0.5);
  const double ex[9] = {1, 1, 0, -1, -1, -1, 0, 1, 1}, ey[9] = {0, 1, 1, 1, 0, -1, -1, -1, 0}, wu[9] = {1, s, 1, s, 1, s, 1, s, 1};
  const double rj[3] = {1, 1, 0}, zj[3] = {L, L + R, L + R}, wv[3] = {1, s, 1};
  std::vector<double> pw(9 * 3 * 4);
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 3; ++j) {
      const double w = wu[i] * wv[j];
      double* q = &pw[(i * 3 + j) * 4];
      q[0] = R * ex[i] * rj[j] * w, q[1] = R * ey[i] * rj[j] * w, q[2] = zj[j] * w, q[3] = w;
    }
  const std::vector<double> ku = {0, 0, 0, 0.25, 0.25, 0.5, 0.5, 0.75, 0.75, 1, 1, 1};
  icam_face_desc f;
  std::memset(&f, 0, sizeof(f));
  f.kind = ICAM_NURBS;
  f.sense = 1.0;  // S_u x S_v points out (along +x at u = 0 on the equator)
  f.du = f.dv = 2;
  f.nu = 9, f.nv = 3;
  f.pw = o->keep(pw);
  f.ku = o->keep(ku);
  f.kv = o->keep({0, 0, 0, 1, 1, 1});
  b.set_trim(f, 0, 0, 1, 1, 0, 0, {}, 0.5, 0.5, true);
  o->faces.push_back(f);
  const uint32_t f_dome = uint32_t(o->faces.size() - 1);
  // the equator: the v = 0 row as four quadratic pieces, u spans of a quarter each
  std::vector<double> row;
  for (int i = 0; i < 9; ++i) row.insert(row.end(), {R * ex[i] * wu[i], R * ey[i] * wu[i], L * wu[i], wu[i]});
  std::vector<double> uv;
  for (int m = 0; m < 4; ++m) uv.insert(uv.end(), {0.25 * m, 0.0, 0.25 * (m + 1), 0.0, NaN, NaN, NaN, NaN});
  b.spline(2, decompose(2, ku, row, 4), f_dome, f_cyl, uv);
  // its seam vertex, both faces' normals (+x) weighted by pi
  b.vert(V{R, 0, L}, V{2 * M_PI, 0, 0});
  o->info = {1.0};
  describe(o->info, {seg(0, 0, R, 0), seg(R, 0, R, L), arc(0, L, R, 0, 0.5 * M_PI), seg(0, L + R, 0, 0)});
  o->finish({-R, -R, 0}, {R, R, L + R}, 1e-9);
  return o;
}

// A cylinder of radius R on a disc at z = 0, cut at the top by the plane
// z = L + tan(a) x: the cut is an ellipse edge (Edge3::Ellipse, curve3.rs:494) of
// semi-axes R / cos(a) along the plane's x and R along y, between the plane and the
// cylinder. The cylinder's loops in its parameters (quarter-arc parameter, height) are
// the bottom line and the cut's curve as 256 cubic pieces (Hermite in the arc
// parameter, within 1e-8 mm); the top plane's loop is the ellipse as 64 cubic pieces.
// Closed forms: the solid is convex (info: 4, R, L, tan a).
Owned* make_oblique(double R, double L, double ta) {
  Owned* o = new Owned();
  Builder b{o};
  const V X{1, 0, 0}, Z{0, 0, 1};
  const double ca = 1.0 / std::sqrt(1.0 + ta * ta), sa = ta * ca, ea = R / ca;
  const V xt{ca, 0, sa}, zt{-sa, 0, ca};  // the cut plane's x axis and normal (its y is Y)
  // the cylinder: the cut's height over the quarter-arc parameter (trim2.rs:1789-1807):
  // on quarter k, u = k + (tan(b / 2) / tan(pi / 8) + 1) / 2 with b the angle from the
  // quarter's middle
  std::vector<Piece> band = {line_piece(0, 0, 4, 0)};
  const double t8 = std::tan(M_PI / 8);
  auto beta = [&](int k, double u) { return 2.0 * std::atan((2.0 * (u - k) - 1.0) * t8); };
  auto theta = [&](int k, double u) { return 0.5 * M_PI * k + 0.25 * M_PI + beta(k, u); };
  auto height = [&](double th) { return L + R * ta * std::cos(th); };
  auto slope = [&](int k, double u) {  // dv / du
    const double c = std::cos(0.5 * beta(k, u));
    return -R * ta * std::sin(theta(k, u)) * 4.0 * t8 * c * c;
  };
  for (int k = 0; k < 4; ++k)
    for (int j = 0; j < 64; ++j) {
      const double ua = k + j / 64.0, ub = k + (j + 1) / 64.0, du = ub - ua;
      const double va = height(theta(k, ua)), vb = height(theta(k, ub));
      add_monotone(band, {ua, va, ua + du / 3, va + du / 3 * slope(k, ua), ub - du / 3, vb - du / 3 * slope(k, ub), ub, vb});
    }
  const double top = L + R * std::fabs(ta);
  const uint32_t cyl = b.analytic(ICAM_CYLINDER, {0, 0, 0}, Z, X, R, 0, 1.0, 0, -1, 4, top + 1, 4, 0, band, 2.0, 0.5 * (L - R * std::fabs(ta)), true);
  std::vector<Piece> disc;
  add_circle(disc, 0, 0, R);
  const uint32_t bot = b.plane({0, 0, 0}, Z, X, -1.0, -R - 1, -R - 1, R + 1, R + 1, disc, -R - 0.5, 0.0123, false);
  std::vector<Piece> unit_circle, ell;
  add_circle(unit_circle, 0, 0, 1.0);
  for (const Piece& c : unit_circle) ell.push_back({ea * c[0], R * c[1], ea * c[2], R * c[3], ea * c[4], R * c[5], ea * c[6], R * c[7]});
  const uint32_t cut = b.plane({0, 0, L}, zt, xt, 1.0, -ea - 1, -R - 1, ea + 1, R + 1, ell, -ea - 0.5, 0.0123, false);
  b.circle({0, 0, 0}, Z, X, R, bot, cyl);
  icam_edge_desc e;
  std::memset(&e, 0, sizeof(e));
  e.kind = ICAM_ELLIPSE;
  put3(e.o, V{0, 0, L}), put3(e.x, xt), put3(e.z, zt);
  e.r = ea, e.r2 = R, e.t0 = 0.0, e.sweep = 2.0 * M_PI;
  e.n_pieces = 1;
  e.faces[0] = cut, e.faces[1] = cyl;
  o->edges.push_back(e);
  // the seam vertices (angle 0), each between two faces at angle pi
  b.vert({R, 0, 0}, V{1, 0, -1} * M_PI);
  b.vert({R, 0, L + R * ta}, (X + zt) * M_PI);
  o->info = {4.0, R, L, ta};
  o->finish({-R, -R, 0}, {R, R, top}, 1e-9);
  return o;
}

// Profiles of turned parts, each with g grooves (concave tori) cut into its
// cylinder; sizes in mm from the seed.
//   0 capsule: hemisphere, cylinder, hemisphere (sphere poles on the axis)
//   1 rod: disc, cylinder, cone narrowing upward, hemisphere cap
//   2 top: cone from its apex on the axis (a cone tip), cylinder, convex torus
//     round, top disc
//   3 washer: annulus, outer cylinder with rounded top edge, inner cylinder
//   4 bullet: disc, cylinder, a lemon nose (a spindle torus whose tube circle's centre
//     lies across the axis) meeting the axis in a tip
//   5 dome: disc, cylinder, an apple top (a spindle torus centred on this side of the
//     axis, its arc running over the top and down into a dimple on the axis)
std::vector<Prof> turned_profile(int type, int grooves, Rng& rng) {
  const double R = 6 + 14 * rng.next(), L = 2 * R + 40 * rng.next();
  std::vector<Prof> p;
  // the outer cylinder from z0 to z1 at radius R with grooves, upward
  auto wall = [&](double z0, double z1) {
    double len_ = z1 - z0;
    int g = grooves;
    while (g > 0 && len_ / (g + 1) < 4.0) --g;
    double step = len_ / (g + 1), gr = std::min(0.3 * step, 0.25 * R);
    double z = z0;
    for (int k = 1; k <= g; ++k) {
      double zg = z0 + step * k;
      p.push_back(seg(R, z, R, zg - gr));
      // a groove: the torus about (R, zg) of radius gr, cut into the material (clockwise)
      p.push_back(arc(R, zg, gr, 1.5 * M_PI, 0.5 * M_PI));
      z = zg + gr;
    }
    p.push_back(seg(R, z, R, z1));
  };
  if (type == 0) {
    p.push_back(arc(0, 0, R, -0.5 * M_PI, 0));
    wall(0, L);
    p.push_back(arc(0, L, R, 0, 0.5 * M_PI));
    p.push_back(seg(0, L + R, 0, -R));
  } else if (type == 1) {
    double h1 = 0.6 * L, r2 = (0.35 + 0.3 * rng.next()) * R, h2 = h1 + (R - r2) * (0.8 + rng.next());
    p.push_back(seg(0, 0, R, 0));
    wall(0, h1);
    p.push_back(seg(R, h1, r2, h2));
    p.push_back(arc(0, h2, r2, 0, 0.5 * M_PI));
    p.push_back(seg(0, h2 + r2, 0, 0));
  } else if (type == 2) {
    double hc = R * (0.6 + 0.8 * rng.next()), f = (0.15 + 0.2 * rng.next()) * R;
    p.push_back(seg(0, 0, R, hc));
    wall(hc, hc + L * 0.5);
    double zt = hc + L * 0.5;
    p.push_back(arc(R - f, zt, f, 0, 0.5 * M_PI));
    p.push_back(seg(R - f, zt + f, 0, zt + f));
    p.push_back(seg(0, zt + f, 0, 0));
  } else if (type == 4) {
    const double d = (0.2 + 0.6 * rng.next()) * R, rad = R + d, pe = std::acos(d / rad);
    p.push_back(seg(0, 0, R, 0));
    wall(0, L);
    p.push_back(arc(-d, L, rad, 0, pe));
    p.push_back(seg(0, L + rad * std::sin(pe), 0, 0));
  } else if (type == 5) {
    const double d = (0.15 + 0.3 * rng.next()) * R, rad = R - d, pe = std::acos(-d / rad);
    p.push_back(seg(0, 0, R, 0));
    wall(0, L);
    p.push_back(arc(d, L, rad, 0, pe));
    p.push_back(seg(0, L + rad * std::sin(pe), 0, 0));
  } else {
    double ri = (0.3 + 0.3 * rng.next()) * R, H = 0.3 * L + 4, f = std::min(0.3 * (R - ri), 0.3 * H);
    p.push_back(seg(ri, 0, R, 0));
    wall(0, H - f);
    p.push_back(arc(R - f, H - f, f, 0, 0.5 * M_PI));
    p.push_back(seg(R - f, H, ri, H));
    p.push_back(seg(ri, H, ri, 0));
  }
  return p;
}

// A plate W x D x T (centred at the origin) with k vertical through holes on a jittered
// grid. A hole of radius r has a plain top edge, a torus fillet of radius f (the
// fillet's circle on the top face at r + f), or a 45-degree countersink of depth c (a
// cone from r at T/2 - c to r + c at T/2); every bottom edge is plain. treat[i]: 0
// plain, 1 fillet, 2 countersink.
struct Hole {
  double x, y, r, e;
  int treat;
};
Owned* make_plate(double W, double D, double T, const std::vector<Hole>& holes) {
  Owned* o = new Owned();
  Builder b{o};
  V X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1};
  const double ax = 0.5 * W, ay = 0.5 * D, az = 0.5 * T;
  auto rect = [](double u0, double v0, double u1, double v1) {
    std::vector<Piece> p;
    add_polyline(p, {{u0, v0}, {u1, v0}, {u1, v1}, {u0, v1}}, true);
    return p;
  };
  std::vector<Piece> top = rect(-ax, -ay, ax, ay), bot = rect(-ax, -ay, ax, ay);
  for (auto& h : holes) add_circle(top, h.x, h.y, h.r + h.e), add_circle(bot, h.x, -h.y, h.r);
  uint32_t f_top = b.plane({0, 0, az}, Z, X, 1.0, -ax - 1, -ay - 1, ax + 1, ay + 1, top, -ax - 0.5, 0.0123, false);
  uint32_t f_bot = b.plane({0, 0, -az}, Z * -1.0, X, 1.0, -ax - 1, -ay - 1, ax + 1, ay + 1, bot, -ax - 0.5, 0.0123, false);
  uint32_t f_xp = b.plane({ax, 0, 0}, X, Y, 1.0, -ay - 1, -az - 1, ay + 1, az + 1, rect(-ay, -az, ay, az), -ay - 0.5, 0.0123, false);
  uint32_t f_xm = b.plane({-ax, 0, 0}, X * -1.0, Y, 1.0, -ay - 1, -az - 1, ay + 1, az + 1, rect(-ay, -az, ay, az), -ay - 0.5, 0.0123, false);
  uint32_t f_yp = b.plane({0, ay, 0}, Y, X, 1.0, -ax - 1, -az - 1, ax + 1, az + 1, rect(-ax, -az, ax, az), -ax - 0.5, 0.0123, false);
  uint32_t f_ym = b.plane({0, -ay, 0}, Y * -1.0, X, 1.0, -ax - 1, -az - 1, ax + 1, az + 1, rect(-ax, -az, ax, az), -ax - 0.5, 0.0123, false);
  V c[2][2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) c[i][j][k] = V{i ? ax : -ax, j ? ay : -ay, k ? az : -az};
  for (int k = 0; k < 2; ++k) {
    uint32_t f = k ? f_top : f_bot;
    b.line(c[0][0][k], c[1][0][k], f, f_ym);
    b.line(c[1][0][k], c[1][1][k], f, f_xp);
    b.line(c[1][1][k], c[0][1][k], f, f_yp);
    b.line(c[0][1][k], c[0][0][k], f, f_xm);
  }
  b.line(c[0][0][0], c[0][0][1], f_xm, f_ym);
  b.line(c[1][0][0], c[1][0][1], f_xp, f_ym);
  b.line(c[1][1][0], c[1][1][1], f_xp, f_yp);
  b.line(c[0][1][0], c[0][1][1], f_xm, f_yp);
  double q = M_PI / 2;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) b.vert(c[i][j][k], V{i ? 1.0 : -1.0, j ? 1.0 : -1.0, k ? 1.0 : -1.0} * q);
  o->info = {2.0, W, D, T, double(holes.size())};
  // each hole's region in its meridian half-plane, counterclockwise: bottom cap, wall,
  // the top treatment, top cap, down the axis
  for (auto& h : holes) {
    std::vector<Prof> p;
    Prof cap_b = seg(0, -az, h.r, -az);
    cap_b.face = false, cap_b.ext = f_bot, cap_b.ext_nz = -1;
    p.push_back(cap_b);
    double zw = az - (h.treat ? h.e : 0.0);
    p.push_back(seg(h.r, -az, h.r, zw));
    if (h.treat == 1) p.push_back(arc(h.r + h.e, zw, h.e, M_PI, 0.5 * M_PI));
    if (h.treat == 2) p.push_back(seg(h.r, zw, h.r + h.e, az));
    Prof cap_t = seg(h.r + h.e, az, 0, az);
    cap_t.face = false, cap_t.ext = f_top, cap_t.ext_nz = 1;
    p.push_back(cap_t);
    Prof axis = seg(0, az, 0, -az);
    p.push_back(axis);
    o->info.push_back(h.x), o->info.push_back(h.y);
    describe(o->info, p);
    revolve(b, h.x, h.y, p, true);
  }
  o->finish({-ax, -ay, -az}, {ax, ay, az}, 1e-9);
  return o;
}

// k holes on a jittered grid over the plate, radii and treatments from the seed, each
// hole's top circle clear of the plate's sides and of its neighbours.
std::vector<Hole> plate_holes(double W, double D, double T, int k, Rng& rng) {
  int nx = std::max(1, int(std::ceil(std::sqrt(k * W / D)))), ny = (k + nx - 1) / nx;
  double cw = W / nx, cd = D / ny, cell = std::min(cw, cd);
  std::vector<Hole> out;
  for (int i = 0; i < k; ++i) {
    int gx = i % nx, gy = i / nx;
    Hole h;
    h.treat = int(rng.next() * 3.0) % 3;
    h.r = cell * (0.12 + 0.12 * rng.next());
    h.e = h.treat ? std::min(cell * (0.04 + 0.06 * rng.next()), 0.3 * T) : 0.0;
    double room = 0.5 * cell - (h.r + h.e) - 0.05 * cell;
    h.x = -0.5 * W + cw * (gx + 0.5) + room * (2 * rng.next() - 1) * 0.5;
    h.y = -0.5 * D + cd * (gy + 0.5) + room * (2 * rng.next() - 1) * 0.5;
    out.push_back(h);
  }
  return out;
}

// ------------------------------------------------------------- C3: the trimmed plate

// A clamped cubic B-spline of 8 control points (5 spans, knots 0 0 0 0 .2 .4 .6 .8 1 1
// 1 1): its 5 Bezier pieces (points of dimension N, flat).
std::vector<std::vector<double>> five_spans(const std::vector<double>& ctrl, int N) {
  std::vector<double> kn = {0, 0, 0, 0, 0.2, 0.4, 0.6, 0.8, 1, 1, 1, 1};
  return decompose(3, kn, ctrl, N);
}

// A plate 200 x 150 x 25 (centred) with 24 pockets and 12 bosses on a jittered 6 x 6
// grid. Each pocket or boss has an outline: a closed curve of F clamped cubic B-spline
// segments of 5 spans each (20 to 60 spans in all), 8 control points a segment from a
// radially perturbed rounded rectangle, consecutive segments meeting with equal
// tangents (the shared control point halfway between its neighbours). The outline is
// a trim loop of the plate's top face and of the pocket's floor or the boss's top;
// each segment is the bottom and top boundary of a wall face: a rational bicubic
// B-spline with an 8 x 8 net, the outline's points along u, heights along v at the
// Greville abscissae, weights varying along v only (its end rows weight 1), so the
// wall is the vertical ruled surface over the outline and its boundary curves are
// the outline's polynomial pieces.
Owned* make_c3(uint64_t seed) {
  Owned* o = new Owned();
  Builder b{o};
  Rng rng{seed ^ 0x3C3C3C3C3C3C3C3Cull};
  for (int k = 0; k < 4; ++k) rng.next();
  V X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1};
  const double ax = 100, ay = 75, az = 12.5;
  // slots: 24 pockets and 12 bosses in random order
  std::vector<int> kind(36);
  for (int i = 0; i < 36; ++i) kind[i] = i < 24 ? 0 : 1;
  for (int i = 35; i > 0; --i) std::swap(kind[i], kind[int(rng.next() * (i + 1)) % (i + 1)]);
  struct Loop {
    int kind;            // 0 pocket, 1 boss
    double z0, z1;       // the wall's height range
    std::vector<std::vector<double>> seg;  // F segments: 8 control points (x, y)
    std::vector<Piece> pieces;             // the loop as monotone trim pieces
    double lo[2], hi[2];
  };
  std::vector<Loop> loops;
  const double cw = 2 * ax / 6, cd = 2 * ay / 6;
  for (int s = 0; s < 36; ++s) {
    Loop L;
    L.kind = kind[s];
    const double cx = -ax + cw * (s % 6 + 0.5) + (rng.next() - 0.5) * 3.0, cy = -ay + cd * (s / 6 + 0.5) + (rng.next() - 0.5) * 2.0;
    const double ha = 0.28 * cw + rng.next() * 0.06 * cw, hb = 0.26 * cd + rng.next() * 0.06 * cd, rc = 2.0 + 2.0 * rng.next();
    if (L.kind == 0) {
      L.z1 = az, L.z0 = az - (4.0 + 8.0 * rng.next());
    } else {
      L.z0 = az, L.z1 = az + (3.0 + 5.0 * rng.next());
    }
    const int F = 4 + int(rng.next() * 9.0) % 9;  // 4..12 segments: 20..60 spans
    const int M = 7 * F;
    // the rounded rectangle by arc length, counterclockwise from its right side's middle
    const double straight_x = 2 * (ha - rc), straight_y = 2 * (hb - rc), quarter = 0.5 * M_PI * rc;
    const double per = 2 * straight_x + 2 * straight_y + 4 * quarter;
    auto rr = [&](double t) {
      double l = t * per;
      // pieces: half right side up, top-right corner, top (leftward), top-left corner,
      // left (down), bottom-left corner, bottom (rightward), bottom-right corner, half
      // right side up
      const double segs[9] = {0.5 * straight_y, quarter, straight_x, quarter, straight_y, quarter, straight_x, quarter, 0.5 * straight_y};
      int k = 0;
      while (k < 8 && l > segs[k]) l -= segs[k], ++k;
      double x = 0, y = 0, nx = 1, ny = 0;
      const double ex = ha - rc, ey = hb - rc;
      switch (k) {
        case 0: x = ha, y = l, nx = 1, ny = 0; break;
        case 1: { double a = l / rc; x = ex + rc * std::cos(a), y = ey + rc * std::sin(a), nx = std::cos(a), ny = std::sin(a); break; }
        case 2: x = ex - l, y = hb, nx = 0, ny = 1; break;
        case 3: { double a = 0.5 * M_PI + l / rc; x = -ex + rc * std::cos(a), y = ey + rc * std::sin(a), nx = std::cos(a), ny = std::sin(a); break; }
        case 4: x = -ha, y = ey - l, nx = -1, ny = 0; break;
        case 5: { double a = M_PI + l / rc; x = -ex + rc * std::cos(a), y = -ey + rc * std::sin(a), nx = std::cos(a), ny = std::sin(a); break; }
        case 6: x = -ex + l, y = -hb, nx = 0, ny = -1; break;
        case 7: { double a = 1.5 * M_PI + l / rc; x = ex + rc * std::cos(a), y = -ey + rc * std::sin(a), nx = std::cos(a), ny = std::sin(a); break; }
        default: x = ha, y = -ey + l, nx = 1, ny = 0; break;
      }
      return std::array<double, 4>{x, y, nx, ny};
    };
    const double a1 = 0.5 + 0.4 * rng.next(), f1 = double(2 + int(rng.next() * 3.0)), p1 = 2 * M_PI * rng.next();
    const double a2 = 0.3 * rng.next(), f2 = double(5 + int(rng.next() * 4.0)), p2 = 2 * M_PI * rng.next();
    std::vector<std::array<double, 2>> pts(M);
    for (int i = 0; i < M; ++i) {
      const double t = double(i) / M;
      auto q = rr(t);
      const double d = a1 * std::sin(2 * M_PI * f1 * t + p1) + a2 * std::sin(2 * M_PI * f2 * t + p2);
      pts[i] = {cx + q[0] + d * q[2], cy + q[1] + d * q[3]};
    }
    // segment k: the joint, 6 points, the next joint; a joint halfway between its
    // neighbours
    auto joint = [&](int k) {
      const auto& a = pts[(7 * k + M - 1) % M];
      const auto& c = pts[(7 * k + 1) % M];
      return std::array<double, 2>{0.5 * (a[0] + c[0]), 0.5 * (a[1] + c[1])};
    };
    L.lo[0] = L.lo[1] = 1e300, L.hi[0] = L.hi[1] = -1e300;
    for (int k = 0; k < F; ++k) {
      std::vector<double> c;
      auto j0 = joint(k), j1 = joint((k + 1) % F);
      c.insert(c.end(), {j0[0], j0[1]});
      for (int m = 1; m <= 6; ++m) c.insert(c.end(), {pts[(7 * k + m) % M][0], pts[(7 * k + m) % M][1]});
      c.insert(c.end(), {j1[0], j1[1]});
      for (int m = 0; m < 8; ++m)
        for (int a = 0; a < 2; ++a) L.lo[a] = std::min(L.lo[a], c[2 * m + a]), L.hi[a] = std::max(L.hi[a], c[2 * m + a]);
      for (auto& pc : five_spans(c, 2)) add_monotone(L.pieces, to_cubic2(pc, 3));
      L.seg.push_back(c);
    }
    loops.push_back(L);
  }
  auto rect = [](double u0, double v0, double u1, double v1) {
    std::vector<Piece> p;
    add_polyline(p, {{u0, v0}, {u1, v0}, {u1, v1}, {u0, v1}}, true);
    return p;
  };
  // the plate
  std::vector<Piece> top = rect(-ax, -ay, ax, ay);
  for (auto& L : loops) top.insert(top.end(), L.pieces.begin(), L.pieces.end());
  uint32_t f_top = b.plane({0, 0, az}, Z, X, 1.0, -ax - 1, -ay - 1, ax + 1, ay + 1, top, -ax - 0.5, 0.0123, false);
  uint32_t f_bot = b.plane({0, 0, -az}, Z * -1.0, X, 1.0, -ax - 1, -ay - 1, ax + 1, ay + 1, rect(-ax, -ay, ax, ay), -ax - 0.5, 0.0123, false);
  uint32_t f_xp = b.plane({ax, 0, 0}, X, Y, 1.0, -ay - 1, -az - 1, ay + 1, az + 1, rect(-ay, -az, ay, az), -ay - 0.5, 0.0123, false);
  uint32_t f_xm = b.plane({-ax, 0, 0}, X * -1.0, Y, 1.0, -ay - 1, -az - 1, ay + 1, az + 1, rect(-ay, -az, ay, az), -ay - 0.5, 0.0123, false);
  uint32_t f_yp = b.plane({0, ay, 0}, Y, X, 1.0, -ax - 1, -az - 1, ax + 1, az + 1, rect(-ax, -az, ax, az), -ax - 0.5, 0.0123, false);
  uint32_t f_ym = b.plane({0, -ay, 0}, Y * -1.0, X, 1.0, -ax - 1, -az - 1, ax + 1, az + 1, rect(-ax, -az, ax, az), -ax - 0.5, 0.0123, false);
  V c[2][2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) c[i][j][k] = V{i ? ax : -ax, j ? ay : -ay, k ? az : -az};
  for (int k = 0; k < 2; ++k) {
    uint32_t f = k ? f_top : f_bot;
    b.line(c[0][0][k], c[1][0][k], f, f_ym);
    b.line(c[1][0][k], c[1][1][k], f, f_xp);
    b.line(c[1][1][k], c[0][1][k], f, f_yp);
    b.line(c[0][1][k], c[0][0][k], f, f_xm);
  }
  b.line(c[0][0][0], c[0][0][1], f_xm, f_ym);
  b.line(c[1][0][0], c[1][0][1], f_xp, f_ym);
  b.line(c[1][1][0], c[1][1][1], f_xp, f_yp);
  b.line(c[0][1][0], c[0][1][1], f_xm, f_yp);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) b.vert(c[i][j][k], V{i ? 1.0 : -1.0, j ? 1.0 : -1.0, k ? 1.0 : -1.0} * (M_PI / 2));
  // the v knots of a wall and their Greville abscissae
  const std::vector<double> kv = {0, 0, 0, 0, 0.2, 0.4, 0.6, 0.8, 1, 1, 1, 1};
  double gv[8];
  for (int j = 0; j < 8; ++j) gv[j] = (kv[j + 1] + kv[j + 2] + kv[j + 3]) / 3.0;
  o->info = {3.0, 2 * ax, 2 * ay, 2 * az, double(loops.size())};
  double zmax = az;
  const V nanv{NaN, NaN, NaN};
  (void)nanv;
  for (auto& L : loops) {
    const int F = int(L.seg.size());
    zmax = std::max(zmax, L.z1);
    o->info.insert(o->info.end(), {double(L.kind), L.z0, L.z1, double(F)});
    for (auto& sgm : L.seg) o->info.insert(o->info.end(), sgm.begin(), sgm.end());
    // the floor (pocket) or the top (boss): a plane facing up, trimmed by the outline
    const double zc = L.kind == 0 ? L.z0 : L.z1;
    uint32_t f_cap = b.plane({0, 0, zc}, Z, X, 1.0, L.lo[0] - 1, L.lo[1] - 1, L.hi[0] + 1, L.hi[1] + 1, L.pieces, L.lo[0] - 0.5, 0.5 * (L.lo[1] + L.hi[1]) + 0.0123,
                             false);
    // the rim on the plate's top face, and the far rim (on the cap)
    const uint32_t f_near = f_top;
    // walls: u along the segment, v up; the wall's normal Su x Sv points out of the
    // outline (counterclockwise), so a pocket's wall faces in (sense -1)
    std::vector<uint32_t> walls;
    for (int k = 0; k < F; ++k) {
      const auto& sgm = L.seg[k];
      std::vector<double> w(8);
      for (int j = 0; j < 8; ++j) w[j] = (j == 0 || j == 7) ? 1.0 : 0.8 + 0.4 * rng.next();
      std::vector<double> pw(8 * 8 * 4);
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
          double* q = &pw[(i * 8 + j) * 4];
          const double z = L.z0 + (L.z1 - L.z0) * gv[j];
          q[0] = sgm[2 * i] * w[j], q[1] = sgm[2 * i + 1] * w[j], q[2] = z * w[j], q[3] = w[j];
        }
      icam_face_desc f;
      std::memset(&f, 0, sizeof(f));
      f.kind = ICAM_NURBS;
      f.sense = L.kind == 0 ? -1.0 : 1.0;
      f.du = f.dv = 3;
      f.nu = f.nv = 8;
      f.pw = o->keep(pw);
      f.ku = o->keep(kv);
      f.kv = o->keep(kv);
      b.set_trim(f, 0, 0, 1, 1, 0, 0, {}, 0.5, 0.5, true);
      o->faces.push_back(f);
      walls.push_back(uint32_t(o->faces.size() - 1));
    }
    for (int k = 0; k < F; ++k) {
      const auto& sgm = L.seg[k];
      for (int end = 0; end < 2; ++end) {
        // end 0: the rim at z0 (v = 0), end 1: at z1 (v = 1)
        const double z = end ? L.z1 : L.z0, v = end ? 1.0 : 0.0;
        std::vector<double> c3;
        for (int i = 0; i < 8; ++i) c3.insert(c3.end(), {sgm[2 * i], sgm[2 * i + 1], z, 1.0});
        auto pcs = five_spans(c3, 4);
        std::vector<double> uv;
        for (int m = 0; m < 5; ++m) uv.insert(uv.end(), {0.2 * m, v, 0.2 * (m + 1), v, NaN, NaN, NaN, NaN});
        const bool on_top = (L.kind == 0) ? end == 1 : end == 0;  // the rim on the plate's top face
        b.spline(3, pcs, walls[k], on_top ? f_near : f_cap, uv);
      }
      // the vertical edge where segment k meets segment k + 1, and its two vertices
      const int k1 = (k + 1) % F;
      V pj{sgm[14], sgm[15], 0};
      const double tx = sgm[14] - sgm[12], ty = sgm[15] - sgm[13], tl = std::hypot(tx, ty);
      V nout{ty / tl, -tx / tl, 0};  // the tangent times z: out of the outline
      V nwall = nout * (L.kind == 0 ? -1.0 : 1.0);
      V a{pj.x, pj.y, L.z0}, e{pj.x, pj.y, L.z1};
      b.line(a, e, walls[k], walls[k1], o->keep({1, 0, 1, 1, 0, 0, 0, 1}));
      // each end: a plane at angle pi (the loop is smooth there), the two walls at pi / 2
      b.vert(a, (Z + nwall) * M_PI);
      b.vert(e, (Z + nwall) * M_PI);
    }
  }
  o->finish({-ax, -ay, -az}, {ax, ay, zmax}, 1e-9);
  return o;
}

// The batched workload's parts: part i of 64 from the seed. Face counts log-uniform in
// 6..400: turned parts (3..14 faces) and NURBS-top blocks (6 faces, nets 4 x 4 to
// 32 x 32) below 15, plates with holes above.
Owned* make_c4(uint64_t seed, int index) {
  Rng rng{(seed ^ (0x9E3779B97F4A7C15ull * uint64_t(index + 1))) | 1};
  for (int k = 0; k < 4; ++k) rng.next();
  const int faces = int(std::lround(std::exp(std::log(6.0) + (std::log(400.0) - std::log(6.0)) * rng.next())));
  if (faces <= 14) {
    if (rng.next() < 0.35) {
      int n = 4 + int(rng.next() * 29.0) % 29;
      double half = 15 + 45 * rng.next();
      return make_c2(rng.next_u64(), false, n, half, 0.04 * half + 1.0);
    }
    int type = int(rng.next() * 4.0) % 4;
    int base = type == 0 ? 3 : type == 1 ? 4 : type == 2 ? 4 : 5;
    int grooves = std::max(0, faces - base);
    return make_revolved(turned_profile(type, std::min(grooves, 8), rng));
  }
  // plates: 6 + k + (treated holes) faces; about two thirds of holes are treated
  int k = std::max(1, int(std::lround((faces - 6) / 1.67)));
  double W = 60 + 140 * rng.next(), D = W * (0.5 + 0.5 * rng.next()), T = 8 + 17 * rng.next();
  return make_plate(W, D, T, plate_holes(W, D, T, k, rng));
}

// ------------------------------------------------------------- C5: the large mixed solid

// A clamped uniform cubic B-spline of n control points (points of dimension N, flat;
// n - 3 spans): its Bezier pieces.
std::vector<std::vector<double>> uniform_spans(const std::vector<double>& ctrl, int N) {
  const int n = int(ctrl.size()) / N, spans = n - 3;
  std::vector<double> kn(4, 0.0);
  for (int k = 1; k < spans; ++k) kn.push_back(double(k) / spans);
  kn.insert(kn.end(), 4, 1.0);
  return decompose(3, kn, ctrl, N);
}

// A rectangle's loop, each side as k pieces.
std::vector<Piece> rect_loop(double u0, double v0, double u1, double v1, int k) {
  std::vector<Piece> p;
  add_line(p, u0, v0, u1, v0, k), add_line(p, u1, v0, u1, v1, k), add_line(p, u1, v1, u0, v1, k), add_line(p, u0, v1, u0, v0, k);
  return p;
}

// A plate 400 x 300 x 40 (centred) with 400 pockets and 264 through holes on a jittered
// 29 x 23 grid (three slots empty), 1,998 faces of which 800 are rational bicubic
// B-splines with 32 x 32 nets and 1,198 analytic, every face bounded by loops of 12 to
// 64 cubic spans.
//  - A pocket (3 to 9 mm deep) has an outline of two clamped cubic B-spline segments of
//    32 control points each (29 spans, 58 in the loop) from a radially perturbed rounded
//    rectangle, meeting with equal tangents; its floor is a plane trimmed by the outline;
//    each segment is a wall, a rational B-spline with a 32 x 32 net: the segment's points
//    along u, heights at the Greville abscissae along v, weights in [0.5, 2] from a
//    smooth profile along v only (the end rows 1), so the wall is the vertical ruled
//    surface over the outline, and its loop is its domain's border, three pieces a side.
//  - A hole of radius r has a 45-degree countersink of depth c at the bottom (a cone), its
//    wall (a cylinder) and a torus fillet of radius f at the top; each face's loops are
//    its two bounding circles in its parameters, eight pieces each.
//  - The plate's top and bottom are trimmed by their rectangle (four pieces a side), the
//    pockets' outlines and the holes' circles (64 pieces each); its sides by their
//    rectangles.
Owned* make_c5(uint64_t seed) {
  Owned* o = new Owned();
  Builder b{o};
  Rng rng{seed ^ 0x5C5C5C5C5C5C5C5Cull};
  for (int k = 0; k < 4; ++k) rng.next();
  V X{1, 0, 0}, Y{0, 1, 0}, Z{0, 0, 1};
  const double ax = 200, ay = 150, az = 20;
  const int NX = 29, NY = 23, NPOCK = 400, NHOLE = 264, NCTL = 32, F = 2;
  std::vector<int> kind(NX * NY);
  for (int i = 0; i < NX * NY; ++i) kind[i] = i < NPOCK ? 0 : i < NPOCK + NHOLE ? 1 : 2;
  for (int i = NX * NY - 1; i > 0; --i) std::swap(kind[i], kind[int(rng.next() * (i + 1)) % (i + 1)]);
  struct Pocket {
    double z0, z1;
    std::vector<std::vector<double>> seg;  // F segments: NCTL control points (x, y)
    std::vector<Piece> pieces;             // the outline as monotone trim pieces
    double lo[2], hi[2];
  };
  struct HoleC5 {
    double x, y, r, f, c;
  };
  std::vector<Pocket> pockets;
  std::vector<HoleC5> holes;
  const double cw = 2 * ax / NX, cd = 2 * ay / NY, cell = std::min(cw, cd);
  for (int s = 0; s < NX * NY; ++s) {
    const double cx = -ax + cw * (s % NX + 0.5) + (rng.next() - 0.5) * 0.1 * cw, cy = -ay + cd * (s / NX + 0.5) + (rng.next() - 0.5) * 0.1 * cd;
    if (kind[s] == 1) {
      HoleC5 h;
      h.x = cx, h.y = cy, h.r = cell * (0.12 + 0.1 * rng.next()), h.f = 0.5 + 0.5 * rng.next(), h.c = 0.6 + 0.6 * rng.next();
      holes.push_back(h);
      continue;
    }
    if (kind[s] != 0) continue;
    Pocket L;
    const double ha = (0.26 + 0.06 * rng.next()) * cw, hb = (0.26 + 0.06 * rng.next()) * cd, rc = 0.8 + 0.8 * rng.next();
    L.z1 = az, L.z0 = az - (3.0 + 6.0 * rng.next());
    const int M = (NCTL - 1) * F;
    const double straight_x = 2 * (ha - rc), straight_y = 2 * (hb - rc), quarter = 0.5 * M_PI * rc;
    const double per = 2 * straight_x + 2 * straight_y + 4 * quarter;
    auto rr = [&](double t) {
      double l = t * per;
      const double segs[9] = {0.5 * straight_y, quarter, straight_x, quarter, straight_y, quarter, straight_x, quarter, 0.5 * straight_y};
      int k = 0;
      while (k < 8 && l > segs[k]) l -= segs[k], ++k;
      double x = 0, y = 0, nx = 1, ny = 0;
      const double ex = ha - rc, ey = hb - rc;
      switch (k) {
        case 0: x = ha, y = l, nx = 1, ny = 0; break;
        case 1: { double a = l / rc; x = ex + rc * std::cos(a), y = ey + rc * std::sin(a), nx = std::cos(a), ny = std::sin(a); break; }
        case 2: x = ex - l, y = hb, nx = 0, ny = 1; break;
        case 3: { double a = 0.5 * M_PI + l / rc; x = -ex + rc * std::cos(a), y = ey + rc * std::sin(a), nx = std::cos(a), ny = std::sin(a); break; }
        case 4: x = -ha, y = ey - l, nx = -1, ny = 0; break;
        case 5: { double a = M_PI + l / rc; x = -ex + rc * std::cos(a), y = -ey + rc * std::sin(a), nx = std::cos(a), ny = std::sin(a); break; }
        case 6: x = -ex + l, y = -hb, nx = 0, ny = -1; break;
        case 7: { double a = 1.5 * M_PI + l / rc; x = ex + rc * std::cos(a), y = -ey + rc * std::sin(a), nx = std::cos(a), ny = std::sin(a); break; }
        default: x = ha, y = -ey + l, nx = 1, ny = 0; break;
      }
      return std::array<double, 4>{x, y, nx, ny};
    };
    const double a1 = 0.15 + 0.15 * rng.next(), f1 = double(2 + int(rng.next() * 3.0)), p1 = 2 * M_PI * rng.next();
    const double a2 = 0.1 * rng.next(), f2 = double(5 + int(rng.next() * 4.0)), p2 = 2 * M_PI * rng.next();
    std::vector<std::array<double, 2>> pts(M);
    for (int i = 0; i < M; ++i) {
      const double t = double(i) / M;
      auto q = rr(t);
      const double d = a1 * std::sin(2 * M_PI * f1 * t + p1) + a2 * std::sin(2 * M_PI * f2 * t + p2);
      pts[i] = {cx + q[0] + d * q[2], cy + q[1] + d * q[3]};
    }
    auto joint = [&](int k) {
      const auto& a = pts[((NCTL - 1) * k + M - 1) % M];
      const auto& c = pts[((NCTL - 1) * k + 1) % M];
      return std::array<double, 2>{0.5 * (a[0] + c[0]), 0.5 * (a[1] + c[1])};
    };
    L.lo[0] = L.lo[1] = 1e300, L.hi[0] = L.hi[1] = -1e300;
    for (int k = 0; k < F; ++k) {
      std::vector<double> c;
      auto j0 = joint(k), j1 = joint((k + 1) % F);
      c.insert(c.end(), {j0[0], j0[1]});
      for (int m = 1; m <= NCTL - 2; ++m) c.insert(c.end(), {pts[((NCTL - 1) * k + m) % M][0], pts[((NCTL - 1) * k + m) % M][1]});
      c.insert(c.end(), {j1[0], j1[1]});
      for (int m = 0; m < NCTL; ++m)
        for (int a = 0; a < 2; ++a) L.lo[a] = std::min(L.lo[a], c[2 * m + a]), L.hi[a] = std::max(L.hi[a], c[2 * m + a]);
      for (auto& pc : uniform_spans(c, 2)) add_monotone(L.pieces, to_cubic2(pc, 3));
      L.seg.push_back(c);
    }
    pockets.push_back(L);
  }
  // the plate
  std::vector<Piece> top = rect_loop(-ax, -ay, ax, ay, 4), bot = rect_loop(-ax, -ay, ax, ay, 4);
  for (auto& L : pockets) top.insert(top.end(), L.pieces.begin(), L.pieces.end());
  for (auto& h : holes) add_circle(top, h.x, h.y, h.r + h.f), add_circle(bot, h.x, -h.y, h.r + h.c);
  uint32_t f_top = b.plane({0, 0, az}, Z, X, 1.0, -ax - 1, -ay - 1, ax + 1, ay + 1, top, -ax - 0.5, 0.0123, false);
  uint32_t f_bot = b.plane({0, 0, -az}, Z * -1.0, X, 1.0, -ax - 1, -ay - 1, ax + 1, ay + 1, bot, -ax - 0.5, 0.0123, false);
  uint32_t f_xp = b.plane({ax, 0, 0}, X, Y, 1.0, -ay - 1, -az - 1, ay + 1, az + 1, rect_loop(-ay, -az, ay, az, 4), -ay - 0.5, 0.0123, false);
  uint32_t f_xm = b.plane({-ax, 0, 0}, X * -1.0, Y, 1.0, -ay - 1, -az - 1, ay + 1, az + 1, rect_loop(-ay, -az, ay, az, 4), -ay - 0.5, 0.0123, false);
  uint32_t f_yp = b.plane({0, ay, 0}, Y, X, 1.0, -ax - 1, -az - 1, ax + 1, az + 1, rect_loop(-ax, -az, ax, az, 4), -ax - 0.5, 0.0123, false);
  uint32_t f_ym = b.plane({0, -ay, 0}, Y * -1.0, X, 1.0, -ax - 1, -az - 1, ax + 1, az + 1, rect_loop(-ax, -az, ax, az, 4), -ax - 0.5, 0.0123, false);
  V c[2][2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) c[i][j][k] = V{i ? ax : -ax, j ? ay : -ay, k ? az : -az};
  for (int k = 0; k < 2; ++k) {
    uint32_t f = k ? f_top : f_bot;
    b.line(c[0][0][k], c[1][0][k], f, f_ym);
    b.line(c[1][0][k], c[1][1][k], f, f_xp);
    b.line(c[1][1][k], c[0][1][k], f, f_yp);
    b.line(c[0][1][k], c[0][0][k], f, f_xm);
  }
  b.line(c[0][0][0], c[0][0][1], f_xm, f_ym);
  b.line(c[1][0][0], c[1][0][1], f_xp, f_ym);
  b.line(c[1][1][0], c[1][1][1], f_xp, f_yp);
  b.line(c[0][1][0], c[0][1][1], f_xm, f_yp);
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) b.vert(c[i][j][k], V{i ? 1.0 : -1.0, j ? 1.0 : -1.0, k ? 1.0 : -1.0} * (M_PI / 2));
  o->info = {5.0, 2 * ax, 2 * ay, 2 * az, double(pockets.size()), double(holes.size())};
  // the walls' knots (29 spans) and their Greville abscissae along v
  std::vector<double> kn(4, 0.0);
  for (int k = 1; k < NCTL - 3; ++k) kn.push_back(double(k) / (NCTL - 3));
  kn.insert(kn.end(), 4, 1.0);
  std::vector<double> gv(NCTL);
  for (int j = 0; j < NCTL; ++j) gv[j] = (kn[j + 1] + kn[j + 2] + kn[j + 3]) / 3.0;
  const int nsp = NCTL - 3;
  for (auto& L : pockets) {
    o->info.insert(o->info.end(), {L.z0, L.z1, double(F)});
    for (auto& sgm : L.seg) o->info.insert(o->info.end(), sgm.begin(), sgm.end());
    uint32_t f_cap = b.plane({0, 0, L.z0}, Z, X, 1.0, L.lo[0] - 1, L.lo[1] - 1, L.hi[0] + 1, L.hi[1] + 1, L.pieces, L.lo[0] - 0.5,
                             0.5 * (L.lo[1] + L.hi[1]) + 0.0123, false);
    std::vector<uint32_t> walls;
    for (int k = 0; k < F; ++k) {
      const auto& sgm = L.seg[k];
      // the weights along v: 2^sin(2 pi (f v + phase)), a smooth profile through [0.5, 2]
      // (f from 0.5 to 1.5 turns over the wall), its end rows 1
      std::vector<double> w(NCTL);
      const double fq = 0.5 + rng.next(), ph = rng.next();
      for (int j = 0; j < NCTL; ++j) w[j] = (j == 0 || j == NCTL - 1) ? 1.0 : std::exp2(std::sin(2 * M_PI * (fq * gv[j] + ph)));
      std::vector<double> pw(size_t(NCTL) * NCTL * 4);
      for (int i = 0; i < NCTL; ++i)
        for (int j = 0; j < NCTL; ++j) {
          double* q = &pw[size_t(i * NCTL + j) * 4];
          const double z = L.z0 + (L.z1 - L.z0) * gv[j];
          q[0] = sgm[2 * i] * w[j], q[1] = sgm[2 * i + 1] * w[j], q[2] = z * w[j], q[3] = w[j];
        }
      icam_face_desc f;
      std::memset(&f, 0, sizeof(f));
      f.kind = ICAM_NURBS;
      f.sense = -1.0;  // Su x Sv points out of the outline: the pocket's wall faces in
      f.du = f.dv = 3;
      f.nu = f.nv = NCTL;
      f.pw = o->keep(pw);
      f.ku = o->keep(kn);
      f.kv = o->keep(kn);
      b.set_trim(f, -0.01, -0.01, 1.01, 1.01, 0, 0, rect_loop(0, 0, 1, 1, 3), 0.5, 0.5, true);
      o->faces.push_back(f);
      walls.push_back(uint32_t(o->faces.size() - 1));
    }
    for (int k = 0; k < F; ++k) {
      const auto& sgm = L.seg[k];
      for (int end = 0; end < 2; ++end) {
        // end 0: the rim on the floor (v = 0), end 1: on the plate's top (v = 1)
        const double z = end ? L.z1 : L.z0, v = end ? 1.0 : 0.0;
        std::vector<double> c3;
        for (int i = 0; i < NCTL; ++i) c3.insert(c3.end(), {sgm[2 * i], sgm[2 * i + 1], z, 1.0});
        auto pcs = uniform_spans(c3, 4);
        std::vector<double> uv;
        for (int m = 0; m < nsp; ++m) uv.insert(uv.end(), {double(m) / nsp, v, double(m + 1) / nsp, v, NaN, NaN, NaN, NaN});
        b.spline(3, pcs, walls[k], end ? f_top : f_cap, uv);
      }
      const int k1 = (k + 1) % F;
      const int e = 2 * (NCTL - 1);
      V pj{sgm[e], sgm[e + 1], 0};
      const double tx = sgm[e] - sgm[e - 2], ty = sgm[e + 1] - sgm[e - 1], tl = std::hypot(tx, ty);
      V nwall = V{ty / tl, -tx / tl, 0} * -1.0;
      V a{pj.x, pj.y, L.z0}, top_v{pj.x, pj.y, L.z1};
      b.line(a, top_v, walls[k], walls[k1], o->keep({1, 0, 1, 1, 0, 0, 0, 1}));
      b.vert(a, (Z + nwall) * M_PI);
      b.vert(top_v, (Z + nwall) * M_PI);
    }
  }
  for (auto& h : holes) {
    o->info.insert(o->info.end(), {h.x, h.y, h.r, h.f, h.c});
    std::vector<Prof> p;
    Prof cap_b = seg(0, -az, h.r + h.c, -az);
    cap_b.face = false, cap_b.ext = f_bot, cap_b.ext_nz = -1;
    p.push_back(cap_b);
    p.push_back(seg(h.r + h.c, -az, h.r, -az + h.c));
    p.push_back(seg(h.r, -az + h.c, h.r, az - h.f));
    p.push_back(arc(h.r + h.f, az - h.f, h.f, M_PI, 0.5 * M_PI));
    Prof cap_t = seg(h.r + h.f, az, 0, az);
    cap_t.face = false, cap_t.ext = f_top, cap_t.ext_nz = 1;
    p.push_back(cap_t);
    p.push_back(seg(0, az, 0, -az));
    revolve(b, h.x, h.y, p, true, 8);
  }
  o->finish({-ax, -ay, -az}, {ax, ay, az}, 1e-9);
  return o;
}

// ------------------------------------------------------------- points near a surface

// The part's surface point at (u, v) of face f (the part model's parameterizations).
// Quarter-arc parameters (cylinder, cone, sphere and torus angles) are turned back into
// angles by the inverse of arc_param_over.
double arc_angle(double u, double a0, double span, int n) {
  const double q = span / n;
  const double k = std::min(std::max(std::floor(u), 0.0), double(n - 1)), t = u - k;
  return a0 + q * k + 0.5 * q + 2.0 * std::atan((2.0 * t - 1.0) * std::tan(0.25 * q));
}
V nurbs_point(const icam_face_desc& f, double u, double v) {
  // de Boor in both directions on the homogeneous net
  auto span_of = [](const double* kn, int n, int p, double x) {
    int s = p;
    while (s < n - 1 && kn[s + 1] <= x) ++s;
    return s;
  };
  const int p = int(f.du), q = int(f.dv), nu = int(f.nu), nv = int(f.nv);
  u = std::min(std::max(u, f.ku[p]), f.ku[nu]), v = std::min(std::max(v, f.kv[q]), f.kv[nv]);
  const int su = span_of(f.ku, nu, p, u), sv = span_of(f.kv, nv, q, v);
  double row[8][4];
  for (int a = 0; a <= p; ++a) {
    const int i = su - p + a;
    double d[8][4];
    for (int b2 = 0; b2 <= q; ++b2)
      for (int k = 0; k < 4; ++k) d[b2][k] = f.pw[size_t((i * nv + sv - q + b2) * 4 + k)];
    for (int r = 1; r <= q; ++r)
      for (int j = q; j >= r; --j) {
        const int idx = sv - q + j;
        const double den = f.kv[idx + q - r + 1] - f.kv[idx], al = den > 0 ? (v - f.kv[idx]) / den : 0.0;
        for (int k = 0; k < 4; ++k) d[j][k] = (1 - al) * d[j - 1][k] + al * d[j][k];
      }
    for (int k = 0; k < 4; ++k) row[a][k] = d[q][k];
  }
  for (int r = 1; r <= p; ++r)
    for (int j = p; j >= r; --j) {
      const int idx = su - p + j;
      const double den = f.ku[idx + p - r + 1] - f.ku[idx], al = den > 0 ? (u - f.ku[idx]) / den : 0.0;
      for (int k = 0; k < 4; ++k) row[j][k] = (1 - al) * row[j - 1][k] + al * row[j][k];
    }
  return V{row[p][0] / row[p][3], row[p][1] / row[p][3], row[p][2] / row[p][3]};
}
V face_point(const icam_face_desc& f, double u, double v) {
  if (f.kind == ICAM_NURBS) return nurbs_point(f, u, v);
  V o{f.o[0], f.o[1], f.o[2]}, x{f.x[0], f.x[1], f.x[2]}, z{f.z[0], f.z[1], f.z[2]}, y = cross(z, x);
  if (f.kind == ICAM_PLANE) return o + x * u + y * v;
  const double a = arc_angle(u, 0.0, 2 * M_PI, 4);
  const V e = x * std::cos(a) + y * std::sin(a);
  if (f.kind == ICAM_CYLINDER) return o + e * f.c[0] + z * v;
  if (f.kind == ICAM_CONE) return o + e * (f.c[0] + v * std::tan(f.c[1])) + z * v;
  if (f.kind == ICAM_SPHERE) {
    const double lat = arc_angle(v, -0.5 * M_PI, M_PI, 2);
    return o + e * (f.c[0] * std::cos(lat)) + z * (f.c[0] * std::sin(lat));
  }
  // ring torus (spindle tori: their sheet's angle range)
  double psi;
  if (f.c[0] > f.c[1]) {
    psi = arc_angle(v, 0.0, 2 * M_PI, 4);
  } else {
    const double pp = std::acos(std::min(std::max(-f.c[0] / f.c[1], -1.0), 1.0));
    psi = arc_angle(v, -pp, 2 * pp, int(std::max(std::ceil(2 * pp / (0.5 * M_PI) - 1e-9), 1.0)));
  }
  return o + e * (f.c[0] + f.c[1] * std::cos(psi)) + z * (f.c[1] * std::sin(psi));
}
// Inside a face's loops, for drawing points: crossing parity along the two-leg path from
// the face's reference point (up or down to the point's v, then across to its u), the
// pieces flattened to 16 chords each (points near a loop may fall on either side), the
// chords bucketed by rows of v for the second leg.
struct FaceTrim {
  double ru = 0, rv = 0, v0 = 0, dv = 1;
  bool rin = true, whole = true;
  std::vector<std::array<double, 4>> chords;
  std::vector<uint32_t> leg1;            // chords whose u range holds ru
  std::vector<std::vector<uint32_t>> row;  // chords whose v range meets each row
};
FaceTrim face_trim(const icam_face_desc& f) {
  FaceTrim t;
  t.ru = f.ref[0], t.rv = f.ref[1], t.rin = f.ref_inside != 0, t.whole = f.n_pieces == 0;
  if (t.whole) return t;
  double vlo = 1e300, vhi = -1e300;
  for (uint32_t k = 0; k < f.n_pieces; ++k) {
    const double* c = f.pieces + 8 * k;
    double pu = c[0], pv = c[1];
    for (int m = 1; m <= 16; ++m) {
      const double x = m / 16.0, y = 1 - x;
      const double b0 = y * y * y, b1 = 3 * x * y * y, b2 = 3 * x * x * y, b3 = x * x * x;
      const double qu = b0 * c[0] + b1 * c[2] + b2 * c[4] + b3 * c[6], qv = b0 * c[1] + b1 * c[3] + b2 * c[5] + b3 * c[7];
      t.chords.push_back({pu, pv, qu, qv});
      vlo = std::min(vlo, std::min(pv, qv)), vhi = std::max(vhi, std::max(pv, qv));
      pu = qu, pv = qv;
    }
  }
  const int R = int(std::min<size_t>(4096, t.chords.size() / 4 + 1));
  t.v0 = vlo, t.dv = std::max((vhi - vlo) / R, 1e-12);
  t.row.resize(size_t(R));
  for (uint32_t i = 0; i < t.chords.size(); ++i) {
    const auto& c = t.chords[i];
    if (std::min(c[0], c[2]) <= t.ru && std::max(c[0], c[2]) >= t.ru) t.leg1.push_back(i);
    const int r0 = std::max(0, int((std::min(c[1], c[3]) - t.v0) / t.dv)), r1 = std::min(R - 1, int((std::max(c[1], c[3]) - t.v0) / t.dv));
    for (int r = r0; r <= r1; ++r) t.row[size_t(r)].push_back(i);
  }
  return t;
}
// segments (a, b) and (c, d) crossing properly
bool seg_cross(double au, double av, double bu, double bv, const std::array<double, 4>& c) {
  const double d1 = (c[2] - c[0]) * (av - c[1]) - (c[3] - c[1]) * (au - c[0]), d2 = (c[2] - c[0]) * (bv - c[1]) - (c[3] - c[1]) * (bu - c[0]);
  const double d3 = (bu - au) * (c[1] - av) - (bv - av) * (c[0] - au), d4 = (bu - au) * (c[3] - av) - (bv - av) * (c[2] - au);
  return (d1 > 0) != (d2 > 0) && (d3 > 0) != (d4 > 0);
}
bool face_inside(const FaceTrim& t, double u, double v) {
  if (t.whole) return true;
  int n = 0;
  for (uint32_t i : t.leg1) n += seg_cross(t.ru, t.rv, t.ru, v, t.chords[i]);
  const int r = int((v - t.v0) / t.dv);
  if (r >= 0 && r < int(t.row.size()))
    for (uint32_t i : t.row[size_t(r)]) n += seg_cross(t.ru, v, u, v, t.chords[i]);
  return t.rin != bool(n & 1);
}
// The domain box a face's points are drawn from (a NURBS face: its knot range).
void face_domain(const icam_face_desc& f, double lo[2], double hi[2]) {
  lo[0] = f.lo[0], lo[1] = f.lo[1], hi[0] = f.hi[0], hi[1] = f.hi[1];
  if (f.kind == ICAM_NURBS) {
    lo[0] = std::max(lo[0], f.ku[f.du]), hi[0] = std::min(hi[0], f.ku[f.nu]);
    lo[1] = std::max(lo[1], f.kv[f.dv]), hi[1] = std::min(hi[1], f.kv[f.nv]);
  }
  if (f.period[0] > 0) lo[0] = 0, hi[0] = f.period[0];
  if (f.period[1] > 0) lo[1] = 0, hi[1] = f.period[1];
}
}  // namespace

extern "C" {

// A synthetic part by name: "can", "bump", "c1", "c2", "c2flat", "c4:<i>" (part i of
// the batched workload), "turned:<type>:<grooves>", "plate:<holes>",
// "nurbs:<net>"; NULL for an unknown name. Free with icam_gen_free.
const icam_part_desc* icam_gen_part(const char* name, uint64_t seed) {
  std::string s(name ? name : "");
  Owned* o = nullptr;
  auto arg = [&](size_t at) { return std::atoi(s.c_str() + at); };
  if (s == "can") o = make_can();
  if (s == "bump") o = make_bump();
  if (s == "c1") o = make_c1();
  if (s == "c2") o = make_c2(seed, false);
  if (s == "c2flat") o = make_c2(seed, true);
  if (s == "c3") o = make_c3(seed);
  if (s.rfind("c4:", 0) == 0) o = make_c4(seed, arg(3));
  if (s == "c5") o = make_c5(seed);
  if (s.rfind("dome:", 0) == 0) {
    Rng r2{(seed ^ (0x51ED270B27FF3A15ull * uint64_t(arg(5) + 1))) | 1};
    const double R = 8 + 12 * r2.next(), L = 10 + 30 * r2.next();
    o = make_dome(R, L);
  }
  if (s.rfind("turned:", 0) == 0) {
    Rng rng{seed | 1};
    size_t c = s.find(':', 7);
    o = make_revolved(turned_profile(arg(7) % 6, c == std::string::npos ? 0 : arg(c + 1), rng));
  }
  if (s.rfind("plate:", 0) == 0) {
    Rng rng{seed | 1};
    o = make_plate(120, 80, 16, plate_holes(120, 80, 16, std::max(1, arg(6)), rng));
  }
  if (s.rfind("nurbs:", 0) == 0) o = make_c2(seed, false, std::min(std::max(arg(6), 4), 32), 40, 3);
  if (s.rfind("crease:", 0) == 0) o = make_c2(seed, false, 17, 60, 5, arg(7) & 3);
  if (s.rfind("quintic:", 0) == 0) o = make_quintic(seed, arg(8) != 0);
  if (s.rfind("oblique:", 0) == 0) {
    Rng r2{(seed ^ (0x9E3779B97F4A7C15ull * uint64_t(arg(8) + 1))) | 1};
    const double R = 5 + 10 * r2.next(), L = 2 * R + 20 * r2.next(), a = (15 + 45 * r2.next()) * M_PI / 180;
    o = make_oblique(R, L, std::tan(a));
  }
  return o ? &o->desc : nullptr;
}

void icam_gen_free(const icam_part_desc* p) { delete reinterpret_cast<const Owned*>(p); }

// What a part was made from, for closed-form checks: [1, n, n elements] for a solid of
// revolution about z, [2, W, D, T, k, then per hole x, y, n, n elements] for a plate,
// [3, W, D, T, loops, then per loop kind (0 pocket, 1 boss), z0, z1, F, then F x 16
// control points (x, y)] for the trimmed plate, [4, R, L, tan a] for the oblique cut,
// [5, p, n, the n + p + 1 knots (both directions), the n x n homogeneous net] for the
// quintic top, [] otherwise; an element is (arc, r0,
// z0, r1, z1, rc, zc, rad, psi0, psi1, face).
// Writes at most cap values and returns how many there are.
size_t icam_gen_describe(const icam_part_desc* p, double* out, size_t cap) {
  const Owned* o = reinterpret_cast<const Owned*>(p);
  for (size_t i = 0; i < o->info.size() && i < cap; ++i) out[i] = o->info[i];
  return o->info.size();
}

// n points within `radius` of the trim curves of the part's planar faces (a planar
// face drawn in proportion to its trim pieces, a piece and its parameter uniformly, the
// offset uniform in the ball), from the reference checks' xorshift64.
void icam_gen_near_trim(const icam_part_desc* p, uint64_t seed, size_t n, double radius, double* xyz) {
  Rng r{seed ? seed : 1};
  std::vector<std::pair<uint32_t, uint32_t>> pieces;  // (face, piece)
  for (uint32_t f = 0; f < p->n_faces; ++f)
    if (p->faces[f].kind == ICAM_PLANE)
      for (uint32_t k = 0; k < p->faces[f].n_pieces; ++k) pieces.push_back({f, k});
  for (size_t i = 0; i < n; ++i) {
    double* out = xyz + 3 * i;
    if (pieces.empty()) {
      out[0] = out[1] = out[2] = 0;
      continue;
    }
    auto fp = pieces[size_t(r.next() * pieces.size()) % pieces.size()];
    const icam_face_desc& f = p->faces[fp.first];
    const double* c = f.pieces + 8 * fp.second;
    const double t = r.next(), s = 1 - t;
    const double b[4] = {s * s * s, 3 * t * s * s, 3 * t * t * s, t * t * t};
    const double u = b[0] * c[0] + b[1] * c[2] + b[2] * c[4] + b[3] * c[6], v = b[0] * c[1] + b[1] * c[3] + b[2] * c[5] + b[3] * c[7];
    V x{f.x[0], f.x[1], f.x[2]}, z{f.z[0], f.z[1], f.z[2]}, y = cross(z, x);
    V q = V{f.o[0], f.o[1], f.o[2]} + x * u + y * v;
    V d;
    do {
      d = V{2 * r.next() - 1, 2 * r.next() - 1, 2 * r.next() - 1};
    } while (dot(d, d) > 1.0);
    q = q + d * radius;
    out[0] = q.x, out[1] = q.y, out[2] = q.z;
  }
}

// n points within `radius` of the part's surface: a face drawn in proportion to its area
// (estimated on a 24 x 24 grid of its domain box, the cells inside its loops), a point
// of it drawn uniformly in its domain box until inside its loops, then offset uniformly
// in the ball of `radius`; from the reference checks' xorshift64.
void icam_gen_near_surface(const icam_part_desc* p, uint64_t seed, size_t n, double radius, double* xyz) {
  Rng r{seed ? seed : 1};
  const uint32_t nf = p->n_faces;
  std::vector<double> cum(nf + 1, 0.0);
  std::vector<FaceTrim> trims(nf);
  for (uint32_t fi = 0; fi < nf; ++fi) {
    const icam_face_desc& f = p->faces[fi];
    trims[fi] = face_trim(f);
    double lo[2], hi[2];
    face_domain(f, lo, hi);
    const int G = 24;
    const double du = (hi[0] - lo[0]) / G, dv = (hi[1] - lo[1]) / G;
    double area = 0;
    for (int i = 0; i < G; ++i)
      for (int j = 0; j < G; ++j) {
        const double u = lo[0] + (i + 0.5) * du, v = lo[1] + (j + 0.5) * dv;
        if (!face_inside(trims[fi], u, v)) continue;
        const V a = face_point(f, u, v), bu = face_point(f, u + 0.5 * du, v) - face_point(f, u - 0.5 * du, v),
                bv = face_point(f, u, v + 0.5 * dv) - face_point(f, u, v - 0.5 * dv);
        (void)a;
        area += len(cross(bu, bv));
      }
    cum[fi + 1] = cum[fi] + area;
  }
  for (size_t i = 0; i < n; ++i) {
    double* out = xyz + 3 * i;
    const double x = r.next() * cum[nf];
    const uint32_t fi = uint32_t(std::min<size_t>(size_t(std::upper_bound(cum.begin() + 1, cum.end(), x) - (cum.begin() + 1)), nf - 1));
    const icam_face_desc& f = p->faces[fi];
    double lo[2], hi[2];
    face_domain(f, lo, hi);
    double u = 0.5 * (lo[0] + hi[0]), v = 0.5 * (lo[1] + hi[1]);
    for (int t = 0; t < 1000; ++t) {
      u = lo[0] + (hi[0] - lo[0]) * r.next(), v = lo[1] + (hi[1] - lo[1]) * r.next();
      if (face_inside(trims[fi], u, v)) break;
    }
    V q = face_point(f, u, v), d;
    do {
      d = V{2 * r.next() - 1, 2 * r.next() - 1, 2 * r.next() - 1};
    } while (dot(d, d) > 1.0);
    q = q + d * radius;
    out[0] = q.x, out[1] = q.y, out[2] = q.z;
  }
}

// n points uniform in [lo, hi] from the reference checks' xorshift64 (x, y, z
// drawn in that order per point).
void icam_gen_uniform(uint64_t seed, size_t n, const double lo[3], const double hi[3], double* xyz) {
  Rng r{seed ? seed : 1};
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) xyz[3 * i + k] = lo[k] + (hi[k] - lo[k]) * r.next();
}

}  // extern "C"
