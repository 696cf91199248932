// Device side of the exact signed distance: one lane per point, every feature's
// closed form or certified Newton, in a compute type T (float or double).
//
// The rules are the reference evaluator's (crates/icam-core/src/solid3.rs:1750-1940)
// and its kernel's (crates/icam-core/src/solid3/gpu.wgsl: face_d :415, edge_foot
// :820, sub_min :679-756, curve_min_on :764-815, node_hit :1794-1906):
//  - an analytic face answers with its closed-form foot when the foot lies inside
//    the face by its trim loops (crossing parity in the cell the foot falls in);
//  - a line or arc edge with its foot when interior; a vertex with itself;
//  - a B-spline sub-patch or curve part is read by projected Newton only where its
//    radius of convexity certifies one minimum (the point's farthest distance to its
//    box under the radius); the minimum counts when it is a critical point inside the
//    face; a part without that certificate leaves its box distance as a lower bound;
//  - signs from the face normal, the two faces' normals summed at an edge point, the
//    vertex pseudo-normal.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/icam_part.h"
#include "layout.hpp"

// The part's tables are read through the constant address space: they do not change
// while a kernel runs, so a read whose address is the same on every lane of a wave
// (a block's candidate, its span's net) becomes a scalar load into SGPRs.
#ifndef ICAM_K
#define ICAM_K __attribute__((address_space(4)))
#endif

namespace icam {

// ------------------------------------------------------------------ math by type
__device__ inline float msqrt(float x) { return __builtin_sqrtf(x); }
__device__ inline double msqrt(double x) { return __builtin_sqrt(x); }
__device__ inline float matan2(float y, float x) { return atan2f(y, x); }
__device__ inline double matan2(double y, double x) { return atan2(y, x); }
__device__ inline float mtan(float x) { return tanf(x); }
__device__ inline double mtan(double x) { return tan(x); }
__device__ inline float msin(float x) { return sinf(x); }
__device__ inline double msin(double x) { return sin(x); }
__device__ inline float mcos(float x) { return cosf(x); }
__device__ inline double mcos(double x) { return cos(x); }
__device__ inline float mfloor(float x) { return floorf(x); }
__device__ inline double mfloor(double x) { return floor(x); }
__device__ inline float mabs(float x) { return fabsf(x); }
__device__ inline double mabs(double x) { return fabs(x); }
template <class T> __device__ inline T mmin(T a, T b) { return a < b ? a : b; }
template <class T> __device__ inline T mmax(T a, T b) { return a > b ? a : b; }
template <class T> __device__ inline T mclamp(T x, T a, T b) { return mmin(mmax(x, a), b); }

template <class T> struct TypeEps;
template <> struct TypeEps<float> {
  static constexpr float ulp = 1.1920929e-7f;
  static constexpr float big = 3.0e38f;
};
template <> struct TypeEps<double> {
  static constexpr double ulp = 2.220446049250313e-16;
  static constexpr double big = 1.0e300;
};

template <class T> struct v3 {
  T x, y, z;
};
template <class T> __device__ inline v3<T> mk3(T a, T b, T c) { return v3<T>{a, b, c}; }
template <class T> __device__ inline v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __device__ inline v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __device__ inline v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> __device__ inline T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __device__ inline v3<T> cross(v3<T> a, v3<T> b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
template <class T> __device__ inline T len(v3<T> a) { return msqrt(dot(a, a)); }
template <class T> __device__ inline v3<T> ld3(const ICAM_K T* p) { return {p[0], p[1], p[2]}; }
template <class T, class S> __device__ inline v3<T> cv3(v3<S> a) { return {T(a.x), T(a.y), T(a.z)}; }

// The tables of one compute type.
template <class T> struct Tab {
  const ICAM_K T* F;
  const ICAM_K uint32_t* U;
  const ICAM_K float* pbox;
  const ICAM_K uint32_t* pcode;
  Layout L;
};

template <class T> __device__ inline T box_dist(v3<T> p, const ICAM_K float* b) {
  T dx = mmax(mmax(T(b[0]) - p.x, p.x - T(b[3])), T(0));
  T dy = mmax(mmax(T(b[1]) - p.y, p.y - T(b[4])), T(0));
  T dz = mmax(mmax(T(b[2]) - p.z, p.z - T(b[5])), T(0));
  return msqrt(dx * dx + dy * dy + dz * dz);
}
// The distance to the farthest point of a box.
template <class T> __device__ inline T far_dist(v3<T> p, const ICAM_K float* b) {
  T dx = mmax(mabs(p.x - T(b[0])), mabs(p.x - T(b[3])));
  T dy = mmax(mabs(p.y - T(b[1])), mabs(p.y - T(b[4])));
  T dz = mmax(mabs(p.z - T(b[2])), mabs(p.z - T(b[5])));
  return msqrt(dx * dx + dy * dy + dz * dz);
}

// ------------------------------------------------------------------ trim loops
// trim2.rs:104-136 a cubic coordinate and its derivative's quadratic
template <class T> __device__ inline T beval3(T c0, T c1, T c2, T c3, T t) {
  if (c0 == c1 && c1 == c2 && c2 == c3) return c0;
  T s = T(1) - t;
  T a0 = s * c0 + t * c1, a1 = s * c1 + t * c2, a2 = s * c2 + t * c3;
  T b0 = s * a0 + t * a1, b1 = s * a1 + t * a2;
  return s * b0 + t * b1;
}
template <class T> __device__ inline T beval2(T c0, T c1, T c2, T t) {
  if (c0 == c1 && c1 == c2) return c0;
  T s = T(1) - t;
  T a0 = s * c0 + t * c1, a1 = s * c1 + t * c2;
  return s * a0 + t * a1;
}

// trim2.rs:1150-1190 the parameter where a monotone cubic coordinate equals x, x in
// its half-open range [min, max): safeguarded Newton. `near` is set when x lies
// within m of an end of that range (the half-open rule decides there).
template <class T> __device__ inline bool crossing(const T a[4], T x, T& out, T m, bool& near) {
  T p0 = a[0], p1 = a[3];
  T lo_ = mmin(p0, p1), hi_ = mmax(p0, p1);
  if (m > T(0) && (mabs(x - lo_) <= m || mabs(x - hi_) <= m) && hi_ > lo_) near = true;
  if (!(lo_ <= x && x < hi_)) return false;
  T lo = 0, hi = 1;
  T flo = beval3(a[0], a[1], a[2], a[3], T(0)) - x;
  if (flo == T(0)) {
    out = 0;
    return true;
  }
  T t = mclamp((x - p0) / (p1 - p0), T(0), T(1));
  const int iters = sizeof(T) == 4 ? 30 : 60;
  for (int it = 0; it < iters; ++it) {
    T ft = beval3(a[0], a[1], a[2], a[3], t) - x;
    if (ft == T(0)) {
      out = t;
      return true;
    }
    if ((ft < T(0)) == (flo < T(0))) {
      lo = t;
      flo = ft;
    } else {
      hi = t;
    }
    T d = T(3) * beval2(a[1] - a[0], a[2] - a[1], a[3] - a[2], t);
    T nt = d != T(0) ? t - ft / d : T(0.5) * (lo + hi);
    if (!(nt > lo && nt < hi)) nt = T(0.5) * (lo + hi);
    if (mabs(nt - t) <= (sizeof(T) == 4 ? T(1e-7) : T(1e-16))) {
      out = nt;
      return true;
    }
    t = nt;
  }
  out = t;
  return true;
}

// trim2.rs:1205-1260 crossing parity of the path a -> (a.u, b.v) -> b, moved up by an
// infinitesimal e and right by e^2. With m > 0, `near` is set when a decision was
// taken within m (the caller then decides again in double precision).
template <class T> __device__ bool parity(const Tab<T>& tb, const T a[2], const T b[2], uint32_t first, uint32_t count, T m, bool& near) {
  unsigned n = 0;
  bool up = a[1] < b[1], right = a[0] < b[0];
  const Layout& L = tb.L;
  for (uint32_t k = 0; k < count; ++k) {
    uint32_t inst = tb.U[L.among_u + first + k];
    uint32_t pc = tb.U[L.inst_u + inst];
    const ICAM_K T* sh = tb.F + L.inst_f + INST_F * inst;
    const ICAM_K T* c = tb.F + L.tpiece_f + TPIECE_F * pc;
    T cu[4] = {c[0] + sh[0], c[2] + sh[0], c[4] + sh[0], c[6] + sh[0]};
    T cv[4] = {c[1] + sh[1], c[3] + sh[1], c[5] + sh[1], c[7] + sh[1]};
    T t;
    if (a[1] != b[1] && crossing(cu, a[0], t, m, near)) {
      T y = beval3(cv[0], cv[1], cv[2], cv[3], t);
      if (m > T(0) && (mabs(y - a[1]) <= m || mabs(y - b[1]) <= m)) near = true;
      if (up ? (y > a[1] && y <= b[1]) : (y > b[1] && y <= a[1])) ++n;
    }
    if (a[0] != b[0] && crossing(cv, b[1], t, m, near)) {
      T x = beval3(cu[0], cu[1], cu[2], cu[3], t);
      if (m > T(0) && (mabs(x - a[0]) <= m || mabs(x - b[0]) <= m)) near = true;
      bool counted;
      if (x == b[0] || x == a[0]) {
        T du = beval2(cu[1] - cu[0], cu[2] - cu[1], cu[3] - cu[2], t);
        T dv = beval2(cv[1] - cv[0], cv[2] - cv[1], cv[3] - cv[2], t);
        if (dv == T(0)) dv = cv[3] - cv[0];
        bool slope_pos = du * dv > T(0);
        counted = x == b[0] ? (!slope_pos) == right : slope_pos == right;
      } else if (right) {
        counted = x > a[0] && x < b[0];
      } else {
        counted = x > b[0] && x < a[0];
      }
      if (counted) ++n;
    }
  }
  return (n & 1u) != 0;
}

// trim2.rs:1433-1458 inside: the query reduced into the domain, the cell it falls in,
// the answer there flipped by the parity from the cell's reference.
// dom = face record words 17..22 (lo, hi, period), root = its first cell.
template <class T> __device__ bool trim_inside(const Tab<T>& tb, const ICAM_K T* dom, uint32_t root, T u, T v, T m, bool& near) {
  const Layout& L = tb.L;
  T q[2] = {u, v};
  for (int k = 0; k < 2; ++k) {
    T lo = dom[k], hi = dom[2 + k], per = dom[4 + k];
    if (per > T(0)) {
      T x = q[k] - lo;
      q[k] = lo + (x - per * mfloor(x / per));
      if (q[k] >= lo + per) q[k] = lo;  // rounding of the reduction
    } else {
      q[k] = mclamp(q[k], lo, hi);
    }
  }
  uint32_t id = root;
  for (int depth = 0; depth < 64; ++depth) {
    const ICAM_K uint32_t* cu = tb.U + L.cell_u + CELL_U * id;
    const ICAM_K T* cf = tb.F + L.cell_f + CELL_F * id;
    uint32_t kind = cu[1];
    if (kind == 0) return cu[0] != 0;
    if (kind == 1) {
      T r[2] = {cf[4], cf[5]};
      return (cu[0] != 0) ^ parity(tb, r, q, cu[2], cu[3], m, near);
    }
    T mu = T(0.5) * (cf[0] + cf[2]), mv = T(0.5) * (cf[1] + cf[3]);
    if (m > T(0) && (mabs(q[0] - mu) <= m || mabs(q[1] - mv) <= m)) {
      // the descent itself is exact whichever child is taken (each child's reference
      // answer is certified); nothing to flag
    }
    id = cu[2] + uint32_t(q[0] >= mu) + 2u * uint32_t(q[1] >= mv);
  }
  near = true;
  return false;
}

// ------------------------------------------------------------------ analytic faces
template <class T> struct Axial {
  T theta, rho, h;
  v3<T> er;
};
// surface3.rs:36-53
template <class T> __device__ inline Axial<T> axial(const ICAM_K T* f, v3<T> p) {
  v3<T> o = ld3(f), x = ld3(f + 3), y = ld3(f + 6), z = ld3(f + 9);
  v3<T> d = p - o;
  T lx = dot(d, x), ly = dot(d, y), lz = dot(d, z);
  T rho = msqrt(lx * lx + ly * ly);
  Axial<T> a;
  a.theta = rho > T(0) ? matan2(ly, lx) : T(0);
  a.rho = rho;
  a.h = lz;
  if (rho > T(0)) {
    a.er = x * (lx / rho) + y * (ly / rho);
  } else {
    a.er = x;
  }
  return a;
}

// surface3.rs:85-128 the foot of p on an analytic surface (kind from the face's U record)
template <class T> __device__ v3<T> foot(uint32_t kind, const ICAM_K T* f, v3<T> p) {
  v3<T> o = ld3(f), z = ld3(f + 9);
  T c0 = f[12], c1 = f[13];
  if (kind == ICAM_PLANE) return p - z * dot(p - o, z);
  if (kind == ICAM_CYLINDER) {
    Axial<T> ax = axial(f, p);
    return o + ax.er * c0 + z * ax.h;
  }
  if (kind == ICAM_CONE) {
    Axial<T> ax = axial(f, p);
    T sa = msin(c1), ca = mcos(c1);
    T ha = -c0 / mtan(c1);
    T s = ax.rho * sa + (ax.h - ha) * ca;
    if (s <= T(0)) return o + z * ha;
    return o + ax.er * (s * sa) + z * (ha + s * ca);
  }
  if (kind == ICAM_SPHERE) {
    v3<T> d = p - o;
    T l = len(d);
    return l > T(0) ? o + d * (c0 / l) : o + ld3(f + 3) * c0;
  }
  // torus
  Axial<T> ax = axial(f, p);
  v3<T> c = o + ax.er * c0;
  v3<T> d = p - c;
  T l = len(d);
  v3<T> q = l > T(0) ? c + d * (c1 / l) : c + ax.er * c1;
  if (dot(q - o, ax.er) >= T(0) || mabs(c0) > c1) return q;
  T h = msqrt(mmax(c1 * c1 - c0 * c0, T(0)));
  return ax.h >= T(0) ? o + z * h : o - z * h;
}

// surface3.rs:131-149 the surface's own unit normal at a point on it
template <class T> __device__ v3<T> snormal(uint32_t kind, const ICAM_K T* f, v3<T> x) {
  v3<T> o = ld3(f), z = ld3(f + 9);
  if (kind == ICAM_PLANE) return z;
  if (kind == ICAM_CYLINDER) return axial(f, x).er;
  if (kind == ICAM_CONE) {
    Axial<T> ax = axial(f, x);
    return ax.er * mcos(f[13]) - z * msin(f[13]);
  }
  if (kind == ICAM_SPHERE) {
    v3<T> d = x - o;
    return d * (T(1) / len(d));
  }
  Axial<T> ax = axial(f, x);
  v3<T> d = x - (o + ax.er * f[12]);
  return d * (T(1) / len(d));
}

// trim2.rs:1789-1807 arc_param_over
template <class T> __device__ inline T arc_param_over(T a, T a0, T span, T n) {
  const T TAU = T(6.283185307179586);
  T q = span / n;
  T x = a - a0;
  if (mabs(span - TAU) < T(1e-6)) {
    x = x - TAU * mfloor(x / TAU);
  } else {
    T y = x - TAU * mfloor((x - T(0.5) * span) / TAU + T(0.5));
    x = mclamp(y, T(0), span);
  }
  T k = mclamp(mfloor(x / q), T(0), n - T(1));
  T alpha = x - q * k - T(0.5) * q;
  T t = T(0.5) * (mtan(T(0.5) * alpha) / mtan(T(0.25) * q) + T(1));
  return k + mclamp(t, T(0), T(1));
}

// trim2.rs:1907-1934 Analytic::uv
template <class T> __device__ void face_uv(uint32_t kind, uint32_t narc, const ICAM_K T* f, v3<T> q, T& u, T& v) {
  const T HALF_PI = T(1.5707963267948966), FULL = T(6.283185307179586);
  if (kind == ICAM_PLANE) {
    v3<T> d = q - ld3(f);
    u = dot(d, ld3(f + 3)), v = dot(d, ld3(f + 6));
    return;
  }
  Axial<T> ax = axial(f, q);
  u = arc_param_over(ax.theta, T(0), FULL, T(4));
  if (kind == ICAM_CYLINDER) {
    v = ax.h;
  } else if (kind == ICAM_CONE) {
    T sa = msin(f[13]), ca = mcos(f[13]);
    T ha = -f[12] / mtan(f[13]);
    T s = mmax(ax.rho * sa + (ax.h - ha) * ca, T(0));
    v = ha + s * ca;
  } else if (kind == ICAM_SPHERE) {
    v = arc_param_over(matan2(ax.h, ax.rho), -HALF_PI, T(2) * HALF_PI, T(2));
  } else {
    v = arc_param_over(matan2(ax.h, ax.rho - f[12]), f[15], f[16], T(narc));
  }
}

// ------------------------------------------------------------------ rational bicubic
template <class T> struct SEval {
  v3<T> p, ps, pt, pss, pst, ptt;
};

template <class T> __device__ inline void bern3(T x, T b[4], T d[4], T e[4]) {
  T y = T(1) - x;
  b[0] = y * y * y, b[1] = T(3) * x * y * y, b[2] = T(3) * x * x * y, b[3] = x * x * x;
  d[0] = T(-3) * y * y, d[1] = T(3) * y * (T(1) - T(3) * x), d[2] = T(3) * x * (T(2) - T(3) * x), d[3] = T(3) * x * x;
  e[0] = T(6) * y, e[1] = T(18) * x - T(12), e[2] = T(6) - T(18) * x, e[3] = T(6) * x;
}

// surface3.rs:322-351 a rational bicubic's point and derivatives at (s, t) from its
// homogeneous net (16 points, i along u)
template <class T> __device__ SEval<T> peval(const ICAM_K T* pw, T s, T t) {
  T bu[4], du[4], eu[4], bv[4], dv[4], ev[4];
  bern3(s, bu, du, eu);
  bern3(t, bv, dv, ev);
  T a[6][4];
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int c = 0; c < 4; ++c) a[k][c] = T(0);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    // the row's sums along v first
    T r0[4], r1[4], r2[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) r0[c] = r1[c] = r2[c] = T(0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const ICAM_K T* h = pw + 16 * i + 4 * j;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        r0[c] += bv[j] * h[c];
        r1[c] += dv[j] * h[c];
        r2[c] += ev[j] * h[c];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      a[0][c] += bu[i] * r0[c];
      a[1][c] += du[i] * r0[c];
      a[2][c] += bu[i] * r1[c];
      a[3][c] += eu[i] * r0[c];
      a[4][c] += du[i] * r1[c];
      a[5][c] += bu[i] * r2[c];
    }
  }
  T iw = T(1) / a[0][3];
  SEval<T> e;
  e.p = mk3(a[0][0], a[0][1], a[0][2]) * iw;
  e.ps = (mk3(a[1][0], a[1][1], a[1][2]) - e.p * a[1][3]) * iw;
  e.pt = (mk3(a[2][0], a[2][1], a[2][2]) - e.p * a[2][3]) * iw;
  e.pss = (mk3(a[3][0], a[3][1], a[3][2]) - e.ps * (T(2) * a[1][3]) - e.p * a[3][3]) * iw;
  e.pst = (mk3(a[4][0], a[4][1], a[4][2]) - e.ps * a[2][3] - e.pt * a[1][3] - e.p * a[4][3]) * iw;
  e.ptt = (mk3(a[5][0], a[5][1], a[5][2]) - e.pt * (T(2) * a[2][3]) - e.p * a[5][3]) * iw;
  return e;
}
// The point and first derivatives alone, summed in peval's order (so equal to its p,
// ps, pt bit for bit); the second derivatives are left unset.
template <class T> __device__ SEval<T> peval1(const ICAM_K T* pw, T s, T t) {
  T bu[4], du[4], eu[4], bv[4], dv[4], ev[4];
  bern3(s, bu, du, eu);
  bern3(t, bv, dv, ev);
  T a[3][4];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < 4; ++c) a[k][c] = T(0);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    T r0[4], r1[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) r0[c] = r1[c] = T(0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const ICAM_K T* h = pw + 16 * i + 4 * j;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        r0[c] += bv[j] * h[c];
        r1[c] += dv[j] * h[c];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      a[0][c] += bu[i] * r0[c];
      a[1][c] += du[i] * r0[c];
      a[2][c] += bu[i] * r1[c];
    }
  }
  T iw = T(1) / a[0][3];
  SEval<T> e;
  e.p = mk3(a[0][0], a[0][1], a[0][2]) * iw;
  e.ps = (mk3(a[1][0], a[1][1], a[1][2]) - e.p * a[1][3]) * iw;
  e.pt = (mk3(a[2][0], a[2][1], a[2][2]) - e.p * a[2][3]) * iw;
  return e;
}

// Projected Newton on |S - p|^2 over the rectangle [s0, s1] x [t0, t1] (after
// surface3.rs:401-452, with the active set decided afresh each step): a parameter on
// its bound is held there while the gradient points out of the rectangle, and Newton
// runs in the other alone. Over a part its radius of convexity certifies, the
// squared distance is strictly convex, so this reaches the part's one minimum,
// inside it or on its border (where two parts share a border both reach it). Returns
// whether it settled (a step under `stop`, or held on both bounds) within `iters`.
template <class T> __device__ bool newton2(const ICAM_K T* pw, v3<T> p, T& s, T& t, T s0, T s1, T t0, T t1, int iters, T stop) {
  for (int it = 0; it < iters; ++it) {
    SEval<T> e = peval(pw, s, t);
    v3<T> r = e.p - p;
    T g0 = dot(e.ps, r), g1 = dot(e.pt, r);
    T jss = dot(e.ps, e.ps), jst = dot(e.ps, e.pt), jtt = dot(e.pt, e.pt);
    T h0 = jss + dot(e.pss, r), h1 = jst + dot(e.pst, r), h2 = jtt + dot(e.ptt, r);
    bool hs = (s <= s0 && g0 > T(0)) || (s >= s1 && g0 < T(0));
    bool ht = (t <= t0 && g1 > T(0)) || (t >= t1 && g1 < T(0));
    T ds, dt;
    if (hs && ht) return true;
    if (hs) {
      ds = 0;
      dt = -g1 / (h2 > T(0) ? h2 : mmax(jtt, TypeEps<T>::ulp));
    } else if (ht) {
      ds = -g0 / (h0 > T(0) ? h0 : mmax(jss, TypeEps<T>::ulp));
      dt = 0;
    } else {
      if (!(h0 > T(0) && h0 * h2 - h1 * h1 > T(0))) h0 = jss, h1 = jst, h2 = jtt;
      T det = h0 * h2 - h1 * h1;
      if (!(mabs(det) > T(0))) return false;
      ds = -(h2 * g0 - h1 * g1) / det;
      dt = -(h0 * g1 - h1 * g0) / det;
    }
    T ns = mclamp(s + ds, s0, s1), nt = mclamp(t + dt, t0, t1);
    T moved = mabs(ns - s) + mabs(nt - t);
    s = ns, t = nt;
    if (moved < stop) return true;
  }
  return false;
}

// A cubic rational curve piece's point and derivatives (curve3.rs:129-156).
template <class T> struct CEval {
  v3<T> p, d1, d2;
};
template <class T> __device__ CEval<T> ceval(const ICAM_K T* pw, T s) {
  T b[4], d[4], e[4];
  bern3(s, b, d, e);
  T a[3][4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    a[0][c] = b[0] * pw[c] + b[1] * pw[4 + c] + b[2] * pw[8 + c] + b[3] * pw[12 + c];
    a[1][c] = d[0] * pw[c] + d[1] * pw[4 + c] + d[2] * pw[8 + c] + d[3] * pw[12 + c];
    a[2][c] = e[0] * pw[c] + e[1] * pw[4 + c] + e[2] * pw[8 + c] + e[3] * pw[12 + c];
  }
  T iw = T(1) / a[0][3];
  CEval<T> r;
  r.p = mk3(a[0][0], a[0][1], a[0][2]) * iw;
  r.d1 = (mk3(a[1][0], a[1][1], a[1][2]) - r.p * a[1][3]) * iw;
  r.d2 = (mk3(a[2][0], a[2][1], a[2][2]) - r.d1 * (T(2) * a[1][3]) - r.p * a[2][3]) * iw;
  return r;
}

// curve3.rs:192-213 Newton on |C - p|^2 over [t0, t1]; *conv (when given) tells whether
// it settled within `iters`.
template <class T> __device__ T newton1(const ICAM_K T* pw, v3<T> p, T s, T t0, T t1, int iters, T stop, bool* conv = nullptr) {
  if (conv) *conv = false;
  for (int it = 0; it < iters; ++it) {
    CEval<T> e = ceval(pw, s);
    v3<T> r = e.p - p;
    T f = dot(e.d1, r);
    T fp = dot(e.d2, r) + dot(e.d1, e.d1);
    if (fp <= T(0)) fp = dot(e.d1, e.d1);
    if (!(fp > T(0))) break;
    T ns = mclamp(s - f / fp, t0, t1);
    if (mabs(ns - s) < stop) {
      s = ns;
      if (conv) *conv = true;
      break;
    }
    s = ns;
  }
  return s;
}

// ------------------------------------------------------------------ the features

// What a lane's search has so far.
template <class T> struct Best {
  T d;            // distance of the best answer (or the bound)
  uint32_t code;  // kind << 28 | record, or ~0 for none
  T a, b;         // its parameters (leaf: s, t; curve leaf: t; edge: t)
  T sgn;          // its sign when known
  T lmin;         // least lower bound among parts not decided
  T d2;           // the runner-up's distance and sign (near-ties)
  T sgn2;
  bool unsure;    // a decision was taken within rounding (mixed precision redoes it)
  bool host;      // a host feature's box came under the answer (its bound is in lmin)
};

// The outward normal of face fi at a point x on (or within rounding of) it. A
// B-spline face takes its parameters from the edge's map (uvmap: u, v at the start
// and end of the piece) at the piece parameter tt.
template <class T> __device__ v3<T> face_normal_at(const Tab<T>& tb, uint32_t fi, v3<T> x, const ICAM_K T* uvm, T tt) {
  const Layout& L = tb.L;
  const ICAM_K T* f = tb.F + L.face_f + FACE_F * fi;
  const ICAM_K uint32_t* fu = tb.U + L.face_u + FACE_U * fi;
  uint32_t kind = fu[0];
  if (kind != ICAM_NURBS) return snormal(kind, f, x) * f[14];
  if (!uvm) return mk3(T(0), T(0), T(0));
  T u = uvm[0] + tt * (uvm[2] - uvm[0]), v = uvm[1] + tt * (uvm[3] - uvm[1]);
  uint32_t span0 = fu[3], nsu = fu[4], nsv = fu[5];
  uint32_t a = 0, b = 0;
  for (uint32_t k = 1; k < nsu; ++k)
    if (tb.F[L.span_f + SPAN_F * (span0 + k * nsv) + 64] <= u) a = k;
  for (uint32_t k = 1; k < nsv; ++k)
    if (tb.F[L.span_f + SPAN_F * (span0 + k) + 66] <= v) b = k;
  const ICAM_K T* sp = tb.F + L.span_f + SPAN_F * (span0 + a * nsv + b);
  T s = mclamp((u - sp[64]) / (sp[65] - sp[64]), T(0), T(1)), t = mclamp((v - sp[66]) / (sp[67] - sp[66]), T(0), T(1));
  // the point is on the surface at these parameters; two Newton steps take up rounding
  newton2(sp, x, s, t, T(0), T(1), T(0), T(1), 2, T(0));
  SEval<T> e = peval(sp, s, t);
  v3<T> n = cross(e.ps, e.pt);
  T l = len(n);
  return l > T(0) ? n * (f[14] / l) : n;
}

// A C0 knot line's pieces (solid3.rs:1818-1886): piece cp of the line of edge record eu
// runs along the span before the line (its far side) and the span after (its near side);
// both spans share the piece's parameter range along the line.
__device__ inline uint32_t crease_before(const Layout& L, const ICAM_K uint32_t* U, uint32_t cp) { return U[L.cpiece_u + CPIECE_U * cp + 3]; }

// The line's point at piece parameter s in the face's parameters: the knot held, the
// other parameter over the span's range.
template <class T> __device__ inline void crease_uv(const ICAM_K T* sb, uint32_t k, T s, T& u, T& v) {
  if (k == 0)
    u = sb[65], v = sb[66] + s * (sb[67] - sb[66]);
  else
    u = sb[64] + s * (sb[65] - sb[64]), v = sb[67];
}

// Whether the line's point at piece parameter s lies inside its face's loops (m: the
// margin counted as near).
template <class T>
__device__ inline bool crease_inside(const Tab<T>& tb, const ICAM_K uint32_t* eu, uint32_t cp, T s, T m, bool& near) {
  const Layout& L = tb.L;
  const uint32_t fi = eu[1];
  const ICAM_K uint32_t* fu = tb.U + L.face_u + FACE_U * fi;
  if (fu[2]) return true;
  const ICAM_K T* sb = tb.F + L.span_f + SPAN_F * crease_before(L, tb.U, cp);
  T u, v;
  crease_uv(sb, eu[6], s, u, v);
  return trim_inside(tb, tb.F + L.face_f + FACE_F * fi + 17, fu[1], u, v, m, near);
}

// solid3.rs:1878-1885: the line's sign vector at piece parameter s: the unit normals of
// the span before it at its far side and of the span after it at its near side, summed,
// times the face's sense.
template <class T> __device__ v3<T> crease_pn(const Tab<T>& tb, const ICAM_K uint32_t* eu, uint32_t cp, T s) {
  const Layout& L = tb.L;
  const uint32_t fi = eu[1], k = eu[6];
  const ICAM_K uint32_t* fu = tb.U + L.face_u + FACE_U * fi;
  const uint32_t before = crease_before(L, tb.U, cp), after = before + (k == 0 ? fu[5] : 1u);
  const SEval<T> eb = k == 0 ? peval(tb.F + L.span_f + SPAN_F * before, T(1), s) : peval(tb.F + L.span_f + SPAN_F * before, s, T(1));
  const SEval<T> ea = k == 0 ? peval(tb.F + L.span_f + SPAN_F * after, T(0), s) : peval(tb.F + L.span_f + SPAN_F * after, s, T(0));
  const v3<T> nb = cross(eb.ps, eb.pt), na = cross(ea.ps, ea.pt);
  const T lb = len(nb), la = len(na);
  const v3<T> zero = mk3(T(0), T(0), T(0));
  return ((lb > T(0) ? nb * (T(1) / lb) : zero) + (la > T(0) ? na * (T(1) / la) : zero)) * tb.F[L.face_f + FACE_F * fi + 14];
}

// An edge's sign vector at its point q (the two faces' outward normals summed).
template <class T> __device__ v3<T> edge_pn(const Tab<T>& tb, uint32_t e, v3<T> q, uint32_t piece, T tt) {
  const Layout& L = tb.L;
  const ICAM_K uint32_t* eu = tb.U + L.edge_u + EDGE_U * e;
  if (eu[0] == EDGE_CREASE) return crease_pn(tb, eu, eu[4] + piece, tt);
  const ICAM_K T* uvm = eu[3] != 0xffffffffu ? tb.F + L.uvmap_f + eu[3] + 8 * piece : nullptr;
  v3<T> n0 = face_normal_at(tb, eu[1], q, uvm, tt);
  v3<T> n1 = face_normal_at(tb, eu[2], q, uvm ? uvm + 4 : nullptr, tt);
  return n0 + n1;
}

template <class T> __device__ inline T sgn_of(T x) { return x < T(0) ? T(-1) : T(1); }

template <class T> __device__ inline void offer(Best<T>& b, T d, uint32_t code, T pa, T pb, T s) {
  if (d < b.d) {
    b.d2 = b.d, b.sgn2 = b.sgn;
    b.d = d, b.code = code, b.a = pa, b.b = pb, b.sgn = s;
  } else if (d < b.d2) {
    b.d2 = d, b.sgn2 = s;
  }
}

// solid3.rs:1911-1925 / gpu.wgsl:415-440: an analytic face's foot inside the face
template <class T> __device__ void try_face(const Tab<T>& tb, uint32_t fi, v3<T> p, Best<T>& b, T tm) {
  const Layout& L = tb.L;
  const ICAM_K T* f = tb.F + L.face_f + FACE_F * fi;
  const ICAM_K uint32_t* fu = tb.U + L.face_u + FACE_U * fi;
  uint32_t kind = fu[0];
  v3<T> q = foot(kind, f, p);
  T d = len(p - q);
  if (!(d < b.d)) return;
  // a cone's apex and a spindle torus's tips (lemon, apple: |major| <= minor, on the
  // axis at +-sqrt(minor^2 - major^2)) are singular: the edges, the vertices and the
  // packed tips (point records) answer there (solid3.rs:1915-1918, 1991-2005)
  const T stol = sizeof(T) == 4 ? T(1e-6) : T(1e-12);
  if (kind == ICAM_CONE) {
    T ha = -f[12] / mtan(f[13]);
    v3<T> apex = ld3(f) + ld3(f + 9) * ha;
    if (len(apex - q) <= stol * (T(1) + len(apex))) return;
  }
  if (kind == ICAM_TORUS && mabs(f[12]) <= f[13]) {
    const T h = msqrt(mmax(f[13] * f[13] - f[12] * f[12], T(0)));
    const v3<T> o = ld3(f), z = ld3(f + 9), t0 = o + z * h, t1 = o - z * h;
    if (len(t0 - q) <= stol * (T(1) + len(t0)) || len(t1 - q) <= stol * (T(1) + len(t1))) return;
  }
  bool inside = true;
  if (!fu[2]) {
    T u, v;
    face_uv(kind, fu[3], f, q, u, v);
    bool near = false;
    inside = trim_inside(tb, f + 17, fu[1], u, v, tm, near);
    if (near) b.unsure = true;
  }
  if (!inside) return;
  v3<T> n = snormal(kind, f, q) * f[14];
  offer(b, d, (P_FACE << 28) | fi, T(0), T(0), sgn_of(dot(p - q, n)));
}

// solid3.rs:1775-1782, curve3.rs:525-562: a line or arc edge's interior foot
template <class T> __device__ void try_edge(const Tab<T>& tb, uint32_t e, v3<T> p, Best<T>& b) {
  const Layout& L = tb.L;
  const ICAM_K T* r = tb.F + L.edge_f + EDGE_F * e;
  const ICAM_K uint32_t* eu = tb.U + L.edge_u + EDGE_U * e;
  v3<T> q;
  T tt;
  if (eu[0] == ICAM_LINE) {
    v3<T> a = ld3(r), ab = ld3(r + 3) - a;
    T l2 = dot(ab, ab);
    T t = l2 > T(0) ? dot(p - a, ab) / l2 : T(0);
    if (!(t > T(0) && t < T(1))) return;
    q = a + ab * t;
    tt = t;
  } else {
    v3<T> o = ld3(r), x = ld3(r + 3), y = ld3(r + 6);
    v3<T> dd = p - o;
    T lx = dot(dd, x), ly = dot(dd, y);
    if (lx == T(0) && ly == T(0)) return;
    T t = matan2(ly, lx);
    T t0 = r[10], sweep = r[11];
    T frac = T(0);
    if (r[12] == T(0)) {
      const T TAU = T(6.283185307179586);
      T dlt = sweep >= T(0) ? t - t0 : t0 - t;
      dlt = dlt - TAU * mfloor(dlt / TAU);
      if (!(dlt <= mabs(sweep))) return;
      frac = dlt / mabs(sweep);
    }
    T rho = msqrt(lx * lx + ly * ly);
    q = o + (x * lx + y * ly) * (r[9] / rho);
    tt = frac;
  }
  T d = len(p - q);
  if (!(d < b.d)) return;
  v3<T> pn = edge_pn(tb, e, q, 0u, tt);
  if (len(pn) < T(1e-6)) b.unsure = true;  // a cusp: the normals cancel
  offer(b, d, (P_EDGE << 28) | e, tt, T(0), sgn_of(dot(p - q, pn)));
}

template <class T> __device__ void try_vert(const Tab<T>& tb, uint32_t vi, v3<T> p, Best<T>& b) {
  const ICAM_K T* r = tb.F + tb.L.vert_f + VERT_F * vi;
  v3<T> x = ld3(r);
  T d = len(p - x);
  if (!(d < b.d)) return;
  offer(b, d, (P_VERT << 28) | vi, T(0), T(0), sgn_of(dot(p - x, ld3(r + 3))));
}

// The critical-point test at a B-spline foot (surface3.rs:564-570): the gradient of
// the squared distance zero to within the precision's rounding.
template <class T> __device__ inline bool is_critical(v3<T> r, v3<T> d, T tol) {
  T l = len(d);
  return mabs(dot(r, d)) <= tol * l;
}

// ------------------------------------------------------------------ settling
// A point past a part's convexity certificate is settled, in double precision, by the
// certified search of the reference's CPU evaluator restated for the part
// (NurbsSurf::closest_trimmed and polish, crates/icam-core/src/surface3.rs:1214-1365;
// a curve piece's search, curve3.rs:408-476): best first over quarters (halves) of
// the part, each read by SETTLE_NEWTON Newton steps from its middle; a critical
// point inside the face answers, a non-critical one is carried across spans by Newton
// first; any point of the face found lowers the point's pruning bound; a quarter is
// dropped when its control points' box, or their slab along the direction from its
// Newton point, keeps it at or past that bound, or at the depth limit, or when the
// squared distance has no critical point on it (surface3.rs:694-731). The pieces of a
// block's points are held in one pool a wave works through together, a piece a lane
// (settle_quarter, settle_half); a point whose pieces do not fit the pool, or that takes
// more than SETTLE_POPS reads a part, is left undecided with its lower bound.
constexpr int SETTLE_DEPTH = 40;     // surface3.rs:1291
constexpr int SETTLE_NEWTON = 12;    // surface3.rs:552
constexpr int SETTLE_POPS = 20000;  // surface3.rs:1230
constexpr double SETTLE_EPS = 1e-9;  // surface3.rs:555

// The homogeneous cubic c over [0, 1] replaced by its piece over [a, b] (de Casteljau:
// the left part at b, then the right part of that at a / b).
__device__ inline void sub_cubic4(double c[4][4], double a, double b) {
  if (b < 1.0) {
    double q[4][4], l[4][4];
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 4; ++k) q[i][k] = c[i][k];
    for (int k = 0; k < 4; ++k) l[0][k] = q[0][k];
    for (int lev = 1; lev <= 3; ++lev) {
      for (int i = 0; i <= 3 - lev; ++i)
        for (int k = 0; k < 4; ++k) q[i][k] = (1.0 - b) * q[i][k] + b * q[i + 1][k];
      for (int k = 0; k < 4; ++k) l[lev][k] = q[0][k];
    }
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 4; ++k) c[i][k] = l[i][k];
  }
  if (a > 0.0) {
    const double t = a / b;
    double q[4][4], r[4][4];
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 4; ++k) q[i][k] = c[i][k];
    for (int k = 0; k < 4; ++k) r[3][k] = q[3][k];
    for (int lev = 1; lev <= 3; ++lev) {
      for (int i = 0; i <= 3 - lev; ++i)
        for (int k = 0; k < 4; ++k) q[i][k] = (1.0 - t) * q[i][k] + t * q[i + 1][k];
      for (int k = 0; k < 4; ++k) r[3 - lev][k] = q[3 - lev][k];
    }
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 4; ++k) c[i][k] = r[i][k];
  }
}

// The homogeneous net of a span over [s0, s1] x [t0, t1] (i along s, j along t).
__device__ inline void sub_net(const ICAM_K double* sp, double s0, double s1, double t0, double t1, double n[4][4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) n[i][j][k] = sp[16 * i + 4 * j + k];
  for (int i = 0; i < 4; ++i) sub_cubic4(n[i], t0, t1);
  for (int j = 0; j < 4; ++j) {
    double c[4][4];
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 4; ++k) c[i][k] = n[i][j][k];
    sub_cubic4(c, s0, s1);
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 4; ++k) n[i][j][k] = c[i][k];
  }
}

// p's distance to the box of the control points, and their least height over p along
// dir (the slab bound, surface3.rs:973-975), of n points of homogeneous net c.
__device__ inline void hull_bounds(const double (*c)[4], int n, v3<double> p, v3<double> dir, double& box, double& slab) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  slab = 1e300;
  for (int k = 0; k < n; ++k) {
    const double iw = 1.0 / c[k][3];
    const double x[3] = {c[k][0] * iw, c[k][1] * iw, c[k][2] * iw};
    for (int a = 0; a < 3; ++a) lo[a] = fmin(lo[a], x[a]), hi[a] = fmax(hi[a], x[a]);
    slab = fmin(slab, (p.x - x[0]) * dir.x + (p.y - x[1]) * dir.y + (p.z - x[2]) * dir.z);
  }
  const double dx = fmax(fmax(lo[0] - p.x, p.x - hi[0]), 0.0), dy = fmax(fmax(lo[1] - p.y, p.y - hi[1]), 0.0),
               dz = fmax(fmax(lo[2] - p.z, p.z - hi[2]), 0.0);
  box = sqrt(dx * dx + dy * dy + dz * dz);
}

// surface3.rs:564-570 critical: the gradient of the squared distance zero to within
// the rounding of the coordinates involved
__device__ inline bool critical64(v3<double> p, v3<double> q, v3<double> r, v3<double> d) {
  auto big = [](v3<double> x) { return fmax(fabs(x.x), fmax(fabs(x.y), fabs(x.z))); };
  const double round = 1024.0 * 2.220446049250313e-16 * (big(q) + big(p) + len(r));
  const double l = len(d);
  return fabs(dot(r, d)) <= (round + 1e-12 * l) * l;
}

// surface3.rs:1323-1365 polish: Newton on whole spans of face fi from (u, v), carried
// into the neighbouring span while the gradient points over a border, up to 16 hops.
__device__ inline bool polish64(const Tab<double>& tb, uint32_t fi, v3<double> p, double u, double v, double& dist, double& ou, double& ov,
                                double& side) {
  const Layout& L = tb.L;
  const ICAM_K uint32_t* fu = tb.U + L.face_u + FACE_U * fi;
  const uint32_t span0 = fu[3], nsu = fu[4], nsv = fu[5];
  uint32_t a = 0, b = 0;
  for (uint32_t k = 1; k < nsu; ++k)
    if (tb.F[L.span_f + SPAN_F * (span0 + k * nsv) + 64] <= u) a = k;
  for (uint32_t k = 1; k < nsv; ++k)
    if (tb.F[L.span_f + SPAN_F * (span0 + k) + 66] <= v) b = k;
  const ICAM_K double* sp = tb.F + L.span_f + SPAN_F * (span0 + a * nsv + b);
  double s = (u - sp[64]) / (sp[65] - sp[64]), t = (v - sp[66]) / (sp[67] - sp[66]);
  for (int hop = 0; hop < 16; ++hop) {
    double ns = mclamp(s, 0.0, 1.0), nt = mclamp(t, 0.0, 1.0);
    newton2(sp, p, ns, nt, 0.0, 1.0, 0.0, 1.0, 40, 1e-13);
    SEval<double> e = peval(sp, ns, nt);
    v3<double> r = e.p - p;
    if (critical64(p, e.p, r, e.ps) && critical64(p, e.p, r, e.pt)) {
      dist = len(r), ou = sp[64] + ns * (sp[65] - sp[64]), ov = sp[66] + nt * (sp[67] - sp[66]);
      const v3<double> n = cross(e.ps, e.pt);
      if (!(len(n) > 0.0)) return false;
      side = sgn_of(-dot(r, n));
      return true;
    }
    const double gs = dot(r, e.ps), gt = dot(r, e.pt);
    bool moved = false;
    uint32_t na = a, nb = b;
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
    sp = tb.F + L.span_f + SPAN_F * (span0 + a * nsv + b);
  }
  return false;
}

// Interval tests of |S - p|^2 over a part for one point (gpu.wgsl:1508-1614), in the
// precision of the search: the rounding allowance is the reference kernel's in single
// precision and a much smaller one in double (the stored bounds are sums of a few
// products, each off by an ulp or two).
template <class T> struct IvTol;
template <> struct IvTol<float> {
  static constexpr float rel = 1e-5f, mag = 2e-6f, tiny = 1e-30f, guard = 1e-4f;
};
template <> struct IvTol<double> {
  static constexpr double rel = 1e-12, mag = 1e-13, tiny = 1e-300, guard = 1e-9;
};
// The interval of (S - p).D with S - p in [el, eh] and D's components in d (x lo, x hi,
// y lo, y hi, z lo, z hi).
template <class T, class P> __device__ inline void dot_iv(const T el[3], const T eh[3], P d, T& lo, T& hi) {
  T rl = T(0), rh = T(0), mg = T(0);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const T q0 = el[k] * d[2 * k], q1 = el[k] * d[2 * k + 1], q2 = eh[k] * d[2 * k], q3 = eh[k] * d[2 * k + 1];
    rl += mmin(mmin(q0, q1), mmin(q2, q3));
    rh += mmax(mmax(q0, q1), mmax(q2, q3));
    mg += mmax(mmax(mabs(q0), mabs(q1)), mmax(mabs(q2), mabs(q3)));
  }
  const T m = IvTol<T>::rel * (mabs(rl) + mabs(rh)) + IvTol<T>::mag * mg + IvTol<T>::tiny;
  lo = rl - m, hi = rh + m;
}
// The greatest |D|^2 over a part.
template <class T, class P> __device__ inline T mag2_hi(P d) {
  T m = T(0);
#pragma unroll
  for (int k = 0; k < 3; ++k) m += mmax(d[2 * k] * d[2 * k], d[2 * k + 1] * d[2 * k + 1]);
  return m * (T(1) + IvTol<T>::rel);
}
// What a sub-patch (bounds cb: lambda_min(G), then S_s, S_t, S_ss, S_st, S_tt) is for p,
// its points within the float box: 1 no interior minimum (a part of the gradient keeps
// one sign, or a diagonal entry of the Hessian is negative throughout: its least point is
// on its border, which its neighbours and the face's edges answer); 2 |S - p|^2 strictly
// convex on it (lambda_min(G) + lambda_min([(S - p).S_ij]) > 0); 0 neither
// (gpu.wgsl:1581-1600).
template <class T, class P> __device__ inline int patch_kind(P cb, const T el[3], const T eh[3]) {
  T lo, hi;
  dot_iv(el, eh, cb + 1, lo, hi);
  if (lo > T(0) || hi < T(0)) return 1;
  dot_iv(el, eh, cb + 7, lo, hi);
  if (lo > T(0) || hi < T(0)) return 1;
  T al, ah, bl, bh, cl, ch;
  dot_iv(el, eh, cb + 13, al, ah);
  dot_iv(el, eh, cb + 19, bl, bh);
  dot_iv(el, eh, cb + 25, cl, ch);
  if (mag2_hi<T>(cb + 1) + ah < T(0) || mag2_hi<T>(cb + 7) + ch < T(0)) return 1;
  const T lg = cb[0], bm = mmax(mabs(bl), mabs(bh));
  const T lm = T(0.5) * ((al + cl) - sqrt((al - cl) * (al - cl) + T(4) * bm * bm));
  return lg > T(0) && lg + lm > IvTol<T>::guard * lg ? 2 : 0;
}
// The box of S - p over a part for the interval tests: in single precision grown by the
// point's own rounding (p is the query rounded to single precision, up to half a unit in
// its last place from the point the answer is for, so a test must hold for every point
// that near) and by an eighth of the search's rounding allowance; exact in double
// precision.
template <class T> __device__ inline void rel_box(const ICAM_K float* box, v3<T> p, T eps, T el[3], T eh[3]) {
  const T w = sizeof(T) == 4 ? T(0.125) * eps + (mabs(p.x) + mabs(p.y) + mabs(p.z)) * T(1.2e-7) : T(0);
  el[0] = T(box[0]) - p.x - w, el[1] = T(box[1]) - p.y - w, el[2] = T(box[2]) - p.z - w;
  eh[0] = T(box[3]) - p.x + w, eh[1] = T(box[4]) - p.y + w, eh[2] = T(box[5]) - p.z + w;
}
template <class T> __device__ inline int leaf_kind(const ICAM_K T* cb, const ICAM_K float* box, v3<T> p, T eps) {
  T el[3], eh[3];
  rel_box(box, p, eps, el, eh);
  return patch_kind<T>(cb, el, eh);
}
// The same for a curve part (bounds cb: |C'|^2 from below, then C', C''; gpu.wgsl:1602-1614).
template <class T, class P> __device__ inline int piece_kind(P cb, const T el[3], const T eh[3]) {
  T lo, hi;
  dot_iv(el, eh, cb + 1, lo, hi);
  if (lo > T(0) || hi < T(0)) return 1;
  T kl, kh;
  dot_iv(el, eh, cb + 7, kl, kh);
  if (mag2_hi<T>(cb + 1) + kh < T(0)) return 1;
  const T l = cb[0];
  return l > T(0) && l + kl > IvTol<T>::guard * l ? 2 : 0;
}
template <class T> __device__ inline int cleaf_kind(const ICAM_K T* cb, const ICAM_K float* box, v3<T> p, T eps) {
  T el[3], eh[3];
  rel_box(box, p, eps, el, eh);
  return piece_kind<T>(cb, el, eh);
}

// The settle's work: pieces of the parts past their certificates, a quarter of a
// sub-patch or a half of a curve part, each with its parameter range as a dyadic step
// of its part's own range (steps ia / 2^da .. (ia + 1) / 2^da of the range along s,
// likewise along t). A part's range is itself a dyadic piece of its span's [0, 1], so
// the bounds are exact and equal the repeated midpoints of the certified search.
struct SPiece {
  double lb;       // lower bound on the point's distance over the piece
  uint64_t ia, ib;
  uint32_t part;   // sub-patch, or curve part with SP_CURVE set
  uint32_t meta;   // bits 0-5 the point's slot, 6-11 depth, 12-17 da, 18-23 db
};
constexpr uint32_t SP_CURVE = 0x80000000u;
__device__ inline int sp_slot(uint32_t m) { return int(m & 63u); }
__device__ inline int sp_depth(uint32_t m) { return int((m >> 6) & 63u); }
__device__ inline int sp_da(uint32_t m) { return int((m >> 12) & 63u); }
__device__ inline int sp_db(uint32_t m) { return int((m >> 18) & 63u); }
__device__ inline uint32_t sp_meta(int slot, int depth, int da, int db) {
  return uint32_t(slot) | (uint32_t(depth) << 6) | (uint32_t(da) << 12) | (uint32_t(db) << 18);
}
__device__ inline void sp_range(double lo, double hi, uint64_t i, int d, double& a0, double& a1) {
  const double w = ldexp(hi - lo, -d);
  a0 = lo + double(i) * w, a1 = lo + double(i + 1) * w;
}
// The pool pieces go to (workgroup memory) and its length, which may run past `cap`:
// a piece past it is not kept, and the caller keeps its lower bound instead.
struct SPool {
  SPiece* items;
  int* count;
  int cap;
};
__device__ inline bool sp_push(const SPool& pl, const SPiece& x) {
  const int k = atomicAdd(pl.count, 1);
  if (k >= pl.cap) return false;
  pl.items[k] = x;
  return true;
}
// A search that meets a part past its certificate (double precision) hands it to the
// settle as a root piece of point `slot`: its lower bound rounded down to single
// precision, the part and the slot, in workgroup memory (`count` may run past `cap`).
struct SSink {
  float* lb;
  uint32_t* part;
  uint8_t* slot_of;
  int* count;
  int cap;
  int slot;
  uint8_t* over;  // per slot: a part was not taken
  double ub;      // the least distance of a point of the part's face (edge) its probe found
};
// What one piece's step found for its point.
struct SOut {
  double sd;    // least distance of a point of the face (edge) found: a bound on the answer
  double lost;  // least lower bound of the pieces the pool could not keep
  double d, pa, pb, sgn;  // the answer, when `ans`
  uint32_t code;
  bool ans, unsure;
  bool over;    // a piece was not kept
};
// Part `part` (past its certificate, lower bound lb) handed to the settle; kept in the
// point's undecided bound when the settle cannot take it.
__device__ inline void settle_root(SSink* sk, double lb, uint32_t part, Best<double>& b) {
  if (sk) {
    const int k = atomicAdd(sk->count, 1);
    if (k < sk->cap) {
      sk->lb[k] = __double2float_rd(lb), sk->part[k] = part, sk->slot_of[k] = uint8_t(sk->slot);
      return;
    }
    sk->over[sk->slot] = 1;
  }
  b.lmin = fmin(b.lmin, lb);
}
// The first step of a part's settle taken when the search meets it: Newton from its
// middle. A point of the face found there bounds the point's answer (sk->ub: the search
// passes over parts at or past it, as the settle prunes its pieces), and a critical one
// answers. Without it the search, past parts it could not read, would read every part
// under the answers of the parts it could.
__device__ inline void settle_probe_leaf(const Tab<double>& tb, uint32_t li, v3<double> p, Best<double>& b, SSink* sk) {
  if (!sk) return;
  const Layout& L = tb.L;
  const ICAM_K double* lf = tb.F + L.leaf_f + LEAF_F * li;
  const ICAM_K uint32_t* lu = tb.U + L.leaf_u + LEAF_U * li;
  const ICAM_K double* sp = tb.F + L.span_f + SPAN_F * lu[0];
  const uint32_t fi = lu[1];
  const ICAM_K double* f = tb.F + L.face_f + FACE_F * fi;
  const ICAM_K uint32_t* fu = tb.U + L.face_u + FACE_U * fi;
  double s = 0.5 * (lf[1] + lf[2]), t = 0.5 * (lf[3] + lf[4]);
  newton2(sp, p, s, t, lf[1], lf[2], lf[3], lf[4], SETTLE_NEWTON, 1e-13 * (lf[2] - lf[1]));
  const SEval<double> e = peval(sp, s, t);
  const v3<double> r = e.p - p;
  const double d = len(r);
  if (!(d < fmin(b.d, sk->ub))) return;
  const double u = sp[64] + s * (sp[65] - sp[64]), v = sp[66] + t * (sp[67] - sp[66]);
  bool near = false;
  if (!(lu[2] != 0 || fu[2] != 0 || trim_inside(tb, f + 17, fu[1], u, v, 0.0, near))) return;
  sk->ub = d;
  const v3<double> nrm = cross(e.ps, e.pt);
  const bool at_pole = lf[LEAF_POLE + 3] != 0.0 && len(e.p - ld3(lf + LEAF_POLE)) <= 1e-9 * (1.0 + len(ld3(lf + LEAF_POLE)));
  if (critical64(p, e.p, r, e.ps) && critical64(p, e.p, r, e.pt) && len(nrm) > 0.0 && !at_pole)
    offer(b, d, (P_LEAF << 28) | li, s, t, sgn_of(dot(r, nrm) * -f[14]));
}
// The same for curve part ci: a point of an edge is a point of the boundary (of a C0
// line, when inside its face's loops).
__device__ inline void settle_probe_curve(const Tab<double>& tb, uint32_t ci, v3<double> p, Best<double>& b, SSink* sk) {
  if (!sk) return;
  const Layout& L = tb.L;
  const ICAM_K double* cf = tb.F + L.cleaf_f + CLEAF_F * ci;
  const ICAM_K uint32_t* cu = tb.U + L.cleaf_u + CLEAF_U * ci;
  const uint32_t cp = cu[0];
  const ICAM_K double* pw = tb.F + L.cpiece_f + CPIECE_F * cp;
  const ICAM_K uint32_t* pu = tb.U + L.cpiece_u + CPIECE_U * cp;
  const ICAM_K uint32_t* eu = tb.U + L.edge_u + EDGE_U * pu[0];
  const bool crease = eu[0] == EDGE_CREASE;
  const double s = newton1(pw, p, 0.5 * (cf[1] + cf[2]), cf[1], cf[2], 40, 1e-15 * (cf[2] - cf[1]));
  const CEval<double> e = ceval(pw, s);
  const v3<double> r = e.p - p;
  const double d = len(r);
  if (!(d < fmin(b.d, sk->ub))) return;
  bool near = false;
  if (crease && !crease_inside(tb, eu, cp, s, 0.0, near)) return;
  sk->ub = d;
  const bool end = !crease && ((pu[1] == 0 && s <= 0.0) || (pu[1] + 1 == pu[2] && s >= 1.0));
  const bool junction = !crease && ((pu[1] > 0 && s <= 0.0) || (pu[1] + 1 < pu[2] && s >= 1.0));
  if (!end && (junction || critical64(p, e.p, r, e.d1))) {
    const v3<double> pn = edge_pn(tb, pu[0], e.p, pu[1], s);
    if (len(pn) < 1e-6) b.unsure = true;
    offer(b, d, (P_CLEAF << 28) | ci, s, 0.0, sgn_of(dot(p - e.p, pn)));
  }
}

// Interval arithmetic for the settle's per-part tests (the bounds the packer keeps for
// a leaf, made here for a quarter from its own net; surface3.rs:866-970).
struct Iv {
  double a, b;
};
__device__ inline Iv iv_hull(Iv r, double x) { return Iv{fmin(r.a, x), fmax(r.b, x)}; }
__device__ inline Iv iv_sub(Iv x, Iv y) { return Iv{x.a - y.b, x.b - y.a}; }
__device__ inline Iv iv_mul(Iv x, Iv y) {
  const double p0 = x.a * y.a, p1 = x.a * y.b, p2 = x.b * y.a, p3 = x.b * y.b;
  return Iv{fmin(fmin(p0, p1), fmin(p2, p3)), fmax(fmax(p0, p1), fmax(p2, p3))};
}
__device__ inline Iv iv_scale(Iv x, double c) { return c >= 0.0 ? Iv{x.a * c, x.b * c} : Iv{x.b * c, x.a * c}; }
__device__ inline Iv iv_divpos(Iv x, Iv w) {  // w.a > 0
  const double p0 = x.a / w.a, p1 = x.a / w.b, p2 = x.b / w.a, p3 = x.b / w.b;
  return Iv{fmin(fmin(p0, p1), fmin(p2, p3)), fmax(fmax(p0, p1), fmax(p2, p3))};
}
__device__ inline double iv_mig(Iv x) { return x.a > 0.0 ? x.a : (x.b < 0.0 ? -x.b : 0.0); }
__device__ inline double iv_mag(Iv x) { return fmax(fabs(x.a), fabs(x.b)); }
// Bounds (PATCH_CERT layout: lambda_min(G), S_s, S_t, S_ss, S_st, S_tt) of a homogeneous
// bicubic net, and the box of its points (relative to p: el, eh); false when a weight
// is not positive.
__device__ inline bool net_bounds(const double n[4][4][4], v3<double> p, double cb[31], double el[3], double eh[3]) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, c[3] = {0.0, 0.0, 0.0};
  Iv wt{1e300, -1e300};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const double w = n[i][j][3];
      if (!(w > 0.0)) return false;
      wt = iv_hull(wt, w);
      for (int k = 0; k < 3; ++k) {
        const double x = n[i][j][k] / w;
        lo[k] = fmin(lo[k], x), hi[k] = fmax(hi[k], x), c[k] += x * (1.0 / 16.0);
      }
    }
  const double pc[3] = {p.x, p.y, p.z};
  for (int k = 0; k < 3; ++k) el[k] = lo[k] - pc[k], eh[k] = hi[k] - pc[k];
  // derivative hulls of the coordinates relative to the points' mean
  Iv A[6][4];
  for (int w = 0; w < 6; ++w)
    for (int k = 0; k < 4; ++k) A[w][k] = Iv{1e300, -1e300};
  auto q = [&](int i, int j, int k) { return k < 3 ? n[i][j][k] - c[k] * n[i][j][3] : n[i][j][3]; };
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        if (i < 3) A[1][k] = iv_hull(A[1][k], 3.0 * (q(i + 1, j, k) - q(i, j, k)));
        if (j < 3) A[2][k] = iv_hull(A[2][k], 3.0 * (q(i, j + 1, k) - q(i, j, k)));
        if (i < 2) A[3][k] = iv_hull(A[3][k], 6.0 * (q(i + 2, j, k) - 2.0 * q(i + 1, j, k) + q(i, j, k)));
        if (i < 3 && j < 3) A[4][k] = iv_hull(A[4][k], 9.0 * (q(i + 1, j + 1, k) - q(i + 1, j, k) - q(i, j + 1, k) + q(i, j, k)));
        if (j < 2) A[5][k] = iv_hull(A[5][k], 6.0 * (q(i, j + 2, k) - 2.0 * q(i, j + 1, k) + q(i, j, k)));
      }
  Iv sv[3], ss[3], st[3], sss[3], sst[3], stt[3];
  for (int k = 0; k < 3; ++k) {
    sv[k] = Iv{lo[k] - c[k], hi[k] - c[k]};
    ss[k] = iv_divpos(iv_sub(A[1][k], iv_mul(sv[k], A[1][3])), wt);
    st[k] = iv_divpos(iv_sub(A[2][k], iv_mul(sv[k], A[2][3])), wt);
  }
  for (int k = 0; k < 3; ++k) {
    sss[k] = iv_divpos(iv_sub(iv_sub(A[3][k], iv_scale(iv_mul(ss[k], A[1][3]), 2.0)), iv_mul(sv[k], A[3][3])), wt);
    sst[k] = iv_divpos(iv_sub(iv_sub(iv_sub(A[4][k], iv_mul(ss[k], A[2][3])), iv_mul(st[k], A[1][3])), iv_mul(sv[k], A[4][3])), wt);
    stt[k] = iv_divpos(iv_sub(iv_sub(A[5][k], iv_scale(iv_mul(st[k], A[2][3]), 2.0)), iv_mul(sv[k], A[5][3])), wt);
  }
  // parameters scaled by the tangents' middle lengths (a congruence: definiteness is
  // unchanged, the entrywise bound below is better conditioned)
  double sa = 0.0, sb = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double ma = 0.5 * (ss[k].a + ss[k].b), mb = 0.5 * (st[k].a + st[k].b);
    sa += ma * ma, sb += mb * mb;
  }
  sa = sa > 0.0 ? 1.0 / sqrt(sa) : 1.0, sb = sb > 0.0 ? 1.0 / sqrt(sb) : 1.0;
  double ga = 0.0, gc = 0.0;
  Iv gb{0.0, 0.0};
  for (int k = 0; k < 3; ++k) {
    ss[k] = iv_scale(ss[k], sa), st[k] = iv_scale(st[k], sb);
    sss[k] = iv_scale(sss[k], sa * sa), sst[k] = iv_scale(sst[k], sa * sb), stt[k] = iv_scale(stt[k], sb * sb);
    ga += iv_mig(ss[k]) * iv_mig(ss[k]), gc += iv_mig(st[k]) * iv_mig(st[k]);
    const Iv m = iv_mul(ss[k], st[k]);
    gb.a += m.a, gb.b += m.b;
  }
  const double gbm = iv_mag(gb);
  const double lg = 0.5 * ((ga + gc) - sqrt((ga - gc) * (ga - gc) + 4.0 * gbm * gbm));
  cb[0] = lg - fabs(lg) * 1e-9 - 1e-300;
  const Iv* vs[5] = {ss, st, sss, sst, stt};
  for (int v = 0; v < 5; ++v)
    for (int k = 0; k < 3; ++k) {
      const double pad = 1e-13 * (fabs(vs[v][k].a) + fabs(vs[v][k].b)) + 1e-300;
      cb[1 + 6 * v + 2 * k] = vs[v][k].a - pad, cb[2 + 6 * v + 2 * k] = vs[v][k].b + pad;
    }
  return true;
}
// The same for a homogeneous cubic's net (PIECE_CERT layout: |C'|^2, C', C'').
__device__ inline bool cnet_bounds(const double n[4][4], v3<double> p, double cb[13], double el[3], double eh[3]) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, c[3] = {0.0, 0.0, 0.0};
  Iv wt{1e300, -1e300};
  for (int i = 0; i < 4; ++i) {
    const double w = n[i][3];
    if (!(w > 0.0)) return false;
    wt = iv_hull(wt, w);
    for (int k = 0; k < 3; ++k) {
      const double x = n[i][k] / w;
      lo[k] = fmin(lo[k], x), hi[k] = fmax(hi[k], x), c[k] += 0.25 * x;
    }
  }
  const double pc[3] = {p.x, p.y, p.z};
  for (int k = 0; k < 3; ++k) el[k] = lo[k] - pc[k], eh[k] = hi[k] - pc[k];
  auto q = [&](int i, int k) { return k < 3 ? n[i][k] - c[k] * n[i][3] : n[i][3]; };
  Iv a1[4], a2[4];
  for (int k = 0; k < 4; ++k) {
    a1[k] = Iv{1e300, -1e300}, a2[k] = Iv{1e300, -1e300};
    for (int i = 0; i < 3; ++i) a1[k] = iv_hull(a1[k], 3.0 * (q(i + 1, k) - q(i, k)));
    for (int i = 0; i < 2; ++i) a2[k] = iv_hull(a2[k], 6.0 * (q(i + 2, k) - 2.0 * q(i + 1, k) + q(i, k)));
  }
  Iv d1[3], d2[3];
  double l = 0.0;
  for (int k = 0; k < 3; ++k) {
    const Iv sv{lo[k] - c[k], hi[k] - c[k]};
    d1[k] = iv_divpos(iv_sub(a1[k], iv_mul(sv, a1[3])), wt);
    d2[k] = iv_divpos(iv_sub(iv_sub(a2[k], iv_scale(iv_mul(d1[k], a1[3]), 2.0)), iv_mul(sv, a2[3])), wt);
    l += iv_mig(d1[k]) * iv_mig(d1[k]);
  }
  cb[0] = l - fabs(l) * 1e-9 - 1e-300;
  for (int k = 0; k < 3; ++k) {
    const double p1 = 1e-13 * (fabs(d1[k].a) + fabs(d1[k].b)) + 1e-300, p2 = 1e-13 * (fabs(d2[k].a) + fabs(d2[k].b)) + 1e-300;
    cb[1 + 2 * k] = d1[k].a - p1, cb[2 + 2 * k] = d1[k].b + p1;
    cb[7 + 2 * k] = d2[k].a - p2, cb[8 + 2 * k] = d2[k].b + p2;
  }
  return true;
}

// out (om x on Bernstein coefficients) += s * a (am x an) * b (bm x bn), om = am + bm,
// on = an + bn: the product of tensor Bernstein polynomials (surface3.rs:593-630).
// C(n, k) for n <= 9: the products below run in scaled Bernstein form (a coefficient
// times its binomials), where a product of two polynomials is a plain convolution.
__device__ constexpr double binom(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * double(n - k + i) / double(i);
  return r;
}
// out (degrees AM + BM, AN + BN) += s a b, every polynomial in scaled form.
template <int AM, int AN, int BM, int BN>
__device__ inline void bern_mul_add(const double* a, const double* b, double s, double* out) {
  constexpr int ON = AN + BN;
#pragma unroll
  for (int i = 0; i <= AM; ++i)
#pragma unroll
    for (int j = 0; j <= AN; ++j) {
      const double x = s * a[i * (AN + 1) + j];
#pragma unroll
      for (int k = 0; k <= BM; ++k)
#pragma unroll
        for (int l = 0; l <= BN; ++l) out[(i + k) * (ON + 1) + j + l] += x * b[k * (BN + 1) + l];
    }
}
// Scaled form and back, for an (M, N) net.
template <int M, int N> __device__ inline void bern_scale(double* x) {
#pragma unroll
  for (int i = 0; i <= M; ++i)
#pragma unroll
    for (int j = 0; j <= N; ++j) x[i * (N + 1) + j] *= binom(M, i) * binom(N, j);
}
template <int M, int N> __device__ inline void bern_unscale(double* x) {
#pragma unroll
  for (int i = 0; i <= M; ++i)
#pragma unroll
    for (int j = 0; j <= N; ++j) x[i * (N + 1) + j] *= 1.0 / (binom(M, i) * binom(N, j));
}

// surface3.rs:694-731 no_critical_point: the squared distance from p has no critical
// point on the closed patch (homogeneous bicubic net n) when the gradient's component
// along (a, b), the gradient's direction at the patch's middle, has all its Bernstein
// coefficients clear of zero: with Q = P - p w (homogeneous), the components are
// w Q.Q_s - w_s Q.Q and w Q.Q_t - w_t Q.Q over w^3.
__device__ inline bool no_critical64(const double n[4][4][4], v3<double> p, double a, double b) {
  double q[3][16], w[16], qu[3][12], qv[3][12], wu[12], wv[12];
  const double pc[3] = {p.x, p.y, p.z};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      w[i * 4 + j] = n[i][j][3];
      if (!(n[i][j][3] > 0.0)) return false;
      for (int k = 0; k < 3; ++k) q[k][i * 4 + j] = n[i][j][k] - pc[k] * n[i][j][3];
    }
  // derivatives: along s (2 x 3 degrees, i < 3), along t (3 x 2, j < 3)
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) {
      wu[i * 4 + j] = 3.0 * (w[(i + 1) * 4 + j] - w[i * 4 + j]);
      for (int k = 0; k < 3; ++k) qu[k][i * 4 + j] = 3.0 * (q[k][(i + 1) * 4 + j] - q[k][i * 4 + j]);
    }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) {
      wv[i * 3 + j] = 3.0 * (w[i * 4 + j + 1] - w[i * 4 + j]);
      for (int k = 0; k < 3; ++k) qv[k][i * 3 + j] = 3.0 * (q[k][i * 4 + j + 1] - q[k][i * 4 + j]);
    }
  auto maxabs = [](const double* x, int c) {
    double m = 0.0;
    for (int i = 0; i < c; ++i) m = fmax(m, fabs(x[i]));
    return m;
  };
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;
  for (int k = 0; k < 3; ++k) q0 = fmax(q0, maxabs(q[k], 16)), q1 = fmax(q1, maxabs(qu[k], 12)), q2 = fmax(q2, maxabs(qv[k], 12));
  const double size = maxabs(w, 16) * 3.0 * q0 * (fabs(a) * q1 + fabs(b) * q2) + (fabs(a) * maxabs(wu, 12) + fabs(b) * maxabs(wv, 12)) * 3.0 * q0 * q0;
  // the products in scaled form, then gs (8, 9) and gt (9, 8) back in Bernstein form
  bern_scale<3, 3>(w);
  bern_scale<2, 3>(wu);
  bern_scale<3, 2>(wv);
#pragma unroll
  for (int k = 0; k < 3; ++k) bern_scale<3, 3>(q[k]), bern_scale<2, 3>(qu[k]), bern_scale<3, 2>(qv[k]);
  double qq[49] = {0}, qqu[42] = {0}, qqv[42] = {0};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    bern_mul_add<3, 3, 3, 3>(q[k], q[k], 1.0, qq);
    bern_mul_add<3, 3, 2, 3>(q[k], qu[k], 1.0, qqu);
    bern_mul_add<3, 3, 3, 2>(q[k], qv[k], 1.0, qqv);
  }
  double gs[90] = {0}, gt[90] = {0};
  bern_mul_add<3, 3, 5, 6>(w, qqu, 1.0, gs);
  bern_mul_add<2, 3, 6, 6>(wu, qq, -1.0, gs);
  bern_mul_add<3, 3, 6, 5>(w, qqv, 1.0, gt);
  bern_mul_add<3, 2, 6, 6>(wv, qq, -1.0, gt);
  bern_unscale<8, 9>(gs);
  bern_unscale<9, 8>(gt);
  // f = a (gs raised along s) + b (gt raised along t), 10 x 10 coefficients
  for (int i = 0; i <= 9; ++i)
    for (int j = 0; j <= 9; ++j) {
      const double al = double(i) / 9.0, be = double(j) / 9.0;
      const double su = (i > 0 ? al * gs[(i - 1) * 10 + j] : 0.0) + (i < 9 ? (1.0 - al) * gs[i * 10 + j] : 0.0);
      const double tv = (j > 0 ? be * gt[i * 9 + j - 1] : 0.0) + (j < 9 ? (1.0 - be) * gt[i * 9 + j] : 0.0);
      if (!(a * su + b * tv > 1e-11 * size)) return false;
    }
  return true;
}

// One quarter of sub-patch it.part read for p, a step of the certified search: Newton
// from its middle; a critical point inside the face answers (`o.ans`), a non-critical
// one is carried across spans by Newton first; any point of the face found lowers the
// pruning bound (`o.sd`); the quarter is dropped when its control points' box, or their
// slab along the direction from its Newton point, keeps it at or past that bound, at
// the depth limit, when its own interval test finds no interior minimum or a single one
// (its Newton point), or when the squared distance has no critical point on it;
// otherwise its four children go to the pool (two, split across the side, at a pole).
// `prune` is the point's bound when the step began, `got` whether the settle answered
// for the point already (a critical point within SETTLE_EPS of a bound no answer set
// yet is taken: a foot on the shared border of two quarters stops the first quarter's
// Newton at the border, not critical, which sets the bound that the next quarter's
// critical point only ties).
__device__ inline void settle_quarter(const Tab<double>& tb, const SPiece& it, v3<double> p, double prune, bool got, const SPool& pool, SOut& o) {
  const Layout& L = tb.L;
  const uint32_t li = it.part;
  const ICAM_K double* lf = tb.F + L.leaf_f + LEAF_F * li;
  const ICAM_K uint32_t* lu = tb.U + L.leaf_u + LEAF_U * li;
  const ICAM_K double* sp = tb.F + L.span_f + SPAN_F * lu[0];
  const uint32_t fi = lu[1];
  const ICAM_K double* f = tb.F + L.face_f + FACE_F * fi;
  const ICAM_K uint32_t* fu = tb.U + L.face_u + FACE_U * fi;
  const bool known_in = lu[2] != 0;
  const int slot = sp_slot(it.meta), depth = sp_depth(it.meta), da = sp_da(it.meta), db = sp_db(it.meta);
  double a0, a1, b0, b1;
  sp_range(lf[1], lf[2], it.ia, da, a0, a1);
  sp_range(lf[3], lf[4], it.ib, db, b0, b1);
  double bd = prune;
  double s = 0.5 * (a0 + a1), t = 0.5 * (b0 + b1);
  newton2(sp, p, s, t, a0, a1, b0, b1, SETTLE_NEWTON, 1e-13 * (a1 - a0));
  SEval<double> e = peval(sp, s, t);
  v3<double> r = e.p - p;
  const double d = len(r);
  v3<double> nrm = cross(e.ps, e.pt);
  const double u = sp[64] + s * (sp[65] - sp[64]), v = sp[66] + t * (sp[67] - sp[66]);
  bool near = false;
  auto inside = [&](double uu, double vv) { return known_in || fu[2] != 0 || trim_inside(tb, f + 17, fu[1], uu, vv, 0.0, near); };
  if (d < bd + (got ? 0.0 : SETTLE_EPS) && inside(u, v)) {
    bd = fmin(bd, d);
    o.sd = fmin(o.sd, d);
    const bool at_pole = lf[LEAF_POLE + 3] != 0.0 && len(e.p - ld3(lf + LEAF_POLE)) <= 1e-9 * (1.0 + len(ld3(lf + LEAF_POLE)));
    if (critical64(p, e.p, r, e.ps) && critical64(p, e.p, r, e.pt)) {
      if (len(nrm) > 0.0 && !at_pole) o.ans = true, o.d = d, o.code = (P_LEAF << 28) | li, o.pa = s, o.pb = t, o.sgn = sgn_of(dot(r, nrm) * -f[14]);
    } else {
      double hd, hu, hv, hs;
      if (polish64(tb, fi, p, u, v, hd, hu, hv, hs) && hd <= d + SETTLE_EPS && inside(hu, hv)) {
        bd = fmin(bd, hd);
        o.sd = fmin(o.sd, hd);
        o.ans = true, o.d = hd, o.code = (P_LEAF << 28) | li, o.pa = s, o.pb = t, o.sgn = hs * f[14];
      }
    }
  }
  double net[4][4][4];
  sub_net(sp, a0, a1, b0, b1, net);
  const v3<double> dir = d > 1e-12 ? (p - e.p) * (1.0 / d) : (len(nrm) > 0.0 ? nrm * (1.0 / len(nrm)) : mk3(0.0, 0.0, 1.0));
  double box, slab;
  hull_bounds(&net[0][0], 16, p, dir, box, slab);
  if (slab >= bd - SETTLE_EPS || depth >= SETTLE_DEPTH) return;
  auto child = [&](uint64_t ia, int cda, uint64_t ib, int cdb) {
    double c0, c1, e0, e1, cn[4][4][4], cb, cs;
    sp_range(lf[1], lf[2], ia, cda, c0, c1);
    sp_range(lf[3], lf[4], ib, cdb, e0, e1);
    sub_net(sp, c0, c1, e0, e1, cn);
    hull_bounds(&cn[0][0], 16, p, dir, cb, cs);
    const double lb = fmax(cb, slab);
    if (lb < bd - SETTLE_EPS && !sp_push(pool, SPiece{lb, ia, ib, li, sp_meta(slot, depth + 1, cda, cdb)})) o.lost = fmin(o.lost, lb), o.over = true;
  };
  // a quarter touching a collapsed side (a pole, answered by its own point record):
  // never cleared by the critical-point test (surface3.rs:1209) and split across the
  // side only, so the quarters at the pole stay one chain instead of doubling
  const int pside = int(lf[LEAF_POLE + 3]);
  const bool touch = (pside == 1 && a0 == 0.0) || (pside == 2 && a1 == 1.0) || (pside == 3 && b0 == 0.0) || (pside == 4 && b1 == 1.0);
  if (touch) {
    for (int c = 0; c < 2; ++c) {
      if (pside >= 3)
        child(it.ia, da, 2 * it.ib + uint64_t(c), db + 1);
      else
        child(2 * it.ia + uint64_t(c), da + 1, it.ib, db);
    }
    return;
  }
  {
    // the quarter's own interval test (gpu.wgsl:1581-1600 on its net): no interior
    // minimum drops it; strictly convex for p, its minimum is the Newton point above
    double cb[31], el[3], eh[3];
    if (net_bounds(net, p, cb, el, eh) && patch_kind<double>((const double*)cb, el, eh) != 0) return;
  }
  {
    // no critical point on the quarter: no minimum of its own to find there
    SEval<double> em = peval(sp, 0.5 * (a0 + a1), 0.5 * (b0 + b1));
    const v3<double> rm = em.p - p;
    const double ga = dot(em.ps, rm) * (a1 - a0), gb = dot(em.pt, rm) * (b1 - b0), gl = sqrt(ga * ga + gb * gb);
    if (gl > 0.0 && no_critical64(net, p, ga / gl, gb / gl)) return;
  }
  for (int c = 0; c < 4; ++c) child(2 * it.ia + uint64_t(c & 1), da + 1, 2 * it.ib + uint64_t((c >> 1) & 1), db + 1);
}

// gpu.wgsl:679-756 / solid3.rs:1904-1910: a sub-patch read by Newton when its radius
// certifies a single minimum for p; the minimum answers when it is a critical point
// inside the face. With `settle` (double precision), a sub-patch past its certificate
// is handed to the settle (sk) instead of leaving its lower bound.
template <class T, bool SETTLE = false>
__device__ void try_leaf(const Tab<T>& tb, uint32_t li, const ICAM_K float* box, T lb, v3<T> p, Best<T>& b, T tm, T eps, SSink* sk = nullptr) {
  const Layout& L = tb.L;
  const ICAM_K T* lf = tb.F + L.leaf_f + LEAF_F * li;
  const ICAM_K uint32_t* lu = tb.U + L.leaf_u + LEAF_U * li;
  T rho = lf[0];
  T fd = far_dist(p, box);
  bool certified = fd < T(0.999) * rho - eps;
  if (!certified) {
    // past its radius: the interval test for this point (gpu.wgsl:1677-1683)
    const int kind = leaf_kind(lf + LEAF_CERT, box, p, eps);
    if (kind == 1) return;
    certified = kind == 2;
  }
  if constexpr (SETTLE && sizeof(T) == 8) {
    if (!certified) {
      settle_probe_leaf(tb, li, p, b, sk);
      settle_root(sk, lb, li, b);
      return;
    }
  }
  const ICAM_K T* sp = tb.F + L.span_f + SPAN_F * lu[0];
  T s0 = lf[1], s1 = lf[2], t0 = lf[3], t1 = lf[4];
  T s = T(0.5) * (s0 + s1), t = T(0.5) * (t0 + t1);
  const bool single = sizeof(T) == 4;
  const bool conv = newton2(sp, p, s, t, s0, s1, t0, t1, single ? 16 : 32, single ? T(1e-5) : T(1e-14));
  SEval<T> e = peval1(sp, s, t);
  v3<T> r = e.p - p;
  T d = len(r);
  if (!certified) b.lmin = mmin(b.lmin, lb);
  T tol = single ? T(1e-4) * d + eps : T(1e-9) * (T(1) + d);
  const bool crit = is_critical(r, e.ps, tol) && is_critical(r, e.pt, tol);
  // a certified part whose single-precision Newton neither settled within its 16 steps
  // nor rests on a critical point: its one minimum is not read, and another part's answer
  // must not stand for it, so the point is searched again in double precision (32 steps;
  // there a step under 1e-14 is not always reached in rounding, and the point it rests on
  // is the minimum)
  if (single && certified && !conv && !crit && lb < b.d) b.unsure = true;
  if (!(d < b.d)) return;
  // a foot at a pole the sub-patch touches: the pole's point record answers there
  // (its normal comes from the fan; the surface's own degenerates)
  if (lf[LEAF_POLE + 3] != T(0)) {
    const v3<T> x = ld3(lf + LEAF_POLE);
    if (len(e.p - x) <= (sizeof(T) == 4 ? T(1e-6) : T(1e-9)) * (T(1) + len(x))) return;
  }
  if (!crit) return;
  const uint32_t fi = lu[1];
  const ICAM_K T* f = tb.F + L.face_f + FACE_F * fi;
  if (!lu[2]) {
    const ICAM_K uint32_t* fu = tb.U + L.face_u + FACE_U * fi;
    T u = sp[64] + s * (sp[65] - sp[64]), v = sp[66] + t * (sp[67] - sp[66]);
    bool near = false;
    bool in = trim_inside(tb, f + 17, fu[1], u, v, tm, near);
    if (near) b.unsure = true;
    if (!in) return;
  }
  v3<T> n = cross(e.ps, e.pt);
  T nl = len(n);
  if (!(nl > T(0))) return;
  offer(b, d, (P_LEAF << 28) | li, s, t, sgn_of(dot(r, n) * -f[14]));
}

// One half of curve part it.part read for p, a step of a curve piece's certified
// search: Newton from its middle; a critical point interior to its edge answers (a
// point where two pieces of the edge meet answers critical or not; a C0 line's point
// only inside its face's loops); the half is dropped at the depth limit or when its own
// interval test finds no interior minimum or a single one; otherwise its two halves go
// to the pool when their hull can come under the bound.
__device__ inline void settle_half(const Tab<double>& tb, const SPiece& it, v3<double> p, double prune, const SPool& pool, SOut& o) {
  const Layout& L = tb.L;
  const uint32_t ci = it.part & ~SP_CURVE;
  const ICAM_K double* cf = tb.F + L.cleaf_f + CLEAF_F * ci;
  const ICAM_K uint32_t* cu = tb.U + L.cleaf_u + CLEAF_U * ci;
  const uint32_t cp = cu[0];
  const ICAM_K double* pw = tb.F + L.cpiece_f + CPIECE_F * cp;
  const ICAM_K uint32_t* pu = tb.U + L.cpiece_u + CPIECE_U * cp;
  const ICAM_K uint32_t* eu = tb.U + L.edge_u + EDGE_U * pu[0];
  const bool crease = eu[0] == EDGE_CREASE;
  const int slot = sp_slot(it.meta), depth = sp_depth(it.meta), da = sp_da(it.meta);
  double a0, a1;
  sp_range(cf[1], cf[2], it.ia, da, a0, a1);
  double bd = prune;
  const double s = newton1(pw, p, 0.5 * (a0 + a1), a0, a1, 40, 1e-15 * (a1 - a0));
  CEval<double> e = ceval(pw, s);
  v3<double> r = e.p - p;
  const double d = len(r);
  if (d < bd) {
    bd = d;
    o.sd = fmin(o.sd, d);
    // an edge's ends are its vertices', a junction of two of its pieces answers critical
    // or not; a C0 line's point answers inside its face's loops
    const bool end = !crease && ((pu[1] == 0 && s <= 0.0) || (pu[1] + 1 == pu[2] && s >= 1.0));
    const bool junction = !crease && ((pu[1] > 0 && s <= 0.0) || (pu[1] + 1 < pu[2] && s >= 1.0));
    bool near = false;
    if (!end && (junction || critical64(p, e.p, r, e.d1)) && (!crease || crease_inside(tb, eu, cp, s, 0.0, near))) {
      v3<double> pn = edge_pn(tb, pu[0], e.p, pu[1], s);
      if (len(pn) < 1e-6) o.unsure = true;
      o.ans = true, o.d = d, o.code = (P_CLEAF << 28) | ci, o.pa = s, o.pb = 0.0, o.sgn = sgn_of(dot(p - e.p, pn));
    }
  }
  if (depth >= SETTLE_DEPTH) return;
  {
    // the half's own interval test (gpu.wgsl:1602-1614 on its net)
    double cc[4][4], cb[13], el[3], eh[3];
    for (int i = 0; i < 4; ++i)
      for (int kk = 0; kk < 4; ++kk) cc[i][kk] = pw[4 * i + kk];
    sub_cubic4(cc, a0, a1);
    if (cnet_bounds(cc, p, cb, el, eh) && piece_kind<double>((const double*)cb, el, eh) != 0) return;
  }
  const v3<double> dir = d > 1e-12 ? (p - e.p) * (1.0 / d) : mk3(0.0, 0.0, 1.0);
  for (int c = 0; c < 2; ++c) {
    const uint64_t ia = 2 * it.ia + uint64_t(c);
    double c0, c1, cc[4][4], cb, cs;
    sp_range(cf[1], cf[2], ia, da + 1, c0, c1);
    for (int i = 0; i < 4; ++i)
      for (int kk = 0; kk < 4; ++kk) cc[i][kk] = pw[4 * i + kk];
    sub_cubic4(cc, c0, c1);
    hull_bounds(cc, 4, p, dir, cb, cs);
    const double lb = fmax(cb, cs);
    if (lb < bd - SETTLE_EPS && !sp_push(pool, SPiece{lb, ia, 0, it.part, sp_meta(slot, depth + 1, da + 1, 0)})) o.lost = fmin(o.lost, lb), o.over = true;
  }
}

// gpu.wgsl:764-815: a part of a curve piece read by Newton when certified; the minimum
// answers when it is critical and interior to its edge. With `settle` (double
// precision), a part past its certificate is handed to the settle (sk).
// An edge's nearest point may also be where two of its pieces meet in a corner, where
// neither piece is critical (curve3.rs:408-476 takes any point interior to the edge): a
// part ending at such a junction answers with its nearest point there, critical or not,
// also when it holds no critical point at all. Such an answer is a point of the edge, so
// it never undercuts the true nearest point, which is read wherever it lies. A C0 knot
// line's points answer only where critical (solid3.rs:1866-1872).
template <class T, bool SETTLE = false>
__device__ void try_cleaf(const Tab<T>& tb, uint32_t ci, const ICAM_K float* box, T lb, v3<T> p, Best<T>& b, T tm, T eps, SSink* sk = nullptr) {
  const Layout& L = tb.L;
  const ICAM_K T* cf = tb.F + L.cleaf_f + CLEAF_F * ci;
  const ICAM_K uint32_t* cu = tb.U + L.cleaf_u + CLEAF_U * ci;
  const uint32_t cp = cu[0];
  const ICAM_K T* pw = tb.F + L.cpiece_f + CPIECE_F * cp;
  const ICAM_K uint32_t* pu = tb.U + L.cpiece_u + CPIECE_U * cp;
  const ICAM_K uint32_t* eu = tb.U + L.edge_u + EDGE_U * pu[0];
  const bool crease = eu[0] == EDGE_CREASE;
  T t0 = cf[1], t1 = cf[2];
  const bool jlo = !crease && pu[1] > 0 && t0 == T(0), jhi = !crease && pu[1] + 1 < pu[2] && t1 == T(1);
  T rho = cf[0];
  bool certified = far_dist(p, box) < T(0.999) * rho - eps;
  if (!certified) {
    const int kind = cleaf_kind(cf + CLEAF_CERT, box, p, eps);
    if (kind == 1) {
      if (!(jlo || jhi)) return;
      // no critical point: the junction end (the nearer if both are)
      T s = jlo ? T(0) : T(1);
      if (jlo && jhi && len(ceval(pw, T(1)).p - p) < len(ceval(pw, T(0)).p - p)) s = T(1);
      const CEval<T> e = ceval(pw, s);
      const T d = len(e.p - p);
      if (d < b.d) offer(b, d, (P_CLEAF << 28) | ci, s, T(0), sgn_of(dot(p - e.p, edge_pn(tb, pu[0], e.p, pu[1], s))));
      return;
    }
    certified = kind == 2;
  }
  if constexpr (SETTLE && sizeof(T) == 8) {
    if (!certified) {
      settle_probe_curve(tb, ci, p, b, sk);
      settle_root(sk, lb, ci | SP_CURVE, b);
      return;
    }
  }
  const bool single = sizeof(T) == 4;
  bool conv = false;
  T s = newton1(pw, p, T(0.5) * (t0 + t1), t0, t1, single ? 10 : 40, single ? T(1e-5) : T(1e-15), &conv);
  CEval<T> e = ceval(pw, s);
  v3<T> r = e.p - p;
  T d = len(r);
  if (!certified) b.lmin = mmin(b.lmin, lb);
  T tol = single ? T(1e-4) * d + eps : T(1e-9) * (T(1) + d);
  const bool crit = is_critical(r, e.d1, tol);
  if (single && certified && !conv && !crit && lb < b.d) b.unsure = true;
  if (!(d < b.d)) return;
  // an end of the edge is its vertex's (a C0 line's ends are in its face's boundary,
  // whose edges answer there)
  if (!crease && ((pu[1] == 0 && s <= T(0)) || (pu[1] + 1 == pu[2] && s >= T(1)))) return;
  const bool junction = (jlo && s <= T(0)) || (jhi && s >= T(1));
  if (!junction && !crit) return;
  if (crease) {
    bool near = false;
    const bool in = crease_inside(tb, eu, cp, s, tm, near);
    if (near) b.unsure = true;
    if (!in) return;
  }
  v3<T> pn = edge_pn(tb, pu[0], e.p, pu[1], s);
  if (len(pn) < T(1e-6)) b.unsure = true;
  offer(b, d, (P_CLEAF << 28) | ci, s, T(0), sgn_of(dot(p - e.p, pn)));
}

// A lower bound on p's distance to a sub-patch from the slab between the two planes
// normal to its middle normal that hold its control points (surface3.rs:973-975 takes
// the same support-plane bound along the direction to a Newton point), less a
// rounding allowance.
template <class T> __device__ inline T slab_dist(const Tab<T>& tb, uint32_t li, v3<T> p, T eps) {
  const ICAM_K T* lf = tb.F + tb.L.leaf_f + LEAF_F * li;
  T x = dot(p, ld3(lf + 5));
  return mmax(mmax(x - lf[9], lf[8] - x) - eps, T(0));
}

// One candidate primitive for one lane.
template <class T> __device__ inline void visit(const Tab<T>& tb, uint32_t code, const ICAM_K float* box, v3<T> p, Best<T>& b, T tm, T eps) {
  T lb = box_dist(p, box);
  if (!(lb < b.d)) return;
  uint32_t kind = code >> 28, id = code & 0x0fffffffu;
  switch (kind) {
    case P_FACE: try_face(tb, id, p, b, tm); break;
    case P_EDGE: try_edge(tb, id, p, b); break;
    case P_VERT: try_vert(tb, id, p, b); break;
    case P_HOST: b.lmin = mmin(b.lmin, lb), b.host = true; break;
    case P_LEAF: try_leaf(tb, id, box, lb, p, b, tm, eps); break;
    default: try_cleaf(tb, id, box, lb, p, b, tm, eps); break;
  }
}

// The last correction in double precision of the answer a single-precision search
// chose: the winner's foot and sign again from the float64 tables.
// The correction of a sub-patch winner: the single-precision winner's parameters are
// within rounding of the foot, where the distance is stationary: one Newton step (more
// only if it moved far) leaves a distance error of the order of the step squared; they
// are first put inside the sub-patch's double-precision bounds (a foot on its edge then
// lies on it exactly).
__device__ inline void refine64_leaf(const Tab<double>& tb, uint32_t id, double pa, double pb, v3<double> p, double& d, double& sg) {
  const Layout& L = tb.L;
  const ICAM_K double* lf = tb.F + L.leaf_f + LEAF_F * id;
  const ICAM_K uint32_t* lu = tb.U + L.leaf_u + LEAF_U * id;
  const ICAM_K double* sp = tb.F + L.span_f + SPAN_F * lu[0];
  double s = mclamp(pa, lf[1], lf[2]), t = mclamp(pb, lf[3], lf[4]);
  newton2(sp, p, s, t, lf[1], lf[2], lf[3], lf[4], 4, 1e-6);
  SEval<double> e = peval1(sp, s, t);
  v3<double> r = e.p - p;
  d = len(r);
  const ICAM_K double* f = tb.F + L.face_f + FACE_F * lu[1];
  sg = sgn_of(dot(r, cross(e.ps, e.pt)) * -f[14]);
}
__device__ inline void refine64_vert(const Tab<double>& tb, uint32_t id, v3<double> p, double& d, double& sg) {
  const ICAM_K double* r = tb.F + tb.L.vert_f + VERT_F * id;
  v3<double> x = ld3(r);
  d = len(p - x);
  sg = sgn_of(dot(p - x, ld3(r + 3)));
}

__device__ inline void refine64(const Tab<double>& tb, uint32_t code, double pa, double pb, v3<double> p, double& d, double& sg) {
  const Layout& L = tb.L;
  uint32_t kind = code >> 28, id = code & 0x0fffffffu;
  if (kind == P_FACE) {
    const ICAM_K double* f = tb.F + L.face_f + FACE_F * id;
    uint32_t k = tb.U[L.face_u + FACE_U * id];
    v3<double> q = foot(k, f, p);
    d = len(p - q);
    sg = sgn_of(dot(p - q, snormal(k, f, q) * f[14]));
  } else if (kind == P_EDGE) {
    Best<double> b;
    b.d = TypeEps<double>::big, b.d2 = b.d, b.code = ~0u, b.lmin = b.d, b.unsure = false, b.host = false, b.sgn = 1, b.sgn2 = 1, b.a = b.b = 0;
    try_edge(tb, id, p, b);
    if (b.code != ~0u) d = b.d, sg = b.sgn;
  } else if (kind == P_VERT) {
    refine64_vert(tb, id, p, d, sg);
  } else if (kind == P_LEAF) {
    refine64_leaf(tb, id, pa, pb, p, d, sg);
  } else {
    const ICAM_K double* cf = tb.F + L.cleaf_f + CLEAF_F * id;
    const ICAM_K uint32_t* cu = tb.U + L.cleaf_u + CLEAF_U * id;
    const ICAM_K double* pw = tb.F + L.cpiece_f + CPIECE_F * cu[0];
    const ICAM_K uint32_t* pu = tb.U + L.cpiece_u + CPIECE_U * cu[0];
    double s = newton1(pw, p, mclamp(pa, cf[1], cf[2]), cf[1], cf[2], 4, 1e-7);
    CEval<double> e = ceval(pw, s);
    d = len(e.p - p);
    sg = sgn_of(dot(p - e.p, edge_pn(tb, pu[0], e.p, pu[1], s)));
  }
}

}  // namespace icam
