// The viewer's ray cast over baked brick levels, on the device: the consumer of the brick
// cache this library bakes (crates/icam-sim/src/cache.wgsl: trace_ray :441-503 over the
// levels, walk :335-420 over a level's bricks, march :270-318 over a brick's cells,
// cell_hit :191-268 and span_root :251-268 for the trilinear interpolant's first zero in a
// cell). One ray a lane, in single precision as the shader reads the cache: a level's
// bricks are crossed by a 3D DDA, air bricks passed, a solid brick met where the ray
// enters it, and in a stored brick the cells crossed by a second DDA; a cell with a corner
// at or under zero has its interpolant along the ray taken as the cubic through four
// samples, split at its turning points and the middle into monotone spans, the first
// sign change found by bisection. Several levels are read finest first: the finest level
// holding the ray at t is walked until the ray leaves it or enters a finer one.
//
// Floating-point contraction is off in this file, so every operation rounds as the CPU
// restatement's (oracle/cast.py) does.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/icam_hip.h"

namespace {

constexpr int CAST_LEVELS = 16;  // levels a cast reads, at most
constexpr float CBIG = 1e30f;
constexpr uint32_t C_AIR = 0xffffffffu, C_SOLID = 0xfffffffeu, C_HOST = 0xfffffffdu, C_WORDS = 365u;

struct CLevel {
  float lo[3];  // the first node
  float h;      // cell side
  int nb[3];    // bricks along each axis
  const uint32_t* map;
  const uint32_t* atlas;
};
struct CastArgs {
  CLevel lv[CAST_LEVELS];
  int count;
};

struct F3 {
  float v[3];
};
__device__ inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// A stored brick's node n (x + 9 y + 81 z) as the shader unpacks it: i16 over 32767, at
// least -1.
__device__ inline float node_value(const uint32_t* atlas, uint32_t slot, int n) {
  const uint32_t w = atlas[size_t(slot) * C_WORDS + uint32_t(n >> 1)];
  const int16_t q = int16_t((n & 1) ? (w >> 16) : (w & 0xffffu));
  return fmaxf(float(q) / 32767.0f, -1.0f);
}

// Where the ray o + t d is inside the box [lo, hi] (inv = 1 / d): enter, leave.
__device__ inline void box_span(const F3& o, const F3& inv, const float lo[3], const float hi[3], float& enter, float& leave) {
  enter = -CBIG, leave = CBIG;
  for (int k = 0; k < 3; ++k) {
    const float a = (lo[k] - o.v[k]) * inv.v[k], b = (hi[k] - o.v[k]) * inv.v[k];
    enter = fmaxf(enter, fminf(a, b));
    leave = fminf(leave, fmaxf(a, b));
  }
}

// The trilinear interpolant of the cell's corners f (index x + 2 y + 4 z) at q.
__device__ inline float trilerp(const float f[8], float qx, float qy, float qz) {
  const float lo = mixf(mixf(f[0], f[1], qx), mixf(f[2], f[3], qx), qy);
  const float hi = mixf(mixf(f[4], f[5], qx), mixf(f[6], f[7], qx), qy);
  return mixf(lo, hi, qz);
}
__device__ inline F3 trigrad(const float f[8], float qx, float qy, float qz) {
  F3 g;
  g.v[0] = mixf(mixf(f[1] - f[0], f[3] - f[2], qy), mixf(f[5] - f[4], f[7] - f[6], qy), qz);
  g.v[1] = mixf(mixf(f[2] - f[0], f[3] - f[1], qx), mixf(f[6] - f[4], f[7] - f[5], qx), qz);
  g.v[2] = mixf(mixf(f[4] - f[0], f[5] - f[1], qx), mixf(f[6] - f[2], f[7] - f[3], qx), qy);
  return g;
}
__device__ inline float poly3(const float c[4], float u) { return ((c[3] * u + c[2]) * u + c[1]) * u + c[0]; }

// The first zero of cubic c in [lo, hi], positive at lo, when it is not positive at hi:
// twenty halvings; else -1.
__device__ inline float first_zero(const float c[4], float lo, float hi) {
  if (hi <= lo || poly3(c, hi) > 0.0f) return -1.0f;
  for (int it = 0; it < 20; ++it) {
    const float m = 0.5f * (lo + hi);
    if (poly3(c, m) > 0.0f)
      lo = m;
    else
      hi = m;
  }
  return 0.5f * (lo + hi);
}

// The first u in [0, 1] where the interpolant along q0 + u dq reaches zero, or -1.
__device__ float cell_zero(const float f[8], const float q0[3], const float dq[3]) {
  const float s0 = trilerp(f, q0[0], q0[1], q0[2]);
  if (s0 <= 0.0f) return 0.0f;
  float s[4];
  s[0] = s0;
  s[1] = trilerp(f, q0[0] + dq[0] / 3.0f, q0[1] + dq[1] / 3.0f, q0[2] + dq[2] / 3.0f);
  s[2] = trilerp(f, q0[0] + dq[0] * (2.0f / 3.0f), q0[1] + dq[1] * (2.0f / 3.0f), q0[2] + dq[2] * (2.0f / 3.0f));
  s[3] = trilerp(f, q0[0] + dq[0], q0[1] + dq[1], q0[2] + dq[2]);
  // the cubic through the four samples, first in x = 3u (forward differences), then in u
  const float d1 = s[1] - s[0], d2 = s[2] - 2.0f * s[1] + s[0], d3 = s[3] - 3.0f * s[2] + 3.0f * s[1] - s[0];
  const float c[4] = {s[0], 3.0f * (d1 - 0.5f * d2 + d3 / 3.0f), 9.0f * (0.5f * d2 - 0.5f * d3), 27.0f * d3 / 6.0f};
  // its turning points (the quadratic's roots in the form that does not cancel), clamped
  float r1 = 0.5f, r2 = 0.5f;
  const float qa = 3.0f * c[3], qb = 2.0f * c[2], disc = qb * qb - 4.0f * qa * c[1];
  if (disc >= 0.0f) {
    const float q = -0.5f * (qb + (qb >= 0.0f ? 1.0f : -1.0f) * sqrtf(disc));
    if (fabsf(qa) > 1e-20f) r1 = q / qa;
    if (fabsf(q) > 1e-20f) r2 = c[1] / q;
  }
  const float x = fminf(fmaxf(fminf(r1, r2), 0.0f), 1.0f), y = fminf(fmaxf(fmaxf(r1, r2), 0.0f), 1.0f);
  // the span's cuts in order, its middle among them: every piece monotone and short
  float cut1, cut2, cut3;
  if (0.5f < x)
    cut1 = 0.5f, cut2 = x, cut3 = y;
  else if (0.5f < y)
    cut1 = x, cut2 = 0.5f, cut3 = y;
  else
    cut1 = x, cut2 = y, cut3 = 0.5f;
  float r = first_zero(c, 0.0f, cut1);
  if (r < 0.0f) r = first_zero(c, cut1, cut2);
  if (r < 0.0f) r = first_zero(c, cut2, cut3);
  if (r < 0.0f) r = first_zero(c, cut3, 1.0f);
  return r;
}

// In the stored brick `slot` whose first node is `base` (cells of side h), the first hit
// between t_in and t_out: t (-1 for none) and the gradient there.
__device__ float brick_hit(const uint32_t* atlas, uint32_t slot, const float base[3], float h, const F3& o, const F3& d, const F3& inv, float t_in,
                           float t_out, F3& grad) {
  int cell[3], step[3];
  float tnext[3], tstep[3];
  for (int k = 0; k < 3; ++k) {
    const float pk = o.v[k] + d.v[k] * t_in;
    cell[k] = int(fminf(fmaxf(floorf((pk - base[k]) / h), 0.0f), 7.0f));
    step[k] = d.v[k] > 0.0f ? 1 : (d.v[k] < 0.0f ? -1 : 0);
    tnext[k] = (base[k] + (float(cell[k]) + (d.v[k] > 0.0f ? 1.0f : 0.0f)) * h - o.v[k]) * inv.v[k];
    tstep[k] = fabsf(h * inv.v[k]);
  }
  float t = t_in;
  for (int s = 0; s < 32; ++s) {
    const float tc = fminf(fminf(tnext[0], tnext[1]), tnext[2]);
    const float te = fminf(tc, t_out);
    const int n0 = cell[0] + 9 * (cell[1] + 9 * cell[2]);
    float f[8];
    float least = CBIG;
    for (int k = 0; k < 8; ++k) {
      f[k] = node_value(atlas, slot, n0 + (k & 1) + 9 * ((k >> 1) & 1) + 81 * (k >> 2));
      least = fminf(least, f[k]);
    }
    if (least <= 0.0f) {
      float q0[3], dq[3];
      for (int k = 0; k < 3; ++k) {
        const float corner = base[k] + float(cell[k]) * h;
        q0[k] = (o.v[k] + d.v[k] * t - corner) / h;
        dq[k] = d.v[k] * (te - t) / h;
      }
      const float u = cell_zero(f, q0, dq);
      if (u >= 0.0f) {
        grad = trigrad(f, q0[0] + dq[0] * u, q0[1] + dq[1] * u, q0[2] + dq[2] * u);
        return t + u * (te - t);
      }
    }
    if (tc >= t_out) break;
    t = tc;
    const int a = (tnext[0] <= tnext[1] && tnext[0] <= tnext[2]) ? 0 : (tnext[1] <= tnext[2] ? 1 : 2);
    cell[a] += step[a];
    tnext[a] += tstep[a];
    if (cell[a] < 0 || cell[a] > 7) break;
  }
  return -1.0f;
}

// The first hit in level L between t0 and t1: t (-1 for none), the gradient, the slot.
__device__ float level_hit(const CLevel& L, const F3& o, const F3& d, const F3& inv, float t0, float t1, F3& grad, uint32_t& slot) {
  const float side = 8.0f * L.h;
  int cell[3] = {0, 0, 0}, step[3];
  float tnext[3] = {0.0f, 0.0f, 0.0f}, tstep[3];
  for (int k = 0; k < 3; ++k) step[k] = d.v[k] > 0.0f ? 1 : (d.v[k] < 0.0f ? -1 : 0), tstep[k] = fabsf(side * inv.v[k]);
  float t = t0;
  bool fresh = true;
  for (int s = 0; s < 4096; ++s) {
    if (fresh) {
      // the brick the ray is in just past t, and where it next crosses a brick face
      for (int k = 0; k < 3; ++k) {
        const float pk = o.v[k] + d.v[k] * (t + 1e-4f * L.h);
        cell[k] = int(fminf(fmaxf(floorf((pk - L.lo[k]) / side), 0.0f), float(L.nb[k]) - 1.0f));
        tnext[k] = (L.lo[k] + (float(cell[k]) + (d.v[k] > 0.0f ? 1.0f : 0.0f)) * side - o.v[k]) * inv.v[k];
      }
      fresh = false;
    }
    const uint32_t entry = L.map[size_t(cell[0]) + size_t(L.nb[0]) * (size_t(cell[1]) + size_t(L.nb[1]) * size_t(cell[2]))];
    const float t_out = fminf(fminf(fminf(tnext[0], tnext[1]), tnext[2]), t1);
    if (entry == C_SOLID || entry == C_HOST) {
      // all material (or left to the caller, who reads it on the CPU): met where the ray
      // enters the brick
      slot = entry;
      grad.v[0] = -d.v[0], grad.v[1] = -d.v[1], grad.v[2] = -d.v[2];
      return t;
    }
    if (entry != C_AIR) {
      const float base[3] = {L.lo[0] + float(cell[0]) * side, L.lo[1] + float(cell[1]) * side, L.lo[2] + float(cell[2]) * side};
      const float r = brick_hit(L.atlas, entry, base, L.h, o, d, inv, t, t_out, grad);
      if (r >= 0.0f) {
        slot = entry;
        return r;
      }
    }
    if (t_out >= t1) break;
    t = t_out;
    const int a = (tnext[0] <= tnext[1] && tnext[0] <= tnext[2]) ? 0 : (tnext[1] <= tnext[2] ? 1 : 2);
    cell[a] += step[a];
    tnext[a] += tstep[a];
    if (cell[a] < 0 || cell[a] >= L.nb[a]) break;
  }
  return -1.0f;
}

__global__ __launch_bounds__(256) void k_cast(CastArgs A, const double* rays, size_t n, float* out, uint32_t* hit) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  F3 o, d, inv;
  for (int k = 0; k < 3; ++k) o.v[k] = float(rays[6 * i + k]), d.v[k] = float(rays[6 * i + 3 + k]);
  for (int k = 0; k < 3; ++k) inv.v[k] = fabsf(d.v[k]) < 1e-12f ? (d.v[k] + 1e-30f >= 0.0f ? CBIG : -CBIG) : 1.0f / d.v[k];
  // where the ray is in each level
  float enter[CAST_LEVELS], leave[CAST_LEVELS];
  float t = CBIG;
  for (int k = 0; k < A.count; ++k) {
    const CLevel& L = A.lv[k];
    const float hi[3] = {L.lo[0] + float(L.nb[0]) * 8.0f * L.h, L.lo[1] + float(L.nb[1]) * 8.0f * L.h, L.lo[2] + float(L.nb[2]) * 8.0f * L.h};
    float a, b;
    box_span(o, inv, L.lo, hi, a, b);
    const bool in = a <= b && b >= 0.0f;
    enter[k] = in ? fmaxf(a, 0.0f) : CBIG;
    leave[k] = in ? b : -CBIG;
    t = fminf(t, enter[k]);
  }
  float rt = -1.0f;
  F3 g{{0.0f, 0.0f, 0.0f}};
  uint32_t slot = C_AIR, level = 0xffffffffu;
  for (int piece = 0; piece < 64 && t < CBIG; ++piece) {
    const float past = t + fmaxf(fabsf(t), 1.0f) * 1e-6f;
    // the finest level holding the ray just past t (levels are finest first)
    int k = A.count;
    for (int j = 0; j < A.count; ++j)
      if (enter[j] <= past && past < leave[j]) {
        k = j;
        break;
      }
    if (k == A.count) {
      float next = CBIG;
      for (int j = 0; j < A.count; ++j)
        if (enter[j] > past) next = fminf(next, enter[j]);
      t = next;
      continue;
    }
    // the piece ends where level k does, or where a finer one begins
    float t1 = leave[k];
    for (int j = 0; j < k; ++j)
      if (enter[j] > past) t1 = fminf(t1, enter[j]);
    uint32_t s = C_AIR;
    const float r = level_hit(A.lv[k], o, d, inv, t, t1, g, s);
    if (r >= 0.0f) {
      rt = r, slot = s, level = uint32_t(k);
      break;
    }
    t = t1;
  }
  out[4 * i] = rt;
  out[4 * i + 1] = rt >= 0.0f ? g.v[0] : 0.0f;
  out[4 * i + 2] = rt >= 0.0f ? g.v[1] : 0.0f;
  out[4 * i + 3] = rt >= 0.0f ? g.v[2] : 0.0f;
  hit[2 * i] = level;
  hit[2 * i + 1] = slot;
}

}  // namespace

// the library's error channel (icam_hip_last_error)
int icam_fail(const std::string& s);

extern "C" int icam_hip_cast(const icam_level* levels, const uint32_t* const* d_maps, const uint32_t* const* d_atlases, int count, const double* d_rays,
                             size_t n, float* d_out, uint32_t* d_hit, void* stream) {
  if (!levels || !d_maps || !d_atlases || (n && (!d_rays || !d_out || !d_hit))) return icam_fail("bad arguments");
  if (count < 1 || count > CAST_LEVELS) return icam_fail("a cast reads 1 to " + std::to_string(CAST_LEVELS) + " levels");
  CastArgs A{};
  A.count = count;
  for (int k = 0; k < count; ++k) {
    const icam_level& l = levels[k];
    if (!(l.h > 0) || l.nb[0] <= 0 || l.nb[1] <= 0 || l.nb[2] <= 0 || !d_maps[k] || !d_atlases[k]) return icam_fail("bad level " + std::to_string(k));
    for (int a = 0; a < 3; ++a) A.lv[k].lo[a] = float(l.lo[a]), A.lv[k].nb[a] = l.nb[a];
    A.lv[k].h = float(l.h), A.lv[k].map = d_maps[k], A.lv[k].atlas = d_atlases[k];
  }
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_cast, dim3(unsigned((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, A, d_rays, n, d_out, d_hit);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : icam_fail(std::string("k_cast: ") + hipGetErrorString(e));
}
