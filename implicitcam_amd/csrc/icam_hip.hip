// libicam_hip: the exact signed distance and the narrow-band brick baker on gfx950.
//
// Kernels (one launch each, on the handle's stream):
//  k_points  one lane a point; the block culls the primitive list against its
//            points' box into LDS, then every lane reads the same candidates in
//            the same order (their records come through the scalar cache);
//  k_decide  one lane a brick centre, searched within the brick's reach
//            (PartCache::decide's test, crates/icam-sim/src/partcache.rs:226-331);
//  k_sign    signs of the bricks no surface comes near, carried along grid lines;
//  k_bricks  one block a brick: its 729 nodes searched within the clamp, far nodes
//            signed by flood fill, the rest bridged, quantized to i16 and packed two
//            to a word (partcache.rs:759-918, 962-970);
//  k_scan    exclusive prefix sums (band bricks, stored slots);
//  k_store   the map and the atlas in map order (partcache.rs:128-146).
// Each search runs in double precision (ICAM_FP64) or in single precision with a
// last correction in double (ICAM_MIXED).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/icam_hip.h"
#include "pack.hpp"
#include "sdf_device.hpp"

using namespace icam;

namespace {

constexpr int WG = 256;
constexpr int MAXC = 4096;
constexpr uint32_t NONE = 0xffffffffu;

thread_local std::string g_error;

int fail(const std::string& s) {
  g_error = s;
  return 1;
}

#define HIP_OK(x)                                                                              \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_));         \
  } while (0)

struct Dev {
  const float* F32;
  const double* F64;
  const uint32_t* U;
  const float* pbox;
  const uint32_t* pcode;
  Layout L;
  double origin[3];
};

template <class T> __device__ inline Tab<T> tab_of(const Dev& d);
template <> __device__ inline Tab<float> tab_of<float>(const Dev& d) { return Tab<float>{d.F32, d.U, d.pbox, d.pcode, d.L}; }
template <> __device__ inline Tab<double> tab_of<double>(const Dev& d) { return Tab<double>{d.F64, d.U, d.pbox, d.pcode, d.L}; }

// ------------------------------------------------------------------ block culling

__device__ inline float box_gap(const float* a, const float* b) {
  float dx = fmaxf(fmaxf(a[0] - b[3], b[0] - a[3]), 0.f);
  float dy = fmaxf(fmaxf(a[1] - b[4], b[1] - a[4]), 0.f);
  float dz = fmaxf(fmaxf(a[2] - b[5], b[2] - a[5]), 0.f);
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// The primitives whose box comes within R of the block's box gb (lo, hi), in list
// order, into cand; returns how many (more than MAXC: the list did not fit).
template <int NT = WG>
__device__ int cull(const Dev& dv, const float* gb, float R, uint32_t* cand, int* wave_tot) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int base = 0;
  for (uint32_t c0 = 0; c0 < dv.L.n_prims; c0 += NT) {
    uint32_t i = c0 + tid;
    bool keep = false;
    if (i < dv.L.n_prims) keep = box_gap(dv.pbox + 6 * i, gb) <= R;
    unsigned long long m = __ballot(keep);
    int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wave; ++w) off += wave_tot[w];
    if (keep && off + before < MAXC) cand[off + before] = i;
    int tot = 0;
    for (int w = 0; w < NT / 64; ++w) tot += wave_tot[w];
    __syncthreads();
    base += tot;
  }
  return base;
}

// Block reduction of a box (lo, hi) and a radius.
template <int NT = WG>
__device__ void block_box(float lo[3], float hi[3], float& r, float* red) {
  const int tid = threadIdx.x;
  for (int k = 0; k < 3; ++k) {
    red[tid] = lo[k];
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] = fminf(red[tid], red[tid + s]);
      __syncthreads();
    }
    lo[k] = red[0];
    __syncthreads();
    red[tid] = hi[k];
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] = fmaxf(red[tid], red[tid + s]);
      __syncthreads();
    }
    hi[k] = red[0];
    __syncthreads();
  }
  red[tid] = r;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmaxf(red[tid], red[tid + s]);
    __syncthreads();
  }
  r = red[0];
  __syncthreads();
}

// ------------------------------------------------------------------ one lane's search

// The answer of one lane: value, flag (0 exact, 1 bound, 2 undecided), feature, and
// whether a double-precision step was taken.
struct Answer {
  double v;
  uint32_t flag;
  uint32_t feat;
  bool refined;
};

template <class T>
__device__ Best<T> search(const Tab<T>& tb, v3<T> p, T bound, const uint32_t* cand, int nc, T tm, T eps) {
  Best<T> b;
  b.d = bound, b.code = NONE, b.a = b.b = T(0), b.sgn = T(1), b.lmin = TypeEps<T>::big, b.d2 = TypeEps<T>::big, b.sgn2 = T(1), b.unsure = false;
  const uint32_t n = nc <= MAXC ? uint32_t(nc) : tb.L.n_prims;
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t i = nc <= MAXC ? cand[k] : k;
    i = __builtin_amdgcn_readfirstlane(i);
    visit(tb, tb.pcode[i], tb.pbox + 6 * i, p, b, tm, eps);
  }
  return b;
}

__device__ inline uint32_t feat_of(const Tab<double>& tb, uint32_t code) {
  uint32_t kind = code >> 28, id = code & 0x0fffffffu;
  if (kind == P_FACE) return id;
  if (kind == P_EDGE) return (1u << 28) | id;
  if (kind == P_VERT) return (2u << 28) | id;
  if (kind == P_LEAF) return tb.U[tb.L.leaf_u + LEAF_U * id + 1];
  return (1u << 28) | tb.U[tb.L.cleaf_u + CLEAF_U * id + 1];
}

template <class T>
__device__ Answer classify(const Best<T>& b, double d, double sg, double bound) {
  Answer a;
  a.refined = false;
  a.feat = NONE;
  if (b.code != NONE && !(b.lmin < b.d) && double(b.d) <= bound) {
    a.v = d * sg, a.flag = 0;
    return a;
  }
  T l = mmin(b.d, b.lmin);
  if (double(l) >= bound || (b.code == NONE && !(b.lmin < b.d))) {
    a.v = bound, a.flag = 1;
    return a;
  }
  a.v = double(l), a.flag = 2;
  return a;
}

// One point: the search in the precision asked, with the last correction in double.
template <bool MIXED>
__device__ Answer answer(const Dev& dv, v3<double> p, double bound, const uint32_t* cand, int nc) {
  Tab<double> t64 = tab_of<double>(dv);
  if (MIXED) {
    Tab<float> t32 = tab_of<float>(dv);
    const float eps = dv.L.eps32;
    float b32 = bound < 3.0e38 ? float(bound) * (1.0f + 1e-6f) + eps : 3.0e38f;
    Best<float> b = search<float>(t32, cv3<float>(p), b32, cand, nc, 4e-5f * float(dv.L.diag) + eps, eps);
    bool tie = b.d2 - b.d <= 8.0f * eps && b.sgn2 != b.sgn;
    if (!b.unsure && !tie) {
      double d = double(b.d), sg = double(b.sgn);
      if (b.code != NONE) refine64(t64, b.code, double(b.a), double(b.b), p, d, sg);
      Answer a = classify(b, d, sg, bound);
      if (a.flag == 0) a.feat = feat_of(t64, b.code);
      return a;
    }
  }
  Best<double> b = search<double>(t64, p, bound < 1e300 ? bound * (1.0 + 1e-12) : bound, cand, nc, 0.0, dv.L.eps64);
  Answer a = classify(b, b.d, b.sgn, bound);
  if (a.flag == 0) a.feat = feat_of(t64, b.code);
  a.refined = MIXED;
  return a;
}

// ------------------------------------------------------------------ kernels

// Points: xyz (world, double), optional per-point bounds.
template <bool MIXED>
__global__ __launch_bounds__(WG) void k_points(Dev dv, const double* xyz, const double* bounds, size_t n, double* out_d, uint32_t* out_flag,
                                              uint32_t* out_feat) {
  __shared__ uint32_t cand[MAXC];
  __shared__ float red[WG];
  __shared__ int wave_tot[WG / 64];
  size_t i = size_t(blockIdx.x) * WG + threadIdx.x;
  bool live = i < n;
  v3<double> p = mk3(0.0, 0.0, 0.0);
  double bound = 1e300;
  if (live) {
    p = mk3(xyz[3 * i] - dv.origin[0], xyz[3 * i + 1] - dv.origin[1], xyz[3 * i + 2] - dv.origin[2]);
    if (bounds) bound = bounds[i];
  }
  // an upper bound on the distance: the nearest of the first vertices (points of the
  // surface), or the bound
  float ub = float(fmin(bound, 3.0e38));
  uint32_t nv = dv.L.n_verts < 16 ? dv.L.n_verts : 16;
  for (uint32_t k = 0; k < nv; ++k) {
    const double* vr = dv.F64 + dv.L.vert_f + VERT_F * k;
    ub = fminf(ub, float(len(p - ld3(vr))) * (1.0f + 1e-5f) + 1e-5f);
  }
  float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
  float r = 0.f;
  if (live) {
    lo[0] = hi[0] = float(p.x), lo[1] = hi[1] = float(p.y), lo[2] = hi[2] = float(p.z);
    r = ub;
  }
  block_box(lo, hi, r, red);
  float gb[6] = {lo[0] - 1e-4f, lo[1] - 1e-4f, lo[2] - 1e-4f, hi[0] + 1e-4f, hi[1] + 1e-4f, hi[2] + 1e-4f};
  int nc = cull(dv, gb, r, cand, wave_tot);
  if (!live) return;
  Answer a = answer<MIXED>(dv, p, bound, cand, nc);
  out_d[i] = a.v;
  if (out_flag) out_flag[i] = a.flag;
  if (out_feat) out_feat[i] = a.feat;
}

// The level's geometry.
struct Lv {
  double lo[3];
  double h;
  int nb[3];
};
__device__ inline v3<double> brick_first(const Lv& lv, const double* origin, long b) {
  long x = b % lv.nb[0], y = (b / lv.nb[0]) % lv.nb[1], z = b / (long(lv.nb[0]) * lv.nb[1]);
  double side = 8.0 * lv.h;
  return mk3(lv.lo[0] - origin[0] + side * x, lv.lo[1] - origin[1] + side * y, lv.lo[2] - origin[2] + side * z);
}

// Brick centres, searched within the brick's reach r + clamp (r the half diagonal of
// the brick grown by the clamp): dc the signed distance or NaN past it; band 1 when
// the decision keeps the brick (|dc| <= reach); sign +-1 when known, else 0. One wave
// a tile of 4 x 4 x 4 bricks, so a tile's candidates are those near the tile.
constexpr int DT = 64;
template <bool MIXED>
__global__ __launch_bounds__(DT) void k_decide(Dev dv, Lv lv, long nbricks, double* dc, uint8_t* band, int8_t* sgn) {
  __shared__ uint32_t cand[MAXC];
  __shared__ float red[DT];
  __shared__ int wave_tot[1];
  const int tx = (lv.nb[0] + 3) / 4, ty = (lv.nb[1] + 3) / 4;
  long tile = blockIdx.x;
  int lane = threadIdx.x;
  long bx = 4 * (tile % tx) + (lane & 3), by = 4 * ((tile / tx) % ty) + ((lane >> 2) & 3), bz = 4 * (tile / (long(tx) * ty)) + (lane >> 4);
  bool live = bx < lv.nb[0] && by < lv.nb[1] && bz < lv.nb[2];
  long b = bx + lv.nb[0] * (by + long(lv.nb[1]) * bz);
  double h = lv.h, clamp = h * 1.7320508075688772;
  double half = 4.0 * h;
  double reach = 0.5 * sqrt(3.0) * (8.0 * h + 2.0 * clamp) + clamp;
  v3<double> c = mk3(0.0, 0.0, 0.0);
  if (live) c = brick_first(lv, dv.origin, b) + mk3(half, half, half);
  float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
  float r = float(reach) * (1.0f + 1e-4f);
  if (live) lo[0] = hi[0] = float(c.x), lo[1] = hi[1] = float(c.y), lo[2] = hi[2] = float(c.z);
  block_box<DT>(lo, hi, r, red);
  float gb[6] = {lo[0] - 1e-4f, lo[1] - 1e-4f, lo[2] - 1e-4f, hi[0] + 1e-4f, hi[1] + 1e-4f, hi[2] + 1e-4f};
  int nc = cull<DT>(dv, gb, r, cand, wave_tot);
  if (!live) return;
  Answer a = answer<MIXED>(dv, c, reach * (1.0 + 1e-9), cand, nc);
  int8_t s = 0;
  if (a.flag == 0) {
    dc[b] = a.v;
    band[b] = fabs(a.v) <= reach ? 1 : 0;
    s = a.v < 0 ? -1 : 1;
  } else {
    dc[b] = __builtin_nan("");
    // an undecided centre keeps its brick unless its lower bound is past the reach
    band[b] = (a.flag == 2 && a.v <= reach) ? 1 : 0;
    // a centre outside the part's box is outside the part
    const float* L0 = dv.L.lo;
    const float* L1 = dv.L.hi;
    if (c.x < L0[0] || c.y < L0[1] || c.z < L0[2] || c.x > L1[0] || c.y > L1[1] || c.z > L1[2]) s = 1;
  }
  sgn[b] = s;
}

// Signs along grid lines of axis `ax`: a brick no surface comes within its reach of
// shares its sign with its neighbours (the reach exceeds the 8 h between centres).
__global__ void k_sign(Lv lv, int ax, const uint8_t* band, int8_t* sgn, int* changed) {
  long nb[3] = {lv.nb[0], lv.nb[1], lv.nb[2]};
  long a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
  long line = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (line >= nb[a1] * nb[a2]) return;
  long i1 = line % nb[a1], i2 = line / nb[a1];
  long stride[3] = {1, nb[0], nb[0] * nb[1]};
  long base = i1 * stride[a1] + i2 * stride[a2];
  int any = 0;
  for (int pass = 0; pass < 2; ++pass) {
    int8_t carry = 0;
    bool carry_ok = false;  // the previous brick may hand its sign on
    for (long k = 0; k < nb[ax]; ++k) {
      long kk = pass == 0 ? k : nb[ax] - 1 - k;
      long id = base + kk * stride[ax];
      int8_t s = sgn[id];
      bool nonband = band[id] == 0;
      if (s == 0 && carry != 0 && (carry_ok || nonband)) {
        sgn[id] = s = carry;
        any = 1;
      }
      carry = s;
      carry_ok = nonband;
    }
  }
  if (any) atomicOr(changed, 1);
}

// Exclusive prefix sum of flags (nonzero counts 1) over n items, one block of 1024.
__global__ __launch_bounds__(1024) void k_scan(const uint8_t* flags, long n, uint32_t* out, uint32_t* total) {
  __shared__ uint32_t s[1024];
  const int tid = threadIdx.x;
  uint32_t base = 0;
  for (long c0 = 0; c0 < n; c0 += 1024) {
    long i = c0 + tid;
    uint32_t f = (i < n && flags[i]) ? 1u : 0u;
    s[tid] = f;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      uint32_t v = tid >= off ? s[tid - off] : 0u;
      __syncthreads();
      s[tid] += v;
      __syncthreads();
    }
    if (i < n) out[i] = base + s[tid] - f;
    uint32_t tot = s[1023];
    __syncthreads();
    base += tot;
  }
  if (tid == 0) *total = base;
}

__global__ void k_band_list(const uint8_t* band, const uint32_t* pos, long n, uint32_t* list) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && band[i]) list[pos[i]] = uint32_t(i);
}

// Node values of a block's brick: 729 nodes, three a thread.
enum { MODE_NODES = 0, MODE_BAKE = 1 };

__device__ inline int nb26(int n, int k) {
  // the k-th of the 26 neighbours of node n (x fastest), or -1 outside the brick
  int d = k < 13 ? k : k + 1;  // skip the centre (13)
  int di = d % 3 - 1, dj = (d / 3) % 3 - 1, dk = d / 9 - 1;
  int i = n % 9 + di, j = (n / 9) % 9 + dj, kk = n / 81 + dk;
  if (i < 0 || i > 8 || j < 0 || j > 8 || kk < 0 || kk > 8) return -1;
  return i + 9 * j + 81 * kk;
}
__device__ inline double nb26_len(int k) {
  int d = k < 13 ? k : k + 1;
  int di = d % 3 - 1, dj = (d / 3) % 3 - 1, dk = d / 9 - 1;
  return sqrt(double(di * di + dj * dj + dk * dk));
}

template <bool MIXED, int MODE>
__global__ __launch_bounds__(WG) void k_bricks(Dev dv, Lv lv, const uint32_t* list, const double* firsts, long nbricks, const double* dcs,
                                               const int8_t* sgns, double* out_nodes, uint32_t* staging, uint32_t* cls,
                                               unsigned long long* counters) {
  __shared__ uint32_t cand[MAXC];
  __shared__ float red[WG];
  __shared__ int wave_tot[WG / 64];
  __shared__ double val[729];
  __shared__ uint8_t st[729];  // 0 exact, 1 far, 2 undecided, 3 filled from a neighbour
  __shared__ int changed;
  const int tid = threadIdx.x;
  long bi = blockIdx.x;
  if (bi >= nbricks) return;
  v3<double> first;
  double h;
  if (firsts) {
    first = mk3(firsts[4 * bi] - dv.origin[0], firsts[4 * bi + 1] - dv.origin[1], firsts[4 * bi + 2] - dv.origin[2]);
    h = firsts[4 * bi + 3];
  } else {
    first = brick_first(lv, dv.origin, long(list[bi]));
    h = lv.h;
  }
  const double clamp = h * 1.7320508075688772;
  float lo[3] = {float(first.x), float(first.y), float(first.z)};
  float hi[3] = {float(first.x + 8 * h), float(first.y + 8 * h), float(first.z + 8 * h)};
  float r = float(clamp) * (1.0f + 1e-4f) + 1e-4f;
  float gb[6] = {lo[0] - 1e-4f, lo[1] - 1e-4f, lo[2] - 1e-4f, hi[0] + 1e-4f, hi[1] + 1e-4f, hi[2] + 1e-4f};
  int nc = cull(dv, gb, r, cand, wave_tot);
  unsigned long long refined = 0;
  for (int n = tid; n < 729; n += WG) {
    v3<double> p = first + mk3(h * (n % 9), h * ((n / 9) % 9), h * (n / 81));
    Answer a = answer<MIXED>(dv, p, clamp, cand, nc);
    refined += a.refined ? 1 : 0;
    if (MODE == MODE_NODES) {
      out_nodes[2 * (729 * bi + n)] = a.v;
      out_nodes[2 * (729 * bi + n) + 1] = double(a.flag);
    } else {
      val[n] = a.v;
      st[n] = uint8_t(a.flag);
    }
  }
  if (MODE == MODE_NODES) return;
  __syncthreads();
  // partcache.rs:825-855 far nodes take the sign of any neighbour known, by flood fill
  auto fill = [&]() {
    for (;;) {
      if (tid == 0) changed = 0;
      __syncthreads();
      for (int n = tid; n < 729; n += WG) {
        if (st[n] != 1) continue;
        for (int k = 0; k < 26; ++k) {
          int m = nb26(n, k);
          if (m < 0) continue;
          uint8_t sm = st[m];
          if ((sm == 0 || sm == 3) && val[m] != 0.0) {
            val[n] = copysign(clamp, val[m]);
            st[n] = 3;
            changed = 1;
            break;
          }
        }
      }
      __syncthreads();
      int c = changed;
      __syncthreads();
      if (!c) break;
    }
  };
  fill();
  // partcache.rs:856-871 far nodes within the empty ball of the centre share its sign
  // the centre's distance; when the decision left it undecided, its sign alone as carried
  // along the grid (a brick no surface comes near hands its sign to its neighbours)
  double dc = dcs ? dcs[list ? list[bi] : bi] : __builtin_nan("");
  const int8_t csgn = sgns ? sgns[list ? list[bi] : bi] : int8_t(0);
  const double half = 4.0 * h;
  if (dc == dc && dc != 0.0) {
    double reach = 0.999 * fmax(clamp, fabs(dc));
    int any = 0;
    for (int n = tid; n < 729; n += WG) {
      if (st[n] != 1) continue;
      v3<double> x = mk3(h * (n % 9) - half, h * ((n / 9) % 9) - half, h * (n / 81) - half);
      if (len(x) < reach) val[n] = copysign(clamp, dc), st[n] = 3, any = 1;
    }
    if (__syncthreads_or(any)) fill();
  }
  // partcache.rs:872-909 what is left is bridged (the mean of the known neighbours held
  // within their Lipschitz brackets), one node after another
  int open = 0;
  for (int n = tid; n < 729; n += WG) open |= (st[n] == 1 || st[n] == 2);
  if (__syncthreads_or(open)) {
    if (tid == 0) {
      unsigned long long bridged = 0;
      int nfar = 0;
      for (int n = 0; n < 729; ++n) nfar += st[n] == 1;
      if (nfar == 729) {
        double s = (dc == dc && dc != 0.0) ? dc : csgn < 0 ? -1.0 : 1.0;
        for (int n = 0; n < 729; ++n) val[n] = copysign(clamp, s), st[n] = 3;
      }
      bool exact[729];
      for (int n = 0; n < 729; ++n) exact[n] = st[n] == 0;
      for (int pass = 0; pass < 729; ++pass) {
        bool left = false, moved = false;
        for (int n = 0; n < 729; ++n) {
          if (st[n] != 1 && st[n] != 2) continue;
          double sum = 0, lo_ = -clamp, hi_ = clamp;
          int cnt = 0;
          for (int k = 0; k < 26; ++k) {
            int m = nb26(n, k);
            if (m < 0) continue;
            if (st[m] == 0 || st[m] == 3 || st[m] == 4) sum += val[m], ++cnt;
            if (exact[m]) lo_ = fmax(lo_, val[m] - nb26_len(k) * h), hi_ = fmin(hi_, val[m] + nb26_len(k) * h);
          }
          if (!cnt) {
            left = true;
            continue;
          }
          double v = sum / cnt;
          val[n] = lo_ <= hi_ ? fmin(fmax(v, lo_), hi_) : 0.5 * (lo_ + hi_);
          st[n] = 4;
          moved = true;
          ++bridged;
        }
        if (!left) break;
        if (!moved) {
          for (int n = 0; n < 729; ++n)
            if (st[n] == 1 || st[n] == 2) val[n] = clamp, st[n] = 4, ++bridged;
          break;
        }
      }
      if (bridged) atomicAdd(&counters[0], bridged);
    }
  }
  __syncthreads();
  // partcache.rs:911-915, 962-970 quantize, class, two values to a word (low half first)
  __shared__ int16_t q[730];
  int pos = 0, nonpos = 0;
  for (int n = tid; n < 729; n += WG) {
    double x = fmin(fmax(val[n] / clamp, -1.0), 1.0) * 32767.0;
    int16_t qq = int16_t(round(x));
    q[n] = qq;
    pos |= qq <= 0, nonpos |= qq > 0;
  }
  if (tid == 0) q[729] = 0;
  int any_nonpos = __syncthreads_or(pos), any_pos = __syncthreads_or(nonpos);
  uint32_t c = !any_nonpos ? 1u : !any_pos ? 2u : 0u;  // 1 AIR, 2 SOLID, 0 words
  if (tid == 0) cls[bi] = c;
  if (c == 0)
    for (int w = tid; w < ICAM_WORDS; w += WG) staging[size_t(bi) * ICAM_WORDS + w] = uint32_t(uint16_t(q[2 * w])) | (uint32_t(uint16_t(q[2 * w + 1])) << 16);
  if (MIXED) {
    // reduce the refined count (one atomic a block)
    red[tid] = float(refined);
    __syncthreads();
    for (int s = WG / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) atomicAdd(&counters[1], (unsigned long long)red[0]);
  }
}

__global__ void k_flags_words(const uint32_t* cls, long n, uint8_t* f) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) f[i] = cls[i] == 0;
}

// partcache.rs:128-146 assemble: the map (AIR, SOLID, or the slot) and the atlas in
// map order (band bricks are listed in map order, so their slots are too).
__global__ void k_store(long nbricks, const uint8_t* band, const int8_t* sgn, uint32_t* map) {
  long b = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nbricks) return;
  if (!band[b]) map[b] = sgn[b] < 0 ? ICAM_SOLID : ICAM_AIR;
}
__global__ void k_store_band(long nband, const uint32_t* list, const uint32_t* cls, const uint32_t* slot, const uint32_t* staging, uint32_t* map,
                             uint32_t* atlas) {
  long bi = blockIdx.x;
  if (bi >= nband) return;
  uint32_t c = cls[bi];
  if (threadIdx.x == 0) map[list[bi]] = c == 1 ? ICAM_AIR : c == 2 ? ICAM_SOLID : slot[bi];
  if (c == 0)
    for (int w = threadIdx.x; w < ICAM_WORDS; w += blockDim.x) atlas[size_t(slot[bi]) * ICAM_WORDS + w] = staging[size_t(bi) * ICAM_WORDS + w];
}

__global__ void k_spin(float* x, int iters) {
  float a = x[threadIdx.x];
  for (int i = 0; i < iters; ++i) a = a * 0.999f + 0.001f;
  x[threadIdx.x] = a;
}

// ------------------------------------------------------------------ host side

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t need(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    return e;
  }
  template <class X> X* as() const { return reinterpret_cast<X*>(p); }
  void free() {
    if (p) hipFree(p);
    p = nullptr, bytes = 0;
  }
};

}  // namespace

struct icam_hip {
  std::mutex mu;
  int device = 0;
  hipStream_t stream = nullptr;
  Packed pk;
  Dev dv{};
  Buf F32, F64, U, pbox, pcode;
  // scratch of queries and bakes
  Buf q_xyz, q_bound, q_d, q_flag, q_feat;
  Buf dc, band, sgn, pos, total, list, staging, cls, wflags, slot, stotal, map, atlas, counters, changed, firsts, nodes_out;
  std::vector<std::pair<const char*, double>> stages;
  std::string name;
};

namespace {

void stage_reset(icam_hip* h) { h->stages.clear(); }

struct Timer {
  hipEvent_t a, b;
  Timer(hipStream_t s) {
    hipEventCreate(&a);
    hipEventCreate(&b);
    hipEventRecord(a, s);
  }
  void stop(hipStream_t s) { hipEventRecord(b, s); }
  double ms() {
    hipEventSynchronize(b);
    float m = 0;
    hipEventElapsedTime(&m, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    return m;
  }
};

template <class K, class... A>
hipError_t launch(icam_hip* h, const char* name, hipStream_t s, dim3 g, dim3 b, std::vector<Timer>* tm, std::vector<const char*>* names, K k, A... args) {
  if (tm) tm->emplace_back(s);
  hipLaunchKernelGGL(k, g, b, 0, s, args...);
  if (tm) tm->back().stop(s), names->push_back(name);
  return hipGetLastError();
}

int upload(Buf& b, const void* src, size_t bytes) {
  HIP_OK(b.need(std::max<size_t>(bytes, 16)));
  if (bytes) HIP_OK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

extern "C" {

const char* icam_hip_last_error(void) { return g_error.c_str(); }

int icam_hip_open(const icam_part_desc* part, double need, icam_hip** out) {
  g_error.clear();
  *out = nullptr;
  if (!part) return fail("no part");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail("no HIP device");
  icam_hip* h = new icam_hip();
  HIP_OK(hipGetDevice(&h->device));
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, h->device));
  h->name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
  if (!pack_part(part, need, h->pk)) {
    std::string e = h->pk.error;
    delete h;
    return fail(e);
  }
  std::vector<float> f32 = to_f32(h->pk.F);
  if (upload(h->F32, f32.data(), f32.size() * 4) || upload(h->F64, h->pk.F.data(), h->pk.F.size() * 8) ||
      upload(h->U, h->pk.U.data(), h->pk.U.size() * 4) || upload(h->pbox, h->pk.pbox.data(), h->pk.pbox.size() * 4) ||
      upload(h->pcode, h->pk.pcode.data(), h->pk.pcode.size() * 4)) {
    std::string e = g_error;
    icam_hip_close(h);
    return fail(e);
  }
  HIP_OK(hipStreamCreate(&h->stream));
  h->dv.F32 = h->F32.as<float>();
  h->dv.F64 = h->F64.as<double>();
  h->dv.U = h->U.as<uint32_t>();
  h->dv.pbox = h->pbox.as<float>();
  h->dv.pcode = h->pcode.as<uint32_t>();
  h->dv.L = h->pk.L;
  for (int k = 0; k < 3; ++k) h->dv.origin[k] = h->pk.origin[k];
  *out = h;
  return 0;
}

void icam_hip_close(icam_hip* h) {
  if (!h) return;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (Buf* b : {&h->F32, &h->F64, &h->U, &h->pbox, &h->pcode, &h->q_xyz, &h->q_bound, &h->q_d, &h->q_flag, &h->q_feat, &h->dc, &h->band, &h->sgn,
                 &h->pos, &h->total, &h->list, &h->staging, &h->cls, &h->wflags, &h->slot, &h->stotal, &h->map, &h->atlas, &h->counters,
                 &h->changed, &h->firsts, &h->nodes_out})
    b->free();
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

int icam_hip_device_name(icam_hip* h, char* buf, size_t cap) {
  if (!h || !buf || !cap) return fail("bad arguments");
  std::snprintf(buf, cap, "%s", h->name.c_str());
  return 0;
}

size_t icam_hip_nodes(icam_hip* h) { return h ? h->pk.nodes : 0; }

int icam_hip_warm(icam_hip* h, double secs) {
  if (!h) return fail("no handle");
  std::lock_guard<std::mutex> g(h->mu);
  Buf x;
  HIP_OK(x.need(4096));
  HIP_OK(hipMemsetAsync(x.p, 0, 4096, h->stream));
  auto t0 = std::chrono::steady_clock::now();
  while (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() < secs) {
    hipLaunchKernelGGL(k_spin, dim3(1024), dim3(1024), 0, h->stream, x.as<float>(), 1 << 16);
    HIP_OK(hipStreamSynchronize(h->stream));
  }
  x.free();
  return 0;
}

int icam_hip_query_device(icam_hip* h, int precision, const double* d_xyz, size_t n, double* d_d, uint32_t* d_flags, uint32_t* d_feat, void* stream) {
  if (!h) return fail("no handle");
  if (!n) return 0;
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  dim3 g(unsigned((n + WG - 1) / WG));
  if (precision == ICAM_MIXED)
    hipLaunchKernelGGL(k_points<true>, g, dim3(WG), 0, s, h->dv, d_xyz, (const double*)nullptr, n, d_d, d_flags, d_feat);
  else
    hipLaunchKernelGGL(k_points<false>, g, dim3(WG), 0, s, h->dv, d_xyz, (const double*)nullptr, n, d_d, d_flags, d_feat);
  HIP_OK(hipGetLastError());
  return 0;
}

int icam_hip_query(icam_hip* h, int precision, const double* xyz, size_t n, double* d, uint32_t* flags, uint32_t* feat) {
  if (!h) return fail("no handle");
  std::lock_guard<std::mutex> g(h->mu);
  if (!n) return 0;
  HIP_OK(h->q_xyz.need(n * 24));
  HIP_OK(h->q_d.need(n * 8));
  HIP_OK(h->q_flag.need(n * 4));
  HIP_OK(h->q_feat.need(n * 4));
  HIP_OK(hipMemcpyAsync(h->q_xyz.p, xyz, n * 24, hipMemcpyHostToDevice, h->stream));
  Timer t(h->stream);
  if (icam_hip_query_device(h, precision, h->q_xyz.as<double>(), n, h->q_d.as<double>(), h->q_flag.as<uint32_t>(), h->q_feat.as<uint32_t>(), h->stream))
    return 1;
  t.stop(h->stream);
  HIP_OK(hipMemcpyAsync(d, h->q_d.p, n * 8, hipMemcpyDeviceToHost, h->stream));
  if (flags) HIP_OK(hipMemcpyAsync(flags, h->q_flag.p, n * 4, hipMemcpyDeviceToHost, h->stream));
  if (feat) HIP_OK(hipMemcpyAsync(feat, h->q_feat.p, n * 4, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  stage_reset(h);
  h->stages.push_back({"points", t.ms()});
  return 0;
}

int icam_hip_points(icam_hip* h, int precision, const icam_point_rec* recs, size_t n, double* out) {
  if (!h) return fail("no handle");
  std::lock_guard<std::mutex> g(h->mu);
  if (!n) return 0;
  std::vector<double> xyz(3 * n), bd(n);
  for (size_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) xyz[3 * i + k] = recs[i].p[k];
    bd[i] = std::isfinite(recs[i].clamp) ? recs[i].clamp : 1e300;
  }
  HIP_OK(h->q_xyz.need(n * 24));
  HIP_OK(h->q_bound.need(n * 8));
  HIP_OK(h->q_d.need(n * 8));
  HIP_OK(h->q_flag.need(n * 4));
  HIP_OK(hipMemcpyAsync(h->q_xyz.p, xyz.data(), n * 24, hipMemcpyHostToDevice, h->stream));
  HIP_OK(hipMemcpyAsync(h->q_bound.p, bd.data(), n * 8, hipMemcpyHostToDevice, h->stream));
  dim3 gr(unsigned((n + WG - 1) / WG));
  if (precision == ICAM_MIXED)
    hipLaunchKernelGGL(k_points<true>, gr, dim3(WG), 0, h->stream, h->dv, h->q_xyz.as<double>(), h->q_bound.as<double>(), n, h->q_d.as<double>(),
                       h->q_flag.as<uint32_t>(), (uint32_t*)nullptr);
  else
    hipLaunchKernelGGL(k_points<false>, gr, dim3(WG), 0, h->stream, h->dv, h->q_xyz.as<double>(), h->q_bound.as<double>(), n, h->q_d.as<double>(),
                       h->q_flag.as<uint32_t>(), (uint32_t*)nullptr);
  HIP_OK(hipGetLastError());
  std::vector<double> d(n);
  std::vector<uint32_t> f(n);
  HIP_OK(hipMemcpyAsync(d.data(), h->q_d.p, n * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipMemcpyAsync(f.data(), h->q_flag.p, n * 4, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  for (size_t i = 0; i < n; ++i) out[2 * i] = d[i], out[2 * i + 1] = double(f[i]);
  return 0;
}

int icam_hip_bricks(icam_hip* h, int precision, const icam_brick_rec* recs, size_t n, double* out) {
  if (!h) return fail("no handle");
  std::lock_guard<std::mutex> g(h->mu);
  if (!n) return 0;
  HIP_OK(h->firsts.need(n * 32));
  HIP_OK(h->nodes_out.need(n * 729 * 16));
  HIP_OK(hipMemcpyAsync(h->firsts.p, recs, n * 32, hipMemcpyHostToDevice, h->stream));
  Lv lv{};
  if (precision == ICAM_MIXED)
    hipLaunchKernelGGL((k_bricks<true, MODE_NODES>), dim3(unsigned(n)), dim3(WG), 0, h->stream, h->dv, lv, (const uint32_t*)nullptr,
                       h->firsts.as<double>(), long(n), (const double*)nullptr, (const int8_t*)nullptr, h->nodes_out.as<double>(), (uint32_t*)nullptr,
                       (uint32_t*)nullptr, (unsigned long long*)nullptr);
  else
    hipLaunchKernelGGL((k_bricks<false, MODE_NODES>), dim3(unsigned(n)), dim3(WG), 0, h->stream, h->dv, lv, (const uint32_t*)nullptr,
                       h->firsts.as<double>(), long(n), (const double*)nullptr, (const int8_t*)nullptr, h->nodes_out.as<double>(), (uint32_t*)nullptr,
                       (uint32_t*)nullptr, (unsigned long long*)nullptr);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(out, h->nodes_out.p, n * 729 * 16, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  return 0;
}

static int bake_impl(icam_hip* h, int precision, const icam_level* lvp, hipStream_t s, icam_bake_stats* stats, bool timed) {
  Lv lv;
  for (int k = 0; k < 3; ++k) lv.lo[k] = lvp->lo[k], lv.nb[k] = lvp->nb[k];
  lv.h = lvp->h;
  if (!(lv.h > 0) || lv.nb[0] <= 0 || lv.nb[1] <= 0 || lv.nb[2] <= 0) return fail("bad level");
  long nb = long(lv.nb[0]) * lv.nb[1] * lv.nb[2];
  HIP_OK(h->dc.need(nb * 8));
  HIP_OK(h->band.need(nb));
  HIP_OK(h->sgn.need(nb));
  HIP_OK(h->pos.need(nb * 4));
  HIP_OK(h->total.need(16));
  HIP_OK(h->list.need(nb * 4));
  HIP_OK(h->map.need(nb * 4));
  HIP_OK(h->counters.need(64));
  HIP_OK(h->changed.need(16));
  HIP_OK(hipMemsetAsync(h->counters.p, 0, 64, s));
  std::vector<Timer> tm;
  std::vector<const char*> names;
  std::vector<Timer>* T = timed ? &tm : nullptr;
  auto* N = timed ? &names : nullptr;
  bool mixed = precision == ICAM_MIXED;
  dim3 gb(unsigned((nb + WG - 1) / WG));
  dim3 gt(unsigned(long((lv.nb[0] + 3) / 4) * ((lv.nb[1] + 3) / 4) * ((lv.nb[2] + 3) / 4)));
  if (mixed)
    HIP_OK(launch(h, "decide", s, gt, dim3(DT), T, N, k_decide<true>, h->dv, lv, nb, h->dc.as<double>(), h->band.as<uint8_t>(), h->sgn.as<int8_t>()));
  else
    HIP_OK(launch(h, "decide", s, gt, dim3(DT), T, N, k_decide<false>, h->dv, lv, nb, h->dc.as<double>(), h->band.as<uint8_t>(), h->sgn.as<int8_t>()));
  // signs of the bricks kept whole: along x, y, z lines, until nothing changes
  for (int round = 0; round < 8; ++round) {
    HIP_OK(hipMemsetAsync(h->changed.p, 0, 4, s));
    for (int ax = 0; ax < 3; ++ax) {
      long lines = long(lv.nb[(ax + 1) % 3]) * lv.nb[(ax + 2) % 3];
      HIP_OK(launch(h, "sign", s, dim3(unsigned((lines + 255) / 256)), dim3(256), T, N, k_sign, lv, ax, (const uint8_t*)h->band.as<uint8_t>(),
                    h->sgn.as<int8_t>(), h->changed.as<int>()));
    }
    int ch = 0;
    HIP_OK(hipMemcpyAsync(&ch, h->changed.p, 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (!ch) break;
  }
  HIP_OK(launch(h, "scan", s, dim3(1), dim3(1024), T, N, k_scan, (const uint8_t*)h->band.as<uint8_t>(), nb, h->pos.as<uint32_t>(),
                h->total.as<uint32_t>()));
  HIP_OK(launch(h, "band_list", s, gb, dim3(WG), T, N, k_band_list, (const uint8_t*)h->band.as<uint8_t>(), (const uint32_t*)h->pos.as<uint32_t>(), nb,
                h->list.as<uint32_t>()));
  uint32_t nband = 0;
  HIP_OK(hipMemcpyAsync(&nband, h->total.p, 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  HIP_OK(h->staging.need(std::max<size_t>(size_t(nband) * ICAM_WORDS * 4, 16)));
  HIP_OK(h->cls.need(std::max<size_t>(size_t(nband) * 4, 16)));
  HIP_OK(h->wflags.need(std::max<size_t>(nband, 16)));
  HIP_OK(h->slot.need(std::max<size_t>(size_t(nband) * 4, 16)));
  HIP_OK(h->stotal.need(16));
  if (nband) {
    if (mixed)
      HIP_OK(launch(h, "bricks", s, dim3(nband), dim3(WG), T, N, k_bricks<true, MODE_BAKE>, h->dv, lv, (const uint32_t*)h->list.as<uint32_t>(),
                    (const double*)nullptr, long(nband), (const double*)h->dc.as<double>(), (const int8_t*)h->sgn.as<int8_t>(), (double*)nullptr,
                    h->staging.as<uint32_t>(),
                    h->cls.as<uint32_t>(), h->counters.as<unsigned long long>()));
    else
      HIP_OK(launch(h, "bricks", s, dim3(nband), dim3(WG), T, N, k_bricks<false, MODE_BAKE>, h->dv, lv, (const uint32_t*)h->list.as<uint32_t>(),
                    (const double*)nullptr, long(nband), (const double*)h->dc.as<double>(), (const int8_t*)h->sgn.as<int8_t>(), (double*)nullptr,
                    h->staging.as<uint32_t>(),
                    h->cls.as<uint32_t>(), h->counters.as<unsigned long long>()));
    HIP_OK(launch(h, "word_flags", s, dim3(unsigned((nband + 255) / 256)), dim3(256), T, N, k_flags_words, (const uint32_t*)h->cls.as<uint32_t>(),
                  long(nband), h->wflags.as<uint8_t>()));
    HIP_OK(launch(h, "scan", s, dim3(1), dim3(1024), T, N, k_scan, (const uint8_t*)h->wflags.as<uint8_t>(), long(nband), h->slot.as<uint32_t>(),
                  h->stotal.as<uint32_t>()));
  } else {
    HIP_OK(hipMemsetAsync(h->stotal.p, 0, 4, s));
  }
  uint32_t nslots = 0;
  HIP_OK(hipMemcpyAsync(&nslots, h->stotal.p, 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  HIP_OK(h->atlas.need(std::max<size_t>(size_t(nslots) * ICAM_WORDS * 4, 16)));
  HIP_OK(launch(h, "store", s, gb, dim3(WG), T, N, k_store, nb, (const uint8_t*)h->band.as<uint8_t>(), (const int8_t*)h->sgn.as<int8_t>(),
                h->map.as<uint32_t>()));
  if (nband)
    HIP_OK(launch(h, "store_band", s, dim3(nband), dim3(128), T, N, k_store_band, long(nband), (const uint32_t*)h->list.as<uint32_t>(),
                  (const uint32_t*)h->cls.as<uint32_t>(), (const uint32_t*)h->slot.as<uint32_t>(), (const uint32_t*)h->staging.as<uint32_t>(),
                  h->map.as<uint32_t>(), h->atlas.as<uint32_t>()));
  if (stats || timed) {
    unsigned long long cnt[2] = {0, 0};
    HIP_OK(hipMemcpyAsync(cnt, h->counters.p, 16, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (stats) {
      stats->bricks = nb, stats->band = nband, stats->slots = nslots;
      stats->undecided = int64_t(cnt[0]), stats->refined = int64_t(cnt[1]);
    }
  }
  if (timed) {
    stage_reset(h);
    for (size_t i = 0; i < tm.size(); ++i) {
      double m = tm[i].ms();
      bool found = false;
      for (auto& st : h->stages)
        if (std::strcmp(st.first, names[i]) == 0) st.second += m, found = true;
      if (!found) h->stages.push_back({names[i], m});
    }
  }
  return 0;
}

int icam_hip_bake_device(icam_hip* h, int precision, const icam_level* lv, uint32_t** d_map, uint32_t** d_atlas, icam_bake_stats* stats, void* stream) {
  if (!h || !lv) return fail("bad arguments");
  std::lock_guard<std::mutex> g(h->mu);
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  if (bake_impl(h, precision, lv, s, stats, true)) return 1;
  if (d_map) *d_map = h->map.as<uint32_t>();
  if (d_atlas) *d_atlas = h->atlas.as<uint32_t>();
  return 0;
}

int icam_hip_bake(icam_hip* h, int precision, const icam_level* lv, uint32_t* map, uint32_t* atlas, size_t atlas_cap, icam_bake_stats* stats) {
  if (!h || !lv) return fail("bad arguments");
  std::lock_guard<std::mutex> g(h->mu);
  icam_bake_stats st{};
  if (bake_impl(h, precision, lv, h->stream, &st, true)) return 1;
  if (stats) *stats = st;
  long nb = long(lv->nb[0]) * lv->nb[1] * lv->nb[2];
  HIP_OK(hipMemcpyAsync(map, h->map.p, nb * 4, hipMemcpyDeviceToHost, h->stream));
  size_t words = size_t(st.slots) * ICAM_WORDS;
  if (words > atlas_cap) {
    HIP_OK(hipStreamSynchronize(h->stream));
    return fail("atlas needs " + std::to_string(words) + " words, the buffer holds " + std::to_string(atlas_cap));
  }
  if (words) HIP_OK(hipMemcpyAsync(atlas, h->atlas.p, words * 4, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  return 0;
}

int icam_hip_stage_times(icam_hip* h, const char** names, double* ms, int cap) {
  if (!h) return fail("no handle");
  int n = 0;
  for (auto& s : h->stages) {
    if (n >= cap) break;
    names[n] = s.first, ms[n] = s.second, ++n;
  }
  return n;
}

}  // extern "C"
