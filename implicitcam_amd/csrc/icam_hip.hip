// libicam_hip: the exact signed distance and the narrow-band brick baker on gfx950.
//
// A query point comes from one of three sources: explicit points, the 729 nodes of
// band bricks, or the centres of a level's bricks. Each is searched by one lane; a
// block culls the primitive list against its points' box into LDS and sorts what it
// keeps nearest first, then every lane reads the same candidates in the same order
// (their records arrive through the scalar cache).
//
// Kernels, one launch each on the handle's stream:
//  k_search<S, float>   the single-precision search: each point's winning feature,
//                        its parameters, distance and sign, and whether a decision was
//                        taken within single-precision rounding;
//  k_refine              the last correction in double precision of each winner; the
//                        points the search was unsure of are queued;
//  k_fallback<S>         the queued points searched again in double precision;
//  k_search<S, double>   the plain double-precision search (ICAM_FP64);
//  k_decide_finish       the level's decision from its brick centres
//                        (PartCache::decide's test, crates/icam-sim/src/partcache.rs:226-331);
//  k_sign                signs of the bricks no surface comes near, along grid lines;
//  k_bricks_finish       a brick's far nodes signed by flood fill, the rest bridged,
//                        quantized to i16 and packed two to a word (partcache.rs:759-918,
//                        962-970);
//  k_scan, k_store*      prefix sums, the map and the atlas in map order (partcache.rs:128-146).
// Splitting the precisions into their own kernels keeps each kernel's registers to
// its own precision's needs.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/icam_hip.h"
#include "brick_rules.hpp"
#include "pack.hpp"
#include "sdf_device.hpp"

using namespace icam;

namespace {

constexpr int WG = 256;
// The settle kernels' grid: as many blocks as the device runs at once (settle_grid,
// from the occupancy the compiled kernel allows, at most one wave a SIMD: the compiler
// gives the settle 256 registers and 256 accumulation registers a lane). Their lanes
// hold the certified search's queue in private memory (10 KB a lane) and run long, so
// blocks past that would wait for a slot and take their share of the queue late, and
// every wave the grid holds is scratch the runtime reserves for the launch.
template <class K> int settle_grid(K kernel, int nt, int cus) {
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, nt, 0) != hipSuccess || per < 1) per = 1;
  per = std::min(per, std::max(1, 4 * 64 / nt));
  return per * std::max(cus, 1);
}
constexpr int SORTC = 1024;  // lists up to this long are sorted nearest first
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int TCAP = 512;    // a tile's list kept for its bricks' nodes, at most
constexpr uint32_t NONE = 0xffffffffu;

thread_local std::string g_error;

int fail(const std::string& s) {
  g_error = s;
  return 1;
}

#define HIP_OK(x)                                                                      \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct Dev {
  const float* F32;
  const double* F64;
  const uint32_t* U;
  const float* pbox;
  const uint32_t* pcode;
  // the primitives' hierarchy (Layout::bvh_*)
  const float* bvh_box;
  const uint32_t* bvh_prim;
  const uint8_t* bvh_holds;
  // the hierarchy's level table in device memory: bvh_off[0..7] then bvh_cnt[0..7] (a
  // batch kernel holds Dev in a private copy, where a level indexed at run time would
  // send the whole copy to scratch memory)
  const uint32_t* bvh_lv;
  Layout L;
  double origin[3];
};

template <class T> __device__ inline Tab<T> tab_of(const Dev& d);
template <> __device__ inline Tab<float> tab_of<float>(const Dev& d) {
  return Tab<float>{(const ICAM_K float*)d.F32, (const ICAM_K uint32_t*)d.U, (const ICAM_K float*)d.pbox, (const ICAM_K uint32_t*)d.pcode, d.L};
}
template <> __device__ inline Tab<double> tab_of<double>(const Dev& d) {
  return Tab<double>{(const ICAM_K double*)d.F64, (const ICAM_K uint32_t*)d.U, (const ICAM_K float*)d.pbox, (const ICAM_K uint32_t*)d.pcode, d.L};
}

// The level's geometry.
struct Lv {
  double lo[3];
  double h;
  int nb[3];
};

// (A level has fewer than 2^31 bricks: brick indices are taken apart in 32 bits.)
__device__ inline v3<double> brick_first(const Lv& lv, const double* origin, long b) {
  const uint32_t ub = uint32_t(b), n0 = uint32_t(lv.nb[0]), n1 = uint32_t(lv.nb[1]);
  const uint32_t yz = ub / n0, x = ub - yz * n0, z = yz / n1, y = yz - z * n1;
  double side = 8.0 * lv.h;
  return mk3(lv.lo[0] - origin[0] + side * x, lv.lo[1] - origin[1] + side * y, lv.lo[2] - origin[2] + side * z);
}

// The reach of a brick: half the diagonal of the brick grown by the clamp, plus the
// clamp (a centre farther than this from the surface decides the brick whole).
__host__ __device__ inline double brick_reach(double h) {
  double clamp = h * 1.7320508075688772;
  return 0.5 * 1.7320508075688772 * (8.0 * h + 2.0 * clamp) + clamp;
}

// ------------------------------------------------------------------ point sources

// Sources of points: scattered points (SRC_POINTS one lane a point; SRC_WIDE four lanes a
// point, for parts with B-spline faces or spline edges, whose unbounded searches reach
// many parts; SRC_FAR a wave a point, four points a block, for the points of blocks whose
// candidates did not fit their list), the 729 nodes of bricks, brick centres.
enum Src { SRC_POINTS = 0, SRC_NODES = 1, SRC_CENTRES = 2, SRC_WIDE = 3, SRC_FAR = 4 };
template <int S> constexpr bool pts() { return S == SRC_POINTS || S == SRC_WIDE || S == SRC_FAR; }

struct Job {
  long n;                 // points
  const double* xyz;      // SRC_POINTS: world points
  const double* bounds;   // SRC_POINTS: per-point bounds (NULL: none)
  const uint32_t* list;   // SRC_NODES: band bricks of the level (NULL: `firsts`)
  const double* firsts;   // SRC_NODES: explicit bricks (first node xyz, h)
  Lv lv;
  // the candidate list of each tile of bricks: written by the centres' search (within
  // the reach of the tile's centres), read by the nodes' search (a brick's nodes'
  // candidates lie within that reach of its centre)
  uint32_t* tile_cand;    // TCAP per tile
  uint32_t* tile_count;   // per tile (more than TCAP: not kept)
  // the primitives within reach of a super-tile's centres (SUPER^3 tiles), in list
  // order, which its tiles cull instead of the whole list
  uint32_t* super_cand;   // SCAP per super-tile
  uint32_t* super_count;  // per super-tile (more than SCAP: not kept)
};

// The first node and the cell side of brick bi of a nodes job, and its n-th node.
__device__ inline v3<double> node_first(const Job& j, const Dev& dv, long bi, double& h) {
  if (j.list) {
    h = j.lv.h;
    return brick_first(j.lv, dv.origin, long(j.list[bi]));
  }
  h = j.firsts[4 * bi + 3];
  return mk3(j.firsts[4 * bi] - dv.origin[0], j.firsts[4 * bi + 1] - dv.origin[1], j.firsts[4 * bi + 2] - dv.origin[2]);
}
__device__ inline v3<double> node_at(v3<double> first, double h, int n) { return first + mk3(h * (n % 9), h * ((n / 9) % 9), h * (n / 81)); }

// Point i of a job (relative to the origin) and the distance within which it must
// be exact.
template <int S> __device__ inline v3<double> job_point(const Job& j, const Dev& dv, long i, double& bound) {
  if (pts<S>()) {
    bound = j.bounds ? j.bounds[i] : 1e300;
    return mk3(j.xyz[3 * i] - dv.origin[0], j.xyz[3 * i + 1] - dv.origin[1], j.xyz[3 * i + 2] - dv.origin[2]);
  }
  if (S == SRC_NODES) {
    const long bi = long(uint64_t(i) / 729u);
    const int n = int(i - 729 * bi);
    double h;
    const v3<double> first = node_first(j, dv, bi, h);
    bound = h * 1.7320508075688772;
    return node_at(first, h, n);
  }
  double h = j.lv.h;
  bound = brick_reach(h) * (1.0 + 1e-9);
  return brick_first(j.lv, dv.origin, i) + mk3(4.0 * h, 4.0 * h, 4.0 * h);
}

// Per-point results. The single-precision search fills code .. info; the final
// answers are val (the signed distance, or the bound / lower bound), flag (0 exact,
// 1 bound, 2 undecided, 3 waiting for the fallback) and feat (kind << 28 | index).
struct Res {
  uint32_t* code;
  float* pa;
  float* pb;
  float* d32;
  uint32_t* info;  // bit 0 found, 1 negative, 2 unsure, 3 undecided
  double* val;
  uint32_t* flag;
  uint32_t* feat;
  uint32_t* queue;    // blocks holding points the single-precision search was unsure of
  uint32_t* qcount;
  uint32_t* qmark;    // per block: queued already
  uint32_t* items;    // points whose parts wait for the dense pass (item_words words each)
  uint32_t* icount;
  uint32_t icap;
  uint32_t* squeue;   // blocks holding points left undecided, for k_settle
  uint32_t* scount;
  uint32_t* smark;
  uint32_t* cqueue;   // scattered points' blocks whose candidates did not fit their list,
  uint32_t* ccount;   // searched again by k_far (NULL: searched over the whole list); [1] the
                      // next of them k_far takes
  uint32_t part;      // the part's number in a batch (0 alone), in queue entries' top byte
  unsigned long long* counters;  // [0] bridged nodes, [1] points searched again in double,
                                 // [3] of them near a trim curve, [4] near-ties of opposite
                                 // sign, [5] points settled by the certified search, [6] search
                                 // blocks whose candidates did not fit their list
};

// ------------------------------------------------------------------ block culling

__device__ inline float box_gap(const float* a, const float* b) {
  float dx = fmaxf(fmaxf(a[0] - b[3], b[0] - a[3]), 0.f);
  float dy = fmaxf(fmaxf(a[1] - b[4], b[1] - a[4]), 0.f);
  float dz = fmaxf(fmaxf(a[2] - b[5], b[2] - a[5]), 0.f);
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// The primitives whose box comes within R of the block's box gb (lo, hi), in list
// order, into cand (C entries); returns how many (more than C: the list did not fit).
template <int NT, int C>
__device__ int cull(const Dev& dv, const float* gb, float R, uint32_t* cand, int* wave_tot, const uint32_t* from = nullptr, uint32_t n_from = 0) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t total = from ? n_from : dv.L.n_prims;
  int base = 0;
  for (uint32_t c0 = 0; c0 < total; c0 += NT) {
    uint32_t k = c0 + tid;
    uint32_t i = 0;
    bool keep = false;
    if (k < total) {
      i = from ? from[k] : k;
      keep = box_gap(dv.pbox + 6 * i, gb) <= R;
    }
    unsigned long long m = __ballot(keep);
    int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wave; ++w) off += wave_tot[w];
    if (keep && off + before < C) cand[off + before] = i;
    int tot = 0;
    for (int w = 0; w < NT / 64; ++w) tot += wave_tot[w];
    __syncthreads();
    base += tot;
  }
  return base;
}

// The kept candidates ordered by their box's distance from the block's box centre,
// nearest first (ties by list order), so a lane's bound drops early: a bitonic sort of
// 64-bit keys (the distance's bits above the list position; distances are not
// negative, so their bits order as the distances do). tmp holds SORTC entries, key
// 2 SORTC words (8-byte aligned).
template <int NT>
__device__ void sort_cand(const Dev& dv, const float* gb, uint32_t* cand, uint32_t* tmp, float* key, int nc) {
  if (nc <= 1 || nc > SORTC) return;
  const int tid = threadIdx.x;
  float c[6] = {0.5f * (gb[0] + gb[3]), 0.5f * (gb[1] + gb[4]), 0.5f * (gb[2] + gb[5]), 0, 0, 0};
  c[3] = c[0], c[4] = c[1], c[5] = c[2];
  uint64_t* k64 = reinterpret_cast<uint64_t*>(key);
  int n = 2;
  while (n < nc) n <<= 1;
  for (int k = tid; k < n; k += NT) {
    if (k < nc) {
      k64[k] = (uint64_t(__float_as_uint(box_gap(dv.pbox + 6 * cand[k], c))) << 32) | uint32_t(k);
      tmp[k] = cand[k];
    } else {
      k64[k] = ~0ull;
    }
  }
  __syncthreads();
  for (int size = 2; size <= n; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < n; i += NT) {
        const int l = i ^ stride;
        if (l > i) {
          const uint64_t a = k64[i], b = k64[l];
          if ((a > b) == ((i & size) == 0)) k64[i] = b, k64[l] = a;
        }
      }
      __syncthreads();
    }
  for (int k = tid; k < nc; k += NT) cand[k] = tmp[uint32_t(k64[k])];
  __syncthreads();
}

// Block reduction of a box (lo, hi) and a radius.
template <int NT>
__device__ void block_box(float lo[3], float hi[3], float& r, float* red) {
  const int tid = threadIdx.x;
  for (int k = 0; k < 7; ++k) {
    bool is_min = k < 3;
    float v = k < 3 ? lo[k] : k < 6 ? hi[k - 3] : r;
    red[tid] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] = is_min ? fminf(red[tid], red[tid + s]) : fmaxf(red[tid], red[tid + s]);
      __syncthreads();
    }
    v = red[0];
    __syncthreads();
    if (k < 3)
      lo[k] = v;
    else if (k < 6)
      hi[k - 3] = v;
    else
      r = v;
  }
}

// The farthest distance between points of two boxes: from a box holding a point of the
// boundary, an upper bound on the distance from every point of the other box to the
// boundary. Faces, edges, vertices and edge pieces hold boundary points; a sub-patch
// only when it lies wholly inside its face's loops (one the packer could not place may
// hold none), a C0 line's piece not at all (it may run outside the loops).
__device__ inline float box_far(const float* a, const float* b) {
  float s = 0.f;
  for (int k = 0; k < 3; ++k) {
    const float g = fmaxf(a[3 + k] - b[k], b[3 + k] - a[k]);
    s += g * g;
  }
  return sqrtf(s);
}
// The primitives whose box comes within R of box gb, found by the block walking the
// hierarchy from its top level down: a level's nodes that pass are listed (FCAP at most),
// and their children are the next level's to test; the primitives that pass go to cand (C
// entries) in the hierarchy's order. Returns how many (more than C: they did not fit; C +
// 1 also when a level's list did not fit). With `reach`, R is first lowered at each level
// to the least farthest distance from gb to a box holding a boundary point (rounded up),
// an upper bound on the distance from every point of gb to the boundary; *reach_out gets
// the final R. la and lb hold FCAP entries each. When the primitives (or a level's list)
// do not fit and `chunks` is not NULL, the level-2 nodes that passed (the grandparents of
// every primitive within R, 4,096 primitives each at most) are copied there and counted
// in *nchunk (0 otherwise).
constexpr int FCAP = 1024;
template <int NT, int C>
__device__ int bvh_cull(const Dev& dv, const float* gb, float R, bool reach, float* reach_out, uint32_t* cand, uint32_t* la, uint32_t* lb, int* wave_tot,
                        float* red, uint32_t* chunks = nullptr, int* nchunk = nullptr) {
  if (nchunk) *nchunk = 0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (reach) {
    // a first bound from one path down the hierarchy, wave 0 taking at each level the
    // child nearest gb among those holding a boundary point: every level's least farthest
    // distance over its holding children bounds the distance from gb to the boundary
    if (wave == 0) {
      uint32_t node = 0;
      float m = 3.0e38f;
      for (int lev = int(dv.L.bvh_top); lev >= 0; --lev) {
        const uint32_t off = dv.bvh_lv[lev], c = 64u * node + uint32_t(lane);
        const bool ok = c < dv.bvh_lv[8 + lev] && dv.bvh_holds[off + c];
        const float* bx = dv.bvh_box + 6 * (size_t(off) + c);
        float g = ok ? box_gap(bx, gb) : 3.0e38f, f = ok ? box_far(bx, gb) : 3.0e38f;
        uint32_t pick = ok ? c : 0xffffffffu;
        for (int o = 1; o < 64; o <<= 1) {
          f = fminf(f, __shfl_xor(f, o));
          const float g2 = __shfl_xor(g, o);
          const uint32_t p2 = uint32_t(__shfl_xor(int(pick), o));
          if (g2 < g || (g2 == g && p2 < pick)) g = g2, pick = p2;
        }
        m = fminf(m, f);
        if (pick == 0xffffffffu) break;
        node = pick;
      }
      if (lane == 0) red[0] = m;
    }
    __syncthreads();
    const float m = red[0];
    __syncthreads();
    if (m < 3.0e38f) R = fminf(R, m * (1.0f + 1e-5f) + 1e-4f);
  }
  uint32_t *par = la, *out = lb;
  if (tid == 0) par[0] = 0u;  // the root, whose children are the top level's nodes
  __syncthreads();
  uint32_t npar = 1;
  int total = 0;
  const uint32_t* p2 = nullptr;  // the level-2 nodes that passed, and how many
  uint32_t n2 = 0;
  for (int lev = int(dv.L.bvh_top); lev >= 0; --lev) {
    const uint32_t off = dv.bvh_lv[lev], cnt = dv.bvh_lv[8 + lev], nk = npar * 64u;
    if (lev == 1) p2 = par, n2 = npar;
    const float* boxes = dv.bvh_box + 6 * size_t(off);
    if (reach) {
      float m = 3.0e38f;
      for (uint32_t k = tid; k < nk; k += NT) {
        const uint32_t node = 64u * par[k >> 6] + (k & 63u);
        if (node < cnt && dv.bvh_holds[off + node]) m = fminf(m, box_far(boxes + 6 * size_t(node), gb));
      }
      red[tid] = m;
      __syncthreads();
      for (int st = NT / 2; st > 0; st >>= 1) {
        if (tid < st) red[tid] = fminf(red[tid], red[tid + st]);
        __syncthreads();
      }
      m = red[0];
      __syncthreads();
      if (m < 3.0e38f) R = fminf(R, m * (1.0f + 1e-5f) + 1e-4f);
    }
    int base = 0;
    for (uint32_t k0 = 0; k0 < nk; k0 += NT) {
      const uint32_t k = k0 + tid;
      uint32_t node = 0;
      bool keep = false;
      if (k < nk) {
        node = 64u * par[k >> 6] + (k & 63u);
        keep = node < cnt && box_gap(boxes + 6 * size_t(node), gb) <= R;
      }
      const unsigned long long m = __ballot(keep);
      const int before = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) wave_tot[wave] = __popcll(m);
      __syncthreads();
      int at = base;
      for (int w = 0; w < wave; ++w) at += wave_tot[w];
      at += before;
      if (keep) {
        if (lev == 0) {
          if (at < C) cand[at] = dv.bvh_prim[node];
        } else if (at < FCAP) {
          out[at] = node;
        }
      }
      int tot = 0;
      for (int w = 0; w < NT / 64; ++w) tot += wave_tot[w];
      __syncthreads();
      base += tot;
    }
    if (lev == 0 || base > FCAP) {
      total = lev == 0 ? base : C + 1;
      if (total > C && chunks && lev <= 1 && dv.L.bvh_top >= 2) {
        for (uint32_t k = tid; k < n2; k += NT) chunks[k] = p2[k];
        *nchunk = int(n2);
        __syncthreads();
      }
      break;
    }
    npar = uint32_t(base);
    uint32_t* t = par;
    par = out, out = t;
  }
  if (reach_out) *reach_out = R;
  return total;
}

// The primitives under level-2 node `node2` whose box comes within R of gb, into cand in
// order (4,096 at most); `sub` holds 64 entries. Returns how many.
template <int NT>
__device__ int bvh_chunk(const Dev& dv, const float* gb, float R, uint32_t node2, uint32_t* cand, uint32_t* sub, int* wave_tot) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // its level-1 children within R
  {
    const uint32_t c = 64u * node2 + uint32_t(lane);
    const bool keep = wave == 0 && c < dv.L.bvh_cnt[1] && box_gap(dv.bvh_box + 6 * (size_t(dv.L.bvh_off[1]) + c), gb) <= R;
    const unsigned long long m = __ballot(keep);
    if (keep) sub[__popcll(m & ((1ull << lane) - 1ull))] = c;
    if (wave == 0 && lane == 0) wave_tot[0] = __popcll(m);
  }
  __syncthreads();
  const uint32_t nsub = uint32_t(wave_tot[0]), cnt = dv.L.bvh_cnt[0], nk = nsub * 64u;
  __syncthreads();
  int base = 0;
  for (uint32_t k0 = 0; k0 < nk; k0 += NT) {
    const uint32_t k = k0 + tid;
    uint32_t node = 0;
    bool keep = false;
    if (k < nk) {
      node = 64u * sub[k >> 6] + (k & 63u);
      keep = node < cnt && box_gap(dv.bvh_box + 6 * size_t(node), gb) <= R;
    }
    const unsigned long long m = __ballot(keep);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = __popcll(m);
    __syncthreads();
    int at = base;
    for (int w = 0; w < wave; ++w) at += wave_tot[w];
    if (keep) cand[at + before] = dv.bvh_prim[node];
    int tot = 0;
    for (int w = 0; w < NT / 64; ++w) tot += wave_tot[w];
    __syncthreads();
    base += tot;
  }
  return base;
}

// Level-2 nodes ordered by their box's distance from gb, nearest first (ties by node): a
// bitonic sort of 64-bit keys in `key` (n rounded up to a power of two, 8-byte aligned).
template <int NT>
__device__ void sort_chunks(const Dev& dv, const float* gb, uint32_t* chunks, int n, unsigned long long* key) {
  if (n <= 1) return;
  const int tid = threadIdx.x;
  const float* boxes = dv.bvh_box + 6 * size_t(dv.L.bvh_off[2]);
  int m = 2;
  while (m < n) m <<= 1;
  for (int k = tid; k < m; k += NT)
    key[k] = k < n ? (uint64_t(__float_as_uint(box_gap(boxes + 6 * size_t(chunks[k]), gb))) << 32) | uint64_t(chunks[k]) : ~0ull;
  __syncthreads();
  for (int size = 2; size <= m; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < m; i += NT) {
        const int l = i ^ stride;
        if (l > i) {
          const uint64_t a = key[i], b = key[l];
          if ((a > b) == ((i & size) == 0)) key[i] = b, key[l] = a;
        }
      }
      __syncthreads();
    }
  for (int k = tid; k < n; k += NT) chunks[k] = uint32_t(key[k] & 0xffffffffu);
  __syncthreads();
}

// ------------------------------------------------------------------ one lane's search

// A point's search over the block's candidates (or the whole list when they did not
// fit), shared by K consecutive lanes of a group: lane `sub` of the group takes the
// candidates sub, sub + K, ... The closed-form features (faces, edges, vertices) go
// first, in list order; then the B-spline sub-patches and curve parts nearest first
// by their lower bound (box and slab), each round every lane its own next one under
// the group's best answer so far, until none can beat it. A lane's parts under the
// bound go once, with their lower bounds, into its column of the block's queue (Q
// entries of NT lanes in LDS), so a round reads the queue and not the tables; a lane
// with more than Q such parts scans its share of the list each round. Equal bounds
// are taken in list order either way.
//
// Every part with a lower bound under the final answer is tried, as in one lane's
// search; a group may try more (its best was still larger when it chose them), and
// those change neither the answer nor its undecided bound. The group's answer is its
// lanes' least distance (ties: the lowest lane); near-ties of opposite signs among all
// its lanes' answers and runners-up are reported as a runner-up at the same distance.
// Lanes of a group whose point is not `valid` take part with no work.
template <class T>
__device__ inline T part_lb(const Tab<T>& tb, uint32_t i, uint32_t code, v3<T> p, T eps) {
  T lb = box_dist(p, tb.pbox + 6 * i);
  if ((code >> 28) == P_LEAF) lb = mmax(lb, slab_dist(tb, code & 0x0fffffffu, p, eps));
  return lb;
}
template <class T, bool SETTLE = false>
__device__ inline void try_part(const Tab<T>& tb, uint32_t i, T lb, v3<T> p, Best<T>& b, T tm, T eps, SSink* sk = nullptr) {
  const uint32_t code = tb.pcode[i], id = code & 0x0fffffffu;
  if ((code >> 28) == P_LEAF)
    try_leaf<T, SETTLE>(tb, id, tb.pbox + 6 * i, lb, p, b, tm, eps, sk);
  else
    try_cleaf<T, SETTLE>(tb, id, tb.pbox + 6 * i, lb, p, b, tm, eps, sk);
}
template <int K, class V> __device__ inline V group_min(V v) {
  for (int o = 1; o < K; o <<= 1) v = mmin(v, V(__shfl_xor(v, o)));
  return v;
}
template <int K> __device__ inline int group_or(int v) {
  for (int o = 1; o < K; o <<= 1) v |= __shfl_xor(v, o);
  return v;
}

// Counters and slots shared by the whole grid, taken one atomic a wave: the active
// lanes that ask are counted by a ballot and the lowest of them adds the count. (A
// per-lane atomic on one address is serialized at the cache, 64 times a wave; the
// compiler merges them only when it can prove the pointer global.)
__device__ inline void wave_count(unsigned long long* p, bool want) {
  const unsigned long long mask = __ballot(want);
  if (mask && __lane_id() == __ffsll(mask) - 1) atomicAdd(p, (unsigned long long)__popcll(mask));
}
// For each asking lane, its own slot of counter *p (the others get 0).
__device__ inline uint32_t wave_slot(uint32_t* p, bool want) {
  const unsigned long long mask = __ballot(want);
  if (!mask) return 0;
  const int leader = __ffsll(mask) - 1;
  uint32_t base = 0;
  if (__lane_id() == leader) base = atomicAdd(p, uint32_t(__popcll(mask)));
  base = uint32_t(__shfl(int(base), leader));
  return base + uint32_t(__popcll(mask & __lanemask_lt()));
}
// A point whose parts under its bound number 1 .. QI can leave its rounds to the
// dense pass (k_leaves): its answer so far and its parts, sorted by lower bound (equal
// bounds in list order), go to a work item, and its Best comes back with code DEFERRED.
// Item words: 0 point, 1 parts, 2 code, 3 unsure; from 4 d, a, b, sgn, lmin, d2, sgn2;
// then the parts' lower bounds (T) and primitives.
constexpr int QI = 8;
constexpr uint32_t DEFERRED = 0xfffffffeu;
template <class T> struct Item {
  static constexpr int W = sizeof(T) == 4 ? 32 : 48;                 // words an item
  static constexpr int LB = 4 + 7 * int(sizeof(T) / 4);
  static constexpr int PR = LB + QI * int(sizeof(T) / 4);
  static_assert(PR + QI <= W, "an item holds its parts");
  static_assert(sizeof(T) == 4 || LB % 2 == 0, "double words aligned");
};
template <class T> __device__ inline void item_put(uint32_t* w, int off, T v) { *reinterpret_cast<T*>(w + off) = v; }
template <class T> __device__ inline T item_get(const uint32_t* w, int off) { return *reinterpret_cast<const T*>(w + off); }
template <class T> __device__ inline void item_state(uint32_t* w, const Best<T>& b) {
  constexpr int u = int(sizeof(T) / 4);
  w[2] = b.code, w[3] = (b.unsure ? 1u : 0u) | (b.host ? 2u : 0u);
  item_put(w, 4, b.d), item_put(w, 4 + u, b.a), item_put(w, 4 + 2 * u, b.b), item_put(w, 4 + 3 * u, b.sgn);
  item_put(w, 4 + 4 * u, b.lmin), item_put(w, 4 + 5 * u, b.d2), item_put(w, 4 + 6 * u, b.sgn2);
}
template <class T> __device__ inline Best<T> item_best(const uint32_t* w) {
  constexpr int u = int(sizeof(T) / 4);
  Best<T> b;
  b.code = w[2], b.unsure = (w[3] & 1u) != 0, b.host = (w[3] & 2u) != 0;
  b.d = item_get<T>(w, 4), b.a = item_get<T>(w, 4 + u), b.b = item_get<T>(w, 4 + 2 * u), b.sgn = item_get<T>(w, 4 + 3 * u);
  b.lmin = item_get<T>(w, 4 + 4 * u), b.d2 = item_get<T>(w, 4 + 5 * u), b.sgn2 = item_get<T>(w, 4 + 6 * u);
  return b;
}

template <class T> __device__ inline Best<T> best_init(T bound) {
  Best<T> b;
  b.d = bound, b.code = NONE, b.a = b.b = T(0), b.sgn = T(1), b.lmin = TypeEps<T>::big, b.d2 = TypeEps<T>::big, b.sgn2 = T(1), b.unsure = false, b.host = false;
  return b;
}
// One candidate list's share of a point's search, the lane's answer b carried in and
// out (a block whose candidates come in chunks calls it once a chunk; every part under
// the final answer is still read, in the chunk that holds it).
template <class T, int NT, int Q, int K, int C, bool SETTLE = false>
__device__ inline void search_in(const Tab<T>& tb, v3<T> p, bool valid, Best<T>& b, const uint32_t* cand, int nc, T tm, T eps, T* qlb, uint16_t* qk,
                                 const Res* defer = nullptr, long point = 0, SSink* sk = nullptr) {
  const bool all = nc > C;
  const uint32_t n = all ? tb.L.n_prims : uint32_t(nc);
  const int lane = threadIdx.x, sub = lane & (K - 1);
  bool parts = false;
  if (valid) {
    for (uint32_t k = sub; k < n; k += K) {
      uint32_t i = all ? k : cand[k];
      if (K == 1) i = __builtin_amdgcn_readfirstlane(i);
      const uint32_t code = tb.pcode[i];
      if ((code >> 28) >= P_LEAF) {
        parts = true;
        continue;
      }
      visit(tb, code, tb.pbox + 6 * i, p, b, tm, eps);
    }
  }
  if (K == 1 && !parts) return;
  int m = 0;
  if (parts && !all) {
    const T bd = b.d;
    for (uint32_t k = sub; k < n; k += K) {
      uint32_t i = cand[k];
      if (K == 1) i = __builtin_amdgcn_readfirstlane(i);
      const uint32_t code = tb.pcode[i];
      if ((code >> 28) < P_LEAF) continue;
      T lb = part_lb(tb, i, code, p, eps);
      if (lb < bd) {
        if (m < Q) qlb[m * NT + lane] = lb, qk[m * NT + lane] = uint16_t(k);
        ++m;
      }
    }
  }
  const bool scan = parts && (all || m > Q);
  const bool leave = K == 1 && defer && !scan && m > 0 && m <= QI;
  uint32_t slot = 0;
  if (K == 1 && defer) slot = wave_slot(defer->icount, leave);
  if (leave) {
    if (slot < defer->icap) {
      // the lane's parts by lower bound, equal bounds in list order
      for (int q = 1; q < m; ++q) {
        T v = qlb[q * NT + lane];
        uint16_t kv = qk[q * NT + lane];
        int r = q;
        for (; r > 0 && qlb[(r - 1) * NT + lane] > v; --r) qlb[r * NT + lane] = qlb[(r - 1) * NT + lane], qk[r * NT + lane] = qk[(r - 1) * NT + lane];
        qlb[r * NT + lane] = v, qk[r * NT + lane] = kv;
      }
      uint32_t* w = defer->items + size_t(slot) * Item<T>::W;
      w[0] = uint32_t(point), w[1] = uint32_t(m) | (defer->part << 8);
      item_state(w, b);
      for (int q = 0; q < m; ++q) item_put(w, Item<T>::LB + q * int(sizeof(T) / 4), qlb[q * NT + lane]), w[Item<T>::PR + q] = cand[qk[q * NT + lane]];
      b.code = DEFERRED;
      return;
    }
  }
  T last_lb = T(-1);
  uint32_t last_k = 0;
  // At most tb.L.max_rounds rounds: past them the parts not read keep their lower bounds
  // (the least becomes lmin), so the point is undecided rather than taken as exact.
  const int max_rounds = int(tb.L.max_rounds);
  for (int round = 0;; ++round) {
    T bd = b.d;
    if constexpr (SETTLE && sizeof(T) == 8) bd = sk ? mmin(bd, T(sk->ub)) : bd;
    T next_lb = K > 1 ? group_min<K>(bd) : bd;
    int next = -1;
    if (parts && !scan) {
      for (int q = 0; q < m; ++q) {
        T v = qlb[q * NT + lane];
        if (v < next_lb) next_lb = v, next = q;
      }
    } else if (scan) {
      for (uint32_t k = sub; k < n; k += K) {
        uint32_t i = all ? k : cand[k];
        if (K == 1) i = __builtin_amdgcn_readfirstlane(i);
        const uint32_t code = tb.pcode[i];
        if ((code >> 28) < P_LEAF) continue;
        T lb = part_lb(tb, i, code, p, eps);
        bool after = lb > last_lb || (lb == last_lb && k > last_k);
        if (after && lb < next_lb) next_lb = lb, next = int(k);
      }
    }
    if (K > 1 ? !__any(next >= 0) : next < 0) break;
    if (round >= max_rounds) {
      if (next >= 0) b.lmin = mmin(b.lmin, next_lb);
      break;
    }
    if (next >= 0) {
      uint32_t k;
      if (!scan) {
        qlb[next * NT + lane] = TypeEps<T>::big;
        k = qk[next * NT + lane];
      } else {
        k = uint32_t(next);
        last_lb = next_lb, last_k = k;
      }
      try_part<T, SETTLE>(tb, all ? k : cand[k], next_lb, p, b, tm, eps, sk);
    }
  }
}
// The K lanes' answers of a group merged (its lanes' least distance, ties: the lowest
// lane; near-ties of opposite signs among all its lanes' answers and runners-up are
// reported as a runner-up at the same distance).
template <class T, int K> __device__ inline void group_merge(Best<T>& b, T eps) {
  const int lane = threadIdx.x, sub = lane & (K - 1);
  {
    const T dmin = group_min<K>(b.d);
    const int wsub = group_min<K>(b.d == dmin ? sub : K);
    const int src = (lane & ~(K - 1)) + wsub;
    const T sw = __shfl(b.sgn, src);
    // runners-up within 8 rounding steps of the answer on the other side
    const T near = T(8) * eps;
    int tie = (b.d2 - dmin <= near && b.sgn2 != sw) || (sub != wsub && b.code != NONE && b.d - dmin <= near && b.sgn != sw);
    tie = group_or<K>(tie);
    b.lmin = group_min<K>(b.lmin);
    b.unsure = group_or<K>(b.unsure ? 1 : 0) != 0;
    b.host = group_or<K>(b.host ? 1 : 0) != 0;
    b.code = __shfl(b.code, src), b.a = __shfl(b.a, src), b.b = __shfl(b.b, src);
    b.d = dmin, b.sgn = sw;
    b.d2 = tie ? dmin : TypeEps<T>::big, b.sgn2 = tie ? -sw : sw;
  }
}
// A point's whole search over one candidate list.
template <class T, int NT, int Q, int K, int C, bool SETTLE = false>
__device__ Best<T> search(const Tab<T>& tb, v3<T> p, bool valid, T bound, const uint32_t* cand, int nc, T tm, T eps, T* qlb, uint16_t* qk,
                          const Res* defer = nullptr, long point = 0, SSink* sk = nullptr) {
  Best<T> b = best_init<T>(bound);
  search_in<T, NT, Q, K, C, SETTLE>(tb, p, valid, b, cand, nc, tm, eps, qlb, qk, defer, point, sk);
  if (K > 1) group_merge<T, K>(b, eps);
  return b;
}

__device__ inline uint32_t feat_of(const Dev& dv, uint32_t code) {
  uint32_t kind = code >> 28, id = code & 0x0fffffffu;
  if (kind == P_FACE) return id;
  if (kind == P_EDGE) return (1u << 28) | id;
  if (kind == P_VERT) {
    // a cone's tip reports its face
    const double tip = dv.F64[dv.L.vert_f + VERT_F * id + 6];
    return tip > 0.0 ? uint32_t(tip) - 1u : (2u << 28) | id;
  }
  if (kind == P_LEAF) return dv.U[dv.L.leaf_u + LEAF_U * id + 1];
  // a C0 knot line reports its face (solid3.rs:1790-1794)
  const uint32_t e = dv.U[dv.L.cleaf_u + CLEAF_U * id + 1];
  const uint32_t* eu = dv.U + dv.L.edge_u + EDGE_U * e;
  return eu[0] == EDGE_CREASE ? eu[1] : (1u << 28) | e;
}

// A finished double-precision search as (value, flag): exact when a feature answered
// and no undecided part had a lower bound under it; a bound when nothing came within
// the bound; left to the caller (flag 3, the least lower bound) when a host feature
// came under the answer; else undecided (flag 2, the least lower bound).
__device__ inline void classify64(const Best<double>& b, double bound, const Dev& dv, double& val, uint32_t& flag, uint32_t& feat) {
  feat = NONE;
  if (b.code != NONE && !(b.lmin < b.d) && b.d <= bound) {
    val = b.d * b.sgn, flag = 0, feat = feat_of(dv, b.code);
    return;
  }
  double l = fmin(b.d, b.lmin);
  if (l >= bound || (b.code == NONE && !(b.lmin < b.d))) {
    val = bound, flag = 1;
    return;
  }
  val = l, flag = b.host ? 3 : 2;
}

// Block shapes: explicit points 256 a block, one lane each; a brick's 729 nodes a
// block of 256 lanes; a tile of TILE_X x TILE_Y x TILE_Z brick centres a block of 256,
// K = 16 lanes a centre (a centre's search reaches far, and the tiles holding the
// surface are few: the 32^3 centres of the 120 mm NURBS-top block at 0.5 mm took 0.43 ms
// at 4 lanes a centre, 0.16 ms at 8, 0.11 ms at 16, 0.10 ms at 32 in single precision but
// more in double).
template <int S> struct Shape {
  static constexpr int NT = 256;
  static constexpr int K = S == SRC_CENTRES ? 16 : S == SRC_WIDE ? 4 : S == SRC_FAR ? 64 : 1;  // lanes a point
  static constexpr int Q = S == SRC_CENTRES || S == SRC_FAR ? 8 : 16;   // queue entries a lane
  static constexpr int C = S == SRC_NODES ? 1024 : 4096;  // a block's candidates, at most
};
constexpr int TILE_X = 4, TILE_Y = 2, TILE_Z = 2;
static_assert(TILE_X * TILE_Y * TILE_Z * Shape<SRC_CENTRES>::K == Shape<SRC_CENTRES>::NT, "a centres block covers its tile");

// Super-tiles of SUPER^3 tiles: the level's count, and a tile's super-tile.
constexpr int SUPER = 2, SCAP = 2048;
__host__ __device__ inline long supers_of(const int nb[3]) {
  const long tx = (nb[0] + TILE_X - 1) / TILE_X, ty = (nb[1] + TILE_Y - 1) / TILE_Y, tz = (nb[2] + TILE_Z - 1) / TILE_Z;
  return ((tx + SUPER - 1) / SUPER) * ((ty + SUPER - 1) / SUPER) * ((tz + SUPER - 1) / SUPER);
}
__device__ inline long super_of_tile(const Lv& lv, long tile) {
  const long tx = (lv.nb[0] + TILE_X - 1) / TILE_X, ty = (lv.nb[1] + TILE_Y - 1) / TILE_Y;
  const long sx = (tx + SUPER - 1) / SUPER, sy = (ty + SUPER - 1) / SUPER;
  return (tile % tx) / SUPER + sx * (((tile / tx) % ty) / SUPER + sy * ((tile / (tx * ty)) / SUPER));
}

// The tiles of a level, the tile a brick lies in, and the g-th brick of a tile (-1
// past the level).
__host__ __device__ inline long tiles_of(const int nb[3]) {
  return long((nb[0] + TILE_X - 1) / TILE_X) * ((nb[1] + TILE_Y - 1) / TILE_Y) * ((nb[2] + TILE_Z - 1) / TILE_Z);
}
__device__ inline long tile_of(const Lv& lv, long b) {
  const long tx = (lv.nb[0] + TILE_X - 1) / TILE_X, ty = (lv.nb[1] + TILE_Y - 1) / TILE_Y;
  long bx = b % lv.nb[0], by = (b / lv.nb[0]) % lv.nb[1], bz = b / (long(lv.nb[0]) * lv.nb[1]);
  return bx / TILE_X + tx * (by / TILE_Y + ty * (bz / TILE_Z));
}
__device__ inline long tile_point(const Lv& lv, long tile, int g) {
  const long tx = (lv.nb[0] + TILE_X - 1) / TILE_X, ty = (lv.nb[1] + TILE_Y - 1) / TILE_Y;
  long bx = TILE_X * (tile % tx) + g % TILE_X, by = TILE_Y * ((tile / tx) % ty) + (g / TILE_X) % TILE_Y,
       bz = TILE_Z * (tile / (tx * ty)) + g / (TILE_X * TILE_Y);
  if (bx >= lv.nb[0] || by >= lv.nb[1] || bz >= lv.nb[2]) return -1;
  return bx + lv.nb[0] * (by + long(lv.nb[1]) * bz);
}

// Block blk's m-th point for lane threadIdx.x (m counts the lane's rounds; the K
// lanes of a group share theirs), or -1; and the block a point belongs to.
template <int S> __device__ inline long block_point(const Job& j, long blk, int m) {
  constexpr int NT = Shape<S>::NT;
  if (pts<S>()) {
    if (m > 0) return -1;
    long i = blk * (NT / Shape<S>::K) + threadIdx.x / Shape<S>::K;
    return i < j.n ? i : -1;
  }
  if (S == SRC_NODES) {
    int n = threadIdx.x + NT * m;
    return n < 729 ? blk * 729 + n : -1;
  }
  return m > 0 ? -1 : tile_point(j.lv, blk, threadIdx.x / Shape<S>::K);
}
template <int S> __device__ inline long block_of(const Job& j, long i) {
  if (pts<S>()) return i / (Shape<S>::NT / Shape<S>::K);
  if (S == SRC_NODES) return i / 729;
  return tile_of(j.lv, i);
}
template <int S> constexpr int rounds() { return S == SRC_NODES ? 3 : 1; }

// The block's candidates: the primitives within reach of its points' box, where the
// reach covers every point's answer (its bound, or for unbounded points the boundary's
// upper bound over the block's box, from the hierarchy). NT threads cull; scattered
// points set the box from the block's own points (only those flagged `only` when
// only_flag is given), or from the pcount points listed in plist (by their place in
// the block) when it is given.
template <int S, int NT = Shape<S>::NT>
__device__ int block_candidates(const Dev& dv, const Job& j, long blk, uint32_t* cand, uint32_t* tmp, float* key, float* red, int* wave_tot,
                                uint32_t* chunks = nullptr, int* nchunk = nullptr, float* gb_out = nullptr, float* r_out = nullptr,
                                const uint32_t* only_flag = nullptr, uint32_t only = 0, const uint16_t* plist = nullptr, int pcount = 0) {
  float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
  float r = 0.f;
  if (!pts<S>()) {
    // the box of a brick's nodes or of a tile's centres from its two extreme points
    // (coordinates grow with the index), computed as each point's own
    double bound;
    v3<double> a, b;
    if (S == SRC_NODES) {
      double h;
      const v3<double> first = node_first(j, dv, blk, h);
      a = node_at(first, h, 0), b = node_at(first, h, 728);
      bound = h * 1.7320508075688772;
    } else {
      const long tx = (j.lv.nb[0] + TILE_X - 1) / TILE_X, ty = (j.lv.nb[1] + TILE_Y - 1) / TILE_Y;
      const long x0 = TILE_X * (blk % tx), y0 = TILE_Y * ((blk / tx) % ty), z0 = TILE_Z * (blk / (tx * ty));
      const long x1 = min(x0 + TILE_X, long(j.lv.nb[0])) - 1, y1 = min(y0 + TILE_Y, long(j.lv.nb[1])) - 1, z1 = min(z0 + TILE_Z, long(j.lv.nb[2])) - 1;
      a = job_point<S>(j, dv, x0 + j.lv.nb[0] * (y0 + long(j.lv.nb[1]) * z0), bound);
      b = job_point<S>(j, dv, x1 + j.lv.nb[0] * (y1 + long(j.lv.nb[1]) * z1), bound);
    }
    lo[0] = float(a.x), lo[1] = float(a.y), lo[2] = float(a.z);
    hi[0] = float(b.x), hi[1] = float(b.y), hi[2] = float(b.z);
    r = float(fmin(bound, 3.0e38)) * (1.0f + 1e-5f) + 1e-4f;
  } else {
  auto add = [&](long i) {
    double bound;
    v3<double> p = job_point<S>(j, dv, i, bound);
    const float ub = float(fmin(bound, 3.0e38));
    lo[0] = fminf(lo[0], float(p.x)), lo[1] = fminf(lo[1], float(p.y)), lo[2] = fminf(lo[2], float(p.z));
    hi[0] = fmaxf(hi[0], float(p.x)), hi[1] = fmaxf(hi[1], float(p.y)), hi[2] = fmaxf(hi[2], float(p.z));
    r = fmaxf(r, ub * (1.0f + 1e-5f) + 1e-4f);
  };
  if (plist) {
    for (int k = threadIdx.x; k < pcount; k += NT) add(blk * long(Shape<S>::NT / Shape<S>::K) + long(plist[k]));
  } else {
    for (int m = 0; m < rounds<S>(); ++m) {
      long i = block_point<S>(j, blk, m);
      // (scattered points searched again: only those flagged `only` set the box)
      if (i < 0 || (only_flag && only_flag[i] != only)) continue;
      add(i);
    }
  }
  block_box<NT>(lo, hi, r, red);
  }
  float gb[6] = {lo[0] - 1e-4f, lo[1] - 1e-4f, lo[2] - 1e-4f, hi[0] + 1e-4f, hi[1] + 1e-4f, hi[2] + 1e-4f};
  int nc;
  if (S == SRC_NODES && j.tile_cand && j.list) {
    // the brick's tile's list, when it was kept, instead of the whole primitive list
    long tile = tile_of(j.lv, long(j.list[blk]));
    uint32_t cnt = j.tile_count[tile];
    nc = cnt <= uint32_t(TCAP) ? cull<NT, Shape<S>::C>(dv, gb, r, cand, wave_tot, j.tile_cand + size_t(tile) * TCAP, cnt)
                               : bvh_cull<NT, Shape<S>::C>(dv, gb, r, false, nullptr, cand, tmp, tmp + FCAP, wave_tot, red);
  } else if (S == SRC_CENTRES && j.super_cand) {
    // the tile's super-tile's list (a superset in the same order), when it was kept
    const long sp = super_of_tile(j.lv, blk);
    const uint32_t cnt = j.super_count[sp];
    nc = cnt <= uint32_t(SCAP) ? cull<NT, Shape<S>::C>(dv, gb, r, cand, wave_tot, j.super_cand + size_t(sp) * SCAP, cnt)
                               : bvh_cull<NT, Shape<S>::C>(dv, gb, r, false, nullptr, cand, tmp, tmp + FCAP, wave_tot, red);
  } else {
    // scattered points: the boundary's own bound over the block's box from the hierarchy
    // (past the list, the level-1 nodes over them, taken nearest first in chunks)
    nc = bvh_cull<NT, Shape<S>::C>(dv, gb, r, pts<S>(), &r, cand, tmp, tmp + FCAP, wave_tot, red, chunks, nchunk);
    if (nchunk && *nchunk > 0) sort_chunks<NT>(dv, gb, chunks, *nchunk, reinterpret_cast<unsigned long long*>(tmp));
  }
  if (gb_out)
    for (int k = 0; k < 6; ++k) gb_out[k] = gb[k];
  if (r_out) *r_out = r;
  sort_cand<NT>(dv, gb, cand, tmp, key, nc);
  if (S == SRC_CENTRES && j.tile_cand) {
    for (int k = threadIdx.x; k < nc && k < TCAP; k += NT) j.tile_cand[size_t(blk) * TCAP + k] = cand[k];
    if (threadIdx.x == 0) j.tile_count[blk] = uint32_t(nc);
  }
  return nc;
}

// Super-tile sp's list: the primitives within reach of its centres, culled from the
// whole list (in list order) by a block of 256.
__device__ __forceinline__ void super_block(const Dev& dv, const Job& j, long sp, uint32_t* cand, int* wave_tot, uint32_t* lists, float* red) {
  const long tx = (j.lv.nb[0] + TILE_X - 1) / TILE_X, ty = (j.lv.nb[1] + TILE_Y - 1) / TILE_Y;
  const long sx = (tx + SUPER - 1) / SUPER, sy = (ty + SUPER - 1) / SUPER;
  const long x0 = TILE_X * SUPER * (sp % sx), y0 = TILE_Y * SUPER * ((sp / sx) % sy), z0 = TILE_Z * SUPER * (sp / (sx * sy));
  const long x1 = min(x0 + TILE_X * SUPER, long(j.lv.nb[0])) - 1, y1 = min(y0 + TILE_Y * SUPER, long(j.lv.nb[1])) - 1,
             z1 = min(z0 + TILE_Z * SUPER, long(j.lv.nb[2])) - 1;
  double bound;
  const v3<double> a = job_point<SRC_CENTRES>(j, dv, x0 + j.lv.nb[0] * (y0 + long(j.lv.nb[1]) * z0), bound);
  const v3<double> b = job_point<SRC_CENTRES>(j, dv, x1 + j.lv.nb[0] * (y1 + long(j.lv.nb[1]) * z1), bound);
  const float r = float(fmin(bound, 3.0e38)) * (1.0f + 1e-5f) + 1e-4f;
  const float gb[6] = {float(a.x) - 1e-4f, float(a.y) - 1e-4f, float(a.z) - 1e-4f, float(b.x) + 1e-4f, float(b.y) + 1e-4f, float(b.z) + 1e-4f};
  const int nc = bvh_cull<256, SCAP>(dv, gb, r, false, nullptr, cand, lists, lists + FCAP, wave_tot, red);
  for (int k = threadIdx.x; k < nc && k < SCAP; k += 256) j.super_cand[size_t(sp) * SCAP + k] = cand[k];
  if (threadIdx.x == 0) j.super_count[sp] = uint32_t(nc);
}
__global__ __launch_bounds__(256) void k_super(Dev dv, Job j) {
  __shared__ uint32_t cand[SCAP];
  __shared__ uint32_t lists[2 * FCAP];
  __shared__ float red[256];
  __shared__ int wave_tot[4];
  super_block(dv, j, blockIdx.x, cand, wave_tot, lists, red);
}

// ------------------------------------------------------------------ search kernels

// A point's finished search: in single precision (T = float) the records k_refine
// reads, in double precision the final answer.
template <class T>
__device__ inline void finish_point(const Dev& dv, const Res& res, long i, double bound, const Best<T>& b, T eps) {
  if constexpr (sizeof(T) == 4) {
    bool tie = b.d2 - b.d <= T(8) * eps && b.sgn2 != b.sgn;
    uint32_t info = (b.code != NONE ? 1u : 0u) | (b.sgn < T(0) ? 2u : 0u) | ((b.unsure || tie) ? 4u : 0u) | ((b.lmin < b.d) ? 8u : 0u) |
                    (b.unsure ? 16u : 0u) | (tie ? 32u : 0u);
    res.code[i] = b.code;
    res.pa[i] = float(b.a);
    res.pb[i] = float(b.b);
    res.d32[i] = float((b.lmin < b.d) ? b.lmin : b.d);
    res.info[i] = info;
  } else {
    double val;
    uint32_t flag, feat;
    classify64(b, bound, dv, val, flag, feat);
    res.val[i] = val;
    res.flag[i] = flag;
    if (res.feat) res.feat[i] = feat;
  }
}
template <class T> __device__ inline T search_bound(double bound, T eps) {
  if constexpr (sizeof(T) == 4) return bound < 3.0e38 ? float(bound) * (1.0f + 1e-6f) + float(eps) : 3.0e38f;
  return bound < 1e300 ? bound * (1.0 + 1e-12) : bound;
}

// Shared memory of a search block: the block's candidates, a reduction buffer, and a
// scratch area holding the candidate sort's keys, then the lanes' queues.
template <int S, class T> struct SearchLds {
  static constexpr int NT = Shape<S>::NT, Q = Shape<S>::Q;
  uint32_t cand[Shape<S>::C];
  float red[NT];
  int wave_tot[NT / 64];
  alignas(16) unsigned char scratch[cmax(cmax(SORTC * 12, 2 * FCAP * 4), Q * NT * (int(sizeof(T)) + 2))];
  uint16_t list[S == SRC_NODES ? 736 : 2];  // a brick's unsure nodes, compacted
  int count;
  uint32_t chunks[pts<S>() ? FCAP : 1];  // scattered points past their list: level-2 nodes
  uint32_t sub[pts<S>() ? 64 : 1];       // a chunk's level-1 nodes
  __device__ uint32_t* tmp() { return (uint32_t*)scratch; }
  __device__ float* key() { return (float*)(scratch + SORTC * 4); }
  __device__ T* qlb() { return (T*)scratch; }
  __device__ uint16_t* qk() { return (uint16_t*)(scratch + Q * NT * sizeof(T)); }
};

// A scattered point's search when its block's candidates did not fit the list: chunk by
// chunk over the hierarchy's level-2 nodes within reach, nearest first, each chunk's
// primitives within reach culled into the list and the search carried on over them; a
// chunk no point of the block can find anything nearer in (its box no nearer than any
// point's answer so far) is passed over.
template <int S, class T, bool SETTLE, int NT = Shape<S>::NT, int K = Shape<S>::K, class L = SearchLds<S, T>>
__device__ Best<T> search_chunks(const Dev& dv, const Tab<T>& tb, v3<T> p, bool valid, T bound, L& sm, int nchunk, const float* gb, float r,
                                 T tm, T eps, SSink* sk = nullptr) {
  constexpr int Q = Shape<S>::Q, C = Shape<S>::C;
  Best<T> b = best_init<T>(bound);
  if constexpr (!pts<S>()) {
    return b;  // (only scattered points are searched in chunks)
  } else {
  static_assert(C >= 4096, "a chunk holds a level-2 node's primitives");
  const float* boxes = dv.bvh_box + 6 * size_t(dv.L.bvh_off[2]);
  for (int c = 0; c < nchunk; ++c) {
    const uint32_t node2 = sm.chunks[c];
    const int need = valid && box_dist(p, (const ICAM_K float*)(boxes + 6 * size_t(node2))) < b.d;
    if (!__syncthreads_or(need)) continue;
    const int cnt = bvh_chunk<NT>(dv, gb, r, node2, sm.cand, sm.sub, sm.wave_tot);
    search_in<T, NT, Q, K, C, SETTLE>(tb, p, valid, b, sm.cand, cnt, tm, eps, sm.qlb(), sm.qk(), nullptr, 0, sk);
    __syncthreads();
  }
  if (K > 1) group_merge<T, K>(b, eps);
  return b;
  }
}

// Block blk of a job searched, in single (T = float: records for k_refine) or double
// precision (T = double: final answers). A brick's nodes leave their parts' rounds to
// the dense pass. Scattered points whose candidates do not fit the list are searched in
// chunks: by this call when CH, else their block is queued for k_chunks (the chunked
// search inlined beside the plain one took the search from two waves a SIMD to one).
template <int S, class T, bool CH = false, int KN = Shape<S>::K>
__device__ __forceinline__ void search_block(const Dev& dv, const Job& j, const Res& res, long blk, SearchLds<S, T>& sm) {
  constexpr int NT = Shape<S>::NT;
  constexpr int Q = Shape<S>::Q;
  int nchunk = 0;
  float gb[6], rr = 0.f;
  const int nc = block_candidates<S>(dv, j, blk, sm.cand, sm.tmp(), sm.key(), sm.red, sm.wave_tot, pts<S>() ? sm.chunks : nullptr, &nchunk, gb, &rr);
  const bool chunked = pts<S>() && nchunk > 0;
  if constexpr (pts<S>() && !CH) {
    if (chunked && res.cqueue) {
      if (threadIdx.x == 0) res.cqueue[atomicAdd(res.ccount, 1u)] = uint32_t(blk);
      return;
    }
  }
  if (threadIdx.x == 0 && nc > Shape<S>::C && !chunked) atomicAdd(&res.counters[6], 1ull);
  const Tab<T> tb = tab_of<T>(dv);
  const T eps = sizeof(T) == 4 ? T(dv.L.eps32) : T(dv.L.eps64);
  const T tm = sizeof(T) == 4 ? T(16) * eps : T(0);
  // lanes a point: a brick's nodes take KN (the tuned lanes a node; a node's parts are
  // split over its lanes, and only nodes searched a lane each leave their parts to the
  // dense pass), every other source its shape's
  static_assert(S == SRC_NODES || KN == Shape<S>::K, "only the nodes' lanes are tuned");
  constexpr int K = S == SRC_NODES ? KN : Shape<S>::K;
  constexpr int PPR = NT / K;  // nodes a round
  constexpr int NRD = S == SRC_NODES ? (729 + PPR - 1) / PPR : rounds<S>();
  const bool owner = (threadIdx.x & (K - 1)) == 0;
  double nh = 0.0;
  const v3<double> nfirst = S == SRC_NODES ? node_first(j, dv, blk, nh) : mk3(0.0, 0.0, 0.0);
  for (int m = 0; m < NRD; ++m) {
    const int nn = int(threadIdx.x) / K + PPR * m;  // the node, for a brick's nodes
    long i = S == SRC_NODES ? (nn < 729 ? blk * 729 + nn : -1) : block_point<S>(j, blk, m);
    if (K == 1 && i < 0 && !chunked) continue;
    const bool valid = i >= 0;
    double bound = 0.0;
    v3<double> p;
    if (S == SRC_NODES)
      p = node_at(nfirst, nh, nn), bound = nh * 1.7320508075688772;
    else
      p = valid ? job_point<S>(j, dv, i, bound) : mk3(0.0, 0.0, 0.0);
    Best<T> b;
    if constexpr (CH) {
      b = chunked ? search_chunks<S, T, false>(dv, tb, cv3<T>(p), valid, search_bound<T>(bound, eps), sm, nchunk, gb, rr, tm, eps)
                  : search<T, NT, Q, K, Shape<S>::C>(tb, cv3<T>(p), valid, search_bound<T>(bound, eps), sm.cand, nc, tm, eps, sm.qlb(), sm.qk());
    } else {
      b = search<T, NT, Q, K, Shape<S>::C>(tb, cv3<T>(p), valid, search_bound<T>(bound, eps), sm.cand, nc, tm, eps, sm.qlb(), sm.qk(),
                                           S == SRC_NODES && K == 1 && res.items ? &res : nullptr, i);
    }
    if (!valid || !owner || b.code == DEFERRED) continue;
    finish_point<T>(dv, res, i, bound, b, eps);
  }
}

template <int S, class T, int KN = Shape<S>::K>
__global__ __launch_bounds__(Shape<S>::NT) void k_search(Dev dv, Job j, Res res) {
  __shared__ SearchLds<S, T> sm;
  search_block<S, T, false, KN>(dv, j, res, blockIdx.x, sm);
}
// The points of the scattered blocks k_search queued (their candidates past the list): a
// block of four points, a wave a point (four points' box reaches far less than their whole
// block's, so their list fits where the block's did not; past it, chunks), searched in
// double precision to their final answers, in the mixed form too, where their records are
// marked final so the last correction leaves them (unbounded points far from the surface,
// each reaching many parts: on the large mixed solid's uniform points the block's chunked
// search took 1.7 s for 2^20 queries). Blocks take the next four points from a counter:
// their searches range from a few parts to thousands.
constexpr uint32_t INFO_FINAL = 0x80000000u;
// A settle mark (smark) per block: SM_QUEUED once the block is queued for the settle,
// SM_FAR when its points were searched in chunks (k_far). A far block's undecided
// points are settled by k_settle_far in SETTLE_SLICES queue entries (the slice in the
// entry's top byte), each taking the points g of the block with g % SETTLE_SLICES ==
// slice: the search before a settle round costs a far point thousands of parts, and one
// wave holding a far block's round was the whole settle's critical path (on the large
// mixed solid's uniform points, one round took 456 of the settle's 459 ms while the
// wave-time summed over all rounds would have filled the GPU for about 40 ms). The
// slices' entries follow the other blocks' in the queue, from entry `blocks` on, counted
// by scount[2] (scount[0] the others, [1] and [3] the kernels' next entries).
constexpr uint32_t SM_QUEUED = 1u, SM_FAR = 2u;
constexpr int SETTLE_SLICES = 16;
template <int S> __device__ inline long settle_blocks(const Job& j) { return (j.n + Shape<S>::NT / Shape<S>::K - 1) / (Shape<S>::NT / Shape<S>::K); }
template <int S>
__global__ __launch_bounds__(256) void k_far(Dev dv, Job j, Res res) {
  static_assert(pts<S>() && S != SRC_FAR, "scattered blocks");
  constexpr int FP = Shape<SRC_FAR>::NT / Shape<SRC_FAR>::K;   // points a far block
  constexpr int SUB = (Shape<S>::NT / Shape<S>::K) / FP;         // far blocks a queued block
  static_assert(SUB * FP == Shape<S>::NT / Shape<S>::K, "a queued block splits into far blocks");
  __shared__ SearchLds<SRC_FAR, double> sm;
  __shared__ uint32_t wq;
  const uint32_t total = *res.ccount * uint32_t(SUB);
  for (;;) {
    if (threadIdx.x == 0) wq = atomicAdd(res.ccount + 1, 1u);
    __syncthreads();
    const uint32_t q = wq;
    if (q >= total) break;
    const long fb = long(res.cqueue[q / SUB]) * SUB + long(q % SUB);
    const long i = fb * FP + long(threadIdx.x) / Shape<SRC_FAR>::K;
    // (the block's undecided points are settled in slices: SM_FAR)
    if (threadIdx.x == 0 && q % SUB == 0) atomicOr(&res.smark[res.cqueue[q / SUB]], SM_FAR);
    if (res.info && (threadIdx.x & (Shape<SRC_FAR>::K - 1)) == 0 && i < j.n) res.info[i] = INFO_FINAL;
    search_block<SRC_FAR, double, true>(dv, j, res, fb, sm);
    __syncthreads();
  }
}

// One deferred point of the dense pass: its parts in order of their lower bounds
// while one can beat its answer.
template <int S, class T>
__device__ __forceinline__ void leaves_item(const Dev& dv, const Job& j, const Res& res, const uint32_t* w) {
  const Tab<T> tb = tab_of<T>(dv);
  const T eps = sizeof(T) == 4 ? T(dv.L.eps32) : T(dv.L.eps64);
  const T tm = sizeof(T) == 4 ? T(16) * eps : T(0);
  const long i = long(w[0]);
  const int m = int(w[1] & 0xffu);
  Best<T> b = item_best<T>(w);
  double bound;
  const v3<T> p = cv3<T>(job_point<S>(j, dv, i, bound));
  for (int q = 0; q < m; ++q) {
    const T lb = item_get<T>(w, Item<T>::LB + q * int(sizeof(T) / 4));
    if (!(lb < b.d)) break;
    try_part(tb, w[Item<T>::PR + q], lb, p, b, tm, eps);
  }
  finish_point<T>(dv, res, i, bound, b, eps);
}

template <int S, class T>
__global__ __launch_bounds__(256) void k_leaves(Dev dv, Job j, Res res) {
  const uint32_t n = min(*res.icount, res.icap);
  for (uint32_t it = blockIdx.x * 256 + threadIdx.x; it < n; it += gridDim.x * 256) leaves_item<S, T>(dv, j, res, res.items + size_t(it) * Item<T>::W);
}

// The last correction in double precision of point i's single-precision winner; a
// point the search was unsure of is marked (flag 3) and its block queued (entries are
// part << 24 | block) for the fallback.
// The first half, from the point's record (info): an unsure point is queued, a point
// with no winner is written; returns whether the point waits for refine_found.
template <int S>
__device__ __forceinline__ bool refine_head(const Dev& dv, const Job& j, const Res& res, long i, uint32_t info) {
  if (info & INFO_FINAL) return false;  // searched in double precision already (k_far)
  // an unsure point that is undecided too (a part past its certificate lies under its
  // answer) goes to the settle, whose search in double precision starts afresh: the
  // fallback's answer would be searched again there (the large mixed solid's bake: 413 k
  // unsure nodes, most of them beside the pockets' rounded corners the settle reads);
  // a brick centre is not settled, so it takes the fallback
  const bool unsure = (info & 4u) && !(S != SRC_CENTRES && (info & 8u));
  wave_count(&res.counters[1], unsure);
  wave_count(&res.counters[3], unsure && (info & 16u));
  wave_count(&res.counters[4], unsure && (info & 32u));
  if (unsure) {
    res.flag[i] = 3;
    long blk = block_of<S>(j, i);
    if (res.qmark[blk] == 0u && atomicOr(&res.qmark[blk], 1u) == 0u) res.queue[atomicAdd(res.qcount, 1u)] = (res.part << 24) | uint32_t(blk);
    return false;
  }
  const bool found = info & 1u, undecided = info & 8u;
  if (found && !undecided) return true;
  double bound;
  job_point<S>(j, dv, i, bound);
  const double d = res.d32[i];
  double val;
  uint32_t flag;
  if (!undecided) {
    val = bound, flag = 1;
  } else if (info & 4u) {
    val = d, flag = 2;  // unsure too: the settle decides, past the bound as well
  } else {
    val = d >= bound ? bound : d;
    flag = d >= bound ? 1 : 2;
  }
  res.val[i] = val;
  res.flag[i] = flag;
  if (res.feat) res.feat[i] = NONE;
  return false;
}
// The second half: the winner's foot and sign again in double precision.
template <int S>
__device__ __forceinline__ void refine_found(const Dev& dv, const Job& j, const Res& res, long i, uint32_t info) {
  double bound;
  v3<double> p = job_point<S>(j, dv, i, bound);
  uint32_t flag = 0, feat = NONE;
  double val, d = res.d32[i];
  const uint32_t code = res.code[i];
  double sg = (info & 2u) ? -1.0 : 1.0;
  refine64(tab_of<double>(dv), code, double(res.pa[i]), double(res.pb[i]), p, d, sg);
  if (d <= bound) {
    val = d * sg, feat = feat_of(dv, code);
  } else {
    val = bound, flag = 1;
  }
  res.val[i] = val;
  res.flag[i] = flag;
  if (res.feat) res.feat[i] = feat;
}
template <int S>
__device__ __forceinline__ void refine_point(const Dev& dv, const Job& j, const Res& res, long i) {
  const uint32_t info = res.info[i];
  if (refine_head<S>(dv, j, res, i, info)) refine_found<S>(dv, j, res, i, info);
}

// A band brick's nodes corrected by one block. The nodes that wait for the correction
// are listed in LDS, sub-patch winners from the front and the other kinds from the
// back, so each wave runs one kind's correction on full lanes: one lane a node left
// the far nodes' lanes idle beside the Newton solves, and mixed the kinds' branches in
// a wave.
struct RefineLds {
  uint16_t list[729];
  uint32_t na, nb;
};
template <int NT>
__device__ __forceinline__ void refine_brick(const Dev& dv, const Job& j, const Res& res, long blk, RefineLds& sm) {
  if (threadIdx.x == 0) sm.na = 0u, sm.nb = 0u;
  __syncthreads();
  for (int n = threadIdx.x; n < 729; n += NT) {
    const long i = blk * 729 + n;
    const uint32_t info = res.info[i];
    const bool wait = refine_head<SRC_NODES>(dv, j, res, i, info);
    const bool leaf = wait && (res.code[i] >> 28) == P_LEAF;
    const uint32_t ka = wave_slot(&sm.na, leaf), kb = wave_slot(&sm.nb, wait && !leaf);
    if (wait) sm.list[leaf ? ka : 728u - kb] = uint16_t(n);
  }
  __syncthreads();
  const int na = int(sm.na), nw = na + int(sm.nb);
  for (int k = threadIdx.x; k < nw; k += NT) {
    const long i = blk * 729 + sm.list[k < na ? k : 728 - (k - na)];
    refine_found<SRC_NODES>(dv, j, res, i, res.info[i]);
  }
}

template <int S>
__global__ __launch_bounds__(256) void k_refine(Dev dv, Job j, Res res) {
  long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i < j.n) refine_point<S>(dv, j, res, i);
}
__global__ __launch_bounds__(256) void k_refine_bricks(Dev dv, Job j, Res res, long nbricks) {
  __shared__ RefineLds sm;
  for (long blk = blockIdx.x; blk < nbricks; blk += gridDim.x) {
    refine_brick<256>(dv, j, res, blk, sm);
    __syncthreads();
  }
}

// A queued block searched again in double precision for its points flagged `want`:
// flag 3, the fallback of the mixed form; flag 4, the settle (parts past their
// convexity certificate settled by the certified search, as the reference settles
// them on the CPU, crates/icam-sim/src/partcache.rs:563-751).
template <int S, bool SETTLE, int RK = 4>
__device__ __forceinline__ void redo_block(const Dev& dv, const Job& j, const Res& res, long blk, SearchLds<S, double>& sm) {
  static_assert(!SETTLE, "the settle is settle_block's");
  const uint32_t want = SETTLE ? 4u : 3u;
  constexpr int NT = Shape<S>::NT;
  int nchunk = 0;
  float gb[6], rr = 0.f;
  const int nc = block_candidates<S>(dv, j, blk, sm.cand, sm.tmp(), sm.key(), sm.red, sm.wave_tot, pts<S>() ? sm.chunks : nullptr, &nchunk, gb, &rr,
                                     pts<S>() ? res.flag : nullptr, want);
  const bool chunked = pts<S>() && nchunk > 0;
  const Tab<double> tb = tab_of<double>(dv);
  if constexpr (S == SRC_NODES && !SETTLE) {
    // the few unsure nodes of a brick (near trim curves, near ties), compacted and
    // searched RK lanes a node: one lane a node left nearly every lane of the block idle.
    // RK follows the nodes' search: 4 where it runs 1 or 2 lanes a node (the NURBS-top
    // block and the 64-part batch: 4 beat 8 and 16), 8 where it runs 4 or 8 (the large
    // mixed solid's fallback 20.8 -> 15.1 ms)
    constexpr int RQ = 8;
    if (threadIdx.x == 0) sm.count = 0;
    __syncthreads();
    for (int n = threadIdx.x; n < 729; n += NT)
      if (res.flag[blk * 729 + n] == want) sm.list[atomicAdd(&sm.count, 1)] = uint16_t(n);
    __syncthreads();
    const int cnt = sm.count;
    const bool lead = (threadIdx.x & (RK - 1)) == 0;
    for (int base = 0; base < cnt; base += NT / RK) {
      const int idx = base + int(threadIdx.x) / RK;
      const bool valid = idx < cnt;
      const long i = valid ? blk * 729 + long(sm.list[idx]) : -1;
      double bound = 0.0;
      const v3<double> p = valid ? job_point<S>(j, dv, i, bound) : mk3(0.0, 0.0, 0.0);
      const Best<double> b = search<double, NT, RQ, RK, Shape<S>::C, false>(tb, p, valid, bound < 1e300 ? bound * (1.0 + 1e-12) : bound, sm.cand, nc,
                                                                            0.0, dv.L.eps64, sm.qlb(), sm.qk());
      if (!valid || !lead) continue;
      double val;
      uint32_t flag, feat;
      classify64(b, bound, dv, val, flag, feat);
      res.val[i] = val;
      res.flag[i] = flag;
      if (res.feat) res.feat[i] = feat;
    }
    return;
  }
  constexpr int K = Shape<S>::K;
  const bool owner = (threadIdx.x & (K - 1)) == 0;
  for (int m = 0; m < rounds<S>(); ++m) {
    long i = block_point<S>(j, blk, m);
    const bool valid = i >= 0 && res.flag[i] == want;
    if (K == 1 && !valid && !chunked) continue;
    double bound = 0.0;
    v3<double> p = valid ? job_point<S>(j, dv, i, bound) : mk3(0.0, 0.0, 0.0);
    const double sb = bound < 1e300 ? bound * (1.0 + 1e-12) : bound;
    Best<double> b = chunked ? search_chunks<S, double, SETTLE>(dv, tb, p, valid, sb, sm, nchunk, gb, rr, 0.0, dv.L.eps64)
                             : search<double, NT, Shape<S>::Q, K, Shape<S>::C, SETTLE>(tb, p, valid, sb, sm.cand, nc, 0.0, dv.L.eps64, sm.qlb(), sm.qk());
    if (!valid || !owner) continue;
    double val;
    uint32_t flag, feat;
    classify64(b, bound, dv, val, flag, feat);
    res.val[i] = val;
    res.flag[i] = flag;
    if (res.feat) res.feat[i] = feat;
  }
}

template <int S, int RK = 4>
__global__ __launch_bounds__(Shape<S>::NT) void k_fallback(Dev dv, Job j, Res res) {
  __shared__ SearchLds<S, double> sm;
  const uint32_t total = *res.qcount;
  for (uint32_t qb = blockIdx.x; qb < total; qb += gridDim.x) {
    redo_block<S, false, RK>(dv, j, res, long(res.queue[qb] & 0xffffffu), sm);
    __syncthreads();
  }
}

// Point i left undecided (flag 2) is marked (flag 4) and its block queued for the
// settle.
template <int S>
__device__ __forceinline__ void mark_point(const Job& j, const Res& res, long i) {
  const bool undecided = res.flag[i] == 2;
  wave_count(&res.counters[5], undecided);
  if (!undecided) return;
  res.flag[i] = 4;
  long blk = block_of<S>(j, i);
  if ((res.smark[blk] & SM_QUEUED) == 0u && (atomicOr(&res.smark[blk], SM_QUEUED) & SM_QUEUED) == 0u) {
    if (pts<S>() && (res.smark[blk] & SM_FAR)) {
      const size_t at = size_t(settle_blocks<S>(j)) + atomicAdd(res.scount + 2, uint32_t(SETTLE_SLICES));
      for (int k = 0; k < SETTLE_SLICES; ++k) res.squeue[at + k] = (uint32_t(k) << 24) | uint32_t(blk);
    } else {
      res.squeue[atomicAdd(res.scount, 1u)] = (res.part << 24) | uint32_t(blk);
    }
  }
}

template <int S>
__global__ __launch_bounds__(256) void k_mark_settle(Job j, Res res) {
  long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i < j.n) mark_point<S>(j, res, i);
}

// The settle, one wave a queued block. The block's marked points (flag 4) are listed in
// LDS and taken NT / K at a time (a round): their candidates are culled for their box
// and each point is searched again in double precision by its K lanes, every part past
// its certificate handed to the round's settle as a root piece instead of being read;
// then the wave works through the round's pool of pieces together, each lane taking one
// piece (a quarter of a sub-patch or a half of a curve part: settle_quarter,
// settle_half) from those nearest their points' bounds, and the lane of each point
// merging its pieces' steps, in lane order, into the point's answer and bound. A
// queued block's marked points are few (a brick's nodes near a concave corner of a wall,
// a block's points near a trim curve) and their searches are long and uneven: with a
// point to a lane, most lanes waited on the longest search, whose queue of pieces lived
// in private memory. Points whose parts or pieces the round could not hold (points on a
// pole's axis see every sub-patch around the pole at nearly one distance) are searched
// again in rounds of a quarter as many points, down to one point a round.
constexpr int SETTLE_NT = 64;
constexpr int SPOOL = 704;  // pieces a round's pool holds
struct SRec {               // a lane's step, for its point's lane to merge
  double sd, lost, d, pa, pb, sgn;
  uint32_t code, flags;     // flags: the point's slot (bits 0-5), SR_*
};
constexpr uint32_t SR_VALID = 1u << 8, SR_ANS = 1u << 9, SR_UNSURE = 1u << 10, SR_OVER = 1u << 11;
template <int S> struct SettleLds {
  static constexpr int NT = SETTLE_NT, Q = Shape<S>::Q, C = Shape<S>::C;
  // lanes a point in the search, which takes a round's points NT / K at a time: a centre
  // its first search's 16 (its reach holds many parts), a brick's node 4 (the large mixed
  // solid's batched settle 39.6 ms at 1, 31.4 at 4, 35.9 at 8), a scattered point
  // 1 -- but scattered points whose candidates come in chunks (far from the surface, each
  // reaching many parts) 4 (settle_block; one lane each took the large mixed solid's 2^20
  // queries from 0.58 to 4.2 s), 16 in a slice of a far block (a few points). A round
  // holds 64 points whatever K is: rounds of 16 at 4 lanes a point left the pool a quarter
  // of the pieces and three lanes of four idle (the trimmed plate's 2^22 queries' settle
  // 72.9 against 45.6 ms at 1 lane a point).
  static constexpr int K = S == SRC_CENTRES ? Shape<S>::K : S == SRC_NODES ? 4 : 1;
  static constexpr int NR = NT;                                // points a round
  // root pieces a round (what workgroup memory leaves for them: a point under a pole
  // has a root in each of a hundred sub-patches around it)
  static constexpr int ROOTS = S == SRC_NODES ? 896 : S == SRC_POINTS ? 448 : S == SRC_CENTRES ? 1024 : 512;
  // the points a queued block holds: a brick's nodes, a tile's centres, a block's points
  static constexpr int NP = S == SRC_NODES ? 729 : S == SRC_CENTRES ? TILE_X * TILE_Y * TILE_Z : Shape<S>::NT / Shape<S>::K;
  union {
    struct {  // a round's search
      uint32_t cand[C];
      alignas(16) unsigned char scratch[cmax(cmax(SORTC * 12, 2 * FCAP * 4), Q * NT * 10)];
      uint32_t chunks[pts<S>() ? FCAP : 1];
      uint32_t sub[pts<S>() ? 64 : 1];
    };
    struct {  // its settle
      SPiece items[SPOOL];
      SRec rec[NT];
      int sel[NT];  // the pieces the lanes take
      int hist[NT];
    };
  };
  float red[NT];
  int wave_tot[1];
  uint16_t list[NP];  // the marked points, by their place in the block
  int count;
  // the round's root pieces
  float rlb[ROOTS];
  uint32_t rpart[ROOTS];
  uint8_t rslot[ROOTS];
  int rcount;
  int pcount;  // the pool's length (may run past SPOOL)
  // the round's points: pruning bound, reads and root pieces, whether the settle
  // answered for it, whether the round could not hold its parts or pieces
  double pr[NR];
  int pops[NR], roots[NR];
  uint8_t got[NR], over[NR];
  __device__ uint32_t* tmp() { return (uint32_t*)scratch; }
  __device__ float* key() { return (float*)(scratch + SORTC * 4); }
  __device__ double* qlb() { return (double*)scratch; }
  __device__ uint16_t* qk() { return (uint16_t*)(scratch + Q * NT * 8); }
};
template <int S, bool FAR = false>
__device__ __forceinline__ void settle_block(const Dev& dv, const Job& j, const Res& res, long blk, SettleLds<S>& sm, int slice = 0, int ns = 1) {
  constexpr int NT = SETTLE_NT, NP = SettleLds<S>::NP, Q = Shape<S>::Q, C = Shape<S>::C;
  const int lane = threadIdx.x;
  // the block's g-th point
  auto point_of = [&](int g) -> long {
    if (S == SRC_NODES) return blk * 729 + g;
    if (S == SRC_CENTRES) return tile_point(j.lv, blk, g);
    return blk * NP + g;
  };
  if (lane == 0) sm.count = 0;
  __syncthreads();
  for (int g = lane; g < NP; g += NT) {
    const long i = point_of(g);
    if (i >= 0 && i < j.n && res.flag[i] == 4u && g % ns == slice) sm.list[atomicAdd(&sm.count, 1)] = uint16_t(g);
  }
  __syncthreads();
  if (sm.count == 0) return;
  int nchunk = 0, nc = 0;
  float gb[6], rr = 0.f;
  bool chunked = false;
  const Tab<double> tb = tab_of<double>(dv);
  const SPool pool{sm.items, &sm.pcount, SPOOL};
  // the rounds, K lanes a point in the search
  auto settle_rounds = [&](auto kc) {
  constexpr int K = decltype(kc)::value, NR = NT;
  for (int R = NR, cnt = sm.count;; R = max(1, R / 4)) {
  int again = 0;  // points to search again in smaller rounds, listed in place
  for (int base = 0; base < cnt; base += R) {
    const int npt = min(R, cnt - base);
    // the candidates, every round (the pool shares their memory): a brick's or a tile's;
    // scattered points' for the round's points alone (a queued block's marked points may
    // lie far apart, far from the surface: the box of all of them could reach past the list)
    __syncthreads();
    nchunk = 0;
    nc = block_candidates<S, NT>(dv, j, blk, sm.cand, sm.tmp(), sm.key(), sm.red, sm.wave_tot, pts<S>() ? sm.chunks : nullptr, &nchunk, gb, &rr,
                                 nullptr, 0u, pts<S>() ? sm.list + base : nullptr, pts<S>() ? npt : 0);
    chunked = pts<S>() && nchunk > 0;
    if (lane == 0) sm.rcount = 0;
    if (lane < NR) sm.roots[lane] = 0, sm.pops[lane] = 0, sm.got[lane] = 0, sm.over[lane] = 0;
    __syncthreads();
    // the search, with the parts past their certificates handed over as root pieces: NT / K
    // points at a time (a sub-round), K lanes a point; lane g < npt then holds point g's
    // answer and its probes' bound (its group's first lane had them)
    Best<double> ob = best_init<double>(0.0);
    double oub = 1e300;
    for (int sub = 0; sub < npt; sub += NT / K) {
      const int g = sub + lane / K;
      const bool valid = g < npt;
      const long i = valid ? point_of(sm.list[base + g]) : -1;
      double bound = 0.0;
      const v3<double> p = valid ? job_point<S>(j, dv, i, bound) : mk3(0.0, 0.0, 0.0);
      const double sb = bound < 1e300 ? bound * (1.0 + 1e-12) : bound;
      SSink sink{sm.rlb, sm.rpart, sm.rslot, &sm.rcount, SettleLds<S>::ROOTS, valid ? g : 0, sm.over, 1e300};
      Best<double> b;
      if constexpr (pts<S>()) {
        b = chunked ? search_chunks<S, double, true, NT, K>(dv, tb, p, valid, sb, sm, nchunk, gb, rr, 0.0, dv.L.eps64, &sink)
                    : search<double, NT, Q, K, C, true>(tb, p, valid, sb, sm.cand, nc, 0.0, dv.L.eps64, sm.qlb(), sm.qk(), nullptr, 0, &sink);
      } else {
        b = search<double, NT, Q, K, C, true>(tb, p, valid, sb, sm.cand, nc, 0.0, dv.L.eps64, sm.qlb(), sm.qk(), nullptr, 0, &sink);
      }
      const double ub = K > 1 ? group_min<K>(sink.ub) : sink.ub;
      if constexpr (K > 1) {
        const int src = ((lane - sub) * K) & (NT - 1);
        Best<double> x;
        x.d = __shfl(b.d, src), x.code = uint32_t(__shfl(int(b.code), src)), x.a = __shfl(b.a, src), x.b = __shfl(b.b, src);
        x.sgn = __shfl(b.sgn, src), x.lmin = __shfl(b.lmin, src), x.d2 = __shfl(b.d2, src), x.sgn2 = __shfl(b.sgn2, src);
        x.unsure = __shfl(int(b.unsure), src) != 0, x.host = __shfl(int(b.host), src) != 0;
        const double xu = __shfl(ub, src);
        if (lane >= sub && lane < sub + NT / K) ob = x, oub = xu;
      } else {
        ob = b, oub = ub;
      }
    }
    const bool own = lane < npt;
    long oi = -1;
    double obound = 0.0, olost = 1e300;
    if (own) {
      oi = point_of(sm.list[base + lane]);
      job_point<S>(j, dv, oi, obound);
      sm.pr[lane] = fmin(ob.d, oub);
    }
    __syncthreads();
    // the root pieces: counted by point, then let into the pool (whose memory the search
    // has left) while it is less than a quarter full, in the order the searches met them
    // (nearest first for each lane); in rounds of one point, one at a time once fewer
    // than NT pieces are left: a point just under a pole has a chain of pieces down to
    // the pole in each of the sub-patches around it, about a hundred pieces each, more
    // together than the pool holds
    const int nroots = min(sm.rcount, SettleLds<S>::ROOTS);
    for (int k = lane; k < nroots; k += NT) atomicAdd(&sm.roots[sm.rslot[k]], 1);
    int n = 0, rnext = 0;
    __syncthreads();
    // a piece's lower bound over its point's bound: the lanes take the lowest
    auto key_of = [&](int k) { return sm.items[k].lb / sm.pr[sp_slot(sm.items[k].meta)]; };
    for (;;) {
      // the pieces that can still come under their point's bound, kept in order
      int m = 0;
      for (int c0 = 0; c0 < n; c0 += NT) {
        const int k = c0 + lane;
        SPiece x{};
        bool keep = false;
        if (k < n) {
          x = sm.items[k];
          const double prb = sm.pr[sp_slot(x.meta)];
          keep = x.lb < prb - SETTLE_EPS && prb > SETTLE_EPS;
        }
        const unsigned long long mask = __ballot(keep);
        __syncthreads();
        if (keep) sm.items[m + int(__popcll(mask & __lanemask_lt()))] = x;
        m += int(__popcll(mask));
        __syncthreads();
      }
      n = m;
      if (rnext < nroots && n < (R > 1 ? SPOOL / 4 : NT)) {
        const int a = R > 1 ? min(nroots - rnext, SPOOL / 2 - n) : 1;
        for (int k = lane; k < a; k += NT)
          sm.items[n + k] = SPiece{double(sm.rlb[rnext + k]), 0ull, 0ull, sm.rpart[rnext + k], sp_meta(sm.rslot[rnext + k], 0, 0, 0)};
        n += a, rnext += a;
        __syncthreads();
      }
      if (n == 0) break;
      // the NT pieces lowest under their bounds, or all: by a histogram of the ratio in NT
      // bins over the pieces' own range of it (near a pole every piece lies within a
      // thousandth of its point's bound: bins over [0, 1) took them in pool order, breadth
      // first, and the pool overflowed)
      const int take = min(n, NT);
      if (n > NT) {
        double kmin = 1e300, kmax = -1e300;
        for (int k = lane; k < n; k += NT) {
          const double x = key_of(k);
          kmin = fmin(kmin, x), kmax = fmax(kmax, x);
        }
        for (int o = 1; o < NT; o <<= 1) kmin = fmin(kmin, __shfl_xor(kmin, o)), kmax = fmax(kmax, __shfl_xor(kmax, o));
        const double scale = kmax > kmin ? double(NT) / (kmax - kmin) : 0.0;
        auto bin_of = [&](int k) {
          const double x = (key_of(k) - kmin) * scale;
          if (!(x >= 0.0)) return 0;
          return x < double(NT - 1) ? int(x) : NT - 1;
        };
        sm.hist[lane] = 0;
        __syncthreads();
        for (int k = lane; k < n; k += NT) atomicAdd(&sm.hist[bin_of(k)], 1);
        __syncthreads();
        int c = sm.hist[lane];
        for (int o = 1; o < NT; o <<= 1) {
          const int t = __shfl_up(c, o);
          if (lane >= o) c += t;
        }
        const int cut = __ffsll((long long)__ballot(c >= NT)) - 1;  // the bin that fills the lanes
        const int below = __shfl(c, cut) - sm.hist[cut];
        int lo = 0, at = 0;
        for (int c0 = 0; c0 < n; c0 += NT) {
          const int k = c0 + lane;
          const int bin = k < n ? bin_of(k) : NT;
          const unsigned long long ml = __ballot(bin < cut), me = __ballot(bin == cut);
          if (bin < cut) sm.sel[lo + int(__popcll(ml & __lanemask_lt()))] = k;
          if (bin == cut) {
            const int q = at + int(__popcll(me & __lanemask_lt()));
            if (q < NT - below) sm.sel[below + q] = k;
          }
          lo += int(__popcll(ml)), at += int(__popcll(me));
        }
      } else {
        sm.sel[lane] = lane;
      }
      __syncthreads();
      const int mine = lane < take ? sm.sel[lane] : -1;
      SPiece it{};
      if (mine >= 0) it = sm.items[mine];
      if (lane == 0) sm.pcount = n;
      __syncthreads();
      // each lane's step (its piece leaves the pool: dropped at the next pass)
      SRec r;
      r.flags = 0, r.sd = 1e300, r.lost = 1e300, r.d = 0.0, r.pa = 0.0, r.pb = 0.0, r.sgn = 1.0, r.code = NONE;
      if (mine >= 0) {
        sm.items[mine].lb = 1e300;
        const int slot = sp_slot(it.meta);
        r.flags = uint32_t(slot) | SR_VALID;
        const double prb = sm.pr[slot];
        if (it.lb < prb - SETTLE_EPS && prb > SETTLE_EPS) {
          if (sm.pops[slot] >= SETTLE_POPS * max(sm.roots[slot], 1)) {
            r.lost = it.lb;  // out of reads: the point is left undecided
          } else {
            SOut o;
            o.sd = 1e300, o.lost = 1e300, o.ans = false, o.unsure = false, o.over = false, o.d = 0.0, o.pa = 0.0, o.pb = 0.0, o.sgn = 1.0, o.code = NONE;
            double pb_;
            const v3<double> pp = job_point<S>(j, dv, point_of(sm.list[base + slot]), pb_);
            if (it.part & SP_CURVE)
              settle_half(tb, it, pp, prb, pool, o);
            else
              settle_quarter(tb, it, pp, prb, sm.got[slot] != 0, pool, o);
            r.sd = o.sd, r.lost = o.lost;
            if (o.ans) r.flags |= SR_ANS, r.d = o.d, r.pa = o.pa, r.pb = o.pb, r.sgn = o.sgn, r.code = o.code;
            if (o.unsure) r.flags |= SR_UNSURE;
            if (o.over) r.flags |= SR_OVER;
          }
        }
      }
      sm.rec[lane] = r;
      __syncthreads();
      // each point's lane merges its pieces' steps in lane order
      if (own) {
        double prl = sm.pr[lane];
        bool gl = sm.got[lane] != 0;
        int reads = 0;
        for (int l = 0; l < NT; ++l) {
          const uint32_t fl = sm.rec[l].flags;
          if (!(fl & SR_VALID) || int(fl & 63u) != lane) continue;
          ++reads;
          prl = fmin(prl, sm.rec[l].sd);
          olost = fmin(olost, sm.rec[l].lost);
          if (fl & SR_ANS) offer(ob, sm.rec[l].d, sm.rec[l].code, sm.rec[l].pa, sm.rec[l].pb, sm.rec[l].sgn), gl = true;
          if (fl & SR_UNSURE) ob.unsure = true;
          if (fl & SR_OVER) sm.over[lane] = 1;
        }
        sm.pr[lane] = prl, sm.got[lane] = gl ? 1 : 0, sm.pops[lane] += reads;
      }
      __syncthreads();
      n = min(sm.pcount, SPOOL);
    }
    // a point the round could not hold, whose dropped parts or pieces could still come
    // under its answer, goes back on the list (its slot there is read); otherwise, or in
    // rounds of one, its answer is final
    const bool retry = own && R > 1 && sm.over[lane] != 0 && fmin(ob.lmin, olost) < ob.d;
    const unsigned long long rm = __ballot(retry);
    const uint16_t og = own ? sm.list[base + lane] : uint16_t(0);
    __syncthreads();
    if (retry) sm.list[again + int(__popcll(rm & __lanemask_lt()))] = og;
    again += int(__popcll(rm));
    if (own && !retry) {
      ob.lmin = fmin(ob.lmin, olost);
      double val;
      uint32_t flag, feat;
      classify64(ob, obound, dv, val, flag, feat);
      res.val[oi] = val;
      res.flag[oi] = flag;
      if (res.feat) res.feat[oi] = feat;
    }
  }
  cnt = again;
  if (cnt == 0) break;
  }
  };
  if constexpr (pts<S>()) {
    // scattered points far from the surface (their candidates in chunks) take 4 lanes a
    // point, 16 in a far block's slice (a few points), the others 1
    __syncthreads();
    nchunk = 0;
    block_candidates<S, NT>(dv, j, blk, sm.cand, sm.tmp(), sm.key(), sm.red, sm.wave_tot, sm.chunks, &nchunk, gb, &rr, nullptr, 0u, sm.list,
                            sm.count);
    if constexpr (FAR) {
      if (nchunk > 0)
        settle_rounds(std::integral_constant<int, 16>{});
      else
        settle_rounds(std::integral_constant<int, 1>{});
    } else {
      if (nchunk > 0)
        settle_rounds(std::integral_constant<int, 4>{});
      else
        settle_rounds(std::integral_constant<int, 1>{});
    }
  } else {
    settle_rounds(std::integral_constant<int, SettleLds<S>::K>{});
  }
}

// Blocks take the next queued block from a counter (scount[1], cleared with the queue's
// length): a queued block's settle costs from one read to thousands, so a fixed share
// each left a few blocks running long after the rest.
template <int S>
__global__ __launch_bounds__(SETTLE_NT) void k_settle(Dev dv, Job j, Res res) {
  static_assert(sizeof(SettleLds<S>) <= 40 * 1024 - 64, "four settle blocks a CU");
  __shared__ SettleLds<S> sm;
  __shared__ uint32_t wslot;
  const uint32_t total = *res.scount;
  for (;;) {
    if (threadIdx.x == 0) wslot = atomicAdd(res.scount + 1, 1u);
    __syncthreads();
    const uint32_t qb = wslot;
    if (qb >= total) break;
    settle_block<S>(dv, j, res, long(res.squeue[qb] & 0xffffffu), sm);
    __syncthreads();
  }
}
// The slices of far blocks (SM_FAR), in a kernel of their own: their rounds at 16 lanes a
// point, compiled into the kernel of the other blocks, slowed those (the trimmed plate's
// 2^24 queries' settle 143 -> 162 ms, with a millisecond of far points).
template <int S>
__global__ __launch_bounds__(SETTLE_NT) void k_settle_far(Dev dv, Job j, Res res) {
  static_assert(pts<S>(), "scattered points");
  __shared__ SettleLds<S> sm;
  __shared__ uint32_t wslot;
  const uint32_t total = res.scount[2];
  const size_t base = size_t(settle_blocks<S>(j));
  for (;;) {
    if (threadIdx.x == 0) wslot = atomicAdd(res.scount + 3, 1u);
    __syncthreads();
    const uint32_t qb = wslot;
    if (qb >= total) break;
    const uint32_t e = res.squeue[base + qb];
    settle_block<S, true>(dv, j, res, long(e & 0xffffffu), sm, int(e >> 24), SETTLE_SLICES);
    __syncthreads();
  }
}

// A band brick whose nodes all lie past the clamp takes its centre's sign. When the
// decision left that centre undecided and no sign came to it along the grid, the
// centre is marked (flag 4) for the settle; centre_sign then takes its settled sign.
__device__ __forceinline__ void need_centre(const Job& jc, const Res& rc, const uint32_t* list, long bi, const uint32_t* nflag, const double* dc,
                                   const int8_t* sgn) {
  int far = 1;
  for (int n = threadIdx.x; n < 729; n += blockDim.x) far &= nflag[729 * bi + n] == 1u;
  far = __syncthreads_and(far);
  if (threadIdx.x == 0 && far) {
    const long b = long(list[bi]);
    if (dc[b] != dc[b] && sgn[b] == 0) mark_point<SRC_CENTRES>(jc, rc, b);
  }
}
__device__ __forceinline__ void centre_sign(const Res& rc, const uint32_t* list, long bi, double* dc, int8_t* sgn) {
  const long b = long(list[bi]);
  if (dc[b] != dc[b] && sgn[b] == 0 && rc.flag[b] == 0) {
    dc[b] = rc.val[b];
    sgn[b] = rc.val[b] < 0.0 ? int8_t(-1) : int8_t(1);
  }
}
__global__ __launch_bounds__(256) void k_need_centre(Job jc, Res rc, const uint32_t* list, long nband, const uint32_t* nflag, const double* dc,
                                                      const int8_t* sgn) {
  for (long bi = blockIdx.x; bi < nband; bi += gridDim.x) {
    need_centre(jc, rc, list, bi, nflag, dc, sgn);
    __syncthreads();
  }
}
__global__ __launch_bounds__(256) void k_centre_sign(Res rc, const uint32_t* list, long nband, double* dc, int8_t* sgn) {
  for (long bi = long(blockIdx.x) * 256 + threadIdx.x; bi < nband; bi += long(gridDim.x) * 256) centre_sign(rc, list, bi, dc, sgn);
}

// ------------------------------------------------------------------ the level

// The decision from the centres' answers: dc the signed distance or NaN; band 1 when
// the brick is kept (|dc| <= reach, or undecided under the reach); sign +-1 when
// known, else 0.
__device__ __forceinline__ void decide_brick(const Dev& dv, const Lv& lv, long b, const double* val, const uint32_t* flag, double* dc, uint8_t* band, int8_t* sgn) {
  double reach = brick_reach(lv.h);
  double v = val[b];
  uint32_t f = flag[b];
  int8_t s = 0;
  if (f == 0) {
    dc[b] = v;
    band[b] = fabs(v) <= reach ? 1 : 0;
    s = v < 0 ? -1 : 1;
  } else {
    dc[b] = __builtin_nan("");
    // undecided under the reach: kept; its centre left to the host (flag 3): kept, and
    // the brick goes to the host whole (band 2)
    band[b] = ((f == 2 || f == 3) && v <= reach) ? (f == 3 ? 2 : 1) : 0;
    // a centre outside the part's box is outside the part
    v3<double> c = brick_first(lv, dv.origin, b) + mk3(4.0 * lv.h, 4.0 * lv.h, 4.0 * lv.h);
    const float* L0 = dv.L.lo;
    const float* L1 = dv.L.hi;
    if (c.x < L0[0] || c.y < L0[1] || c.z < L0[2] || c.x > L1[0] || c.y > L1[1] || c.z > L1[2]) s = 1;
  }
  sgn[b] = s;
}
__global__ void k_decide_finish(Dev dv, Lv lv, long nbricks, const double* val, const uint32_t* flag, double* dc, uint8_t* band, int8_t* sgn) {
  long b = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b < nbricks) decide_brick(dv, lv, b, val, flag, dc, band, sgn);
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
// The same sweeps, x then y then z, repeated until a round changes nothing (at most
// SIGN_ROUNDS rounds), by one block with the level's bricks in LDS (one byte each:
// the band bit 4, the sign + 1 in bits 0-1), for levels that fit SIGN_LDS bytes. Rows
// along x are padded to an odd number of dwords, and the lines of a sweep are dealt
// x fastest (or y, for the x sweep), so a wave's lanes read distinct LDS banks.
constexpr int SIGN_T = 1024, SIGN_LDS = 96 * 1024, SIGN_ROUNDS = 8;
__host__ __device__ inline long sign_pitch(int nb0) { return 4L * (2L * ((nb0 + 7) / 8) + 1); }
__device__ __forceinline__ void sign_block(const Lv& lv, const uint8_t* band, int8_t* sgn, uint8_t* g, int& changed) {
  const int nb[3] = {lv.nb[0], lv.nb[1], lv.nb[2]};
  const int P = int(sign_pitch(lv.nb[0]));
  const int rows = nb[1] * nb[2];  // row r = y + nb[1] z: nb[0] bricks along x
  for (int r = threadIdx.x; r < rows; r += SIGN_T)
    for (int x = 0; x < nb[0]; ++x) {
      const long i = long(r) * nb[0] + x;
      g[P * r + x] = uint8_t((band[i] ? 16 : 0) | (sgn[i] + 1));
    }
  const int stride[3] = {1, P, P * nb[1]};
  for (int round = 0; round < SIGN_ROUNDS; ++round) {
    if (threadIdx.x == 0) changed = 0;
    __syncthreads();
    int any = 0;
    for (int ax = 0; ax < 3; ++ax) {
      // the two other axes, the faster one first
      const int f = ax == 0 ? 1 : 0, o = ax == 2 ? 1 : 2;
      for (int line = threadIdx.x; line < nb[f] * nb[o]; line += SIGN_T) {
        const int base = (line % nb[f]) * stride[f] + (line / nb[f]) * stride[o];
        for (int pass = 0; pass < 2; ++pass) {
          int carry = 0;
          bool carry_ok = false;
          const int step = pass == 0 ? stride[ax] : -stride[ax];
          int id = pass == 0 ? base : base + (nb[ax] - 1) * stride[ax];
          for (int k = 0; k < nb[ax]; ++k, id += step) {
            const uint8_t v = g[id];
            int sg = int(v & 3) - 1;
            const bool nonband = !(v & 16);
            if (sg == 0 && carry != 0 && (carry_ok || nonband)) {
              sg = carry;
              g[id] = uint8_t((v & 16) | (sg + 1));
              any = 1;
            }
            carry = sg;
            carry_ok = nonband;
          }
        }
      }
      __syncthreads();
    }
    if (any) changed = 1;
    __syncthreads();
    const int c = changed;
    __syncthreads();
    if (!c) break;
  }
  for (int r = threadIdx.x; r < rows; r += SIGN_T)
    for (int x = 0; x < nb[0]; ++x) sgn[long(r) * nb[0] + x] = int8_t(int(g[P * r + x] & 3) - 1);
}
__global__ __launch_bounds__(SIGN_T) void k_sign_block(Lv lv, const uint8_t* band, int8_t* sgn) {
  __shared__ uint8_t g[SIGN_LDS];
  __shared__ int changed;
  sign_block(lv, band, sgn, g, changed);
}

// Positions of the bricks a predicate keeps, in order, by one block: each thread
// counts a run of consecutive entries (a multiple of 16, read 16 bytes at a time), the
// block scans the counts, and each thread writes its run's positions. Kept entries
// get out[i] (their rank) and, when list is not NULL, list[rank] = i; total gets the
// count. The predicate: band[i] != 0 (band), or cls[i] == 0 (cls: a band brick stored
// with words).
constexpr int SCAN_T = 1024;
__device__ __forceinline__ void scan_block(const uint8_t* band, const uint32_t* cls, long n, uint32_t* out, uint32_t* list, uint32_t* total, uint32_t* s) {
  const int tid = threadIdx.x;
  const long run = ((n + SCAN_T - 1) / SCAN_T + 15) / 16 * 16, i0 = tid * run, i1 = i0 + run < n ? i0 + run : n;
  // the run's keep bits, 16 entries a word
  auto bits16 = [&](long i) -> uint32_t {
    uint32_t m = 0;
    if (band && i + 16 <= n && (reinterpret_cast<uintptr_t>(band + i) & 15) == 0) {
      const uint4 v = *reinterpret_cast<const uint4*>(band + i);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
      for (int q = 0; q < 16; ++q) m |= ((w[q >> 2] >> (8 * (q & 3))) & 0xffu) ? (1u << q) : 0u;
      return m;
    }
    for (int q = 0; q < 16 && i + q < n; ++q) m |= (band ? band[i + q] != 0 : cls[i + q] == 0u) ? (1u << q) : 0u;
    return m;
  };
  uint32_t c = 0;
  for (long i = i0; i < i1; i += 16) c += __popc(bits16(i));
  s[tid] = c;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {
    uint32_t v = tid >= off ? s[tid - off] : 0u;
    __syncthreads();
    s[tid] += v;
    __syncthreads();
  }
  uint32_t pos = s[tid] - c;
  for (long i = i0; i < i1; i += 16) {
    uint32_t m = bits16(i);
    while (m) {
      const int q = __ffs(m) - 1;
      m &= m - 1;
      if (out) out[i + q] = pos;
      if (list) list[pos] = uint32_t(i + q);
      ++pos;
    }
  }
  if (tid == SCAN_T - 1) *total = s[SCAN_T - 1];
}
__global__ __launch_bounds__(SCAN_T) void k_scan(const uint8_t* band, const uint32_t* cls, long n, uint32_t* out, uint32_t* list, uint32_t* total) {
  __shared__ uint32_t s[SCAN_T];
  scan_block(band, cls, n, out, list, total, s);
}

// A band brick from its nodes' answers: far nodes signed by flood fill from the nodes
// known, then from the centre's empty ball, what is left bridged; quantized, classed
// and packed.
struct FinishLds {
  double val[729];
  uint8_t st[729];  // 0 exact, 1 far, 2 undecided, 3 filled, 4 bridged (5 during a pass)
  int changed;
  int16_t q[730];
  // the fill's 729-bit masks, node n at bit n % 64 of word n / 64: the nodes known
  // positive and negative, the far ones, two dilation stages of each sign, the nodes
  // filled of each sign, and the sides (x = 0, x = 8, y = 0, y = 8, inside the brick)
  uint64_t sgn[2][12], far[12], ta[2][12], tb[2][12], filled[2][12], side[5][12];
};

// One sign's 3 x 3 x 3 dilation of a 729-bit mask, a word a lane, in the three passes the
// caller separates by barriers: along x (+-1 bit, not across the x = 0 / 8 sides), along y
// (+-9 bits, not across y = 0 / 8), along z (+-81 = 64 + 17 bits), the last kept inside.
__device__ inline uint64_t mword(const uint64_t* x, int w) { return w >= 0 && w < 12 ? x[w] : 0ull; }
__device__ inline uint64_t dilate_x(const uint64_t* x, const uint64_t* lo, const uint64_t* hi, int w) {
  const uint64_t up = ((x[w] & ~hi[w]) << 1) | ((mword(x, w - 1) & ~mword(hi, w - 1)) >> 63);
  const uint64_t dn = ((x[w] & ~lo[w]) >> 1) | ((mword(x, w + 1) & ~mword(lo, w + 1)) << 63);
  return x[w] | up | dn;
}
__device__ inline uint64_t dilate_y(const uint64_t* x, const uint64_t* lo, const uint64_t* hi, int w) {
  const uint64_t up = ((x[w] & ~hi[w]) << 9) | ((mword(x, w - 1) & ~mword(hi, w - 1)) >> 55);
  const uint64_t dn = ((x[w] & ~lo[w]) >> 9) | ((mword(x, w + 1) & ~mword(lo, w + 1)) << 55);
  return x[w] | up | dn;
}
__device__ inline uint64_t dilate_z(const uint64_t* x, const uint64_t* inside, int w) {
  const uint64_t up = (mword(x, w - 1) << 17) | (mword(x, w - 2) >> 47);
  const uint64_t dn = (mword(x, w + 1) >> 17) | (mword(x, w + 2) << 47);
  return (x[w] | up | dn) & inside[w];
}
__device__ __forceinline__ void finish_brick(const Lv& lv, const uint32_t* list, long bi, const double* dcs, const int8_t* sgns, const uint8_t* bands,
                                    const double* val_in, const uint32_t* flag_in, uint32_t* staging, uint32_t* cls, unsigned long long* counters,
                                    FinishLds& sm) {
  double* val = sm.val;
  uint8_t* st = sm.st;
  int& changed = sm.changed;
  int16_t* q = sm.q;
  const int tid = threadIdx.x;
  const double h = lv.h, clamp = h * 1.7320508075688772, half = 4.0 * h;
  int host = 0;
  for (int n = tid; n < 729; n += WG) {
    val[n] = val_in[729 * bi + n];
    uint32_t f = flag_in[729 * bi + n];
    st[n] = uint8_t(f > 2 ? 2 : f);
    host |= f == 3u;
  }
  // gpu.rs:714-716, partcache.rs:794-797 a brick with a node (or its centre) left to
  // the host is read there whole: class 3, no words
  if (__syncthreads_or(host | (bands[list[bi]] == 2))) {
    if (tid == 0) cls[bi] = 3u, atomicAdd(&counters[7], 1ull);
    return;
  }
  // partcache.rs:825-855 far nodes take the sign of any neighbour known, by flood fill. A
  // far node is farther from the boundary than from any of its neighbours, so all its
  // known neighbours share its sign, and the fill is the dilation of the nodes known
  // positive (negative) within the far ones, repeated until nothing changes: on 729-bit
  // masks, a lane a word (a node with known neighbours of both signs, which consistent
  // answers never give, is left to the bridging).
  static_assert(WG == 256, "the fill's masks take a wave's ballot as a word");
  const int lane = tid & 63, wv = tid >> 6;
  for (int k = 0; k < 3; ++k) {
    const int n = tid + 256 * k;  // wave wv of pass k: nodes of word 4 k + wv
    const bool in = n < 729;
    const uint64_t b0 = __ballot(in && n % 9 == 0), b1 = __ballot(in && n % 9 == 8);
    const uint64_t b2 = __ballot(in && (n / 9) % 9 == 0), b3 = __ballot(in && (n / 9) % 9 == 8), b4 = __ballot(in);
    if (lane == 0) sm.side[0][4 * k + wv] = b0, sm.side[1][4 * k + wv] = b1, sm.side[2][4 * k + wv] = b2, sm.side[3][4 * k + wv] = b3, sm.side[4][4 * k + wv] = b4;
  }
  auto fill = [&]() {
    for (int k = 0; k < 3; ++k) {
      const int n = tid + 256 * k;
      const bool in = n < 729;
      const uint8_t sn = in ? st[n] : uint8_t(255);
      const double v = in ? val[n] : 0.0;
      const bool known = (sn == 0 || sn == 3) && v != 0.0;
      const uint64_t bp = __ballot(known && v > 0.0), bn = __ballot(known && v < 0.0), bf = __ballot(sn == 1);
      if (lane == 0) {
        const int w = 4 * k + wv;
        sm.sgn[0][w] = bp, sm.sgn[1][w] = bn, sm.far[w] = bf, sm.filled[0][w] = 0ull, sm.filled[1][w] = 0ull;
      }
    }
    __syncthreads();
    const int sg = tid / 12, w = tid % 12;
    for (;;) {
      if (tid < 24) sm.ta[sg][w] = dilate_x(sm.sgn[sg], sm.side[0], sm.side[1], w);
      __syncthreads();
      if (tid < 24) sm.tb[sg][w] = dilate_y(sm.ta[sg], sm.side[2], sm.side[3], w);
      __syncthreads();
      if (tid < 24) sm.ta[sg][w] = dilate_z(sm.tb[sg], sm.side[4], w) & sm.far[w];
      if (tid == 0) changed = 0;
      __syncthreads();
      if (tid < 12) {
        const uint64_t np = sm.ta[0][tid] & ~sm.ta[1][tid], nn = sm.ta[1][tid] & ~sm.ta[0][tid];
        sm.sgn[0][tid] |= np, sm.sgn[1][tid] |= nn, sm.far[tid] &= ~(np | nn);
        sm.filled[0][tid] |= np, sm.filled[1][tid] |= nn;
        if (np | nn) changed = 1;
      }
      __syncthreads();
      if (!changed) break;
    }
    for (int n = tid; n < 729; n += WG) {
      const uint64_t b = 1ull << (n & 63);
      if (sm.filled[0][n >> 6] & b) val[n] = clamp, st[n] = 3;
      if (sm.filled[1][n >> 6] & b) val[n] = -clamp, st[n] = 3;
    }
    __syncthreads();
  };
  fill();
  // partcache.rs:856-871 far nodes within the empty ball of the centre share its sign
  const long b = list[bi];
  const double dc = dcs[b];
  const int8_t csgn = sgns[b];
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
  // partcache.rs:872-909 what is left is bridged (bridge_nodes: in passes, the mean of
  // the neighbours known when the pass began, held within the Lipschitz brackets of the
  // nodes known exactly or by sign); a brick all past the clamp takes its centre's sign
  // first (the centre's own, or as carried along the grid when the decision left it
  // undecided)
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
      bridged += (unsigned long long)bridge_nodes(val, st, h, clamp);
      if (bridged) atomicAdd(&counters[0], bridged);
    }
  }
  __syncthreads();
  // partcache.rs:911-915, 962-970 quantize, class, two values to a word (low half first)
  int any_pos = 0, any_nonpos = 0;
  for (int n = tid; n < 729; n += WG) {
    double x = fmin(fmax(val[n] / clamp, -1.0), 1.0) * 32767.0;
    int16_t qq = int16_t(round(x));
    q[n] = qq;
    any_nonpos |= qq <= 0, any_pos |= qq > 0;
  }
  if (tid == 0) q[729] = 0;
  any_nonpos = __syncthreads_or(any_nonpos);
  any_pos = __syncthreads_or(any_pos);
  uint32_t c = !any_nonpos ? 1u : !any_pos ? 2u : 0u;  // 1 AIR, 2 SOLID, 0 words (3 host)
  if (tid == 0) cls[bi] = c;
  if (c == 0)
    for (int w = tid; w < ICAM_WORDS; w += WG) staging[size_t(bi) * ICAM_WORDS + w] = uint32_t(uint16_t(q[2 * w])) | (uint32_t(uint16_t(q[2 * w + 1])) << 16);
}
__global__ __launch_bounds__(WG) void k_bricks_finish(Lv lv, const uint32_t* list, long nbricks, const double* dcs, const int8_t* sgns,
                                                      const uint8_t* bands, const double* val_in, const uint32_t* flag_in, uint32_t* staging, uint32_t* cls,
                                                      unsigned long long* counters) {
  __shared__ FinishLds sm;
  if (long(blockIdx.x) < nbricks) finish_brick(lv, list, blockIdx.x, dcs, sgns, bands, val_in, flag_in, staging, cls, counters, sm);
}


// partcache.rs:128-146 assemble: the map (AIR, SOLID, or the slot) and the atlas in
// map order (band bricks are listed in map order, so their slots are too). A brick
// kept whole whose sign was never found is counted (counters[2]).
__device__ __forceinline__ void store_brick(long b, const uint8_t* band, const int8_t* sgn, uint32_t* map, unsigned long long* counters) {
  if (!band[b]) {
    map[b] = sgn[b] < 0 ? ICAM_SOLID : ICAM_AIR;
    if (sgn[b] == 0) atomicAdd(&counters[2], 1ull);
  }
}
__global__ void k_store(long nbricks, const uint8_t* band, const int8_t* sgn, uint32_t* map, unsigned long long* counters) {
  long b = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b < nbricks) store_brick(b, band, sgn, map, counters);
}
// Band brick bi's map entry and atlas words (its slot's), by a block of any size.
__device__ __forceinline__ void store_band_brick(long bi, const uint32_t* list, const uint32_t* cls, const uint32_t* slot, const uint32_t* staging, uint32_t* map,
                                        uint32_t* atlas) {
  uint32_t c = cls[bi];
  if (threadIdx.x == 0) map[list[bi]] = c == 1 ? ICAM_AIR : c == 2 ? ICAM_SOLID : c == 3 ? ICAM_HOST : slot[bi];
  if (c == 0)
    for (int w = threadIdx.x; w < ICAM_WORDS; w += blockDim.x) atlas[size_t(slot[bi]) * ICAM_WORDS + w] = staging[size_t(bi) * ICAM_WORDS + w];
}
__global__ void k_store_band(long nband, const uint32_t* list, const uint32_t* cls, const uint32_t* slot, const uint32_t* staging, uint32_t* map,
                             uint32_t* atlas) {
  if (long(blockIdx.x) < nband) store_band_brick(blockIdx.x, list, cls, slot, staging, map, atlas);
}

__global__ void k_spin(float* x, int iters) {
  float a = x[threadIdx.x];
  for (int i = 0; i < iters; ++i) a = a * 0.999f + 0.001f;
  x[threadIdx.x] = a;
}

// ------------------------------------------------------------------ batches
//
// Many parts' levels baked by one sequence of launches (captured once in a hipGraph):
// each stage is one launch over every part's work. The block stages (the centres' and
// the nodes' searches, the fallback and settle, the band bricks' finish and store) are
// persistent work queues: a fixed grid whose blocks take the next work item with an
// atomic counter until none is left, the item's part found from prefix offsets; the
// point stages stride over all parts' points; the per-part scans take a block each.
// Parts are ordered heaviest first by the host, so the queue serves the long bricks
// early (longest-processing-time order).

// A part of a batch: its tables, level, jobs and results, and its level's arrays.
struct BPart {
  Dev dv;
  Lv lv;
  Job jc, jn;  // the centres' and the band nodes' jobs
  Res rc, rn;
  long nb;     // bricks
  uint32_t cap;  // band bricks it has room for
  double* dc;
  uint8_t* band;
  int8_t* sgn;
  uint32_t* list;
  uint32_t* staging;
  uint32_t* cls;
  uint32_t* slot;
  uint32_t* map;
  uint32_t* atlas;
  unsigned long long* counters;
};

// The control words of a batch (zeroed before each run).
enum BCtl : uint32_t {
  BC_NEXT = 0,      // work counters of the persistent stages (16)
  BC_QC = 16,       // centres' fallback queue count
  BC_SC = 17,       // centres' settle queue count
  BC_QN = 18,       // nodes' fallback queue count
  BC_SN = 19,       // nodes' settle queue count
  BC_ITEMS = 20,    // deferred items
  BC_OVER = 21,     // a part whose band outgrew its room
  BC_WORDS = 32,
};
enum BStage : int { BS_CSEARCH = 0, BS_CFALL, BS_CSETTLE, BS_NSEARCH, BS_NFALL, BS_NSETTLE, BS_FINISH };

struct BatchArgs {
  const BPart* parts;
  int n;
  const long* brick_off;  // n + 1: bricks of the parts before p
  const long* tile_off;   // n + 1: tiles
  const long* super_off;  // n + 1: super-tiles
  uint32_t* band_off;     // n + 1: band bricks (written by kb_plan)
  uint32_t* btotal;       // per part: band bricks its scan found
  uint32_t* stotal;       // per part: slots
  uint32_t* ctl;
  const int* wide;        // the parts whose level is too large for kb_sign's workgroup memory
  int nwide;
};

// The part whose range [off[p], off[p + 1]) holds w.
template <class O> __device__ inline int part_of(const O* off, int n, O w) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= w)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}
__device__ inline const ICAM_K BPart& bpart(const BatchArgs& a, int p) { return ((const ICAM_K BPart*)a.parts)[p]; }
// A record copied out of the constant address space, a dword at a time (the loads are
// uniform, so they are scalar; fields the caller never reads are dropped).
template <class X> __device__ inline X ldk(const ICAM_K X& src) {
  static_assert(sizeof(X) % 4 == 0, "dword records");
  X out;
  const ICAM_K uint32_t* s = reinterpret_cast<const ICAM_K uint32_t*>(&src);
  uint32_t* d = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < int(sizeof(X) / 4); ++i) d[i] = s[i];
  return out;
}

// The next work item of stage st for the whole block (uniform), or `total` and past.
__device__ inline uint32_t next_item(const BatchArgs& a, int st, uint32_t* slot) {
  if (threadIdx.x == 0) *slot = atomicAdd(&a.ctl[BC_NEXT + st], 1u);
  __syncthreads();
  const uint32_t w = __builtin_amdgcn_readfirstlane(*slot);
  __syncthreads();
  return w;
}

// Up to six buffers cleared by one launch (a fill kernel, not hipMemsetAsync: the
// batch's launches are captured in a graph).
struct FillList {
  uint32_t* p[6];
  size_t n[6];
};
__global__ void kb_fill(FillList f) {
  for (int k = 0; k < 6; ++k)
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < f.n[k]; i += size_t(gridDim.x) * blockDim.x) f.p[k][i] = 0u;
}

// The searches, persistent: centres over every part's tiles, nodes over every part's
// band bricks.
// The single-precision node search at three waves a SIMD (168 registers, a few spilled):
// 0.41 -> 0.37 ms on the NURBS-top block's 6,889 band bricks; the centres' search and
// four waves a SIMD lose. At several lanes a node the registers are the compiler's choice
// (at three waves a SIMD the large mixed solid's search at 4 lanes a node spilled 20 GB a
// launch against the single-level kernel's 2.7 GB).
template <int S, class T, int KN = Shape<S>::K>
__global__ __launch_bounds__(Shape<S>::NT) __attribute__((amdgpu_waves_per_eu((S == SRC_NODES && sizeof(T) == 4 && KN == 1) ? 3 : 1, 8))) void kb_search(BatchArgs a) {
  __shared__ SearchLds<S, T> sm;
  __shared__ uint32_t wslot;
  const uint32_t total = S == SRC_CENTRES ? uint32_t(a.tile_off[a.n]) : a.band_off[a.n];
  for (;;) {
    const uint32_t w = next_item(a, S == SRC_CENTRES ? BS_CSEARCH : BS_NSEARCH, &wslot);
    if (w >= total) break;
    int p;
    long blk;
    // the part and the block are the same on every lane: said so, so that the part's
    // records and tables are read by scalar loads
    if (S == SRC_CENTRES) {
      p = __builtin_amdgcn_readfirstlane(part_of<long>(a.tile_off, a.n, long(w)));
      blk = long(w) - long(__builtin_amdgcn_readfirstlane(uint32_t(a.tile_off[p])));
    } else {
      p = __builtin_amdgcn_readfirstlane(part_of<uint32_t>(a.band_off, a.n, w));
      blk = long(w - __builtin_amdgcn_readfirstlane(a.band_off[p]));
    }
    const ICAM_K BPart& P = bpart(a, p);
    const Dev dv = ldk(P.dv);
    const Job j = S == SRC_CENTRES ? ldk(P.jc) : ldk(P.jn);
    const Res r = S == SRC_CENTRES ? ldk(P.rc) : ldk(P.rn);
    search_block<S, T, false, KN>(dv, j, r, blk, sm);
    __syncthreads();
  }
}

// Every part's super-tile lists, a block each.
__global__ __launch_bounds__(256) void kb_super(BatchArgs a) {
  __shared__ uint32_t cand[SCAP];
  __shared__ uint32_t lists[2 * FCAP];
  __shared__ float red[256];
  __shared__ int wave_tot[4];
  const long total = a.super_off[a.n];
  for (long w = blockIdx.x; w < total; w += gridDim.x) {
    const int p = __builtin_amdgcn_readfirstlane(part_of<long>(a.super_off, a.n, w));
    const ICAM_K BPart& P = bpart(a, p);
    const Dev dv = ldk(P.dv);
    const Job j = ldk(P.jc);
    super_block(dv, j, w - a.super_off[p], cand, wave_tot, lists, red);
    __syncthreads();
  }
}

// The points of every part: centres (static counts) or band nodes (from the plan).
template <int S> __device__ inline long batch_points(const BatchArgs& a) { return S == SRC_CENTRES ? a.brick_off[a.n] : long(a.band_off[a.n]) * 729; }
template <int S> __device__ inline int point_part(const BatchArgs& a, long g, long& i) {
  if (S == SRC_CENTRES) {
    int p = part_of<long>(a.brick_off, a.n, g);
    i = g - a.brick_off[p];
    return p;
  }
  const uint32_t bi = uint32_t(g / 729);
  int p = part_of<uint32_t>(a.band_off, a.n, bi);
  i = g - long(a.band_off[p]) * 729;
  return p;
}

template <int S>
__global__ __launch_bounds__(256) void kb_refine(BatchArgs a) {
  const long total = batch_points<S>(a);
  for (long g = long(blockIdx.x) * 256 + threadIdx.x; g < total; g += long(gridDim.x) * 256) {
    long i;
    const int p = point_part<S>(a, g, i);
    const ICAM_K BPart& P = bpart(a, p);
    const Dev dv = ldk(P.dv);
    const Job j = S == SRC_CENTRES ? ldk(P.jc) : ldk(P.jn);
    const Res r = S == SRC_CENTRES ? ldk(P.rc) : ldk(P.rn);
    refine_point<S>(dv, j, r, i);
  }
}
// The band bricks' nodes of every part, a block a brick (the part is the same on every
// lane, so its records are read by scalar loads).
__global__ __launch_bounds__(256) void kb_refine_bricks(BatchArgs a) {
  __shared__ RefineLds sm;
  const uint32_t total = a.band_off[a.n];
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    const int p = __builtin_amdgcn_readfirstlane(part_of<uint32_t>(a.band_off, a.n, w));
    const ICAM_K BPart& P = bpart(a, p);
    const Dev dv = ldk(P.dv);
    const Job j = ldk(P.jn);
    const Res r = ldk(P.rn);
    refine_brick<256>(dv, j, r, long(w - __builtin_amdgcn_readfirstlane(a.band_off[p])), sm);
    __syncthreads();
  }
}

// The band bricks' nodes left undecided, marked for the settle a block a brick.
__global__ __launch_bounds__(256) void kb_mark_bricks(BatchArgs a) {
  const uint32_t total = a.band_off[a.n];
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    const int p = __builtin_amdgcn_readfirstlane(part_of<uint32_t>(a.band_off, a.n, w));
    const ICAM_K BPart& P = bpart(a, p);
    const Job j = ldk(P.jn);
    const Res r = ldk(P.rn);
    const long blk = long(w - __builtin_amdgcn_readfirstlane(a.band_off[p]));
    for (int n = threadIdx.x; n < 729; n += 256) mark_point<SRC_NODES>(j, r, blk * 729 + n);
  }
}

// The fallback (want 3) or the settle (want 4) over the batch's shared queue, whose
// entries are part << 24 | block.
template <int S, bool SETTLE, int RK = 4>
__global__ __launch_bounds__(Shape<S>::NT) void kb_redo(BatchArgs a) {
  static_assert(!SETTLE, "the batch's settle is kb_settle");
  __shared__ SearchLds<S, double> sm;
  __shared__ uint32_t wslot;
  const bool settle = SETTLE;
  const int st = S == SRC_CENTRES ? (settle ? BS_CSETTLE : BS_CFALL) : (settle ? BS_NSETTLE : BS_NFALL);
  const uint32_t total = a.ctl[S == SRC_CENTRES ? (settle ? BC_SC : BC_QC) : (settle ? BC_SN : BC_QN)];
  for (;;) {
    const uint32_t w = next_item(a, st, &wslot);
    if (w >= total) break;
    const ICAM_K BPart& P0 = bpart(a, 0);
    const Res r0 = S == SRC_CENTRES ? ldk(P0.rc) : ldk(P0.rn);
    const uint32_t e = __builtin_amdgcn_readfirstlane((settle ? r0.squeue : r0.queue)[w]);
    const ICAM_K BPart& P = bpart(a, __builtin_amdgcn_readfirstlane(int(e >> 24)));
    const Dev dv = ldk(P.dv);
    const Job j = S == SRC_CENTRES ? ldk(P.jc) : ldk(P.jn);
    const Res r = S == SRC_CENTRES ? ldk(P.rc) : ldk(P.rn);
    redo_block<S, SETTLE, RK>(dv, j, r, long(e & 0xffffffu), sm);
    __syncthreads();
  }
}

// Every part's settle queue, one wave a queued block (settle_block).
template <int S>
__global__ __launch_bounds__(SETTLE_NT) void kb_settle(BatchArgs a) {
  __shared__ SettleLds<S> sm;
  __shared__ uint32_t wslot;
  const int st = S == SRC_CENTRES ? BS_CSETTLE : BS_NSETTLE;
  const uint32_t total = a.ctl[S == SRC_CENTRES ? BC_SC : BC_SN];
  for (;;) {
    const uint32_t w = next_item(a, st, &wslot);
    if (w >= total) break;
    const ICAM_K BPart& P0 = bpart(a, 0);
    const Res r0 = S == SRC_CENTRES ? ldk(P0.rc) : ldk(P0.rn);
    const uint32_t e = __builtin_amdgcn_readfirstlane(r0.squeue[w]);
    const ICAM_K BPart& P = bpart(a, __builtin_amdgcn_readfirstlane(int(e >> 24)));
    const Dev dv = ldk(P.dv);
    const Job j = S == SRC_CENTRES ? ldk(P.jc) : ldk(P.jn);
    const Res r = S == SRC_CENTRES ? ldk(P.rc) : ldk(P.rn);
    settle_block<S>(dv, j, r, long(e & 0xffffffu), sm);
    __syncthreads();
  }
}

// The dense pass over every part's deferred points (the part in an item's word 1).
template <class T>
__global__ __launch_bounds__(256) void kb_leaves(BatchArgs a) {
  const ICAM_K BPart& P0 = bpart(a, 0);
  const uint32_t n = min(a.ctl[BC_ITEMS], P0.rn.icap);
  const uint32_t* items = P0.rn.items;
  for (uint32_t it = blockIdx.x * 256 + threadIdx.x; it < n; it += gridDim.x * 256) {
    const uint32_t* w = items + size_t(it) * Item<T>::W;
    const ICAM_K BPart& P = bpart(a, int(w[1] >> 8));
    const Dev dv = ldk(P.dv);
    const Job j = ldk(P.jn);
    const Res r = ldk(P.rn);
    leaves_item<SRC_NODES, T>(dv, j, r, w);
  }
}

__global__ void kb_decide(BatchArgs a) {
  const long total = a.brick_off[a.n];
  for (long g = long(blockIdx.x) * blockDim.x + threadIdx.x; g < total; g += long(gridDim.x) * blockDim.x) {
    const int p = part_of<long>(a.brick_off, a.n, g);
    const ICAM_K BPart& P = bpart(a, p);
    const Dev dv = ldk(P.dv);
    const Lv lv = ldk(P.lv);
    decide_brick(dv, lv, g - a.brick_off[p], P.rc.val, P.rc.flag, P.dc, P.band, P.sgn);
  }
}

// One block a part: the signs of its bricks kept whole, then its band list.
__global__ __launch_bounds__(SIGN_T) void kb_sign(BatchArgs a) {
  __shared__ uint8_t g[SIGN_LDS];
  __shared__ int changed;
  const ICAM_K BPart& P = bpart(a, blockIdx.x);
  const Lv lv = ldk(P.lv);
  if (sign_pitch(lv.nb[0]) * lv.nb[1] * lv.nb[2] > SIGN_LDS) return;  // kb_sign_wide's
  sign_block(lv, P.band, P.sgn, g, changed);
}
// The levels too large for kb_sign's workgroup memory, grid-wide: one round's sweeps along
// axis ax of every such part's lines (blockIdx.y the part among them), as k_sign. The graph
// runs SIGN_ROUNDS rounds with no check in between; a brick still without a sign after them
// is counted by kb_store and fails the part's results.
__global__ __launch_bounds__(256) void kb_sign_wide(BatchArgs a, int ax) {
  const ICAM_K BPart& P = bpart(a, a.wide[blockIdx.y]);
  const Lv lv = ldk(P.lv);
  const long nb[3] = {lv.nb[0], lv.nb[1], lv.nb[2]};
  const long a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
  const long line = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (line >= nb[a1] * nb[a2]) return;
  const long stride[3] = {1, nb[0], nb[0] * nb[1]};
  const long base = (line % nb[a1]) * stride[a1] + (line / nb[a1]) * stride[a2];
  const uint8_t* band = P.band;
  int8_t* sgn = P.sgn;
  for (int pass = 0; pass < 2; ++pass) {
    int8_t carry = 0;
    bool carry_ok = false;  // the previous brick may hand its sign on
    for (long k = 0; k < nb[ax]; ++k) {
      const long id = base + (pass == 0 ? k : nb[ax] - 1 - k) * stride[ax];
      int8_t sg = sgn[id];
      const bool nonband = band[id] == 0;
      if (sg == 0 && carry != 0 && (carry_ok || nonband)) sgn[id] = sg = carry;
      carry = sg;
      carry_ok = nonband;
    }
  }
}
__global__ __launch_bounds__(SCAN_T) void kb_band_list(BatchArgs a) {
  __shared__ uint32_t s[SCAN_T];
  const ICAM_K BPart& P = bpart(a, blockIdx.x);
  scan_block(P.band, nullptr, P.nb, nullptr, P.list, &a.btotal[blockIdx.x], s);
}
// The band bricks' offsets (each part's count held to its room; a part past it is
// flagged).
__global__ void kb_plan(BatchArgs a) {
  if (threadIdx.x != 0) return;
  uint32_t acc = 0;
  for (int p = 0; p < a.n; ++p) {
    a.band_off[p] = acc;
    const uint32_t cap = bpart(a, p).cap, t = a.btotal[p];
    if (t > cap) a.ctl[BC_OVER] = uint32_t(p) + 1u;
    acc += t < cap ? t : cap;
  }
  a.band_off[a.n] = acc;
}

__global__ __launch_bounds__(256) void kb_need_centre(BatchArgs a) {
  const uint32_t total = a.band_off[a.n];
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    const int p = __builtin_amdgcn_readfirstlane(part_of<uint32_t>(a.band_off, a.n, w));
    const ICAM_K BPart& P = bpart(a, p);
    const Job jc = ldk(P.jc);
    const Res rc = ldk(P.rc);
    need_centre(jc, rc, P.list, long(w - __builtin_amdgcn_readfirstlane(a.band_off[p])), P.rn.flag, P.dc, P.sgn);
    __syncthreads();
  }
}
__global__ __launch_bounds__(256) void kb_centre_sign(BatchArgs a) {
  const uint32_t total = a.band_off[a.n];
  for (uint32_t w = blockIdx.x * 256 + threadIdx.x; w < total; w += gridDim.x * 256) {
    const int p = part_of<uint32_t>(a.band_off, a.n, w);
    const ICAM_K BPart& P = bpart(a, p);
    const Res rc = ldk(P.rc);
    centre_sign(rc, P.list, long(w - a.band_off[p]), P.dc, P.sgn);
  }
}

__global__ __launch_bounds__(WG) void kb_finish(BatchArgs a) {
  __shared__ FinishLds sm;
  const uint32_t total = a.band_off[a.n];
  // bricks in strides (their costs are alike; a queue item each cost an atomic a brick)
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    const int p = __builtin_amdgcn_readfirstlane(part_of<uint32_t>(a.band_off, a.n, w));
    const ICAM_K BPart& P = bpart(a, p);
    const Lv lv = ldk(P.lv);
    finish_brick(lv, P.list, long(w - __builtin_amdgcn_readfirstlane(a.band_off[p])), P.dc, P.sgn, P.band, P.rn.val, P.rn.flag, P.staging, P.cls,
                 P.counters, sm);
    __syncthreads();
  }
}
__global__ __launch_bounds__(SCAN_T) void kb_slots(BatchArgs a) {
  __shared__ uint32_t s[SCAN_T];
  const ICAM_K BPart& P = bpart(a, blockIdx.x);
  const uint32_t nb = min(a.btotal[blockIdx.x], P.cap);
  scan_block(nullptr, P.cls, long(nb), P.slot, nullptr, &a.stotal[blockIdx.x], s);
}
__global__ void kb_store(BatchArgs a) {
  const long total = a.brick_off[a.n];
  for (long g = long(blockIdx.x) * blockDim.x + threadIdx.x; g < total; g += long(gridDim.x) * blockDim.x) {
    const int p = part_of<long>(a.brick_off, a.n, g);
    const ICAM_K BPart& P = bpart(a, p);
    store_brick(g - a.brick_off[p], P.band, P.sgn, P.map, P.counters);
  }
}
// Every band brick costs the same here, so the blocks stride over them (a queue's
// round trip per brick would cost more than its copy).
__global__ __launch_bounds__(128) void kb_store_band(BatchArgs a) {
  const uint32_t total = a.band_off[a.n];
  for (uint32_t w = blockIdx.x; w < total; w += gridDim.x) {
    const int p = __builtin_amdgcn_readfirstlane(part_of<uint32_t>(a.band_off, a.n, w));
    const ICAM_K BPart& P = bpart(a, p);
    store_band_brick(long(w - __builtin_amdgcn_readfirstlane(a.band_off[p])), P.list, P.cls, P.slot, P.staging, P.map, P.atlas);
  }
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

// The scratch of icam_hip_query_device_noalloc: one static array a device (the
// runtime gives every device that loads the library its own copy, held until the process
// ends). Host code uses the device address hipGetSymbolAddress gives for it
// (noalloc_scratch), never the array's name.
constexpr size_t NOALLOC_POINTS = size_t(1) << 18;
constexpr size_t NOALLOC_BYTES = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
std::mutex g_noalloc_mu;
constexpr int MAX_DEVICES = 64;
void* g_noalloc[MAX_DEVICES] = {};  // per device: the scratch array's device address

// Looks the array's address up once per device, with that device current.
int noalloc_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return fail("device number out of range");
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (g_noalloc[device]) return 0;
  void* ptr = nullptr;
  HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fail("no device address for the scratch array");
  g_noalloc[device] = ptr;
  return 0;
}
void* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  return device >= 0 && device < MAX_DEVICES ? g_noalloc[device] : nullptr;
}
}  // namespace

struct icam_hip {
  std::mutex mu;
  int device = 0;
  int settle_grid = 256;
  // lanes a node in a brick's nodes' search, double and mixed precision (1, 2, 4 or 8:
  // icam_hip_tune, ICAM_NODE_LANES). Untuned: 4 for parts of more than 65,536 sub-patches
  // and curve parts, whose nodes each reach many (the large mixed solid's level: 1, 2, 4
  // and 8 took 70, 53, 38 and 48 ms single-precision, 172, 128, 104 and 131 double), else
  // 1 (the 64 parts' nodes, few parts each, lose at 4: 67 -> 92 ms baked one by one)
  int kn[2] = {1, 1};
  // the brick centres' decision in double precision in the mixed form too (icam_hip_tune):
  // the large mixed solid's single-precision centres' search and its fallback took 4.0 +
  // 5.1 ms against 7.7 in double, the NURBS-top block's 0.14 against 0.12 ms
  bool c64 = true;
  hipStream_t stream = nullptr;
  Packed pk;
  Dev dv{};
  Buf F32, F64, U, pbox, pcode, bvh_box, bvh_prim, bvh_holds, bvh_lv;
  // per-point results of the centres' and the nodes' searches, and host-call staging
  struct ResBufs {
    Buf code, pa, pb, d32, info, val, flag, queue, qcount, qmark, items, icount, squeue, scount, smark, cqueue, ccount;
    void free() {
      for (Buf* b : {&code, &pa, &pb, &d32, &info, &val, &flag, &queue, &qcount, &qmark, &items, &icount, &squeue, &scount, &smark, &cqueue, &ccount})
        b->free();
    }
  } centres, nodes;
  Buf q_xyz, q_bound, q_d, q_flag, q_feat;
  // scattered queries taken in Morton order: the part's box (world), the sort's scratch,
  // the keys before and after the sort, the indices before and after, and the points
  // and results in that order
  double box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};
  Buf o_count, o_cell, o_rank, o_idx0, o_idx, o_xyz, o_d, o_flag, o_feat;
  // the level
  Buf dc, band, sgn, total, list, staging, cls, slot, stotal, map, atlas, counters, changed, firsts, tile_cand, tile_count, super_cand, super_count;
  std::vector<std::pair<const char*, double>> stages;
  std::string name;
  // the end of the last call's work: calls share the scratch above, and a call that
  // returns before its work is done (query_device, bake_device) may have run it on
  // another stream, so each call's work waits for this event first
  hipEvent_t done = nullptr;
  bool pending = false;
};

namespace {

struct Timer {
  hipEvent_t a, b;
  explicit Timer(hipStream_t s) {
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

// Per-stage HIP events, on the stream the kernels run on.
struct Stages {
  std::vector<Timer> tm;
  std::vector<const char*> names;
};

template <class K, class... A>
hipError_t launch(Stages* st, const char* name, hipStream_t s, dim3 g, dim3 b, K k, A... args) {
  if (st) st->tm.emplace_back(s);
  hipLaunchKernelGGL(k, g, b, 0, s, args...);
  if (st) st->tm.back().stop(s), st->names.push_back(name);
  return hipGetLastError();
}

// A call's work after the previous call's (whatever stream that ran on), and its end
// recorded for the next.
int after_previous(icam_hip* h, hipStream_t s) {
  if (h->pending) HIP_OK(hipStreamWaitEvent(s, h->done, 0));
  return 0;
}
int mark_end(icam_hip* h, hipStream_t s) {
  HIP_OK(hipEventRecord(h->done, s));
  h->pending = true;
  return 0;
}

void collect(icam_hip* h, Stages& st) {
  h->stages.clear();
  for (size_t i = 0; i < st.tm.size(); ++i) {
    double m = st.tm[i].ms();
    bool found = false;
    for (auto& x : h->stages)
      if (std::strcmp(x.first, st.names[i]) == 0) x.second += m, found = true;
    if (!found) h->stages.push_back({st.names[i], m});
  }
}

int upload(Buf& b, const void* src, size_t bytes) {
  HIP_OK(b.need(std::max<size_t>(bytes, 16)));
  if (bytes) HIP_OK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
  return 0;
}

// Result buffers for n points from one of the handle's two sets (centres or nodes);
// val, flag and feat may be the caller's.
int results(icam_hip* h, bool nodes, long n, Res& r, double* val, uint32_t* flag, uint32_t* feat, bool mixed, bool items = false) {
  size_t m = size_t(std::max<long>(n, 1));
  icam_hip::ResBufs& B = nodes ? h->nodes : h->centres;
  Buf &code = B.code, &pa = B.pa, &pb = B.pb, &d32 = B.d32, &info = B.info, &bval = B.val, &bflag = B.flag, &queue = B.queue, &qcount = B.qcount;
  if (mixed) {
    HIP_OK(code.need(m * 4));
    HIP_OK(pa.need(m * 4));
    HIP_OK(pb.need(m * 4));
    HIP_OK(d32.need(m * 4));
    HIP_OK(info.need(m * 4));
    HIP_OK(queue.need(m * 4));
    HIP_OK(B.qmark.need(m * 4));
  }
  HIP_OK(qcount.need(16));
  HIP_OK(B.squeue.need((m + SETTLE_SLICES) * 4));
  HIP_OK(B.smark.need(m * 4));
  HIP_OK(B.scount.need(16));
  HIP_OK(h->counters.need(64));
  if (!val) HIP_OK(bval.need(m * 8));
  if (!flag) HIP_OK(bflag.need(m * 4));
  r.code = code.as<uint32_t>(), r.pa = pa.as<float>(), r.pb = pb.as<float>(), r.d32 = d32.as<float>();
  r.info = info.as<uint32_t>(), r.queue = queue.as<uint32_t>(), r.qcount = qcount.as<uint32_t>(), r.qmark = B.qmark.as<uint32_t>();
  r.val = val ? val : bval.as<double>();
  r.flag = flag ? flag : bflag.as<uint32_t>();
  r.feat = feat;
  r.counters = h->counters.as<unsigned long long>();
  r.squeue = B.squeue.as<uint32_t>(), r.scount = B.scount.as<uint32_t>(), r.smark = B.smark.as<uint32_t>();
  r.cqueue = nullptr, r.ccount = nullptr;
  if (!nodes) {
    // (points and brick centres share these buffers; only points are searched in chunks)
    HIP_OK(B.cqueue.need(m * 4));
    HIP_OK(B.ccount.need(16));
    r.cqueue = B.cqueue.as<uint32_t>(), r.ccount = B.ccount.as<uint32_t>();
  }
  r.items = nullptr, r.icount = nullptr, r.icap = 0;
  if (items) {
    // room for an eighth of the points; points past it run their rounds in place
    const size_t cap = m / 8 + 1024;
    HIP_OK(B.items.need(cap * Item<double>::W * 4));
    HIP_OK(B.icount.need(16));
    r.items = B.items.as<uint32_t>(), r.icount = B.icount.as<uint32_t>(), r.icap = uint32_t(cap);
  }
  return 0;
}

// Runs the search pipeline for a job of source S: single precision, the last
// correction and the fallback (mixed), or double precision (fp64).
template <int S>
int run_search(icam_hip* h, bool mixed, const Job& j, Res& r, hipStream_t s, Stages* st, long blocks) {
  if (j.n <= 0) return 0;
  constexpr int NT = Shape<S>::NT;
  // (the far points' kernel runs one block a CU at a time; a larger grid only reserves
  // more scratch for its spills)
  const unsigned gfar = unsigned(std::min<long>(blocks * 16, 1024));
  static const char* names[4][8] = {{"points_search32", "points_refine64", "points_fallback64", "points_search64", "points_leaves32", "points_leaves64",
                                     "points_mark", "points_settle64"},
                                    {"nodes_search32", "nodes_refine64", "nodes_fallback64", "nodes_search64", "nodes_leaves32", "nodes_leaves64",
                                     "nodes_mark", "nodes_settle64"},
                                    {"centres_search32", "centres_refine64", "centres_fallback64", "centres_search64", "centres_leaves32",
                                     "centres_leaves64", "centres_mark", "centres_settle64"},
                                    {"points_search32", "points_refine64", "points_fallback64", "points_search64", "points_leaves32", "points_leaves64",
                                     "points_mark", "points_settle64"}};
  const unsigned lb = r.items ? unsigned(std::min<long>((long(r.icap) + 255) / 256, 2048)) : 0;
  {
    // the queues' counts and marks, cleared by one fill launch (no hipMemsetAsync: a
    // caller may capture these launches in a graph)
    FillList f{{r.items ? r.icount : nullptr, mixed ? r.qcount : nullptr, mixed ? r.qmark : nullptr, r.scount, r.smark, r.ccount},
               {r.items ? 1u : 0u, mixed ? 1u : 0u, mixed ? size_t(blocks) : 0u, 4u, size_t(blocks), r.ccount ? 2u : 0u}};
    HIP_OK(launch(st, "zero", s, dim3(unsigned(std::min<long>((blocks + 255) / 256, 4096))), dim3(256), kb_fill, f));
  }
  // a brick's nodes searched kn lanes a node (1, 2, 4 or 8)
  auto search_kernel = [&](auto t) {
    using T = decltype(t);
    if constexpr (S == SRC_NODES) {
      const int kn = h->kn[mixed ? 1 : 0];
      return kn == 8 ? k_search<S, T, 8> : kn == 4 ? k_search<S, T, 4> : kn == 2 ? k_search<S, T, 2> : k_search<S, T, 1>;
    } else {
      return k_search<S, T>;
    }
  };
  if (mixed) {
    HIP_OK(launch(st, names[S][0], s, dim3(unsigned(blocks)), dim3(NT), search_kernel(0.f), h->dv, j, r));
    if constexpr (pts<S>())
      if (r.cqueue) HIP_OK(launch(st, "points_far64", s, dim3(gfar), dim3(256), k_far<S>, h->dv, j, r));
    if (lb) HIP_OK(launch(st, names[S][4], s, dim3(lb), dim3(256), k_leaves<S, float>, h->dv, j, r));
    if constexpr (S == SRC_NODES)
      HIP_OK(launch(st, names[S][1], s, dim3(unsigned(std::min<long>(blocks, 8192))), dim3(256), k_refine_bricks, h->dv, j, r, blocks));
    else
      HIP_OK(launch(st, names[S][1], s, dim3(unsigned((j.n + 255) / 256)), dim3(256), k_refine<S>, h->dv, j, r));
    unsigned fb = unsigned(std::min<long>(blocks, 2048));
    // (a brick's unsure nodes 8 lanes a node where its search runs 4 or more)
    HIP_OK(launch(st, names[S][2], s, dim3(fb), dim3(NT), S == SRC_NODES && h->kn[1] >= 4 ? k_fallback<S, 8> : k_fallback<S, 4>, h->dv, j, r));
  } else {
    HIP_OK(launch(st, names[S][3], s, dim3(unsigned(blocks)), dim3(NT), search_kernel(0.0), h->dv, j, r));
    if constexpr (pts<S>())
      if (r.cqueue) HIP_OK(launch(st, "points_far64", s, dim3(gfar), dim3(256), k_far<S>, h->dv, j, r));
    if (lb) HIP_OK(launch(st, names[S][5], s, dim3(lb), dim3(256), k_leaves<S, double>, h->dv, j, r));
  }
  // what is still undecided: settled by the certified search (not the brick centres:
  // an undecided centre under its reach keeps its brick, whose nodes are read)
  if (S == SRC_CENTRES) return 0;
  HIP_OK(launch(st, names[S][6], s, dim3(unsigned((j.n + 255) / 256)), dim3(256), k_mark_settle<S>, j, r));
  HIP_OK(launch(st, names[S][7], s, dim3(unsigned(std::min<long>(blocks, h->settle_grid))), dim3(SETTLE_NT), k_settle<S>, h->dv, j, r));
  if constexpr (pts<S>())
    if (r.cqueue) HIP_OK(launch(st, "points_settle_far64", s, dim3(unsigned(h->settle_grid)), dim3(SETTLE_NT), k_settle_far<S>, h->dv, j, r));
  return 0;
}

// ------------------------------------------------------------------ point order
//
// Scattered queries are searched in the Morton order of their positions in the part's box
// (10 bits an axis, points outside it clamped to its border): a block's points then lie
// together and cull to a short candidate list, where points in the caller's order may
// each lie anywhere. The keys are sorted with their indices by rocPRIM's radix sort (a
// plain library sort), the points gathered in that order, and the results put back in
// the caller's order.
constexpr int ORDER_BITS = 10;
constexpr size_t ORDER_MIN = 8192;  // fewer points are searched as given

__global__ void k_order_keys(const double* xyz, long n, double lx, double ly, double lz, double sx, double sy, double sz, uint32_t* key, uint32_t* idx) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const double lo[3] = {lx, ly, lz}, sc[3] = {sx, sy, sz};
  uint32_t code = 0;
  for (int k = 0; k < 3; ++k) {
    const double f = (xyz[3 * i + k] - lo[k]) * sc[k];
    const uint32_t q = f >= 0.0 ? (f < double((1 << ORDER_BITS) - 1) ? uint32_t(f) : uint32_t((1 << ORDER_BITS) - 1)) : 0u;
    for (int b = 0; b < ORDER_BITS; ++b) code |= ((q >> b) & 1u) << (3 * b + k);
  }
  key[i] = code;
  idx[i] = uint32_t(i);
}

__global__ void k_order_place(const double* xyz, long n, const uint32_t* idx, double* oxyz) {
  const long k = long(blockIdx.x) * 256 + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = idx[k];
  for (int c = 0; c < 3; ++c) oxyz[3 * size_t(k) + c] = xyz[3 * size_t(i) + c];
}

// The radix sort's scratch for n keys with their indices (bytes).
size_t order_sort_bytes(size_t n) {
  size_t bytes = 0;
  rocprim::radix_sort_pairs(nullptr, bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr, n, 0, 3 * ORDER_BITS);
  return bytes;
}

__global__ void k_order_back(long n, const uint32_t* idx, const double* od, const uint32_t* of, const uint32_t* oft, double* d, uint32_t* f, uint32_t* ft) {
  const long k = long(blockIdx.x) * 256 + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = idx[k];
  if (d) d[i] = od[k];
  if (f) f[i] = of[k];
  if (ft) ft[i] = oft[k];
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
  h->name = std::string(prop.name[0] ? prop.name : "AMD GPU") + " (" + prop.gcnArchName + ", " + std::to_string(prop.multiProcessorCount) + " CUs)";
  h->settle_grid = settle_grid(k_settle<SRC_WIDE>, SETTLE_NT, prop.multiProcessorCount);
  if (noalloc_setup(h->device)) {
    std::string e = g_error;
    delete h;
    return fail(e);
  }
  if (!pack_part(part, need, h->pk)) {
    std::string e = h->pk.error;
    delete h;
    return fail(e);
  }
  std::vector<float> f32 = to_f32(h->pk.F);
  uint32_t lvtab[16];
  for (int l = 0; l < 8; ++l) lvtab[l] = h->pk.L.bvh_off[l], lvtab[8 + l] = h->pk.L.bvh_cnt[l];
  if (upload(h->bvh_lv, lvtab, sizeof(lvtab)) || upload(h->F32, f32.data(), f32.size() * 4) || upload(h->F64, h->pk.F.data(), h->pk.F.size() * 8) ||
      upload(h->U, h->pk.U.data(), h->pk.U.size() * 4) || upload(h->pbox, h->pk.pbox.data(), h->pk.pbox.size() * 4) ||
      upload(h->pcode, h->pk.pcode.data(), h->pk.pcode.size() * 4) || upload(h->bvh_box, h->pk.bvh_box.data(), h->pk.bvh_box.size() * 4) ||
      upload(h->bvh_prim, h->pk.bvh_prim.data(), h->pk.bvh_prim.size() * 4) || upload(h->bvh_holds, h->pk.bvh_holds.data(), h->pk.bvh_holds.size())) {
    std::string e = g_error;
    icam_hip_close(h);
    return fail(e);
  }
  HIP_OK(hipStreamCreate(&h->stream));
  HIP_OK(hipEventCreateWithFlags(&h->done, hipEventDisableTiming));
  // ICAM_NODE_LANES: lanes a node in a brick's nodes' search (1, 2, 4 or 8)
  if (h->pk.L.n_leaves + h->pk.L.n_cleaves > 65536) h->kn[0] = h->kn[1] = 4;
  if (const char* e = std::getenv("ICAM_NODE_LANES")) {
    const int v = std::atoi(e);
    if (v == 1 || v == 2 || v == 4 || v == 8) h->kn[0] = h->kn[1] = v;
  }
  // ICAM_SEARCH_ROUNDS: the searches' best-first rounds over parts, at most (4096)
  if (const char* e = std::getenv("ICAM_SEARCH_ROUNDS")) {
    const long r = std::atol(e);
    if (r > 0) h->pk.L.max_rounds = uint32_t(r);
  }
  h->dv.F32 = h->F32.as<float>();
  h->dv.F64 = h->F64.as<double>();
  h->dv.U = h->U.as<uint32_t>();
  h->dv.pbox = h->pbox.as<float>();
  h->dv.pcode = h->pcode.as<uint32_t>();
  h->dv.bvh_box = h->bvh_box.as<float>(), h->dv.bvh_prim = h->bvh_prim.as<uint32_t>(), h->dv.bvh_holds = h->bvh_holds.as<uint8_t>();
  h->dv.bvh_lv = h->bvh_lv.as<uint32_t>();
  h->dv.L = h->pk.L;
  for (int k = 0; k < 3; ++k) h->dv.origin[k] = h->pk.origin[k];
  for (int k = 0; k < 3; ++k) h->box_lo[k] = 1e300, h->box_hi[k] = -1e300;
  for (size_t i = 0; i + 6 <= h->pk.pbox.size(); i += 6)
    for (int k = 0; k < 3; ++k) {
      h->box_lo[k] = std::min(h->box_lo[k], double(h->pk.pbox[i + k]) + h->pk.origin[k]);
      h->box_hi[k] = std::max(h->box_hi[k], double(h->pk.pbox[i + 3 + k]) + h->pk.origin[k]);
    }
  *out = h;
  return 0;
}

void icam_hip_close(icam_hip* h) {
  if (!h) return;
  if (h->pending) hipEventSynchronize(h->done);
  if (h->stream) hipStreamSynchronize(h->stream);
  h->centres.free();
  h->nodes.free();
  for (Buf* b : {&h->F32, &h->F64, &h->U, &h->pbox, &h->pcode, &h->bvh_box, &h->bvh_prim, &h->bvh_holds, &h->bvh_lv, &h->q_xyz, &h->q_bound, &h->q_d, &h->q_flag, &h->q_feat, &h->dc, &h->band, &h->sgn, &h->total,
                 &h->list, &h->staging, &h->cls, &h->slot, &h->stotal, &h->map, &h->atlas, &h->counters, &h->changed, &h->firsts,
                 &h->tile_cand, &h->tile_count, &h->super_cand, &h->super_count, &h->o_count, &h->o_cell, &h->o_rank, &h->o_idx0, &h->o_idx, &h->o_xyz, &h->o_d, &h->o_flag, &h->o_feat})
    b->free();
  if (h->stream) hipStreamDestroy(h->stream);
  if (h->done) hipEventDestroy(h->done);
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
  if (after_previous(h, h->stream)) return 1;
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

// The scratch of a query of n points: the cell order's counts, cells, ranks, order,
// points and results in that order; the single-precision records; the fallback's and
// the settle's block queues, their counts and marks; the statistics.
struct QBufs {
  uint32_t *key, *key2, *idx0, *idx;  // the Morton keys and indices, before and after the sort
  void* sort;                          // the sort's scratch (sort_bytes)
  size_t sort_bytes;
  double* xyz;
  double* d;
  uint32_t *flag, *feat;
  uint32_t *code, *info;
  float *pa, *pb, *d32;
  uint32_t *queue, *qcount, *qmark, *squeue, *scount, *smark, *cqueue, *ccount;
  unsigned long long* counters;
};
static bool query_order(size_t n, bool bounds) { return n >= ORDER_MIN && !bounds && n < (size_t(1) << 31); }

// From the handle's buffers, grown as needed.
static int query_bufs(icam_hip* h, size_t n, bool mixed, bool order, bool feat, QBufs& q) {
  q = QBufs{};
  const size_t m = std::max<size_t>(n, 1);
  if (order) {
    q.sort_bytes = order_sort_bytes(m);
    HIP_OK(h->o_count.need(std::max<size_t>(q.sort_bytes, 16)));
    HIP_OK(h->o_cell.need(m * 4));
    HIP_OK(h->o_rank.need(m * 4));
    HIP_OK(h->o_idx0.need(m * 4));
    HIP_OK(h->o_idx.need(m * 4));
    HIP_OK(h->o_xyz.need(m * 24));
    HIP_OK(h->o_d.need(m * 8));
    HIP_OK(h->o_flag.need(m * 4));
    if (feat) HIP_OK(h->o_feat.need(m * 4));
    q.sort = h->o_count.p, q.key = h->o_cell.as<uint32_t>(), q.key2 = h->o_rank.as<uint32_t>(), q.idx0 = h->o_idx0.as<uint32_t>(), q.idx = h->o_idx.as<uint32_t>();
    q.xyz = h->o_xyz.as<double>(), q.d = h->o_d.as<double>(), q.flag = h->o_flag.as<uint32_t>(), q.feat = feat ? h->o_feat.as<uint32_t>() : nullptr;
  }
  icam_hip::ResBufs& B = h->nodes;
  if (mixed) {
    HIP_OK(B.code.need(m * 4));
    HIP_OK(B.pa.need(m * 4));
    HIP_OK(B.pb.need(m * 4));
    HIP_OK(B.d32.need(m * 4));
    HIP_OK(B.info.need(m * 4));
    HIP_OK(B.queue.need(m * 4));
    HIP_OK(B.qmark.need(m * 4));
    q.code = B.code.as<uint32_t>(), q.pa = B.pa.as<float>(), q.pb = B.pb.as<float>(), q.d32 = B.d32.as<float>(), q.info = B.info.as<uint32_t>();
    q.queue = B.queue.as<uint32_t>(), q.qmark = B.qmark.as<uint32_t>();
  }
  HIP_OK(B.qcount.need(16));
  HIP_OK(B.squeue.need((m + SETTLE_SLICES) * 4));
  HIP_OK(B.smark.need(m * 4));
  HIP_OK(B.scount.need(16));
  HIP_OK(B.cqueue.need(m * 4));
  HIP_OK(B.ccount.need(16));
  HIP_OK(h->counters.need(64));
  q.qcount = B.qcount.as<uint32_t>(), q.squeue = B.squeue.as<uint32_t>(), q.smark = B.smark.as<uint32_t>(), q.scount = B.scount.as<uint32_t>();
  q.cqueue = B.cqueue.as<uint32_t>(), q.ccount = B.ccount.as<uint32_t>();
  q.counters = h->counters.as<unsigned long long>();
  return 0;
}

// The same carved from device memory at `base` (cap bytes); the bytes it takes (more
// than cap: it does not fit, and q is not usable).
static size_t query_carve(char* base, size_t cap, size_t n, bool mixed, bool order, bool feat, QBufs& q) {
  q = QBufs{};
  size_t at = 0;
  auto take = [&](size_t bytes) -> char* {
    char* p = base + at;
    at = (at + bytes + 255) / 256 * 256;
    return at <= cap ? p : nullptr;
  };
  const size_t m = std::max<size_t>(n, 1), blk = m / 64 + 1;  // blocks of 64 points or more
  if (order) {
    q.sort_bytes = order_sort_bytes(m);
    q.sort = take(std::max<size_t>(q.sort_bytes, 16)), q.key = (uint32_t*)take(m * 4), q.key2 = (uint32_t*)take(m * 4), q.idx0 = (uint32_t*)take(m * 4);
    q.idx = (uint32_t*)take(m * 4);
    q.xyz = (double*)take(m * 24), q.d = (double*)take(m * 8), q.flag = (uint32_t*)take(m * 4);
    q.feat = feat ? (uint32_t*)take(m * 4) : nullptr;
  }
  if (mixed) {
    q.code = (uint32_t*)take(m * 4), q.pa = (float*)take(m * 4), q.pb = (float*)take(m * 4), q.d32 = (float*)take(m * 4), q.info = (uint32_t*)take(m * 4);
    q.queue = (uint32_t*)take(blk * 4), q.qmark = (uint32_t*)take(blk * 4);
  }
  q.qcount = (uint32_t*)take(16), q.squeue = (uint32_t*)take(blk * 4 * (SETTLE_SLICES + 1)), q.smark = (uint32_t*)take(blk * 4), q.scount = (uint32_t*)take(16);
  q.cqueue = (uint32_t*)take(blk * 4), q.ccount = (uint32_t*)take(16);
  q.counters = (unsigned long long*)take(64);
  return at;
}

static int query_impl(icam_hip* h, int precision, const double* d_xyz, const double* d_bounds, size_t n, double* d_d, uint32_t* d_flags, uint32_t* d_feat,
               hipStream_t s, Stages* st, const QBufs& q) {
  if (!n) return 0;
  const bool mixed = precision == ICAM_MIXED;
  const unsigned blocks = unsigned((n + 255) / 256);
  Job j{};
  j.n = long(n), j.xyz = d_xyz, j.bounds = d_bounds;
  // scattered points without bounds of their own: searched in cell order
  const bool order = query_order(n, d_bounds != nullptr);
  {
    FillList f{{(uint32_t*)q.counters, nullptr, nullptr, nullptr, nullptr, nullptr}, {16u, 0u, 0u, 0u, 0u, 0u}};
    HIP_OK(launch(st, "zero", s, dim3(1), dim3(256), kb_fill, f));
  }
  if (order) {
    double sc[3];
    for (int k = 0; k < 3; ++k) {
      const double side = h->box_hi[k] - h->box_lo[k];
      sc[k] = side > 0.0 ? double(1 << ORDER_BITS) / side : 0.0;
    }
    HIP_OK(launch(st, "order", s, dim3(blocks), dim3(256), k_order_keys, d_xyz, long(n), h->box_lo[0], h->box_lo[1], h->box_lo[2], sc[0], sc[1], sc[2],
                  q.key, q.idx0));
    size_t bytes = q.sort_bytes;
    HIP_OK(rocprim::radix_sort_pairs(q.sort, bytes, q.key, q.key2, q.idx0, q.idx, n, 0, 3 * ORDER_BITS, s));
    HIP_OK(launch(st, "order", s, dim3(blocks), dim3(256), k_order_place, d_xyz, long(n), (const uint32_t*)q.idx, q.xyz));
    j.xyz = q.xyz;
  }
  Res r{};
  r.code = q.code, r.pa = q.pa, r.pb = q.pb, r.d32 = q.d32, r.info = q.info;
  r.queue = q.queue, r.qcount = q.qcount, r.qmark = q.qmark, r.squeue = q.squeue, r.scount = q.scount, r.smark = q.smark;
  r.cqueue = q.cqueue, r.ccount = q.ccount;
  r.counters = q.counters;
  r.val = order ? q.d : d_d, r.flag = order ? q.flag : d_flags, r.feat = order ? (d_feat ? q.feat : nullptr) : d_feat;
  // a part with B-spline faces or spline edges: four lanes a point (on the trimmed plate's
  // uniform points 2.2x faster than one; its points at the trim curves 1.2x slower)
  const bool wide = h->dv.L.n_leaves + h->dv.L.n_cleaves > 0;
  const long ppb = wide ? Shape<SRC_WIDE>::NT / Shape<SRC_WIDE>::K : Shape<SRC_POINTS>::NT;
  const long sblocks = (long(n) + ppb - 1) / ppb;
  if (wide ? run_search<SRC_WIDE>(h, mixed, j, r, s, st, sblocks) : run_search<SRC_POINTS>(h, mixed, j, r, s, st, sblocks)) return 1;
  if (order)
    HIP_OK(launch(st, "order", s, dim3(blocks), dim3(256), k_order_back, long(n), (const uint32_t*)q.idx, (const double*)q.d, (const uint32_t*)q.flag,
                  (const uint32_t*)(d_feat ? q.feat : nullptr), d_d, d_flags, d_feat));
  return 0;
}

// A query on the handle's own scratch.
static int query_handle(icam_hip* h, int precision, const double* d_xyz, const double* d_bounds, size_t n, double* d_d, uint32_t* d_flags, uint32_t* d_feat,
                 hipStream_t s, Stages* st) {
  if (!n) return 0;
  if (precision != ICAM_FP64 && precision != ICAM_MIXED) return fail("unknown precision");
  QBufs q;
  if (query_bufs(h, n, precision == ICAM_MIXED, query_order(n, d_bounds != nullptr), d_feat != nullptr, q)) return 1;
  return query_impl(h, precision, d_xyz, d_bounds, n, d_d, d_flags, d_feat, s, st, q);
}

int icam_hip_query_device(icam_hip* h, int precision, const double* d_xyz, size_t n, double* d_d, uint32_t* d_flags, uint32_t* d_feat, void* stream) {
  if (!h) return fail("no handle");
  std::lock_guard<std::mutex> g(h->mu);
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  if (after_previous(h, s) || query_handle(h, precision, d_xyz, nullptr, n, d_d, d_flags, d_feat, s, nullptr)) return 1;
  return mark_end(h, s);
}

int icam_hip_query_device_noalloc(icam_hip* h, int precision, const double* d_xyz, size_t n, double* d_d, uint32_t* d_flags, uint32_t* d_feat,
                                  void* stream) {
  if (!h) return fail("no handle");
  if (precision != ICAM_FP64 && precision != ICAM_MIXED) return fail("unknown precision");
  if (n > NOALLOC_POINTS) return fail("icam_hip_query_device_noalloc takes at most " + std::to_string(NOALLOC_POINTS) + " points a call, not " + std::to_string(n));
  if (!n) return 0;
  std::lock_guard<std::mutex> g(h->mu);
  char* base = static_cast<char*>(noalloc_scratch(h->device));
  if (!base) return fail("the scratch array of device " + std::to_string(h->device) + " is not set up");
  QBufs q;
  const size_t bytes = query_carve(base, NOALLOC_BYTES, n, precision == ICAM_MIXED, query_order(n, false), d_feat != nullptr, q);
  if (bytes > NOALLOC_BYTES) return fail("the query needs " + std::to_string(bytes) + " bytes of scratch, the array holds " + std::to_string(NOALLOC_BYTES));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  if (after_previous(h, s) || query_impl(h, precision, d_xyz, nullptr, n, d_d, d_flags, d_feat, s, nullptr, q)) return 1;
  return mark_end(h, s);
}

int icam_hip_query(icam_hip* h, int precision, const double* xyz, size_t n, double* d, uint32_t* flags, uint32_t* feat) {
  if (!h) return fail("no handle");
  std::lock_guard<std::mutex> g(h->mu);
  if (!n) return 0;
  HIP_OK(h->q_xyz.need(n * 24));
  HIP_OK(h->q_d.need(n * 8));
  HIP_OK(h->q_flag.need(n * 4));
  HIP_OK(h->q_feat.need(n * 4));
  if (after_previous(h, h->stream)) return 1;
  HIP_OK(hipMemcpyAsync(h->q_xyz.p, xyz, n * 24, hipMemcpyHostToDevice, h->stream));
  Stages st;
  if (query_handle(h, precision, h->q_xyz.as<double>(), nullptr, n, h->q_d.as<double>(), h->q_flag.as<uint32_t>(), h->q_feat.as<uint32_t>(), h->stream,
                   &st))
    return 1;
  HIP_OK(hipMemcpyAsync(d, h->q_d.p, n * 8, hipMemcpyDeviceToHost, h->stream));
  if (flags) HIP_OK(hipMemcpyAsync(flags, h->q_flag.p, n * 4, hipMemcpyDeviceToHost, h->stream));
  if (feat) HIP_OK(hipMemcpyAsync(feat, h->q_feat.p, n * 4, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  h->pending = false;
  collect(h, st);
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
  if (after_previous(h, h->stream)) return 1;
  HIP_OK(hipMemcpyAsync(h->q_xyz.p, xyz.data(), n * 24, hipMemcpyHostToDevice, h->stream));
  HIP_OK(hipMemcpyAsync(h->q_bound.p, bd.data(), n * 8, hipMemcpyHostToDevice, h->stream));
  if (query_handle(h, precision, h->q_xyz.as<double>(), h->q_bound.as<double>(), n, h->q_d.as<double>(), h->q_flag.as<uint32_t>(), nullptr,
                   h->stream, nullptr))
    return 1;
  std::vector<double> d(n);
  std::vector<uint32_t> f(n);
  HIP_OK(hipMemcpyAsync(d.data(), h->q_d.p, n * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipMemcpyAsync(f.data(), h->q_flag.p, n * 4, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  h->pending = false;
  for (size_t i = 0; i < n; ++i) out[2 * i] = d[i], out[2 * i + 1] = double(f[i]);
  return 0;
}

int icam_hip_bricks(icam_hip* h, int precision, const icam_brick_rec* recs, size_t n, double* out) {
  if (!h) return fail("no handle");
  std::lock_guard<std::mutex> g(h->mu);
  if (!n) return 0;
  if (precision != ICAM_FP64 && precision != ICAM_MIXED) return fail("unknown precision");
  HIP_OK(h->firsts.need(n * 32));
  if (after_previous(h, h->stream)) return 1;
  HIP_OK(hipMemcpyAsync(h->firsts.p, recs, n * 32, hipMemcpyHostToDevice, h->stream));
  Res r{};
  long np = long(n) * 729;
  if (results(h, true, np, r, nullptr, nullptr, nullptr, precision == ICAM_MIXED, true)) return 1;
  Job j{};
  j.n = np, j.firsts = h->firsts.as<double>();
  if (run_search<SRC_NODES>(h, precision == ICAM_MIXED, j, r, h->stream, nullptr, long(n))) return 1;
  std::vector<double> v(np);
  std::vector<uint32_t> f(np);
  HIP_OK(hipMemcpyAsync(v.data(), r.val, np * 8, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipMemcpyAsync(f.data(), r.flag, np * 4, hipMemcpyDeviceToHost, h->stream));
  HIP_OK(hipStreamSynchronize(h->stream));
  h->pending = false;
  for (long i = 0; i < np; ++i) out[2 * i] = v[i], out[2 * i + 1] = double(f[i]);
  return 0;
}

static int bake_impl(icam_hip* h, int precision, const icam_level* lvp, hipStream_t s, icam_bake_stats* stats) {
  Lv lv;
  for (int k = 0; k < 3; ++k) lv.lo[k] = lvp->lo[k], lv.nb[k] = lvp->nb[k];
  lv.h = lvp->h;
  if (!(lv.h > 0) || lv.nb[0] <= 0 || lv.nb[1] <= 0 || lv.nb[2] <= 0) return fail("bad level");
  if (double(lv.nb[0]) * lv.nb[1] * lv.nb[2] >= 2147483648.0) return fail("a level holds fewer than 2^31 bricks");
  if (precision != ICAM_FP64 && precision != ICAM_MIXED) return fail("unknown precision");
  const bool mixed = precision == ICAM_MIXED;
  long nb = long(lv.nb[0]) * lv.nb[1] * lv.nb[2];
  HIP_OK(h->dc.need(nb * 8));
  HIP_OK(h->band.need(nb));
  HIP_OK(h->sgn.need(nb));
  HIP_OK(h->total.need(16));
  HIP_OK(h->list.need(nb * 4));
  HIP_OK(h->map.need(nb * 4));
  HIP_OK(h->counters.need(64));
  HIP_OK(h->changed.need(16));
  Stages st;
  dim3 gb(unsigned((nb + WG - 1) / WG));
  // the brick centres
  Res r{};
  const bool cmixed = mixed && !h->c64;
  if (results(h, false, nb, r, nullptr, nullptr, nullptr, cmixed)) return 1;
  HIP_OK(hipMemsetAsync(h->counters.p, 0, 64, s));
  Job jc{};
  jc.n = nb, jc.lv = lv;
  long tiles = tiles_of(lv.nb);
  HIP_OK(h->tile_cand.need(size_t(tiles) * TCAP * 4));
  HIP_OK(h->tile_count.need(size_t(tiles) * 4));
  jc.tile_cand = h->tile_cand.as<uint32_t>(), jc.tile_count = h->tile_count.as<uint32_t>();
  const long supers = supers_of(lv.nb);
  HIP_OK(h->super_cand.need(size_t(supers) * SCAP * 4));
  HIP_OK(h->super_count.need(size_t(supers) * 4));
  jc.super_cand = h->super_cand.as<uint32_t>(), jc.super_count = h->super_count.as<uint32_t>();
  HIP_OK(launch(&st, "centres_lists", s, dim3(unsigned(supers)), dim3(256), k_super, h->dv, jc));
  if (run_search<SRC_CENTRES>(h, cmixed, jc, r, s, &st, tiles)) return 1;
  HIP_OK(launch(&st, "decide", s, gb, dim3(WG), k_decide_finish, h->dv, lv, nb, (const double*)r.val, (const uint32_t*)r.flag, h->dc.as<double>(),
                h->band.as<uint8_t>(), h->sgn.as<int8_t>()));
  // signs of the bricks kept whole: along x, y, z lines, until nothing changes
  if (sign_pitch(lv.nb[0]) * lv.nb[1] * lv.nb[2] <= SIGN_LDS) {
    HIP_OK(launch(&st, "sign", s, dim3(1), dim3(SIGN_T), k_sign_block, lv, (const uint8_t*)h->band.as<uint8_t>(), h->sgn.as<int8_t>()));
  } else for (int round = 0; round < SIGN_ROUNDS; ++round) {
    HIP_OK(hipMemsetAsync(h->changed.p, 0, 4, s));
    for (int ax = 0; ax < 3; ++ax) {
      long lines = long(lv.nb[(ax + 1) % 3]) * lv.nb[(ax + 2) % 3];
      HIP_OK(launch(&st, "sign", s, dim3(unsigned((lines + 255) / 256)), dim3(256), k_sign, lv, ax, (const uint8_t*)h->band.as<uint8_t>(),
                    h->sgn.as<int8_t>(), h->changed.as<int>()));
    }
    int ch = 0;
    HIP_OK(hipMemcpyAsync(&ch, h->changed.p, 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (!ch) break;
  }
  HIP_OK(launch(&st, "band_list", s, dim3(1), dim3(SCAN_T), k_scan, (const uint8_t*)h->band.as<uint8_t>(), (const uint32_t*)nullptr, nb,
                (uint32_t*)nullptr, h->list.as<uint32_t>(), h->total.as<uint32_t>()));
  uint32_t nband = 0;
  HIP_OK(hipMemcpyAsync(&nband, h->total.p, 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  HIP_OK(h->staging.need(std::max<size_t>(size_t(nband) * ICAM_WORDS * 4, 16)));
  HIP_OK(h->cls.need(std::max<size_t>(size_t(nband) * 4, 16)));
  HIP_OK(h->slot.need(std::max<size_t>(size_t(nband) * 4, 16)));
  HIP_OK(h->stotal.need(16));
  if (nband) {
    // the band bricks' nodes
    Res rn{};
    long np = long(nband) * 729;
    if (results(h, true, np, rn, nullptr, nullptr, nullptr, mixed, true)) return 1;
    Job jn{};
    jn.n = np, jn.list = h->list.as<uint32_t>(), jn.lv = lv;
    jn.tile_cand = h->tile_cand.as<uint32_t>(), jn.tile_count = h->tile_count.as<uint32_t>();
    if (run_search<SRC_NODES>(h, mixed, jn, rn, s, &st, long(nband))) return 1;
    // the centres of all-far band bricks with no sign: settled
    HIP_OK(hipMemsetAsync(r.scount, 0, 8, s));
    HIP_OK(hipMemsetAsync(r.smark, 0, size_t(tiles) * 4, s));
    HIP_OK(launch(&st, "centres_need", s, dim3(unsigned(std::min<long>(nband, 4096))), dim3(256), k_need_centre, jc, r, (const uint32_t*)h->list.as<uint32_t>(),
                  long(nband), (const uint32_t*)rn.flag, (const double*)h->dc.as<double>(), (const int8_t*)h->sgn.as<int8_t>()));
    HIP_OK(launch(&st, "centres_settle64", s, dim3(unsigned(std::min<long>(tiles, h->settle_grid))), dim3(SETTLE_NT), k_settle<SRC_CENTRES>,
                  h->dv, jc, r));
    HIP_OK(launch(&st, "centres_sign", s, dim3(unsigned(std::min<long>((long(nband) + 255) / 256, 4096))), dim3(256), k_centre_sign, r,
                  (const uint32_t*)h->list.as<uint32_t>(), long(nband), h->dc.as<double>(), h->sgn.as<int8_t>()));
    HIP_OK(launch(&st, "bricks_finish", s, dim3(nband), dim3(WG), k_bricks_finish, lv, (const uint32_t*)h->list.as<uint32_t>(), long(nband),
                  (const double*)h->dc.as<double>(), (const int8_t*)h->sgn.as<int8_t>(), (const uint8_t*)h->band.as<uint8_t>(),
                  (const double*)rn.val, (const uint32_t*)rn.flag,
                  h->staging.as<uint32_t>(), h->cls.as<uint32_t>(), h->counters.as<unsigned long long>()));
    HIP_OK(launch(&st, "slots", s, dim3(1), dim3(SCAN_T), k_scan, (const uint8_t*)nullptr, (const uint32_t*)h->cls.as<uint32_t>(), long(nband),
                  h->slot.as<uint32_t>(), (uint32_t*)nullptr, h->stotal.as<uint32_t>()));
  } else {
    HIP_OK(hipMemsetAsync(h->stotal.p, 0, 4, s));
  }
  uint32_t nslots = 0;
  HIP_OK(hipMemcpyAsync(&nslots, h->stotal.p, 4, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  HIP_OK(h->atlas.need(std::max<size_t>(size_t(nslots) * ICAM_WORDS * 4, 16)));
  HIP_OK(launch(&st, "store", s, gb, dim3(WG), k_store, nb, (const uint8_t*)h->band.as<uint8_t>(), (const int8_t*)h->sgn.as<int8_t>(),
                h->map.as<uint32_t>(), h->counters.as<unsigned long long>()));
  if (nband)
    HIP_OK(launch(&st, "store_band", s, dim3(nband), dim3(128), k_store_band, long(nband), (const uint32_t*)h->list.as<uint32_t>(),
                  (const uint32_t*)h->cls.as<uint32_t>(), (const uint32_t*)h->slot.as<uint32_t>(), (const uint32_t*)h->staging.as<uint32_t>(),
                  h->map.as<uint32_t>(), h->atlas.as<uint32_t>()));
  unsigned long long cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIP_OK(hipMemcpyAsync(cnt, h->counters.p, 64, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  collect(h, st);
  if (cnt[2]) return fail(std::to_string(cnt[2]) + " bricks kept whole with no sign found (no surface near them and none outside the part's box)");
  if (stats) {
    stats->bricks = nb, stats->band = nband, stats->slots = nslots;
    stats->undecided = int64_t(cnt[0]), stats->refined = int64_t(cnt[1]);
    stats->unsure_trim = int64_t(cnt[3]), stats->unsure_tie = int64_t(cnt[4]), stats->settled = int64_t(cnt[5]);
    stats->overflowed = int64_t(cnt[6]), stats->host = int64_t(cnt[7]);
  }
  return 0;
}

int icam_hip_bake_device(icam_hip* h, int precision, const icam_level* lv, uint32_t** d_map, uint32_t** d_atlas, icam_bake_stats* stats, void* stream) {
  if (!h || !lv) return fail("bad arguments");
  std::lock_guard<std::mutex> g(h->mu);
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  icam_bake_stats tmp{};
  if (after_previous(h, s) || bake_impl(h, precision, lv, s, stats ? stats : &tmp) || mark_end(h, s)) return 1;
  if (d_map) *d_map = h->map.as<uint32_t>();
  if (d_atlas) *d_atlas = h->atlas.as<uint32_t>();
  return 0;
}

int icam_hip_bake(icam_hip* h, int precision, const icam_level* lv, uint32_t* map, uint32_t* atlas, size_t atlas_cap, icam_bake_stats* stats) {
  if (!h || !lv) return fail("bad arguments");
  std::lock_guard<std::mutex> g(h->mu);
  icam_bake_stats st{};
  if (after_previous(h, h->stream) || bake_impl(h, precision, lv, h->stream, &st) || mark_end(h, h->stream)) return 1;
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

int icam_hip_tune(icam_hip* h, int precision, const icam_level* lv, char* report, size_t cap) {
  if (!h || !lv || (precision != ICAM_FP64 && precision != ICAM_MIXED)) return fail("bad arguments");
  std::lock_guard<std::mutex> g(h->mu);
  const bool mixed = precision == ICAM_MIXED;
  int& kn = h->kn[mixed ? 1 : 0];
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  // the level's bake timed: the least of two runs after a warm-up
  auto time_bake = [&](double& t) -> int {
    t = 1e300;
    for (int r = 0; r < 3; ++r) {
      icam_bake_stats st{};
      if (after_previous(h, h->stream)) return 1;
      HIP_OK(hipEventRecord(e0, h->stream));
      if (bake_impl(h, precision, lv, h->stream, &st)) return 1;
      HIP_OK(hipEventRecord(e1, h->stream));
      if (mark_end(h, h->stream)) return 1;
      HIP_OK(hipEventSynchronize(e1));
      float m = 0.f;
      HIP_OK(hipEventElapsedTime(&m, e0, e1));
      if (r > 0) t = std::min(t, double(m));
    }
    return 0;
  };
  std::string rep;
  char line[80];
  int best = kn;
  double best_ms = 1e300;
  for (int c : {1, 2, 4, 8}) {
    kn = c;
    double t;
    if (time_bake(t)) return 1;
    std::snprintf(line, sizeof line, "%slanes a node %d: %.3f ms", rep.empty() ? "" : "; ", c, t);
    rep += line;
    if (t < best_ms) best_ms = t, best = c;
  }
  kn = best;
  rep += "; chosen " + std::to_string(best);
  if (mixed) {
    // the centres' decision in single precision (with its double-precision fallback) or
    // in double
    const bool keep = h->c64;
    h->c64 = !keep;
    double t;
    if (time_bake(t)) return 1;
    std::snprintf(line, sizeof line, "; centres in %s: %.3f ms", h->c64 ? "double" : "single", t);
    rep += line;
    if (!(t < best_ms)) h->c64 = keep;
    rep += h->c64 ? "; centres chosen double" : "; centres chosen single";
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (report && cap) std::snprintf(report, cap, "%s", rep.c_str());
  return 0;
}

int icam_hip_query_stats(icam_hip* h, icam_bake_stats* stats) {
  if (!h || !stats) return fail("bad arguments");
  std::lock_guard<std::mutex> g(h->mu);
  unsigned long long cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIP_OK(hipDeviceSynchronize());
  if (h->counters.p) HIP_OK(hipMemcpy(cnt, h->counters.p, 64, hipMemcpyDeviceToHost));
  *stats = icam_bake_stats{};
  stats->refined = int64_t(cnt[1]), stats->unsure_trim = int64_t(cnt[3]), stats->unsure_tie = int64_t(cnt[4]), stats->settled = int64_t(cnt[5]);
  stats->overflowed = int64_t(cnt[6]);
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

// ------------------------------------------------------------------ batches (host)

struct icam_batch {
  int precision = ICAM_MIXED;
  int n = 0;
  std::vector<icam_hip*> parts;  // in batch order (heaviest first)
  std::vector<int> caller;       // batch position -> the caller's index
  std::vector<int> pos;          // the caller's index -> batch position
  std::vector<BPart> host;       // the records uploaded
  std::vector<long> brick_off, tile_off, super_off;
  std::vector<size_t> cap_off;   // band room before each part
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  Buf bparts, brick_off_d, tile_off_d, super_off_d, band_off, btotal, stotal, ctl, counters, super_cand, super_count, wide_d;
  long wide_lines = 0;  // the most lines along one axis of a level signed grid-wide
  Buf c_code, c_pa, c_pb, c_d32, c_info, c_val, c_flag, c_qmark, c_smark, c_queue, c_squeue;
  Buf n_code, n_pa, n_pb, n_d32, n_info, n_val, n_flag, n_qmark, n_smark, n_queue, n_squeue, items;
  Buf tile_cand, tile_count, dc, band, sgn, list, staging, cls, slot, map, atlas;
  long total_bricks = 0, total_tiles = 0, total_supers = 0;
  size_t total_cap = 0;
  // the launch shapes: persistent grids of the centres' search, the nodes' search
  // and the band bricks' finish (blocks)
  int gc = 2304, gn = 1024, gf = 2048;
  int kn = 1;    // lanes a node in the nodes' search
  int c64 = 0;   // the mixed form's centres' decision in double precision
  int gs = 256;  // the settle stages' grid
  std::string tuning;
  BatchArgs args{};
  ~icam_batch() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    for (Buf* b : {&wide_d, &bparts, &brick_off_d, &tile_off_d, &super_off_d, &super_cand, &super_count, &band_off, &btotal, &stotal, &ctl, &counters, &c_code, &c_pa, &c_pb, &c_d32, &c_info, &c_val,
                   &c_flag, &c_qmark, &c_smark, &c_queue, &c_squeue, &n_code, &n_pa, &n_pb, &n_d32, &n_info, &n_val, &n_flag, &n_qmark, &n_smark,
                   &n_queue, &n_squeue, &items, &tile_cand, &tile_count, &dc, &band, &sgn, &list, &staging, &cls, &slot, &map, &atlas})
      b->free();
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

// Room for a part's band bricks: the bricks whose centre lies within the reach of
// some primitive's box (a band brick's centre is within its reach of the surface, or
// of a part's lower bound, and every surface point and bound lies in a primitive's
// box).
uint32_t band_room(const icam_hip* h, const Lv& lv) {
  const long nb0 = lv.nb[0], nb1 = lv.nb[1], nb2 = lv.nb[2];
  std::vector<uint8_t> mark(size_t(nb0 * nb1 * nb2), 0);
  const double reach = brick_reach(lv.h) * (1.0 + 1e-6) + 1e-6 * h->pk.L.diag, side = 8.0 * lv.h;
  const long nbs[3] = {nb0, nb1, nb2};
  const std::vector<float>& bx = h->pk.pbox;
  for (size_t k = 0; k < bx.size() / 6; ++k) {
    long a[3], b[3];
    bool empty = false;
    for (int ax = 0; ax < 3; ++ax) {
      const double c0 = lv.lo[ax] + 4.0 * lv.h;  // the first centre
      const double lo = double(bx[6 * k + ax]) + h->pk.origin[ax] - reach, hi = double(bx[6 * k + 3 + ax]) + h->pk.origin[ax] + reach;
      a[ax] = std::max(long(std::ceil((lo - c0) / side)), 0L);
      b[ax] = std::min(long(std::floor((hi - c0) / side)), nbs[ax] - 1);
      empty |= a[ax] > b[ax];
    }
    if (empty) continue;
    for (long z = a[2]; z <= b[2]; ++z)
      for (long y = a[1]; y <= b[1]; ++y)
        for (long x = a[0]; x <= b[0]; ++x) mark[size_t(x + nb0 * (y + nb1 * z))] = 1;
  }
  size_t c = 0;
  for (uint8_t m : mark) c += m;
  return uint32_t(c);
}

// The batch's launches on stream s (with per-stage HIP events when st is not NULL).
int batch_enqueue(icam_batch* B, hipStream_t s, Stages* st = nullptr) {
  const BatchArgs& a = B->args;
  const bool mixed = B->precision == ICAM_MIXED;
  {
    // control words, counters and the queues' marks
    FillList f{{B->ctl.as<uint32_t>(), (uint32_t*)B->counters.as<unsigned long long>(), B->c_qmark.as<uint32_t>(), B->c_smark.as<uint32_t>(),
                B->n_qmark.as<uint32_t>(), B->n_smark.as<uint32_t>()},
               {BC_WORDS, size_t(B->n) * 16, size_t(B->total_tiles), size_t(B->total_tiles), B->total_cap, B->total_cap}};
    size_t most = 0;
    for (size_t n : f.n) most = std::max(most, n);
    HIP_OK(launch(st, "zero", s, dim3(unsigned(std::min<size_t>((most + 255) / 256, 4096))), dim3(256), kb_fill, f));
  }
  const unsigned pts_c = unsigned(std::min<long>((B->total_bricks + 255) / 256, 8192));
  const dim3 ntc(Shape<SRC_CENTRES>::NT), ntn(Shape<SRC_NODES>::NT);
  // the brick centres: the level's decision (the super-tiles' lists first)
  HIP_OK(launch(st, "centres_lists", s, dim3(unsigned(std::min<long>(std::max<long>(B->total_supers, 1), 4096))), dim3(256), kb_super, a));
  if (mixed && !B->c64) {
    HIP_OK(launch(st, "centres_search32", s, dim3(B->gc), ntc, kb_search<SRC_CENTRES, float>, a));
    HIP_OK(launch(st, "centres_refine64", s, dim3(pts_c), dim3(256), kb_refine<SRC_CENTRES>, a));
    HIP_OK(launch(st, "centres_fallback64", s, dim3(B->gc), ntc, kb_redo<SRC_CENTRES, false>, a));
  } else {
    HIP_OK(launch(st, "centres_search64", s, dim3(B->gc), ntc, kb_search<SRC_CENTRES, double>, a));
  }
  HIP_OK(launch(st, "decide", s, dim3(pts_c), dim3(256), kb_decide, a));
  HIP_OK(launch(st, "sign", s, dim3(B->n), dim3(SIGN_T), kb_sign, a));
  if (a.nwide)
    for (int round = 0; round < SIGN_ROUNDS; ++round)
      for (int ax = 0; ax < 3; ++ax)
        HIP_OK(launch(st, "sign", s, dim3(unsigned((B->wide_lines + 255) / 256), unsigned(a.nwide)), dim3(256), kb_sign_wide, a, ax));
  HIP_OK(launch(st, "band_list", s, dim3(B->n), dim3(SCAN_T), kb_band_list, a));
  HIP_OK(launch(st, "plan", s, dim3(1), dim3(64), kb_plan, a));
  // the band bricks' nodes
  const unsigned gl = unsigned(std::min<size_t>((B->total_cap * 729 / 8 + 1024 + 255) / 256, 2048));
  if (mixed) {
    HIP_OK(launch(st, "nodes_search32", s, dim3(B->gn), ntn, B->kn == 8 ? kb_search<SRC_NODES, float, 8> : B->kn == 4 ? kb_search<SRC_NODES, float, 4> : B->kn == 2 ? kb_search<SRC_NODES, float, 2> : kb_search<SRC_NODES, float, 1>, a));
    HIP_OK(launch(st, "nodes_leaves32", s, dim3(gl), dim3(256), kb_leaves<float>, a));
    HIP_OK(launch(st, "nodes_refine64", s, dim3(unsigned(std::min<size_t>(std::max<size_t>(B->total_cap, 1), 8192))), dim3(256), kb_refine_bricks, a));
    HIP_OK(launch(st, "nodes_fallback64", s, dim3(B->gn), ntn, B->kn >= 4 ? kb_redo<SRC_NODES, false, 8> : kb_redo<SRC_NODES, false, 4>, a));
  } else {
    HIP_OK(launch(st, "nodes_search64", s, dim3(B->gn), ntn, B->kn == 8 ? kb_search<SRC_NODES, double, 8> : B->kn == 4 ? kb_search<SRC_NODES, double, 4> : B->kn == 2 ? kb_search<SRC_NODES, double, 2> : kb_search<SRC_NODES, double, 1>, a));
    HIP_OK(launch(st, "nodes_leaves64", s, dim3(gl), dim3(256), kb_leaves<double>, a));
  }
  HIP_OK(launch(st, "nodes_mark", s, dim3(unsigned(std::min<size_t>(std::max<size_t>(B->total_cap, 1), 8192))), dim3(256), kb_mark_bricks, a));
  HIP_OK(launch(st, "nodes_settle64", s, dim3(unsigned(B->gs)), dim3(SETTLE_NT), kb_settle<SRC_NODES>, a));
  // the centres of all-far band bricks with no sign: settled
  HIP_OK(launch(st, "centres_need", s, dim3(B->gf), dim3(256), kb_need_centre, a));
  HIP_OK(launch(st, "centres_settle64", s, dim3(unsigned(B->gs)), dim3(SETTLE_NT), kb_settle<SRC_CENTRES>, a));
  HIP_OK(launch(st, "centres_sign", s, dim3(unsigned(std::min<size_t>((B->total_cap + 255) / 256, 4096))), dim3(256), kb_centre_sign, a));
  // fill, quantize, pack, number and store
  HIP_OK(launch(st, "bricks_finish", s, dim3(B->gf), dim3(WG), kb_finish, a));
  HIP_OK(launch(st, "slots", s, dim3(B->n), dim3(SCAN_T), kb_slots, a));
  HIP_OK(launch(st, "store", s, dim3(pts_c), dim3(256), kb_store, a));
  HIP_OK(launch(st, "store_band", s, dim3(unsigned(std::min<size_t>(std::max<size_t>(B->total_cap, 1), 65535))), dim3(128), kb_store_band, a));
  return 0;
}

int batch_time(icam_batch* B, double& ms) {
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, B->stream));
  if (batch_enqueue(B, B->stream)) return 1;
  HIP_OK(hipEventRecord(e1, B->stream));
  HIP_OK(hipEventSynchronize(e1));
  float m = 0;
  HIP_OK(hipEventElapsedTime(&m, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  ms = m;
  return 0;
}

// The autotuner: the whole batch timed with each candidate of one knob at a time (the
// others held), the fastest kept, knob after knob: the persistent grids of the centres'
// search, the nodes' search and the finish (blocks a CU), then the lanes a node of the
// nodes' search (points a wave: 64, 32, 16 or 8), then, in the mixed form, the precision of the
// centres' decision (single with its fallback, or double).
int batch_tune(icam_batch* B, int cus) {
  double ms;
  if (batch_time(B, ms)) return 1;  // warm-up
  char line[160];
  std::snprintf(line, sizeof line, "start gc %d gn %d gf %d kn %d c64 %d: %.3f ms;", B->gc, B->gn, B->gf, B->kn, B->c64, ms);
  B->tuning = line;
  int* knobs[5] = {&B->gc, &B->gn, &B->gf, &B->kn, &B->c64};
  const char* names[5] = {"gc", "gn", "gf", "kn", "c64"};
  const int cand[5][4] = {{2 * cus, 4 * cus, 9 * cus, 18 * cus}, {cus, 2 * cus, 4 * cus, 8 * cus}, {2 * cus, 4 * cus, 8 * cus, 16 * cus}, {1, 2, 4, 8},
                          {0, 1, -1, -1}};
  const int knob_count = B->precision == ICAM_MIXED ? 5 : 4;
  for (int k = 0; k < knob_count; ++k) {
    int best = *knobs[k];
    double best_ms = 1e300;
    for (int c = 0; c < 4 && cand[k][c] >= 0; ++c) {
      *knobs[k] = cand[k][c];
      double t0, t1;
      if (batch_time(B, t0) || batch_time(B, t1)) return 1;
      const double t = std::min(t0, t1);
      std::snprintf(line, sizeof line, " %s %d: %.3f ms;", names[k], *knobs[k], t);
      B->tuning += line;
      if (t < best_ms) best_ms = t, best = *knobs[k];
    }
    *knobs[k] = best;
  }
  std::snprintf(line, sizeof line, " chosen gc %d gn %d gf %d kn %d c64 %d", B->gc, B->gn, B->gf, B->kn, B->c64);
  B->tuning += line;
  return 0;
}

}  // namespace

extern "C" {

int icam_hip_batch_open(icam_hip* const* parts, const icam_level* levels, size_t n, int precision, int tune, icam_batch** out) {
  g_error.clear();
  if (!out) return fail("bad arguments");
  *out = nullptr;
  if (!parts || !levels || n == 0) return fail("an empty batch");
  if (n > 255) return fail("a batch holds at most 255 parts");
  if (precision != ICAM_FP64 && precision != ICAM_MIXED) return fail("unknown precision");
  std::unique_ptr<icam_batch> B(new icam_batch());
  B->precision = precision;
  B->n = int(n);
  int dev = parts[0] ? parts[0]->device : 0;
  // levels and room, then the order: heaviest first (band room times the part's
  // primitives, a proxy for its bricks' work)
  std::vector<Lv> lvs(n);
  std::vector<uint32_t> caps(n);
  std::vector<double> work(n);
  for (size_t i = 0; i < n; ++i) {
    if (!parts[i]) return fail("part " + std::to_string(i) + " is not open");
    if (parts[i]->device != dev) return fail("the parts of a batch live on one device");
    Lv& lv = lvs[i];
    for (int k = 0; k < 3; ++k) lv.lo[k] = levels[i].lo[k], lv.nb[k] = levels[i].nb[k];
    lv.h = levels[i].h;
    if (!(lv.h > 0) || lv.nb[0] <= 0 || lv.nb[1] <= 0 || lv.nb[2] <= 0) return fail("bad level " + std::to_string(i));
    if (double(lv.nb[0]) * lv.nb[1] * lv.nb[2] >= double(1 << 24)) return fail("level " + std::to_string(i) + " has 2^24 bricks or more");
    caps[i] = band_room(parts[i], lv);
    // ICAM_BATCH_ROOM: band bricks a part has room for, at most (a smaller room makes a
    // band outgrow it, which the results then report)
    if (const char* e = std::getenv("ICAM_BATCH_ROOM")) {
      const long r = std::atol(e);
      if (r > 0) caps[i] = std::min<uint32_t>(caps[i], uint32_t(r));
    }
    work[i] = double(caps[i]) * (1.0 + double(parts[i]->pk.L.n_prims));
  }
  B->caller.resize(n);
  for (size_t i = 0; i < n; ++i) B->caller[i] = int(i);
  std::stable_sort(B->caller.begin(), B->caller.end(), [&](int x, int y) { return work[x] > work[y]; });
  B->pos.resize(n);
  for (size_t k = 0; k < n; ++k) B->pos[B->caller[k]] = int(k);
  // offsets
  B->brick_off.assign(n + 1, 0), B->tile_off.assign(n + 1, 0), B->super_off.assign(n + 1, 0), B->cap_off.assign(n + 1, 0);
  for (size_t k = 0; k < n; ++k) {
    const int i = B->caller[k];
    B->parts.push_back(parts[i]);
    const Lv& lv = lvs[i];
    B->brick_off[k + 1] = B->brick_off[k] + long(lv.nb[0]) * lv.nb[1] * lv.nb[2];
    B->tile_off[k + 1] = B->tile_off[k] + tiles_of(lv.nb);
    B->super_off[k + 1] = B->super_off[k] + supers_of(lv.nb);
    B->cap_off[k + 1] = B->cap_off[k] + caps[i];
  }
  B->total_bricks = B->brick_off[n], B->total_tiles = B->tile_off[n], B->total_supers = B->super_off[n], B->total_cap = B->cap_off[n];
  const size_t nbk = size_t(std::max<long>(B->total_bricks, 1)), ntl = size_t(std::max<long>(B->total_tiles, 1));
  const size_t ncap = std::max<size_t>(B->total_cap, 1), nnode = ncap * 729;
  const bool mixed = precision == ICAM_MIXED;
  const size_t icap = nnode / 8 + 1024;
  HIP_OK(hipSetDevice(dev));
  HIP_OK(hipStreamCreate(&B->stream));
  // device memory
  for (auto [buf, bytes] : std::initializer_list<std::pair<Buf*, size_t>>{
           {&B->bparts, n * sizeof(BPart)}, {&B->brick_off_d, (n + 1) * 8}, {&B->tile_off_d, (n + 1) * 8}, {&B->super_off_d, (n + 1) * 8},
           {&B->super_cand, size_t(std::max<long>(B->total_supers, 1)) * SCAP * 4}, {&B->super_count, size_t(std::max<long>(B->total_supers, 1)) * 4},
           {&B->band_off, (n + 1) * 4},
           {&B->btotal, n * 4}, {&B->stotal, n * 4}, {&B->ctl, BC_WORDS * 4}, {&B->counters, n * 64},
           {&B->c_val, nbk * 8}, {&B->c_flag, nbk * 4}, {&B->c_qmark, ntl * 4}, {&B->c_smark, ntl * 4}, {&B->c_queue, ntl * 4},
           {&B->c_squeue, ntl * 4}, {&B->n_val, nnode * 8}, {&B->n_flag, nnode * 4}, {&B->n_qmark, ncap * 4}, {&B->n_smark, ncap * 4},
           {&B->n_queue, ncap * 4}, {&B->n_squeue, ncap * 4}, {&B->items, icap * Item<double>::W * 4}, {&B->tile_cand, ntl * TCAP * 4},
           {&B->tile_count, ntl * 4}, {&B->dc, nbk * 8}, {&B->band, nbk}, {&B->sgn, nbk}, {&B->list, nbk * 4}, {&B->staging, ncap * ICAM_WORDS * 4},
           {&B->cls, ncap * 4}, {&B->slot, ncap * 4}, {&B->map, nbk * 4}, {&B->atlas, ncap * ICAM_WORDS * 4}})
    HIP_OK(buf->need(std::max<size_t>(bytes, 16)));
  if (mixed)
    for (auto [buf, bytes] : std::initializer_list<std::pair<Buf*, size_t>>{{&B->c_code, nbk * 4}, {&B->c_pa, nbk * 4}, {&B->c_pb, nbk * 4},
                                                                             {&B->c_d32, nbk * 4}, {&B->c_info, nbk * 4}, {&B->n_code, nnode * 4},
                                                                             {&B->n_pa, nnode * 4}, {&B->n_pb, nnode * 4}, {&B->n_d32, nnode * 4},
                                                                             {&B->n_info, nnode * 4}})
      HIP_OK(buf->need(std::max<size_t>(bytes, 16)));
  // the parts' records
  uint32_t* ctl = B->ctl.as<uint32_t>();
  for (size_t k = 0; k < n; ++k) {
    const int i = B->caller[k];
    icam_hip* h = parts[i];
    BPart P{};
    const long bo = B->brick_off[k], to = B->tile_off[k], so = B->super_off[k];
    const size_t co = B->cap_off[k];
    P.dv = h->dv, P.lv = lvs[i], P.nb = B->brick_off[k + 1] - bo, P.cap = caps[i];
    P.dc = B->dc.as<double>() + bo, P.band = B->band.as<uint8_t>() + bo, P.sgn = B->sgn.as<int8_t>() + bo, P.list = B->list.as<uint32_t>() + bo;
    P.staging = B->staging.as<uint32_t>() + co * ICAM_WORDS, P.cls = B->cls.as<uint32_t>() + co, P.slot = B->slot.as<uint32_t>() + co;
    P.map = B->map.as<uint32_t>() + bo, P.atlas = B->atlas.as<uint32_t>() + co * ICAM_WORDS;
    P.counters = B->counters.as<unsigned long long>() + 8 * k;
    P.jc.n = P.nb, P.jc.lv = P.lv, P.jc.tile_cand = B->tile_cand.as<uint32_t>() + size_t(to) * TCAP, P.jc.tile_count = B->tile_count.as<uint32_t>() + to;
    P.jc.super_cand = B->super_cand.as<uint32_t>() + size_t(so) * SCAP, P.jc.super_count = B->super_count.as<uint32_t>() + so;
    P.jn = P.jc;
    P.jn.n = long(P.cap) * 729, P.jn.list = P.list;
    Res& rc = P.rc;
    if (mixed)
      rc.code = B->c_code.as<uint32_t>() + bo, rc.pa = B->c_pa.as<float>() + bo, rc.pb = B->c_pb.as<float>() + bo, rc.d32 = B->c_d32.as<float>() + bo,
      rc.info = B->c_info.as<uint32_t>() + bo;
    rc.val = B->c_val.as<double>() + bo, rc.flag = B->c_flag.as<uint32_t>() + bo;
    rc.queue = B->c_queue.as<uint32_t>(), rc.qcount = ctl + BC_QC, rc.qmark = B->c_qmark.as<uint32_t>() + to;
    rc.squeue = B->c_squeue.as<uint32_t>(), rc.scount = ctl + BC_SC, rc.smark = B->c_smark.as<uint32_t>() + to;
    rc.counters = P.counters, rc.part = uint32_t(k);
    Res& rn = P.rn;
    const size_t no = co * 729;
    if (mixed)
      rn.code = B->n_code.as<uint32_t>() + no, rn.pa = B->n_pa.as<float>() + no, rn.pb = B->n_pb.as<float>() + no, rn.d32 = B->n_d32.as<float>() + no,
      rn.info = B->n_info.as<uint32_t>() + no;
    rn.val = B->n_val.as<double>() + no, rn.flag = B->n_flag.as<uint32_t>() + no;
    rn.queue = B->n_queue.as<uint32_t>(), rn.qcount = ctl + BC_QN, rn.qmark = B->n_qmark.as<uint32_t>() + co;
    rn.squeue = B->n_squeue.as<uint32_t>(), rn.scount = ctl + BC_SN, rn.smark = B->n_smark.as<uint32_t>() + co;
    rn.items = B->items.as<uint32_t>(), rn.icount = ctl + BC_ITEMS, rn.icap = uint32_t(icap);
    rn.counters = P.counters, rn.part = uint32_t(k);
    B->host.push_back(P);
  }
  HIP_OK(hipMemcpy(B->bparts.p, B->host.data(), n * sizeof(BPart), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B->brick_off_d.p, B->brick_off.data(), (n + 1) * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B->tile_off_d.p, B->tile_off.data(), (n + 1) * 8, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(B->super_off_d.p, B->super_off.data(), (n + 1) * 8, hipMemcpyHostToDevice));
  BatchArgs& a = B->args;
  // the levels too large for the sign pass in workgroup memory, signed grid-wide
  std::vector<int> wide;
  for (size_t k = 0; k < n; ++k) {
    const Lv& lv = B->host[k].lv;
    if (sign_pitch(lv.nb[0]) * lv.nb[1] * lv.nb[2] <= SIGN_LDS) continue;
    wide.push_back(int(k));
    for (int ax = 0; ax < 3; ++ax) B->wide_lines = std::max(B->wide_lines, long(lv.nb[(ax + 1) % 3]) * lv.nb[(ax + 2) % 3]);
  }
  if (!wide.empty()) {
    HIP_OK(B->wide_d.need(wide.size() * 4));
    HIP_OK(hipMemcpy(B->wide_d.p, wide.data(), wide.size() * 4, hipMemcpyHostToDevice));
  }
  a.wide = B->wide_d.as<int>(), a.nwide = int(wide.size());
  a.parts = B->bparts.as<BPart>(), a.n = int(n), a.brick_off = B->brick_off_d.as<long>(), a.tile_off = B->tile_off_d.as<long>(), a.super_off = B->super_off_d.as<long>();
  a.band_off = B->band_off.as<uint32_t>(), a.btotal = B->btotal.as<uint32_t>(), a.stotal = B->stotal.as<uint32_t>(), a.ctl = ctl;
  // launch shapes: defaults from the device's compute units, then the autotuner
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, dev));
  const int cus = std::max(prop.multiProcessorCount, 1);
  B->gc = 9 * cus, B->gn = 4 * cus, B->gf = 8 * cus;
  // untuned lanes a node as a single handle's: 4 when a part holds more than 65,536
  // sub-patches and curve parts (the large mixed solid alone: 120.9 ms at 1, 100.7 at 4)
  for (const icam_hip* hp : B->parts)
    if (hp->pk.L.n_leaves + hp->pk.L.n_cleaves > 65536) B->kn = 4;
  B->gs = settle_grid(kb_settle<SRC_NODES>, SETTLE_NT, cus);
  if (tune && batch_tune(B.get(), cus)) return 1;
  // the captured graph
  HIP_OK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeRelaxed));
  if (batch_enqueue(B.get(), B->stream)) {
    hipGraph_t g = nullptr;
    hipStreamEndCapture(B->stream, &g);
    if (g) hipGraphDestroy(g);
    return 1;
  }
  HIP_OK(hipStreamEndCapture(B->stream, &B->graph));
  HIP_OK(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  *out = B.release();
  return 0;
}

int icam_hip_batch_run(icam_batch* b, void* stream) {
  if (!b) return fail("no batch");
  HIP_OK(hipGraphLaunch(b->exec, stream ? (hipStream_t)stream : b->stream));
  return 0;
}

int icam_hip_batch_sync(icam_batch* b) {
  if (!b) return fail("no batch");
  HIP_OK(hipStreamSynchronize(b->stream));
  return 0;
}

int icam_hip_batch_result(icam_batch* b, size_t i, const uint32_t** d_map, const uint32_t** d_atlas, icam_bake_stats* stats) {
  if (!b || i >= size_t(b->n)) return fail("bad arguments");
  const int k = b->pos[i];
  const BPart& P = b->host[k];
  // the failures are read on every call, before any result is handed out: a part
  // whose band outgrew its room, a brick kept whole with no sign
  uint32_t over = 0;
  unsigned long long cnt[8];
  HIP_OK(hipMemcpy(&over, b->ctl.as<uint32_t>() + BC_OVER, 4, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(cnt, P.counters, 64, hipMemcpyDeviceToHost));
  if (over) return fail("the band of the batch's part " + std::to_string(b->caller[over - 1]) + " outgrew its room");
  if (cnt[2]) return fail(std::to_string(cnt[2]) + " bricks kept whole with no sign found (no surface near them and none outside the part's box)");
  if (d_map) *d_map = P.map;
  if (d_atlas) *d_atlas = P.atlas;
  if (stats) {
    uint32_t bt = 0, sl = 0;
    HIP_OK(hipMemcpy(&bt, b->btotal.as<uint32_t>() + k, 4, hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(&sl, b->stotal.as<uint32_t>() + k, 4, hipMemcpyDeviceToHost));
    stats->bricks = P.nb, stats->band = bt, stats->slots = sl;
    stats->undecided = int64_t(cnt[0]), stats->refined = int64_t(cnt[1]);
    stats->unsure_trim = int64_t(cnt[3]), stats->unsure_tie = int64_t(cnt[4]), stats->settled = int64_t(cnt[5]);
    stats->overflowed = int64_t(cnt[6]), stats->host = int64_t(cnt[7]);
  }
  return 0;
}

int icam_hip_batch_download(icam_batch* b, size_t i, uint32_t* map, uint32_t* atlas, size_t atlas_cap, icam_bake_stats* stats) {
  if (!b || i >= size_t(b->n)) return fail("bad arguments");
  HIP_OK(hipStreamSynchronize(b->stream));
  icam_bake_stats st{};
  const uint32_t *dm, *da;
  if (icam_hip_batch_result(b, i, &dm, &da, &st)) return 1;
  if (stats) *stats = st;
  if (map) HIP_OK(hipMemcpy(map, dm, size_t(st.bricks) * 4, hipMemcpyDeviceToHost));
  const size_t words = size_t(st.slots) * ICAM_WORDS;
  if (words > atlas_cap) return fail("atlas needs " + std::to_string(words) + " words, the buffer holds " + std::to_string(atlas_cap));
  if (atlas && words) HIP_OK(hipMemcpy(atlas, da, words * 4, hipMemcpyDeviceToHost));
  return 0;
}

int icam_hip_batch_stage_times(icam_batch* b, const char** names, double* ms, int cap) {
  if (!b || !names || !ms) return fail("bad arguments");
  Stages st;
  if (batch_enqueue(b, b->stream, &st)) return 1;
  HIP_OK(hipStreamSynchronize(b->stream));
  std::vector<std::pair<const char*, double>> acc;
  for (size_t i = 0; i < st.tm.size(); ++i) {
    double m = st.tm[i].ms();
    bool found = false;
    for (auto& x : acc)
      if (std::strcmp(x.first, st.names[i]) == 0) x.second += m, found = true;
    if (!found) acc.push_back({st.names[i], m});
  }
  int n = 0;
  for (auto& x : acc) {
    if (n >= cap) break;
    names[n] = x.first, ms[n] = x.second, ++n;
  }
  return n;
}

int icam_hip_batch_tuning(icam_batch* b, char* buf, size_t cap) {
  if (!b || !buf || !cap) return fail("bad arguments");
  std::snprintf(buf, cap, "%s", b->tuning.c_str());
  return 0;
}

void icam_hip_batch_close(icam_batch* b) {
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  delete b;
}

}  // extern "C"

// The error channel for the library's other translation units (cast.hip).
int icam_fail(const std::string& s) { return fail(s); }
