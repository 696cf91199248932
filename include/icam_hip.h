/*
 * icam_hip.h -- the C ABI of libicam_hip.so: the exact signed distance of a
 * B-rep part and its sparse narrow-band brick baker on an MI355X (gfx950).
 *
 * It replaces the reference's GPU runner `PartGpu` (crates/icam-sim/src/partgpu.rs,
 * binary-private) and the device half of its brick cache `PartCache`
 * (crates/icam-sim/src/partcache.rs). A Rust host binds it through a thin
 * `extern "C"` crate (see the binding in the repository's integration notes).
 *
 * Conventions
 *  - Return 0 on success; otherwise a nonzero code, and icam_hip_last_error()
 *    (thread-local) says what failed, as `PartGpu` returns Err(String).
 *  - The caller owns every host array for the duration of the call; the
 *    library owns device memory from icam_hip_open until icam_hip_close.
 *  - A handle may be used from several threads: calls on one handle are
 *    serialized by its own lock (PartCache builds two levels at once through
 *    one runner, crates/icam-sim/src/play.rs:979-997).
 *  - World coordinates in float64; the tables are packed relative to the part's
 *    centre, in float32 and float64 copies.
 *  - precision: ICAM_FP64 runs every step in double precision; ICAM_MIXED runs
 *    the search in single precision and the last correction of the nearest
 *    feature, the trim test near a trim curve and near-ties in double.
 *
 * Output flags per node or point (the reference kernel's, gpu.wgsl:1913-1922):
 *   0  the value is the exact signed distance;
 *   1  only a lower bound past the record's bound is known (the value is that
 *      bound); the sign is not known;
 *   2  undecided (a convexity certificate was missing within the bound);
 *   3  left to the caller: a host feature (one the device does not read: a B-spline
 *      face past bicubic, a spline edge past cubic, an ellipse edge on a B-spline
 *      face, an edge of such a face) came within the answer, so the point is read
 *      on the CPU, as a brick holding the reference's Code::Cpu is
 *      (crates/icam-core/src/solid3/gpu.rs:714-716); the value is a lower bound.
 */
#ifndef ICAM_HIP_H
#define ICAM_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "icam_part.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct icam_hip icam_hip;

enum icam_precision { ICAM_FP64 = 0, ICAM_MIXED = 1 };

/* PartGpu::new(&Exact, need) (partgpu.rs:132-186) with Exact::gpu_pack(origin,
 * need) (crates/icam-core/src/solid3/gpu.rs:259-701): packs the part relative to
 * its centre, builds the trim cell trees, the B-spline sub-patch trees and the
 * curve-piece trees, certified for points within `need` of each part (mm), and
 * uploads them. Fails when no device is present. Features past the degrees the
 * device reads are packed as host features (flag 3 above; ICAM_HOST bricks). */
int icam_hip_open(const icam_part_desc* part, double need, icam_hip** out);

/* Drop(PartGpu): frees the device tables. */
void icam_hip_close(icam_hip* h);

/* The last error on this thread (empty when none). */
const char* icam_hip_last_error(void);

/* PartGpu.adapter (partgpu.rs:84): the device's name. */
int icam_hip_device_name(icam_hip* h, char* buf, size_t cap);

/* PartGpu::warm(secs) (partgpu.rs:188-225): keeps the device busy for about
 * `secs` seconds so its clocks are up before timing. */
int icam_hip_warm(icam_hip* h, double secs);

/* PartGpu.nodes (partgpu.rs:86): sub-patches and curve pieces packed. */
size_t icam_hip_nodes(icam_hip* h);

/* A point record (PartGpu::points, partgpu.rs:424-434; the reference's Brick in
 * point mode, gpu.wgsl:11-18): the point, and the distance past which a lower
 * bound will do (infinity: always exact). */
typedef struct icam_point_rec {
  double p[3];
  double clamp;
} icam_point_rec;

/* PartGpu::points / points_refined (partgpu.rs:424-434): out[2 i] the signed
 * distance (or bound), out[2 i + 1] the flag. */
int icam_hip_points(icam_hip* h, int precision, const icam_point_rec* recs, size_t n, double* out);

/* A brick record (PartGpu::bricks, partgpu.rs:227-233): the first node and the
 * cell side; its 729 nodes are first + h (n % 9, n / 9 % 9, n / 81). */
typedef struct icam_brick_rec {
  double first[3];
  double h;
} icam_brick_rec;

/* PartGpu::bricks (partgpu.rs:227-233): the 729 nodes of each brick, exact within
 * the clamp h sqrt 3 (flag 0), else flagged 1 with the clamp as value. out holds
 * 729 (value, flag) pairs per brick. */
int icam_hip_bricks(icam_hip* h, int precision, const icam_brick_rec* recs, size_t n, double* out);

/* Signed distances at n world points (unbounded), flags as above, and the nearest
 * feature as kind << 28 | index (0 face, 1 edge, 2 vertex) when feat is not NULL.
 * Host arrays. (Exact::sdf, solid3.rs:1750.) */
int icam_hip_query(icam_hip* h, int precision, const double* xyz, size_t n, double* d, uint32_t* flags, uint32_t* feat);

/* The same on device arrays already resident in HBM, enqueued on `stream`
 * (a hipStream_t; NULL: the default stream); returns without waiting. */
int icam_hip_query_device(icam_hip* h, int precision, const double* d_xyz, size_t n, double* d_d, uint32_t* d_flags, uint32_t* d_feat, void* stream);

/* No counterpart in the reference: icam_hip_query_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). It allocates nothing: its scratch (cell order, single-precision
 * records, queues, counters) is carved from one static device array of 32 MiB a GPU,
 * held until the process ends, so a call takes at most 2^18 points (more fail). The
 * arguments and results are icam_hip_query_device's; it returns without waiting. Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (from different
 * handles, or on streams not ordered with each other) race. Calls on one handle are
 * ordered after each other like the handle's other calls. */
int icam_hip_query_device_noalloc(icam_hip* h, int precision, const double* d_xyz, size_t n, double* d_d, uint32_t* d_flags, uint32_t* d_feat, void* stream);

/* Statistics of a bake (or of a query: icam_hip_query_stats). */
typedef struct icam_bake_stats {
  int64_t bricks;      /* bricks in the level */
  int64_t band;        /* bricks whose nodes were read (the decision's) */
  int64_t slots;       /* bricks stored with words */
  int64_t undecided;   /* nodes left undecided (bridged) */
  int64_t refined;     /* points searched again in double precision (ICAM_MIXED) */
  int64_t unsure_trim; /* of them, for a trim decision taken near a trim curve */
  int64_t unsure_tie;  /* of them, for a near-tie of opposite signs */
  int64_t settled;     /* points past a convexity certificate settled by the certified
                          search (left undecided, then bridged, when it does not finish) */
  int64_t overflowed;  /* search blocks whose candidates did not fit their list (those
                          scan every primitive each round: slow, not wrong) */
  int64_t host;        /* bricks left to the caller (ICAM_HOST map entries: a node or the
                          centre flagged 3), read on the CPU as partcache.rs:794-797 does */
} icam_bake_stats;

/* Statistics of the last query on this handle (waits for the device): the points
 * searched again in double precision (refined, unsure_trim, unsure_tie) and the points
 * settled by the certified search; the other fields are zero. */
int icam_hip_query_stats(icam_hip* h, icam_bake_stats* stats);

/* PartCache::build (partcache.rs:458-486) on the device: the level's brick map
 * (nb[0] nb[1] nb[2] entries, x fastest: ICAM_AIR, ICAM_SOLID or a slot) and its
 * atlas (ICAM_WORDS words a slot, slots in map order), bit-compatible with
 * crates/icam-sweep/src/cache.rs:74-101. Host arrays; atlas_cap in words.
 * Returns nonzero when the atlas does not fit (stats->slots says how many).
 * A part with host features may leave bricks as ICAM_HOST (stats->host of them,
 * no slot): the caller reads those on the CPU, as PartCache does the bricks whose
 * candidates hold Code::Cpu (partcache.rs:794-797). */
int icam_hip_bake(icam_hip* h, int precision, const icam_level* lv, uint32_t* map, uint32_t* atlas, size_t atlas_cap, icam_bake_stats* stats);

/* The same with the map and atlas left in device memory the library owns until
 * the next bake or close; *d_map and *d_atlas receive their device addresses.
 * Enqueued on `stream`; the statistics need the stream synchronized first, so
 * they are written when stats is not NULL (which waits). */
int icam_hip_bake_device(icam_hip* h, int precision, const icam_level* lv, uint32_t** d_map, uint32_t** d_atlas, icam_bake_stats* stats, void* stream);

/* Batches: many parts' levels baked by one launch of a captured hipGraph. The
 * reference bakes one level a call (PartCache::build, partcache.rs:458-486; the
 * viewer runs two at once, play.rs:979-997); these entry points are additive. Each
 * stage of the bake is one launch over every part's work, the block stages being
 * persistent work queues that serve the parts heaviest first.
 *
 * icam_hip_batch_open: part i (an open handle, kept open while the batch lives) at
 * level i, n <= 255, on one device; levels of at most 32 x 32 x 32 bricks (more
 * along one axis if fewer along the others: the sign pass keeps a level in LDS).
 * With `tune` nonzero the autotuner times the batch with candidate persistent grid
 * sizes of its queue stages and keeps the fastest (icam_hip_batch_tuning reports
 * what it tried). Device memory for every part's band is sized from its primitives'
 * boxes, once. */
typedef struct icam_batch icam_batch;
int icam_hip_batch_open(icam_hip* const* parts, const icam_level* levels, size_t n, int precision, int tune, icam_batch** out);

/* One run: the graph launched on `stream` (NULL: the batch's own); no waiting. */
int icam_hip_batch_run(icam_batch* b, void* stream);

/* Waits for the batch's own stream. */
int icam_hip_batch_sync(icam_batch* b);

/* After the run has finished: part i's map and atlas in device memory (valid until
 * the next run or close) and its statistics (stats may be NULL). Fails when a
 * part's band outgrew its room or a brick kept whole found no sign. */
int icam_hip_batch_result(icam_batch* b, size_t i, const uint32_t** d_map, const uint32_t** d_atlas, icam_bake_stats* stats);

/* The same copied to host arrays (waits for the batch's own stream first). */
int icam_hip_batch_download(icam_batch* b, size_t i, uint32_t* map, uint32_t* atlas, size_t atlas_cap, icam_bake_stats* stats);

/* One run outside the graph with HIP events between its launches (waits): the
 * device time (ms) of each stage, names[i] (static strings) and ms[i] for i below the
 * returned count (at most cap). */
int icam_hip_batch_stage_times(icam_batch* b, const char** names, double* ms, int cap);

/* The autotuner's report: the candidates it timed and its choice. */
int icam_hip_batch_tuning(icam_batch* b, char* buf, size_t cap);

void icam_hip_batch_close(icam_batch* b);

/* The autotuner of the single-level bake (no counterpart in the reference, which
 * bakes with one fixed kernel shape): the level is baked with a brick's nodes searched
 * 1, 2, 4 and 8 lanes a node (a wave's 64 lanes on 64, 32, 16 or 8 nodes), each timed
 * twice after a warm-up on the handle's stream; the fastest is kept for later bakes in
 * `precision` on this handle. report (cap bytes, may be NULL) receives each
 * candidate's time and the choice. The level's results are left as a bake's. */
int icam_hip_tune(icam_hip* h, int precision, const icam_level* level, char* report, size_t cap);

/* The viewer's ray cast over baked brick levels (crates/icam-sim/src/cache.wgsl:441-503
 * trace_ray, :335-420 walk, :270-318 march, :191-268 cell_hit), the consumer of the
 * brick cache: n rays (origin xyz, unit direction xyz: 6 doubles a ray, device memory)
 * against `count` levels (1 to 16, finest first), each with the device map and atlas a
 * bake left (icam_hip_bake_device, icam_hip_batch_result). For each ray, d_out gets 4
 * floats (t along the ray, -1 when it meets nothing; the gradient of the trilinear
 * interpolant of the level's nodes where it meets it, or -direction at a solid brick)
 * and d_hit 2 words (the level met, ~0 for none; the brick's slot, or ICAM_SOLID, or
 * ICAM_HOST for a brick left to the caller, met where the ray enters it). Single
 * precision, as the shader reads the cache. Enqueued on `stream` (NULL: the default
 * stream); needs no handle. */
int icam_hip_cast(const icam_level* levels, const uint32_t* const* d_maps, const uint32_t* const* d_atlases, int count, const double* d_rays, size_t n,
                  float* d_out, uint32_t* d_hit, void* stream);

/* Average device time (ms) of each kernel stage of the last bake or query on
 * this handle, measured with HIP events on the stream the kernels ran on:
 * names[i] (static strings) and ms[i] for i < returned count (at most cap). */
int icam_hip_stage_times(icam_hip* h, const char** names, double* ms, int cap);

#ifdef __cplusplus
}
#endif
#endif
