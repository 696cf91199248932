/*
 * oracle.h -- FP64 CPU oracle of the exact signed distance and its brick baker.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the HIP library is compared
 * against, and the timed CPU baseline. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links or calls it.
 *
 * It restates the reference's double-precision evaluator `solid3::Exact`
 * (crates/icam-core/src/solid3.rs:1528-1940) with its geometry
 * (surface3.rs, curve3.rs, trim2.rs) and the CPU rules of the brick cache
 * `PartCache` (crates/icam-sim/src/partcache.rs:226-331, 962-1058). Each
 * function in oracle.cpp cites the reference lines it follows.
 */
#ifndef ICAM_ORACLE_H
#define ICAM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/icam_part.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_exact orc_exact;

/* Builds the evaluator over a part (Exact::new, solid3.rs:1586-1652); NULL when a
 * surface or spline edge is past degree 7. */
orc_exact* orc_open(const icam_part_desc* part);
void orc_close(orc_exact*);

/* Signed distance (negative inside) at n points, each searched within bound[i]
 * (NULL: unbounded; Exact::sdf_within, solid3.rs:1750-1816). feat[i] is the
 * nearest feature as kind << 28 | index (0 face, 1 edge, 2 vertex); flags[i] bit 0
 * is `unsure` (a capped search), bit 1 a cusp edge. threads <= 0: all cores. */
void orc_sdf(orc_exact*, const double* xyz, size_t n, const double* bound, double* d, uint32_t* feat, uint32_t* flags, int threads);

/* The level's brick map (nb[0] nb[1] nb[2] entries, x fastest: AIR, SOLID or a
 * slot) and its atlas (ICAM_WORDS words a slot, slots in map order), as
 * PartCache::build makes them on the CPU (partcache.rs:458-486, decide :226,
 * exact_brick :979). Returns the slot count, or -1 when atlas_cap_words is too
 * small (the map is still written). bricks_read (may be NULL) gets how many bricks
 * the decision left to read. */
long orc_bake(orc_exact*, const icam_level* lv, uint32_t* map, uint32_t* atlas, size_t atlas_cap_words, long* bricks_read, int threads);

/* One brick's 729 node values (i16, x fastest) read exactly on the CPU, and its
 * class: 0 words, 1 AIR, 2 SOLID (exact_brick + packed, partcache.rs:962-1058). */
int orc_brick(orc_exact*, const double first_node[3], double h, int16_t* q);

/* The bricks the level's decision leaves to read (partcache.rs:226-331): writes
 * their indices (map order) to out (capacity cap) and returns how many. */
long orc_decide(orc_exact*, const icam_level* lv, int64_t* out, size_t cap, int threads);

/* The primitives executed by every call since the last reset, summed over threads
 * (out[k], k < cap), optionally reset; returns how many counts there are (ORC_COUNTS).
 * The useful work of a run is these counts times the device code's operation count
 * per primitive. */
enum {
  ORC_COUNT_PEVAL = 0,   /* rational tensor Bezier point, 1st and 2nd derivatives */
  ORC_COUNT_NEWTON2 = 1, /* surface Newton step (its evaluation excluded) */
  ORC_COUNT_CEVAL = 2,   /* rational Bezier curve point, 1st and 2nd derivatives */
  ORC_COUNT_NEWTON1 = 3, /* curve Newton step (its evaluation excluded) */
  ORC_COUNT_FOOT = 4,    /* analytic foot: plane, cylinder, cone, sphere, torus (4..8) */
  ORC_COUNT_UV = 9,      /* analytic parameters, same kinds (9..13) */
  ORC_COUNT_CROSS = 14,  /* trim crossing Newton iteration */
  ORC_COUNT_EDGE_LINE = 15,
  ORC_COUNT_EDGE_ARC = 16,
  ORC_COUNT_VERT = 17,   /* vertex, pole or tip distance */
  ORC_COUNTS = 18
};
int orc_counts(uint64_t* out, int cap, int reset);

/* Pieces of the evaluator, for tests. */
int orc_face_inside(orc_exact*, uint32_t face, double u, double v);
void orc_face_uv(orc_exact*, uint32_t face, const double q[3], double uv[2]);
void orc_face_foot(orc_exact*, uint32_t face, const double p[3], double q[3]);
/* The closest point of a NURBS face's untrimmed surface, searched within cutoff:
 * returns 0 when none is found, else writes distance and the point. */
int orc_nurbs_closest(orc_exact*, uint32_t face, const double p[3], double cutoff, double* dist, double q[3]);
/* NURBS face point at (u, v). */
void orc_nurbs_eval(orc_exact*, uint32_t face, double u, double v, double q[3]);
/* Edge closest point: returns its distance, the point, and 1 when interior. */
int orc_edge_closest(orc_exact*, uint32_t edge, const double p[3], double* dist, double q[3]);

#ifdef __cplusplus
}
#endif
#endif
