/*
 * icam_part.h -- the part a signed-distance evaluator is built on: a B-rep solid
 * whose faces lie on exact surfaces and are bounded by certified trim loops in
 * their surfaces' parameter domains, with its edges and vertices.
 *
 * This is the data model the evaluator consumes (the reference's `Solid3` after
 * import: crates/icam-core/src/solid3.rs:44-278, with `Face3.trim2` loops from
 * crates/icam-core/src/trim2.rs:51-90 and `Face3.analytic2` parameter forms from
 * trim2.rs:1680-1934). It is plain C: every array is caller-owned, read during
 * the call that takes the part and never kept.
 *
 * Conventions (all FP64, millimetres, world coordinates):
 *  - A frame is (o, x, z); its y axis is z cross x.
 *  - Analytic surfaces use the STEP parameterizations (surface3.rs:1-12):
 *      plane     o + u x + v y
 *      cylinder  o + r e(u) + v z
 *      cone      o + (r + v tan a) e(u) + v z            (a = semi-angle)
 *      sphere    o + r cos v e(u) + r sin v z
 *      torus     o + (R + r cos v) e(u) + r sin v z
 *    and the surface's own normal points away from the axis / centre / tube
 *    centre (plane: z).  `sense` is +1 when the solid's outward normal is the
 *    surface's own normal, -1 otherwise.
 *  - Trim loops live in the face's parameter domain. For an analytic face those
 *    are the parameters of its exact rational form (trim2.rs:1680-1710): plane
 *    (x, y) in its frame; cylinder and cone (quarter-arc parameter in [0, 4),
 *    height); sphere (longitude arc parameter, latitude arc parameter in [0, 2]);
 *    torus (arc parameter about the axis, arc parameter about the tube). For a
 *    NURBS face they are its knot parameters.
 *  - A trim piece is a polynomial cubic in the domain, 4 control points
 *    (u0 v0 u1 v1 u2 v2 u3 v3), monotone in both parameters (trim2.rs:51-58).
 *    A face with no pieces is its whole domain.
 *  - `ref` is a domain point whose answer (inside the face or not) is known; the
 *    loops' answer everywhere else follows from crossing parity against it
 *    (it stands in for the offset reference of trim2.rs:1346-1420).
 */
#ifndef ICAM_PART_H
#define ICAM_PART_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum icam_surf_kind {
  ICAM_PLANE = 0,
  ICAM_CYLINDER = 1,
  ICAM_CONE = 2,
  ICAM_SPHERE = 3,
  ICAM_TORUS = 4,
  ICAM_NURBS = 5
};

enum icam_edge_kind { ICAM_LINE = 0, ICAM_ARC = 1, ICAM_SPLINE = 2, ICAM_ELLIPSE = 3 };

typedef struct icam_face_desc {
  uint32_t kind;      /* icam_surf_kind */
  double sense;       /* +1 / -1 */
  double o[3], x[3], z[3]; /* analytic frame */
  double c[2];        /* cylinder, sphere: r; cone: r, a; torus: major, minor */
  /* NURBS: degrees, control counts, homogeneous net (x w, y w, z w, w) with
   * point (i, j) at i * nv + j (i along u), clamped knot vectors. */
  uint32_t du, dv, nu, nv;
  const double* pw;
  const double* ku; /* nu + du + 1 */
  const double* kv; /* nv + dv + 1 */
  /* trim loops */
  double lo[2], hi[2];  /* domain */
  double period[2];     /* 0: not periodic */
  uint32_t n_pieces;
  const double* pieces; /* n_pieces * 8 */
  double ref[2];
  uint32_t ref_inside;
  /* vertex loops: points of the face bounded by a lone vertex (a cone's apex),
   * n_points * 3; they lie in the face's box (Face3.points, solid3.rs:1226-1231).
   * The vertex itself is one of the part's vertices. */
  uint32_t n_points;
  const double* points;
} icam_face_desc;

typedef struct icam_edge_desc {
  uint32_t kind;       /* icam_edge_kind */
  double a[3], b[3];   /* line from a to b */
  double o[3], x[3], z[3]; /* arc frame */
  double r, t0, sweep; /* arc: o + r (cos t x + sin t y), t from t0 through sweep */
  /* spline: rational Bezier pieces of one degree, in traversal order,
   * homogeneous control points, n_pieces * (degree + 1) * 4 */
  uint32_t degree, n_pieces;
  const double* pw;
  /* The faces using the edge (two in a closed shell). */
  uint32_t faces[2];
  /* For each piece and each of the two faces, where the piece runs in that
   * face's parameters when the face is a NURBS face: (u, v) at the piece's
   * start and end, n_pieces * 2 * 4 values (a seed for the face normal at an
   * edge point; the point itself is on the surface). NULL when no face is NURBS.
   * Line and arc edges count as one piece. */
  const double* uvmap;
  /* ellipse (Edge3::Ellipse, curve3.rs:494): o + r cos t x + r2 sin t y in the arc
   * frame, t from t0 through sweep (r along x, r2 along y = z x x) */
  double r2;
} icam_edge_desc;

typedef struct icam_vert_desc {
  double p[3];
  double pn[3]; /* angle-weighted pseudo-normal (need not be unit) */
} icam_vert_desc;

typedef struct icam_part_desc {
  uint32_t n_faces, n_edges, n_verts;
  const icam_face_desc* faces;
  const icam_edge_desc* edges;
  const icam_vert_desc* verts;
  double lo[3], hi[3]; /* bounding box of the solid */
  double tol;          /* model tolerance */
} icam_part_desc;

/* A level of the brick cache (crates/icam-sweep/src/cache.rs:98-101): bricks of
 * 8 x 8 x 8 cells of side h, 9 x 9 x 9 nodes, the first node of brick (i, j, k)
 * at lo + 8 h (i, j, k); nb bricks along each axis, x fastest in the map. */
typedef struct icam_level {
  double lo[3];
  double h;
  int32_t nb[3];
} icam_level;

/* Brick map entries (cache.rs:84-85) and words per brick (cache.rs:92). */
#define ICAM_AIR 0xffffffffu
#define ICAM_SOLID 0xfffffffeu
/* Not a reference map value: a brick the device left to the caller (a host feature
 * within its reach); the caller reads it on the CPU and stores it as the reference's
 * PartCache does a brick whose candidates hold Code::Cpu (partcache.rs:794-797). */
#define ICAM_HOST 0xfffffffdu
#define ICAM_WORDS 365
#define ICAM_NODES 729

#ifdef __cplusplus
}
#endif
#endif
