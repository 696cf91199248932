// Record layout of the packed part, shared by the host packer and the kernels.
//
// The part lives in HBM as two parallel record tables, one of float64 and one of
// float32 words with the same offsets (`F`), and one table of uint32 words (`U`)
// for indices and kinds. Each record kind is a contiguous section; a record is a
// fixed number of words, read whole by the lane that needs it (a brick's lanes
// all read the same candidate, so those reads are broadcast). Coordinates are
// relative to the part's centre (`origin`).
//
// A flat list of primitives, each with an axis-aligned box (float32, rounded
// outward) and a code kind << 28 | record, is what the kernels cull against a
// brick or a tile of points. Kinds follow the reference's candidate codes
// (crates/icam-core/src/solid3/gpu.rs:47-51): analytic face, line or arc edge,
// vertex, B-spline sub-patch (leaf of its span's tree), curve-piece leaf; and a
// feature the device does not read (a host feature, the reference's Code::Cpu,
// gpu.rs:55-58, 714-716): a B-spline face past bicubic, a spline edge past cubic, an
// ellipse edge on a B-spline face, or an edge of a host face. A host feature's box
// within a point's answer leaves the point to the caller (flag 3).
#pragma once
#include <stdint.h>

namespace icam {

enum PrimKind : uint32_t { P_FACE = 0, P_EDGE = 1, P_VERT = 2, P_HOST = 3, P_LEAF = 4, P_CLEAF = 5 };
// a host feature's code: P_HOST << 28 | the face, or HOST_EDGE | the edge
constexpr uint32_t HOST_EDGE = 1u << 27;

// FACE: analytic face (F words)
enum : uint32_t {
  FACE_F = 24,   // o 0-2, x 3-5, y 6-8, z 9-11, c0 12, c1 13, sense 14, psi0 15, span 16,
                 // domain lo 17-18, hi 19-20, period 21-22
  FACE_U = 8,    // kind 0, root cell 1, whole 2 (no pieces), narc 3
};
// EDGE: line (a 0-2, b 3-5) or arc (o 0-2, x 3-5, y 6-8, r 9, t0 10, sweep 11, full 12);
// U: kind 0, face0 1, face1 2, uvmap offset 3 (F), first cpiece 4, n 5. Spline and
// ellipse edges are searched through their pieces (an ellipse as its rational quadratic
// arcs of at most a quarter turn, curve3.rs:986-1008). After the part's
// edges, one record of kind EDGE_CREASE per interior C0 knot line of a B-spline face:
// knot 0; U: both faces the line's face, its pieces (4, 5) as a spline edge's, the
// direction held 6 (0: u); each of its pieces holds the span before the line in U word 3
enum : uint32_t { EDGE_F = 16, EDGE_U = 8, EDGE_CREASE = 4 };
// VERT: p 0-2, pn 3-5
enum : uint32_t { VERT_F = 8 };
// LEAF (sub-patch of a span): rho 0, s0 1, s1 2, t0 3, t1 4, slab normal 5-7, slab lo 8, hi 9
// (the sub-patch lies between the planes n.x = lo and n.x = hi); from LEAF_CERT its bounds
// for the per-point convexity test: lambda_min(G) from below, then the intervals (x lo, x hi,
// y lo, y hi, z lo, z hi) of S_s, S_t, S_ss, S_st, S_tt in the parameters scaled by the
// middle's tangent lengths; from LEAF_POLE: a pole the sub-patch touches (x, y, z, then the
// collapsed side in its span's parameters: 1 s = 0, 2 s = 1, 3 t = 0, 4 t = 1; all zero
// when none); U: span 0, face 1, known inside 2
enum : uint32_t { LEAF_CERT = 12, LEAF_POLE = 44, LEAF_F = 48, LEAF_U = 4 };
// SPAN: 16 homogeneous points (i along u, 4 words each) 0-63, u0 64, u1 65, v0 66, v1 67;
// U: face 0, trimmed 1 (0 whole surface), root cell 2
enum : uint32_t { SPAN_F = 72, SPAN_U = 4 };
// NURBS face: sense 0; U: first span 0, spans along u 1, along v 2, root cell 3, whole 4
// (it sits in the FACE section too, kind 5)
// CLEAF (leaf of a curve piece's tree): rho 0, t0 1, t1 2; from CLEAF_CERT: |C'|^2 from
// below, then the intervals of C' and C''; U: cpiece 0, edge 1
enum : uint32_t { CLEAF_CERT = 4, CLEAF_F = 17, CLEAF_U = 2 };
// CPIECE: 4 homogeneous points 0-15; U: edge 0, piece number in its edge 1, pieces 2
enum : uint32_t { CPIECE_F = 16, CPIECE_U = 4 };
// CELL (trim cell tree): lo 0-1, hi 2-3, ref 4-5; U: inside 0, kind 1 (0 leaf, 1 crossed,
// 2 split), a 2, b 3 (crossed: first in `among`, count; split: first of 4 children)
enum : uint32_t { CELL_F = 6, CELL_U = 4 };
// INST (shifted copy of a trim piece): shift u 0, shift v 1; U: piece 0
enum : uint32_t { INST_F = 2, INST_U = 1 };
// TPIECE: 4 control points (u, v) 0-7
enum : uint32_t { TPIECE_F = 8 };

struct Layout {
  uint32_t n_prims;
  uint32_t n_faces, n_edges, n_verts, n_spans, n_leaves, n_cleaves, n_cpieces;
  // section offsets in F and U
  uint32_t face_f, face_u, edge_f, edge_u, vert_f;
  uint32_t leaf_f, leaf_u, span_f, span_u, cleaf_f, cleaf_u, cpiece_f, cpiece_u;
  uint32_t cell_f, cell_u, inst_f, inst_u, tpiece_f, among_u, uvmap_f;
  // part extent (relative), the model tolerance, rounding allowances
  float lo[3], hi[3];
  float tol;
  float eps32;   // rounding allowance on single-precision lengths
  double eps64;  // on double-precision lengths
  double diag;   // the part's box diagonal
  uint32_t max_rounds;  // a search's best-first rounds over parts, at most
  // The primitives' hierarchy: an implicit 64-ary tree over the primitives sorted by the
  // Morton code of their box centres. Level 0 holds the primitives' boxes in that order
  // (bvh_prim[k] the k-th one's index); node j of level l + 1 bounds nodes 64 j .. 64 j + 63
  // of level l. Level l's boxes (6 floats, lo then hi) and flags (a boundary point inside)
  // start at node bvh_off[l] of the hierarchy's arrays; bvh_cnt[l] nodes; level bvh_top
  // holds at most 64.
  uint32_t bvh_top;
  uint32_t bvh_off[8], bvh_cnt[8];
};

}  // namespace icam
