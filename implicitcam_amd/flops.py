"""Floating-point operations per primitive of the device search (csrc/sdf_device.hpp),
counted by hand from its straight-line code, for the useful-work rate of a run:

    useful FLOP = sum over primitives of (executed count x operations per primitive)

where the executed counts are those of the double-precision evaluator restated on the
CPU over the same inputs (its work counters), so the rate measures the answers produced,
not the operations this implementation happens to spend on them.

Counting rule: an add, subtract or multiply is 1, a fused multiply-add written as
`a * b + c` is 2, a divide, square root, floor or transcendental (atan2, sin, cos, tan)
is 1; comparisons, min, max, abs and clamps are not counted. Constants computed once
per record (a cone's sin, cos and tan of its half angle are evaluated per call, so they
are counted) follow the code as written. The same code runs in single and double
precision, so one table serves both.

Per primitive, where it is counted from:

- patch_eval 709 (`peval`): two cubic Bernstein bases with first and second
  derivatives (2 x 29); the 16 homogeneous points summed along v with three weights
  (4 x 4 x 4 x 3 multiply-adds, 384); the rows along u into six sums (4 x 4 x 6
  multiply-adds, 192); the rational quotient rules for the point and its five
  derivatives (1 + 3 + 9 + 9 + 16 + 21 + 16 = 75).
- surface_newton_step 62 (`newton2`, its evaluation excluded): residual 3, gradient
  10, Gauss-Newton matrix 15, Hessian 18, determinant 3, the two step components 8,
  the clamped update 2, the moved length 3.
- curve_eval 142 (`ceval`): one cubic basis with derivatives 29; 12 homogeneous sums of
  four terms (12 x 7); quotient rules 1 + 3 + 9 + 16.
- curve_newton_step 23 (`newton1`, evaluation excluded): residual 3, first derivative 5,
  second 11, update 2, moved 2.
- foot_plane 14, foot_cylinder 46, foot_cone 57, foot_sphere 16, foot_torus 64
  (`foot`; `axial` is 34: offset 3, three dots 15, radius 4, atan2 1, radial
  direction 11).
- uv_plane 13, uv_cylinder 55, uv_cone 65, uv_sphere 80, uv_torus 81 (`face_uv`;
  `arc_param_over` is 21 over a full turn and 24 over a partial sweep).
- trim_crossing_step 37 (`crossing`): the cubic coordinate 19 and its offset 1, the
  derivative's quadratic 13 and its factor 1, the Newton update 2, the moved length 1.
- edge_line 32, edge_arc 49 (`try_edge`: the foot and its distance 9).
- vertex 9 (`try_vert`: the offset and its length).

A C0 knot line's or a pole's search is counted through the primitives it runs (curve
evaluations and Newton steps; a pole as a vertex).
"""

FLOPS = {
    "patch_eval": 709,
    "surface_newton_step": 62,
    "curve_eval": 142,
    "curve_newton_step": 23,
    "foot_plane": 14,
    "foot_cylinder": 46,
    "foot_cone": 57,
    "foot_sphere": 16,
    "foot_torus": 64,
    "uv_plane": 13,
    "uv_cylinder": 55,
    "uv_cone": 65,
    "uv_sphere": 80,
    "uv_torus": 81,
    "trim_crossing_step": 37,
    "edge_line": 32,
    "edge_arc": 49,
    "vertex": 9,
}


def useful_flops(counts):
    """Useful floating-point operations of a run from its executed primitive counts
    ({name: count}, the names of FLOPS); unknown names are an error."""
    unknown = set(counts) - set(FLOPS)
    if unknown:
        raise KeyError(f"no operation count for {sorted(unknown)}")
    return sum(FLOPS[k] * int(v) for k, v in counts.items())
