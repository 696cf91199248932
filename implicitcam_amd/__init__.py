"""Exact signed distance of B-rep parts and its narrow-band brick baker on MI355X.

A from-scratch MI355X implementation of ImplicitCAM's `solid3::Exact` evaluator and
`PartCache` brick baker: HIP kernels for gfx950 behind the C ABI of include/icam_hip.h.
"""
from .exact import FP64, MIXED, Batch, PartGpu, cast_device  # noqa: F401
from .part import Part, c4_grid, c4_level, level_for, near_surface_points, near_trim_points, uniform_points  # noqa: F401
