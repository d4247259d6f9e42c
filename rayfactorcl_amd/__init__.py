"""rayfactorcl_amd -- MI355X (gfx950) Monte Carlo view-factor engine.

Drop-in for the OpenCL kernel dispatch of RayFactorCL (RayFactorCL/main.cpp:54-251):
the C ABI in include/rayfactor.h, implemented by hand-written HIP kernels in
``csrc/``.  This package is the Python host mirror of that interface.
"""
from ._lib import (BV_DTYPE, PRIM_DTYPE, CYLINDER, FRUSTUM, SPHERE, ANNULUS, RECTANGLE, DISC, NONE,
                   RayFactorError, lib)
from .solver import (Context, DeviceScene, SolveResult, Stream, areas, autotune, batch_solve, fail_solve, make_opts,
                     normalise, nthreads, run_cl_gpurng, same_tallies, scene_footprint, selftest_math, selftest_sincos, solve, trace,
                     vf_report, write_output)
from .scenes import SceneData, load_xml, parse_xml, read_rfscene

__all__ = [
    "BV_DTYPE", "PRIM_DTYPE", "CYLINDER", "FRUSTUM", "SPHERE", "ANNULUS", "RECTANGLE", "DISC", "NONE",
    "RayFactorError", "lib", "Context", "DeviceScene", "SolveResult", "Stream", "areas", "autotune", "batch_solve",
    "fail_solve", "make_opts", "normalise", "same_tallies", "nthreads", "run_cl_gpurng", "scene_footprint", "selftest_math", "selftest_sincos", "solve", "trace", "vf_report",
    "write_output",
    "SceneData", "load_xml", "parse_xml", "read_rfscene",
]
