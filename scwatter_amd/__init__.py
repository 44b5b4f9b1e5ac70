"""scwatter_amd — MI355X-native plane-wave field evaluator for scwatter.

E(r) = sum_w E0_w exp(i k_w . r) over the waves of the layer region holding r,
evaluated by hand-written HIP kernels for gfx950 (direct form and separable
MFMA GEMM form) behind the C ABI in include/scwatter_amd.h.
"""
from ._native import DTYPES, FORMS, ScwError, lib  # noqa: F401
from .cw import CoupledWaveStructure, HomogeneousLayer, WaveTables, load_tables  # noqa: F401
from .evaluator import (BatchExec, Exec, Plan, axis_cpu, axis_gpu, cpu_cw_evaluate, eval_planes,  # noqa: F401
                        eval_planes_noalloc, gpu_cw_evaluate)

__all__ = ["CoupledWaveStructure", "HomogeneousLayer", "WaveTables", "load_tables", "Plan", "Exec", "BatchExec",
           "gpu_cw_evaluate", "eval_planes", "eval_planes_noalloc", "cpu_cw_evaluate", "axis_cpu", "axis_gpu", "ScwError", "FORMS", "DTYPES", "lib"]
