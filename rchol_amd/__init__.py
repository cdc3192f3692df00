"""rchol_amd: MI355X-native solve engine for rchol's preconditioned CG.

Reference interfaces mirrored:
  pcg(A, b, epsilon, maxitr, G, Gt) -> (x, relres, niters)   python/rchol/util.py:32-61
  Solver: device-resident A/G (the residency layer behind c++/util/pcg.cpp)
"""
from .solver import (FP32, FP64, Solver, analyze, device_count, pcg, stack_systems,
                     stored_from_lower)

__all__ = ["FP32", "FP64", "Solver", "analyze", "device_count", "pcg", "stack_systems",
           "stored_from_lower"]
