"""Sailfish split-explicit ocean hot path on AMD Instinct MI355X (gfx950).

Python host code mirrors the reference's drop-in interface (step2dPredictor /
step2dCorrector / rhs3d / omega / ... and the mod_ocean / mod_grid layouts) and calls
hand-written HIP kernels in ``libsfhip.so`` through the C ABI of
``include/sailfish_hip.h``.
"""
__version__ = "0.1.0"
