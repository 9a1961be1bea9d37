"""us_lab4a_amd — MI355X-native STA/TFM delay-and-sum imaging (mista).

The compute path is libmista_hip.so (hand-written gfx950 HIP kernels behind the
C ABI in include/mista.h).  This package holds the Python side of that
boundary: the reference-shaped processor (sta.HipSTAProcessor), host helpers
(imaging), seeded synthetic scenes (synth, workloads) and the serving loop.
"""
from .imaging import XYZ_VALUE_FACTOR, ArrayAcquisition, STAAcquisition, rectangular_grid, window  # noqa: F401
from .sta import CoherenceFactor, HipSTAProcessor, STAConfiguration  # noqa: F401

__all__ = ["HipSTAProcessor", "STAConfiguration", "CoherenceFactor", "ArrayAcquisition", "STAAcquisition",
           "rectangular_grid", "window", "XYZ_VALUE_FACTOR"]
