"""simstack_amd — MI355X-native batched Rosenbrock stiff-chemistry solver.

Host side (mechanisms, zone blocks, synthetic workloads) lives in
libsimk_host.so; the solver itself runs only in libsimk_gpu.so (hand-written
HIP for gfx950). There is no CPU fallback: gpu.GpuModule raises when the
module or a device is missing.
"""
from . import _ffi, problem  # noqa: F401
