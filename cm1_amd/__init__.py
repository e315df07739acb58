"""cm1_amd: MI355X-native CM1 dynamical-core step (RK3 + 5th-order advection
+ implicit acoustic solver + Kessler + boundary fills) behind a C ABI.

The compute path is ``libcm1gpu.so`` (HIP, gfx950); this package is the
host-side mirror of the reference driver around it.
"""
from ._native import CM1GError, Config, FID, KID, lib  # noqa: F401
from .model import Model, make_config, setup  # noqa: F401
