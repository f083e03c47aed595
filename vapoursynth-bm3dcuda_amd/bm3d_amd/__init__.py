"""bm3d_amd: MI355X-native BM3D / V-BM3D (host side over libbm3d_mi355x.so)."""
import importlib

from . import _capi, synth  # noqa: F401
from ._capi import BM3DError  # noqa: F401

_LAZY = {"core", "Plan", "aggregate", "vaggregate", "unpack_groups"}


def __getattr__(name):
    # torch-dependent parts load on first use, so that the CPU-only test
    # suite can import the package without initialising a device.
    if name in _LAZY:
        core = importlib.import_module(".core", __name__)
        return core if name == "core" else getattr(core, name)
    raise AttributeError(name)
