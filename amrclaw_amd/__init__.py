"""amrclaw_amd -- MI355X-native per-level patch update for AMRClaw-style AMR.

The compute path is libamrgpu.so (HIP kernels for gfx950 behind the C ABI in
include/amrgpu.h); `capi` is its ctypes binding.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
