"""mdacp_amd — MI355X-native short-range Lennard-Jones step of MDACP.

The compute path is libmdx.so (HIP kernels for gfx950 behind the C ABI of
include/mdx.h); :mod:`mdacp_amd.mdx` binds it with ctypes and
:mod:`mdacp_amd.workloads` builds the seeded input states.
"""
from . import workloads  # noqa: F401

__all__ = ["workloads", "mdx"]
