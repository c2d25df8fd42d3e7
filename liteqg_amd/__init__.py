"""MI355X (gfx950) implementation of LiteQG's query path.

The compute path is liteqg_amd/liblqg.so (HIP kernels behind the C ABI of
include/liteqg.h); ``liteqg_amd.symphonyqg.Index`` is the reference's Python
API on top of it.
"""
__version__ = "0.1.0"
