"""``import symphonyqg`` drop-in: the reference module's name, this library's Index."""
from liteqg_amd.symphonyqg import UINT_MAX, Index, LqgError  # noqa: F401

__all__ = ["Index", "LqgError", "UINT_MAX"]
