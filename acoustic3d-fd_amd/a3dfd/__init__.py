"""Python side of the MI355X acoustic forward modeller.

- `_lib`: ctypes binding of libacoustic3d.so (Modeller, ShotDesc, make_desc).
- `models`: seeded synthetic models of the benchmark configurations and the
  stability (CFL) check.
- `formats`: reference model / shot / snapshot file formats.
"""
from ._lib import (A3DError, Batch, LEGACY_NAMES, LIB_DIR, LIB_PATH, Modeller, OP_2ND, OP_ISO, OP_NAMES, OP_TTI,
                   OP_VTI, REFERENCE_DEFAULTS, ShotDesc, Window, desc_dict, load, make_desc)

__all__ = [
    "A3DError", "Batch", "LEGACY_NAMES", "LIB_DIR", "LIB_PATH", "Modeller", "OP_2ND", "OP_ISO", "OP_NAMES", "OP_TTI",
    "OP_VTI", "REFERENCE_DEFAULTS", "ShotDesc", "Window", "desc_dict", "load", "make_desc",
]
