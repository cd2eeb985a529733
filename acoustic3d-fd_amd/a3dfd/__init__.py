"""Python side of the MI355X acoustic forward modeller.

- `_lib`: ctypes binding of libacoustic3d.so (Modeller for one shot at a time,
  Batch for many uneven windows of one model, ShotDesc, make_desc, Mlp and
  strip_collect for the learned absorbing strip).
- `strip`: the learned absorbing strip's training data, training, judge
  (reflected energy against an extended-domain run) and PML fallback.
- `models`: seeded synthetic models of the benchmark configurations and the
  stability (CFL) check.
The reference's model / shot / snapshot file formats are handled by the legacy
entry points and by `readmodel.py` beside this package.
"""
from ._lib import (A3DError, Batch, LEGACY_NAMES, LIB_DIR, LIB_PATH, STRIP_NB, Mlp, Modeller, OP_2ND, OP_ISO, OP_NAMES,
                   OP_TTI, OP_VTI, REFERENCE_DEFAULTS, ShotDesc, Window, desc_dict, load, make_desc, strip_collect)

__all__ = [
    "A3DError", "Batch", "LEGACY_NAMES", "LIB_DIR", "LIB_PATH", "Mlp", "Modeller", "OP_2ND", "OP_ISO", "OP_NAMES",
    "OP_TTI", "OP_VTI", "REFERENCE_DEFAULTS", "STRIP_NB", "ShotDesc", "Window", "desc_dict", "load", "make_desc",
    "strip_collect",
]
