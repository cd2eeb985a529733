"""The C ABI library: loads on a CPU-only host, exports every symbol the public
header declares, carries the reference's four library names, and the ctypes
mirror of a3d_shot_desc matches the header's field list."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "acoustic3d.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*([a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_library_loads_and_exports_every_declared_symbol():
    from a3dfd import LIB_PATH, load
    lib = load()
    assert lib.a3d_version().startswith(b"acoustic3d")
    decl = declared_functions()
    assert {"cuda_3dfd", "cuda_3dfd_1order", "cuda_3dfd_vti", "cuda_3dfd_tti", "a3d_run"} <= set(decl)
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (\S+)$", out, flags=re.M))
    missing = [n for n in decl if n not in exported]
    assert not missing, missing
    for n in decl:
        getattr(lib, n)


def test_reference_library_names_resolve_to_the_library():
    from a3dfd import LEGACY_NAMES, LIB_DIR, LIB_PATH
    for name in LEGACY_NAMES.values():
        p = os.path.join(LIB_DIR, name)
        assert os.path.exists(p), p
        assert os.path.realpath(p) == os.path.realpath(LIB_PATH)


def test_shot_desc_mirror_matches_header():
    from a3dfd import ShotDesc
    text = open(HEADER).read()
    body = text[text.index("typedef struct a3d_shot_desc {"):text.index("} a3d_shot_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for typ, names in re.findall(r"^\s*(int|float)\s+([^;]+);", body, flags=re.M):
        fields += [(n.strip(), typ) for n in names.split(",")]
    mirror = [(n, "float" if t.__name__ == "c_float" else "int") for n, t in ShotDesc._fields_]
    assert fields == mirror


def test_errors_without_a_gpu_are_reported_not_fatal():
    """On a host without a GPU the library returns an error code; it never exits."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from a3dfd import A3DError, Modeller, make_desc
    with pytest.raises(A3DError) as e:
        Modeller(make_desc("vti", 8, 8, 8, 10, 10, 10, 4, 4, 2, 10, 1e-3, 25.0, nb=6))
    assert e.value.code in (-1, -2)


def test_legacy_symbol_reports_missing_files(tmp_path, capfd):
    """A missing model file is reported and the call returns (reference: printf + crash)."""
    from ctypes import CDLL, c_bool, c_char_p, c_float, c_int
    from a3dfd import LIB_PATH
    lib = CDLL(LIB_PATH)
    lib.cuda_3dfd_tti.restype = c_int
    s = lambda x: c_char_p(str(x).encode())  # noqa: E731
    shot = tmp_path / "shot.bin"
    rc = lib.cuda_3dfd_tti(s(tmp_path / "missing.bin"), s(""), s("e"), s("d"), s("t"), s("p"), s(tmp_path), s(shot),
                           c_int(0), c_int(1), c_int(8), c_int(8), c_int(8), c_float(10), c_float(10), c_float(10),
                           c_int(4), c_int(4), c_int(2), c_int(1), c_int(1), c_int(1),
                           c_float(10), c_float(10), c_float(0.001), c_int(10), c_float(0.001), c_float(20.0),
                           c_bool(False), c_bool(False), c_int(1), c_int(0))
    assert rc < 0
    err = capfd.readouterr().err
    assert "missing.bin" in err
    bad = tmp_path / "no_such_dir" / "shot.bin"
    rc = lib.cuda_3dfd_tti(s("v"), s(""), s("e"), s("d"), s("t"), s("p"), s(tmp_path), s(bad),
                           c_int(0), c_int(1), c_int(8), c_int(8), c_int(8), c_float(10), c_float(10), c_float(10),
                           c_int(4), c_int(4), c_int(2), c_int(1), c_int(1), c_int(1),
                           c_float(10), c_float(10), c_float(0.001), c_int(10), c_float(0.001), c_float(20.0),
                           c_bool(False), c_bool(False), c_int(1), c_int(0))
    assert rc == 1   # reference tti:429-432


def test_setters_reject_bad_arguments_without_touching_a_gpu():
    """Argument checks of the newer setters run before any device call: a null handle or an
    out-of-range value gives A3D_EINVAL (-1) and a message."""
    from ctypes import c_int, c_void_p
    from a3dfd import load
    lib = load()
    null = c_void_p()
    assert lib.a3d_set_strip_head(null, 1) == -1
    assert lib.a3d_batch_set_zseg(null, 2, 2) == -1
    assert lib.a3d_batch_get_zseg(null, None, None) == -1
    v, s = c_int(), c_int()
    assert lib.a3d_batch_get_zseg(null, v, s) == -1
    assert lib.a3d_last_error()
