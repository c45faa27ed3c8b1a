"""CPU tests of the C ABI library: it loads and exports every function that
include/vc_api.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\bint\s+(vc_\w+)\s*\(", txt))
    return sorted(names)


def test_header_declares_expected_entry_points():
    import voxel_ma_amd
    assert _declared() == sorted(voxel_ma_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    import voxel_ma_amd
    lib = ctypes.CDLL(voxel_ma_amd.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    voxel_ma_amd.library()  # argtypes bind for every export


def test_library_is_gfx950_hip_code():
    import voxel_ma_amd
    data = open(voxel_ma_amd.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    assert b"__hip_fatbin" in data or b".hip_fatbin" in data


def test_status_codes_match_header():
    import voxel_ma_amd
    txt = open(os.path.join(ROOT, "include", "vc_api.h")).read()
    for name in ("VC_OK", "VC_ERR_INVALID", "VC_ERR_HIP", "VC_ERR_CAPACITY", "VC_ERR_EMPTY_VOL",
                 "VC_ERR_NOMEM", "VC_ERR_STATE", "VC_ERR_GUARD"):
        m = re.search(r"#define %s \(?(-?\d+)\)?" % name, txt)
        assert m and int(m.group(1)) == getattr(voxel_ma_amd, name), name
