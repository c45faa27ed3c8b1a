"""ctypes access to the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product library.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# VC_ORACLE_LIBRARY: another build of the oracle (the sanitizer build)
LIB_PATH = os.environ.get("VC_ORACLE_LIBRARY",
                          os.path.join(ROOT, "oracle", "_build", "liboracle.so"))

_lib = None


class OrEuler(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in
                ("V", "E", "F", "T", "C", "euler", "V_bnd", "E_bnd", "V_int", "E_int", "F_int")]


class OrThinStats(ctypes.Structure):
    _fields_ = [("removed", ctypes.c_int64 * 64), ("sweeps", ctypes.c_int64 * 64),
                ("tests", ctypes.c_int64)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run __graft_entry__.build() (or make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32, i64 = ctypes.c_int, ctypes.c_int64
        L.or_set_threads.argtypes = [i32]
        L.or_set_threads.restype = i32
        L.or_edt_ft.argtypes = [P, i32, i32, i32, P, P]
        L.or_edt_brute.argtypes = [P, i32, i32, i32, P, P]
        L.or_envelope_1d.argtypes = [P, i32, P, P]
        L.or_lambda2.argtypes = [P, P, i32, i32, i32, P]
        L.or_level_threshold.argtypes = [ctypes.c_float]
        L.or_level_threshold.restype = i64
        L.or_levels.argtypes = [P, P, i64, P, i32, P]
        L.or_is_simple.argtypes = [ctypes.c_uint32]
        L.or_is_simple_topo.argtypes = [ctypes.c_uint32]
        L.or_simple_lut.argtypes = [P]
        L.or_simple_exhaustive.argtypes = [ctypes.POINTER(i64), i64, i64, ctypes.POINTER(i64),
                                           ctypes.POINTER(i64)]
        L.or_simple_exhaustive.restype = i64
        L.or_thin.argtypes = [P, P, i32, i32, i32, i32, ctypes.POINTER(OrThinStats)]
        L.or_euler_voxels.argtypes = [P, i32, i32, i32]
        L.or_euler_voxels.restype = i64
        L.or_components26.argtypes = [P, i32, i32, i32]
        L.or_components26.restype = i64
        L.or_bg_components6.argtypes = [P, i32, i32, i32]
        L.or_bg_components6.restype = i64
        L.or_cube_components.argtypes = [P]
        L.or_boundary.argtypes = [P, i32, i32, i32, P, i64, ctypes.POINTER(i64), P, i64,
                                  ctypes.POINTER(i64), ctypes.POINTER(OrEuler)]
        L.or_cc_finalize.argtypes = [i32, i32, P, i32, P, P, P, i32, P]
        L.or_cc_prune.argtypes = [i32, i32, P, i32, P, P, P, P, ctypes.c_float, ctypes.c_float,
                                  P, P, P, P]
        L.or_simple_triangulate.argtypes = [P, i32, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def edt_ft(occ):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    d2 = np.empty(occ.shape, np.int32)
    ft = np.empty(occ.shape, np.int32)
    rc = lib().or_edt_ft(_p(occ), *occ.shape, _p(d2), _p(ft))
    assert rc == 0
    return d2, ft


def edt_brute(occ):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    d2 = np.empty(occ.shape, np.int32)
    ft = np.empty(occ.shape, np.int32)
    lib().or_edt_brute(_p(occ), *occ.shape, _p(d2), _p(ft))
    return d2, ft


def envelope_1d(f):
    f = np.ascontiguousarray(f, dtype=np.int64)
    D = np.empty(len(f), np.int64)
    a = np.empty(len(f), np.int32)
    lib().or_envelope_1d(_p(f), len(f), _p(D), _p(a))
    return D, a


def lambda2(occ, ft):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    ft = np.ascontiguousarray(ft, dtype=np.int32)
    out = np.empty(occ.shape, np.int32)
    lib().or_lambda2(_p(occ), _p(ft), *occ.shape, _p(out))
    return out


def level_thresholds(levels):
    return np.array([lib().or_level_threshold(float(l)) for l in levels], dtype=np.int64)


def levels(occ, lam2, lv):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    lam2 = np.ascontiguousarray(lam2, dtype=np.int32)
    T = level_thresholds(lv)
    out = np.empty(occ.shape, np.uint8)
    lib().or_levels(_p(occ), _p(lam2), occ.size, _p(T), len(T), _p(out))
    return out


def is_simple(nb):
    return bool(lib().or_is_simple(int(nb)))


def is_simple_topo(nb):
    return bool(lib().or_is_simple_topo(int(nb)))


def simple_lut():
    out = np.empty(1 << 21, np.uint32)
    lib().or_simple_lut(_p(out))
    return out


def thin(occ, lvl, nlev):
    label = np.ascontiguousarray(occ, dtype=np.uint8).copy()
    lvl = np.ascontiguousarray(lvl, dtype=np.uint8)
    st = OrThinStats()
    rc = lib().or_thin(_p(label), _p(lvl), *label.shape, int(nlev), ctypes.byref(st))
    assert rc == 0
    return label, st


def euler_voxels(occ):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    return lib().or_euler_voxels(_p(occ), *occ.shape)


def components26(occ):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    return lib().or_components26(_p(occ), *occ.shape)


def bg_components6(occ):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    return lib().or_bg_components6(_p(occ), *occ.shape)


def betti(occ):
    """(b0, b1, b2, chi) of the inside voxels, (26,6) adjacency."""
    chi = euler_voxels(occ)
    b0 = components26(occ)
    b2 = bg_components6(occ) - 1
    return b0, b0 + b2 - chi, b2, chi


def simple_exhaustive(topo_stride=0, topo_offset=0):
    """(BFS-vs-table mismatches over all 2^26 codes, simple codes,
    topological-definition mismatches, codes checked topologically)."""
    ns, tb, tc = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    bad = lib().or_simple_exhaustive(ctypes.byref(ns), int(topo_stride), int(topo_offset),
                                     ctypes.byref(tb), ctypes.byref(tc))
    return int(bad), ns.value, tb.value, tc.value


def cube_components():
    t = np.empty(256, np.int32)
    lib().or_cube_components(_p(t))
    return t


def boundary(occ, with_euler=True):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    ns, nq = ctypes.c_int64(), ctypes.c_int64()
    lib().or_boundary(_p(occ), *occ.shape, None, 0, ctypes.byref(ns), None, 0,
                      ctypes.byref(nq), None)
    sites = np.empty((ns.value, 3), np.float32)
    quads = np.empty((nq.value, 4), np.int32)
    eu = OrEuler()
    rc = lib().or_boundary(_p(occ), *occ.shape, _p(sites), ns.value, ctypes.byref(ns), _p(quads),
                           nq.value, ctypes.byref(nq), ctypes.byref(eu) if with_euler else None)
    assert rc == 0
    e = {n: getattr(eu, n) for n, _ in OrEuler._fields_} if with_euler else None
    return sites, quads, e


def cc_finalize(nv, edges, faces):
    """edges: list of (a, b); faces: list of vertex lists.  Returns (edge
    array (ne, 2), face offsets, face edge ids)."""
    e_in = np.ascontiguousarray(np.array(edges, np.int32).reshape(-1, 2))
    foff = np.zeros(len(faces) + 1, np.int64)
    for i, f in enumerate(faces):
        foff[i + 1] = foff[i] + len(f)
    fv = np.ascontiguousarray(np.concatenate([np.array(f, np.int32) for f in faces])
                              if faces else np.zeros(0, np.int32))
    cap = len(e_in) + len(fv) + 1
    eout = np.zeros((cap, 2), np.int32)
    fe = np.zeros(max(len(fv), 1), np.int32)
    ne = lib().or_cc_finalize(nv, len(e_in), _p(e_in), len(faces), _p(foff), _p(fv), _p(eout),
                              cap, _p(fe))
    assert ne >= 0
    return eout[:ne].copy(), foff, fe[:len(fv)].copy()


def cc_prune(nv, edges, foff, fe, em, fm, f_t, l_t):
    edges = np.ascontiguousarray(edges, np.int32)
    ne, nf = len(edges), len(foff) - 1
    em = np.ascontiguousarray(em, np.float32)
    fm = np.ascontiguousarray(fm, np.float32)
    fe = np.ascontiguousarray(fe, np.int32)
    rv, re, rf = np.zeros(max(nv, 1), np.uint8), np.zeros(max(ne, 1), np.uint8), np.zeros(max(nf, 1), np.uint8)
    ref = np.zeros(max(ne, 1), np.int32)
    rc = lib().or_cc_prune(nv, ne, _p(edges), nf, _p(foff), _p(fe), _p(em), _p(fm), f_t, l_t,
                           _p(rv), _p(re), _p(rf), _p(ref))
    return rc, rv[:nv], re[:ne], rf[:nf], ref[:ne]


def simple_triangulate(poly):
    p = np.ascontiguousarray(poly, np.int32)
    out = np.zeros((max(len(p) - 2, 1), 3), np.int32)
    n = lib().or_simple_triangulate(_p(p), len(p), _p(out))
    return out[:n]


def set_threads(n):
    """Threads of the oracle's OpenMP loops (1 = sequential); returns the
    number in use.  Results do not depend on it."""
    return int(lib().or_set_threads(int(n)))
