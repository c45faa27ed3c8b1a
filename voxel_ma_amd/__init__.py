"""voxel_ma_amd -- MI355X-native Voxel Cores volumetric hot path.

Python view of the C ABI in include/vc_api.h (libvoxelcore.so, gfx950).  The
product is the HIP library; this module only binds it with ctypes.  There is
no CPU fallback: if the library is missing or no GPU is present every call
raises.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libvoxelcore.so")

VC_OK = 0
VC_ERR_INVALID = -1
VC_ERR_HIP = -2
VC_ERR_CAPACITY = -3
VC_ERR_EMPTY_VOL = -4
VC_ERR_NOMEM = -5
VC_ERR_STATE = -6
VC_ERR_GUARD = -7
VC_ERR_IO = -8
VC_MAX_LEVELS = 64

BUF_OCC, BUF_EDT, BUF_FT, BUF_LAM2, BUF_LABEL, BUF_LEVEL = range(6)
_BUF_DTYPE = {BUF_OCC: np.uint8, BUF_EDT: np.int32, BUF_FT: np.int32,
              BUF_LAM2: np.uint32, BUF_LABEL: np.uint8, BUF_LEVEL: np.uint8}

# every symbol declared in include/vc_api.h
EXPORTS = [
    "vc_ctx_create", "vc_ctx_destroy", "vc_last_error", "vc_ctx_stream",
    "vc_ctx_synchronize", "vc_simple_table_host", "vc_thin_timing",
    "vc_thin_timing_read", "vc_volume_upload",
    "vc_volume_from_device", "vc_volume_release", "vc_volume_dims", "vc_edt",
    "vc_measure", "vc_set_measure_f32", "vc_thin_levels", "vc_boundary_extract",
    "vc_get", "vc_get_device", "vc_synth_blurred_noise", "vc_synth_solids", "vc_cc_finalize",
    "vc_cc_prune", "vc_cc_prune_device", "vc_cc_finalize_device", "vc_skeleton_noalloc",
    # include/vc_cc.h
    "vc_cc_create", "vc_cc_release", "vc_cc_sizes", "vc_cc_get", "vc_cc_topology",
    "vc_cc_subset_components", "vc_ccthin_create", "vc_ccthin_release", "vc_ccthin_assign",
    "vc_ccthin_prune", "vc_ccthin_removed", "vc_ccthin_remaining", "vc_ccthin_remained",
    "vc_ccthin_remaining_euler", "vc_cc_closed_components", "vc_euler_char",
    # include/vc_batch.h
    "vc_batch_create", "vc_batch_destroy", "vc_batch_size", "vc_batch_edt", "vc_batch_measure",
    "vc_batch_thin", "vc_batch_run", "vc_schedule_defaults", "vc_batch_get_schedule",
    "vc_batch_set_schedule", "vc_batch_tune",
    # include/vc_surrogate.h
    "vc_surrogate_create", "vc_surrogate_destroy", "vc_surrogate_get_params",
    "vc_surrogate_set_params", "vc_surrogate_save", "vc_surrogate_load",
    "vc_surrogate_gradients", "vc_surrogate_train_step", "vc_surrogate_infer",
    "vc_surrogate_measure", "vc_surrogate_activation",
    # include/vc_voro.h
    "vc_voro_tag_points", "vc_voro_measures", "vc_voro_extract_inside",
    "vc_voro_extract_inside_device", "vc_voro_merge_close", "vc_voro_merge_close_device",
    "vc_voro_radii", "vc_voro_part_inside", "vc_voro_from_ft",
    # include/vc_vdiag.h
    "vc_vdiag_from_ft", "vc_vdiag_release", "vc_vdiag_info_get", "vc_vdiag_radii",
    "vc_vdiag_tag", "vc_vdiag_merge_close", "vc_vdiag_get", "vc_vdiag_extract_inside",
    # include/vc_octree.h
    "vc_octree_parse", "vc_octree_free", "vc_octree_value", "vc_volume_from_octree",
    "vc_octree_leaf_centers",
    # include/vc_export.h
    "vc_write_ply", "vc_write_radii", "vc_write_dotma", "vc_write_off", "vc_write_tetnodes",
    "vc_xform_apply",
]
NOALLOC_SCRATCH_BYTES = 64 << 20
NOALLOC_MAX_VOXELS = 1 << 22
SURROGATE_PARAMS = 678786
SURROGATE_HALO = 56


class VcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("vc error %d: %s" % (code, msg))
        self.code = code


class ThinStats(ctypes.Structure):
    _fields_ = [("removed", ctypes.c_int64 * VC_MAX_LEVELS), ("passes", ctypes.c_int64),
                ("tests", ctypes.c_int64), ("launches", ctypes.c_int64),
                ("voxel_passes", ctypes.c_int64)]


class Euler(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in
                ("V", "E", "F", "T", "C", "euler", "V_bnd", "E_bnd", "V_int", "E_int", "F_int",
                 "chi")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Schedule(ctypes.Structure):
    """vc_schedule: launch parameters of a batch (results never depend on it)."""
    _fields_ = [("edt_lds_kb", ctypes.c_int), ("edt_grid_per_cu", ctypes.c_int),
                ("thin_grid_per_cu", ctypes.c_int), ("sweeps_per_sync", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("tail_sweep", ctypes.c_int),
                ("tile_sweep", ctypes.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class SurrogateStats(ctypes.Structure):
    _fields_ = [("inside", ctypes.c_int64), ("accepted", ctypes.c_int64),
                ("fallback", ctypes.c_int64), ("ms_infer", ctypes.c_float),
                ("ms_exact", ctypes.c_float), ("ms_mix", ctypes.c_float)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class VdiagInfo(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int64), ("ne", ctypes.c_int64), ("nf", ctypes.c_int64),
                ("ns", ctypes.c_int64), ("n_sides", ctypes.c_int64),
                ("tagged", ctypes.c_int32), ("radii", ctypes.c_int32),
                ("n_valid", ctypes.c_int64), ("bb_min", ctypes.c_float * 3),
                ("bb_max", ctypes.c_float * 3)]


class CcTopo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("V", "E", "F", "T", "C", "euler",
                                              "face_components")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class OctreeInfo(ctypes.Structure):
    _fields_ = [("resolution", ctypes.c_int32), ("n_internal", ctypes.c_int64),
                ("n_empty", ctypes.c_int64), ("n_leaf", ctypes.c_int64),
                ("lower_left", ctypes.c_float * 3), ("bbox_len", ctypes.c_float)]


class CcView(ctypes.Structure):
    _fields_ = [("nv", ctypes.c_int32), ("ne", ctypes.c_int32), ("nf", ctypes.c_int32),
                ("edges", ctypes.c_void_p), ("face_off", ctypes.c_void_p),
                ("face_edges", ctypes.c_void_p)]


_lib = None


def library():
    """Load libvoxelcore.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    # VC_LIBRARY names another build of the C ABI to load instead (the
    # sanitizer build of the host-only entry points); only the entry points
    # it has are bound
    path = os.environ.get("VC_LIBRARY", LIB_PATH)
    if not os.path.exists(path):
        raise RuntimeError("%s is not built (run __graft_entry__.build())" % path)
    L = ctypes.CDLL(path)
    P, I, I64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    PP = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "vc_ctx_create": [I, PP],
        "vc_ctx_destroy": [P],
        "vc_last_error": [P, ctypes.c_char_p, ctypes.c_size_t],
        "vc_ctx_stream": [P, PP],
        "vc_ctx_synchronize": [P],
        "vc_simple_table_host": [P, P],
        "vc_thin_timing": [P, I],
        "vc_thin_timing_read": [P, ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_double)],
        "vc_volume_upload": [P, P, I, I, I, PP],
        "vc_volume_from_device": [P, P, I, I, I, PP],
        "vc_volume_release": [P, P],
        "vc_volume_dims": [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)],
        "vc_edt": [P, P, I],
        "vc_measure": [P, P],
        "vc_set_measure_f32": [P, P, P],
        "vc_thin_levels": [P, P, P, I, ctypes.POINTER(ThinStats)],
        "vc_boundary_extract": [P, P, P, I64, ctypes.POINTER(I64), P, I64,
                                ctypes.POINTER(I64), ctypes.POINTER(Euler)],
        "vc_get": [P, P, I, P, I64],
        "vc_get_device": [P, P, I, P, I64],
        "vc_synth_blurred_noise": [P, P, I, I, I, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_uint64, I, ctypes.POINTER(ctypes.c_float)],
        "vc_synth_solids": [P, P, I, I, I, P, I],
        "vc_cc_finalize": [I, I, P, I, P, P, P, I, ctypes.POINTER(I), P],
        "vc_cc_finalize_device": [P, I, I, P, I, P, P, P, I, ctypes.POINTER(I), P],
        "vc_cc_prune": [ctypes.POINTER(CcView), P, P, ctypes.c_float, ctypes.c_float, P, P, P, P],
        "vc_cc_prune_device": [P, ctypes.POINTER(CcView), P, P, ctypes.c_float, ctypes.c_float,
                               P, P, P, P, ctypes.POINTER(ctypes.c_int64)],
        "vc_skeleton_noalloc": [P, P, I, I, I, P, I, P, ctypes.POINTER(ThinStats)],
        "vc_cc_create": [P, I, I, P, I, P, P, PP],
        "vc_cc_release": [P, P],
        "vc_cc_sizes": [P, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I),
                        ctypes.POINTER(I64)],
        "vc_cc_get": [P, P, P, P],
        "vc_cc_topology": [P, P, ctypes.POINTER(CcTopo)],
        "vc_cc_subset_components": [P, P, P, ctypes.POINTER(I64)],
        "vc_ccthin_create": [P, P, PP],
        "vc_ccthin_release": [P, P],
        "vc_ccthin_assign": [P, P, P, P],
        "vc_ccthin_prune": [P, P, ctypes.c_float, ctypes.c_float, ctypes.POINTER(I64)],
        "vc_ccthin_removed": [P, P, P, P, P, P],
        "vc_ccthin_remaining": [P, P, P, I64, ctypes.POINTER(I64), P, I64,
                                ctypes.POINTER(I64), P, P, I64, ctypes.POINTER(I64)],
        "vc_ccthin_remained": [P, P, P, I64, ctypes.POINTER(I64), P, I64, ctypes.POINTER(I64),
                               P, I64, ctypes.POINTER(I64)],
        "vc_ccthin_remaining_euler": [P, P, ctypes.POINTER(CcTopo)],
        "vc_cc_closed_components": [P, P, ctypes.POINTER(I64)],
        "vc_euler_char": [P, I, P, I64, P, I64, ctypes.POINTER(CcTopo)],
        "vc_batch_create": [P, P, I, PP],
        "vc_batch_destroy": [P, P],
        "vc_batch_size": [P, ctypes.POINTER(I), ctypes.POINTER(I64)],
        "vc_batch_edt": [P, P, I],
        "vc_batch_measure": [P, P],
        "vc_batch_thin": [P, P, P, I, ctypes.POINTER(ThinStats)],
        "vc_batch_run": [P, P, P, I, ctypes.POINTER(ThinStats)],
        "vc_schedule_defaults": [ctypes.POINTER(Schedule)],
        "vc_batch_get_schedule": [P, ctypes.POINTER(Schedule)],
        "vc_batch_set_schedule": [P, P, ctypes.POINTER(Schedule)],
        "vc_batch_tune": [P, P, P, I, I, ctypes.POINTER(Schedule), ctypes.POINTER(ctypes.c_float)],
        "vc_surrogate_create": [P, ctypes.c_uint64, PP],
        "vc_surrogate_destroy": [P, P],
        "vc_surrogate_get_params": [P, P, P, I64],
        "vc_surrogate_set_params": [P, P, P, I64],
        "vc_surrogate_save": [P, P, ctypes.c_char_p],
        "vc_surrogate_load": [P, P, ctypes.c_char_p],
        "vc_surrogate_gradients": [P, P, P, P, ctypes.POINTER(ctypes.c_float)],
        "vc_surrogate_train_step": [P, P, P, I, ctypes.c_float, ctypes.POINTER(ctypes.c_float)],
        "vc_surrogate_infer": [P, P, P, I, P, P],
        "vc_surrogate_measure": [P, P, P, ctypes.c_float, I, ctypes.POINTER(SurrogateStats)],
        "vc_surrogate_activation": [P, P, I, P, I64],
        "vc_vdiag_from_ft": [P, P, PP],
        "vc_vdiag_release": [P, P],
        "vc_vdiag_info_get": [P, ctypes.POINTER(VdiagInfo)],
        "vc_vdiag_radii": [P, P],
        "vc_vdiag_tag": [P, P, P],
        "vc_vdiag_merge_close": [P, P, ctypes.c_float, I, ctypes.POINTER(I)],
        "vc_vdiag_get": [P, P, P, P, P, P, P, P, P, P],
        "vc_vdiag_extract_inside": [P, P, I, P, P, I64, ctypes.POINTER(I64), P, P, I64,
                                    ctypes.POINTER(I64), P, P, P, I64, ctypes.POINTER(I64),
                                    PP],
        "vc_octree_parse": [P, ctypes.c_size_t, I, PP, ctypes.POINTER(OctreeInfo)],
        "vc_octree_free": [P],
        "vc_octree_value": [P, I, I, I, ctypes.POINTER(ctypes.c_double)],
        "vc_volume_from_octree": [P, P, PP],
        "vc_octree_leaf_centers": [P, P, I64, ctypes.POINTER(I64)],
        "vc_voro_tag_points": [P, P, P, I64, P],
        "vc_voro_measures": [P, I, I, I, P, I, P, P, P, P, P, P, P, P, P],
        "vc_voro_extract_inside": [I, I, I, P, P, P, P, P, P, P, I64, ctypes.POINTER(I64), P, P,
                                   I64, ctypes.POINTER(I64), P, P, P, I64, ctypes.POINTER(I64)],
        "vc_voro_extract_inside_device": [P, I, I, I, P, P, P, P, P, P, P, I64,
                                          ctypes.POINTER(I64), P, P, I64, ctypes.POINTER(I64),
                                          P, P, P, I64, ctypes.POINTER(I64)],
        "vc_voro_merge_close": [I, I, I, P, P, P, P, P, P, P, ctypes.c_float, I, P, P,
                                ctypes.POINTER(I), P, P, P, P, I64, ctypes.POINTER(I64), P],
        "vc_voro_merge_close_device": [P, I, I, I, P, P, P, P, P, P, P, ctypes.c_float, I, P, P,
                                ctypes.POINTER(I), P, P, P, P, I64, ctypes.POINTER(I64), P, ctypes.POINTER(I)],
        "vc_voro_radii": [P, I, I, P, P, P, P, I, P, P],
        "vc_voro_from_ft": [P, P, P, I64, ctypes.POINTER(I64), P, P, I64,
                            ctypes.POINTER(I64), P, I64, ctypes.POINTER(I64)],
        "vc_voro_part_inside": [I, I, P, P, P, P, I64, ctypes.POINTER(I64), P, I64,
                                ctypes.POINTER(I64), P, I64, ctypes.POINTER(I64)],
        "vc_write_ply": [ctypes.c_char_p, P, I64, P, I64, P, I64],
        "vc_write_radii": [ctypes.c_char_p, P, P, I64, P],
        "vc_write_dotma": [ctypes.c_char_p, P, P, I64, P, I64, P, I64],
        "vc_write_off": [ctypes.c_char_p, P, I64, P, I64, P],
        "vc_write_tetnodes": [ctypes.c_char_p, P, I64],
        "vc_xform_apply": [P, P, I64, P],
    }
    for name, args in sig.items():
        if path != LIB_PATH and not hasattr(L, name):
            continue
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _Hip:
    """The HIP runtime libvoxelcore.so itself links (one runtime per process)."""

    def __init__(self):
        library()
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        self.h = h


_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = _Hip()
    return _hip.h


class DeviceBuffer:
    """Raw device memory for handing device pointers across the C ABI."""

    H2D, D2H = 1, 2

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise VcError(VC_ERR_NOMEM, "hipMalloc(%d) failed (%d)" % (self.nbytes, rc))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        rc = hip().hipMemcpy(ctypes.c_void_p(self.ptr), _p(a), a.nbytes, self.H2D)
        if rc != 0:
            raise VcError(VC_ERR_HIP, "hipMemcpy H2D failed (%d)" % rc)

    def download(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        hip().hipDeviceSynchronize()
        rc = hip().hipMemcpy(_p(out), ctypes.c_void_p(self.ptr), out.nbytes, self.D2H)
        if rc != 0:
            raise VcError(VC_ERR_HIP, "hipMemcpy D2H failed (%d)" % rc)
        return out

    def free(self):
        if getattr(self, "ptr", None):
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One HIP device + stream + simple-point table (vc_ctx)."""

    def __init__(self, device=0):
        self._lib = library()
        h = ctypes.c_void_p()
        rc = self._lib.vc_ctx_create(int(device), ctypes.byref(h))
        if rc != VC_OK:
            raise VcError(rc, "vc_ctx_create(device=%d) failed" % device)
        self.handle = h

    def last_error(self):
        buf = ctypes.create_string_buffer(1024)
        self._lib.vc_last_error(self.handle, buf, 1024)
        return buf.value.decode(errors="replace")

    def check(self, rc):
        if rc != VC_OK:
            raise VcError(rc, self.last_error())
        return rc

    def stream(self):
        s = ctypes.c_void_p()
        self.check(self._lib.vc_ctx_stream(self.handle, ctypes.byref(s)))
        return s.value

    def synchronize(self):
        self.check(self._lib.vc_ctx_synchronize(self.handle))

    def thin_timing(self, enable):
        """Bracket every tile-sweep launch of this context with HIP events
        (True) or stop (False); either resets the totals."""
        self.check(self._lib.vc_thin_timing(self.handle, 1 if enable else 0))

    def thin_timing_read(self):
        """(launches timed, summed duration in ms) since thin_timing()."""
        n, ms = ctypes.c_int64(), ctypes.c_double()
        self.check(self._lib.vc_thin_timing_read(self.handle, ctypes.byref(n),
                                                 ctypes.byref(ms)))
        return int(n.value), float(ms.value)

    def simple_table(self):
        out = np.empty(1 << 21, np.uint32)
        self.check(self._lib.vc_simple_table_host(self.handle, _p(out)))
        return out

    def synth_blurred_noise(self, shape, sigma=12.0, fill=0.30, seed=3, shell=True):
        """Blurred-noise occupancy generated on the device (vc_synth.h);
        returns (DeviceBuffer, threshold)."""
        nx, ny, nz = (int(s) for s in shape)
        buf = DeviceBuffer(nx * ny * nz)
        thr = ctypes.c_float()
        self.check(self._lib.vc_synth_blurred_noise(self.handle, ctypes.c_void_p(buf.ptr), nx, ny,
                                                    nz, float(sigma), float(fill), int(seed),
                                                    1 if shell else 0, ctypes.byref(thr)))
        return buf, thr.value

    def synth_solids(self, shape, prims):
        """Union of primitives (vc_synth_solids) into a new DeviceBuffer."""
        nx, ny, nz = (int(s) for s in shape)
        pr = np.ascontiguousarray(prims, dtype=np.float64).reshape(-1, 16)
        buf = DeviceBuffer(nx * ny * nz)
        self.check(self._lib.vc_synth_solids(self.handle, ctypes.c_void_p(buf.ptr), nx, ny, nz,
                                             _p(pr), len(pr)))
        return buf

    def solids_volume(self, shape, prims):
        """A Volume holding the union of primitives, generated on the device."""
        buf = self.synth_solids(shape, prims)
        vol = self.from_device(buf.ptr, shape)
        self.synchronize()
        buf.free()
        return vol

    def upload(self, occ):
        occ = np.ascontiguousarray(occ, dtype=np.uint8)
        if occ.ndim != 3:
            raise ValueError("occupancy must be 3-D [x][y][z]")
        return Volume(self, occ.shape, host=occ)

    def from_device(self, dev_ptr, shape):
        return Volume(self, shape, dev_ptr=dev_ptr)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.vc_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Volume:
    """A volume resident in HBM plus its pipeline results (vc_vol)."""

    def __init__(self, ctx, shape, host=None, dev_ptr=None):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        h = ctypes.c_void_p()
        L = ctx._lib
        if host is not None:
            rc = L.vc_volume_upload(ctx.handle, _p(host), *self.shape, ctypes.byref(h))
        else:
            rc = L.vc_volume_from_device(ctx.handle, ctypes.c_void_p(dev_ptr), *self.shape,
                                         ctypes.byref(h))
        ctx.check(rc)
        self.handle = h

    @property
    def N(self):
        return self.shape[0] * self.shape[1] * self.shape[2]

    def edt(self):
        self.ctx.check(self.ctx._lib.vc_edt(self.ctx.handle, self.handle, 0))
        return self

    def measure(self):
        self.ctx.check(self.ctx._lib.vc_measure(self.ctx.handle, self.handle))
        return self

    def set_measure_f32(self, dev_ptr):
        self.ctx.check(self.ctx._lib.vc_set_measure_f32(self.ctx.handle, self.handle,
                                                        ctypes.c_void_p(dev_ptr)))
        return self

    def thin_levels(self, levels, stats=True):
        lv = np.ascontiguousarray(levels, dtype=np.float32)
        st = ThinStats() if stats else None
        self.ctx.check(self.ctx._lib.vc_thin_levels(self.ctx.handle, self.handle, _p(lv), len(lv),
                                                    ctypes.byref(st) if stats else None))
        return st

    def boundary(self, with_euler=True):
        L, c = self.ctx._lib, self.ctx
        ns, nq = ctypes.c_int64(), ctypes.c_int64()
        rc = L.vc_boundary_extract(c.handle, self.handle, None, 0, ctypes.byref(ns), None, 0,
                                   ctypes.byref(nq), None)
        if rc not in (VC_OK, VC_ERR_CAPACITY):
            c.check(rc)
        sites = np.empty((ns.value, 3), np.float32)
        quads = np.empty((nq.value, 4), np.int32)
        eu = Euler()
        c.check(L.vc_boundary_extract(c.handle, self.handle, _p(sites), ns.value, ctypes.byref(ns),
                                      _p(quads), nq.value, ctypes.byref(nq),
                                      ctypes.byref(eu) if with_euler else None))
        return sites, quads, (eu.as_dict() if with_euler else None)

    def get(self, which):
        out = np.empty(self.shape, _BUF_DTYPE[which])
        self.ctx.check(self.ctx._lib.vc_get(self.ctx.handle, self.handle, which, _p(out),
                                            out.nbytes))
        return out

    def get_device(self, which, dev_ptr, nbytes):
        self.ctx.check(self.ctx._lib.vc_get_device(self.ctx.handle, self.handle, which,
                                                   ctypes.c_void_p(dev_ptr), int(nbytes)))

    def release(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.vc_volume_release(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Batch:
    """Several volumes of one context processed together (vc_batch): one
    launch per pass over all of them.  Results stay in the member volumes
    (Volume.get)."""

    def __init__(self, ctx, volumes):
        self.ctx = ctx
        self.volumes = list(volumes)
        arr = (ctypes.c_void_p * len(self.volumes))(*[v.handle.value for v in self.volumes])
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.vc_batch_create(ctx.handle, arr, len(self.volumes), ctypes.byref(h)))
        self.handle = h

    def _levels(self, levels):
        lv = np.ascontiguousarray(levels, dtype=np.float32)
        if lv.ndim == 1:
            lv = np.ascontiguousarray(np.broadcast_to(lv, (len(self.volumes), lv.size)))
        if lv.shape[0] != len(self.volumes):
            raise ValueError("levels: one row per volume")
        return lv

    def size(self):
        n, vox = ctypes.c_int(), ctypes.c_int64()
        self.ctx.check(self.ctx._lib.vc_batch_size(self.handle, ctypes.byref(n),
                                                   ctypes.byref(vox)))
        return n.value, vox.value

    def edt(self):
        self.ctx.check(self.ctx._lib.vc_batch_edt(self.ctx.handle, self.handle, 0))
        return self

    def measure(self):
        self.ctx.check(self.ctx._lib.vc_batch_measure(self.ctx.handle, self.handle))
        return self

    def thin(self, levels, stats=True):
        lv = self._levels(levels)
        st = ThinStats() if stats else None
        self.ctx.check(self.ctx._lib.vc_batch_thin(self.ctx.handle, self.handle, _p(lv),
                                                   lv.shape[1],
                                                   ctypes.byref(st) if stats else None))
        return st

    def run(self, levels, stats=True):
        lv = self._levels(levels)
        st = ThinStats() if stats else None
        self.ctx.check(self.ctx._lib.vc_batch_run(self.ctx.handle, self.handle, _p(lv),
                                                  lv.shape[1],
                                                  ctypes.byref(st) if stats else None))
        return st

    def schedule(self):
        sc = Schedule()
        self.ctx.check(self.ctx._lib.vc_batch_get_schedule(self.handle, ctypes.byref(sc)))
        return sc

    def set_schedule(self, **kw):
        sc = self.schedule()
        for k, v in kw.items():
            setattr(sc, k, int(v))
        self.ctx.check(self.ctx._lib.vc_batch_set_schedule(self.ctx.handle, self.handle,
                                                           ctypes.byref(sc)))
        return sc

    def tune(self, levels, reps=3):
        """Autotune the schedule on this batch; returns (Schedule, (edt+measure
        ms, thinning ms))."""
        lv = self._levels(levels)
        sc = Schedule()
        t = (ctypes.c_float * 2)()
        self.ctx.check(self.ctx._lib.vc_batch_tune(self.ctx.handle, self.handle, _p(lv),
                                                   lv.shape[1], int(reps), ctypes.byref(sc), t))
        return sc, (t[0], t[1])

    def release(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.vc_batch_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Surrogate:
    """The learned measure surrogate (vc_surrogate): 3-D U-Net on BF16 MFMA,
    trained by the library's own forward / backward kernels."""

    def __init__(self, ctx, seed=1, path=None):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.vc_surrogate_create(ctx.handle, int(seed), ctypes.byref(h)))
        self.handle = h
        if path:
            self.load(path)

    def params(self):
        out = np.empty(SURROGATE_PARAMS, np.float32)
        self.ctx.check(self.ctx._lib.vc_surrogate_get_params(self.ctx.handle, self.handle,
                                                             _p(out), out.size))
        return out

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        self.ctx.check(self.ctx._lib.vc_surrogate_set_params(self.ctx.handle, self.handle,
                                                             _p(th), th.size))

    def save(self, path):
        self.ctx.check(self.ctx._lib.vc_surrogate_save(self.ctx.handle, self.handle,
                                                       os.fsencode(path)))

    def load(self, path):
        self.ctx.check(self.ctx._lib.vc_surrogate_load(self.ctx.handle, self.handle,
                                                       os.fsencode(path)))

    def gradients(self, vol):
        g = np.empty(SURROGATE_PARAMS, np.float32)
        loss = ctypes.c_float()
        self.ctx.check(self.ctx._lib.vc_surrogate_gradients(self.ctx.handle, self.handle,
                                                            vol.handle, _p(g),
                                                            ctypes.byref(loss)))
        return loss.value, g

    def train_step(self, vols, lr):
        arr = (ctypes.c_void_p * len(vols))(*[v.handle.value for v in vols])
        loss = ctypes.c_float()
        self.ctx.check(self.ctx._lib.vc_surrogate_train_step(self.ctx.handle, self.handle, arr,
                                                             len(vols), float(lr),
                                                             ctypes.byref(loss)))
        return loss.value

    def infer(self, vol, tile_core=0):
        lam = np.empty(vol.shape, np.float32)
        rel = np.empty(vol.shape, np.float32)
        self.ctx.check(self.ctx._lib.vc_surrogate_infer(self.ctx.handle, self.handle, vol.handle,
                                                        int(tile_core), _p(lam), _p(rel)))
        return lam, rel

    def measure(self, vol, max_rel_err=0.01, tile_core=0):
        st = SurrogateStats()
        self.ctx.check(self.ctx._lib.vc_surrogate_measure(self.ctx.handle, self.handle,
                                                          vol.handle, float(max_rel_err),
                                                          int(tile_core), ctypes.byref(st)))
        return st

    def activation(self, which, shape):
        out = np.empty(shape, np.float32)
        self.ctx.check(self.ctx._lib.vc_surrogate_activation(self.ctx.handle, self.handle,
                                                             int(which), _p(out), out.size))
        return out

    def release(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.vc_surrogate_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# ---------------------------------------------------------------- cell complex
def cc_finalize(nv, edges, faces, ctx=None):
    """cellcomplex::finalize through vc_cc_finalize (host), or through
    vc_cc_finalize_device when a context is given.  edges: list of (a, b);
    faces: list of vertex lists.  Returns (edges (ne, 2), face offsets, face
    side edge ids)."""
    e_in = np.ascontiguousarray(np.array(edges, np.int32).reshape(-1, 2))
    foff, fv = _faces_csr(faces)
    fv = fv[:foff[-1]] if foff[-1] < len(fv) else fv
    cap = len(e_in) + len(fv) + 1
    eout = np.zeros((cap, 2), np.int32)
    fe = np.zeros(max(len(fv), 1), np.int32)
    ne = ctypes.c_int()
    if ctx is None:
        rc = library().vc_cc_finalize(nv, len(e_in), _p(e_in), len(faces), _p(foff), _p(fv),
                                      _p(eout), cap, ctypes.byref(ne), _p(fe))
    else:
        rc = library().vc_cc_finalize_device(ctx.handle, nv, len(e_in), _p(e_in), len(faces),
                                             _p(foff), _p(fv), _p(eout), cap, ctypes.byref(ne),
                                             _p(fe))
    if rc != VC_OK:
        raise VcError(rc, "vc_cc_finalize")
    return eout[:ne.value].copy(), foff, fe[:len(fv)].copy()


def cc_prune(nv, edges, foff, fe, em, fm, f_t, l_t):
    """CellComplexThinning::prune through vc_cc_prune.  Returns (rc,
    removed_v, removed_e, removed_f, edge ref counts)."""
    edges = np.ascontiguousarray(edges, np.int32)
    foff = np.ascontiguousarray(foff, np.int64)
    fe = np.ascontiguousarray(fe, np.int32)
    ne, nf = len(edges), len(foff) - 1
    em = np.ascontiguousarray(em, np.float32)
    fm = np.ascontiguousarray(fm, np.float32)
    rv, re, rf = (np.zeros(max(nv, 1), np.uint8), np.zeros(max(ne, 1), np.uint8),
                  np.zeros(max(nf, 1), np.uint8))
    ref = np.zeros(max(ne, 1), np.int32)
    view = CcView(nv, ne, nf, edges.ctypes.data, foff.ctypes.data, fe.ctypes.data)
    rc = library().vc_cc_prune(ctypes.byref(view), _p(em), _p(fm), float(f_t), float(l_t),
                               _p(rv), _p(re), _p(rf), _p(ref))
    return rc, rv[:nv], re[:ne], rf[:nf], ref[:ne]


def cc_prune_device(ctx, nv, edges, foff, fe, em, fm, f_t, l_t):
    """The same collapse on the device (vc_cc_prune_device).  Returns (rc,
    removed_v, removed_e, removed_f, edge ref counts, queue generations)."""
    edges = np.ascontiguousarray(edges, np.int32)
    foff = np.ascontiguousarray(foff, np.int64)
    fe = np.ascontiguousarray(fe, np.int32)
    ne, nf = len(edges), len(foff) - 1
    em = np.ascontiguousarray(em, np.float32)
    fm = np.ascontiguousarray(fm, np.float32)
    rv, re, rf = (np.zeros(max(nv, 1), np.uint8), np.zeros(max(ne, 1), np.uint8),
                  np.zeros(max(nf, 1), np.uint8))
    ref = np.zeros(max(ne, 1), np.int32)
    gens = ctypes.c_int64()
    view = CcView(nv, ne, nf, edges.ctypes.data, foff.ctypes.data, fe.ctypes.data)
    rc = library().vc_cc_prune_device(ctx.handle, ctypes.byref(view), _p(em), _p(fm),
                                      float(f_t), float(l_t), _p(rv), _p(re), _p(rf),
                                      _p(ref), ctypes.byref(gens))
    return rc, rv[:nv], re[:ne], rf[:nf], ref[:ne], gens.value


class CellComplex:
    """A 2-complex finalized on the device and kept there (include/vc_cc.h):
    cellcomplex(vts, edges, faces) + finalize, its topology checks, and the
    collapse states over it (CellComplexThin)."""

    def __init__(self, ctx, nv, edges, faces):
        self.ctx = ctx
        self.nv = int(nv)
        e_in = np.ascontiguousarray(np.array(edges, np.int32).reshape(-1, 2))
        foff, fv = _faces_csr(faces)
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.vc_cc_create(ctx.handle, self.nv, len(e_in), _p(e_in),
                                        len(foff) - 1, _p(foff), _p(fv), ctypes.byref(h)))
        self.handle = h
        self.face_off = foff
        self.face_vts = fv[:foff[-1]]

    def sizes(self):
        a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        self.ctx.check(self.ctx._lib.vc_cc_sizes(self.handle, ctypes.byref(a), ctypes.byref(b),
                                                 ctypes.byref(c), ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    def finalized(self):
        """(edges (ne, 2), face side edge ids)"""
        _, ne, _, ns = self.sizes()
        e = np.zeros((max(ne, 1), 2), np.int32)
        fe = np.zeros(max(ns, 1), np.int32)
        self.ctx.check(self.ctx._lib.vc_cc_get(self.ctx.handle, self.handle, _p(e), _p(fe)))
        return e[:ne], fe[:ns]

    def topology(self):
        t = CcTopo()
        self.ctx.check(self.ctx._lib.vc_cc_topology(self.ctx.handle, self.handle,
                                                    ctypes.byref(t)))
        return t.as_dict()

    def subset_components(self, face_in):
        f = np.ascontiguousarray(np.asarray(face_in, np.uint8))
        n = ctypes.c_int64()
        self.ctx.check(self.ctx._lib.vc_cc_subset_components(self.ctx.handle, self.handle,
                                                             _p(f), ctypes.byref(n)))
        return n.value

    def closed_components(self):
        n = ctypes.c_int64()
        self.ctx.check(self.ctx._lib.vc_cc_closed_components(self.ctx.handle, self.handle,
                                                             ctypes.byref(n)))
        return n.value

    def thinning(self):
        return CellComplexThin(self)

    def release(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.vc_cc_release(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class VoroDiagram:
    """The discrete Voronoi diagram of a volume kept on the device
    (include/vc_vdiag.h): built from the feature transform, given radii,
    tagged, merged and its inside part extracted without leaving the device."""

    def __init__(self, vol):
        self.ctx = vol.ctx
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx._lib.vc_vdiag_from_ft(self.ctx.handle, vol.handle,
                                                      ctypes.byref(h)))
        self.handle = h

    def info(self):
        i = VdiagInfo()
        self.ctx.check(self.ctx._lib.vc_vdiag_info_get(self.handle, ctypes.byref(i)))
        return {"nv": i.nv, "ne": i.ne, "nf": i.nf, "ns": i.ns, "n_sides": i.n_sides,
                "tagged": i.tagged, "radii": i.radii, "n_valid": i.n_valid,
                "bb_min": list(i.bb_min), "bb_max": list(i.bb_max)}

    def radii(self):
        self.ctx.check(self.ctx._lib.vc_vdiag_radii(self.ctx.handle, self.handle))
        return self

    def tag(self, vol):
        self.ctx.check(self.ctx._lib.vc_vdiag_tag(self.ctx.handle, self.handle, vol.handle))
        return self

    def merge_close(self, eps, only_inside):
        on = ctypes.c_int()
        self.ctx.check(self.ctx._lib.vc_vdiag_merge_close(self.ctx.handle, self.handle,
                                                          float(eps), int(bool(only_inside)),
                                                          ctypes.byref(on)))
        return bool(on.value)

    def get(self):
        """dict of host copies: vts (nv, 3), edges (ne, 2), face_off, face_vts,
        face_sites (nf, 2), sites (ns, 3), radii (or None), valid"""
        i = self.info()
        nv, ne, nf, ns, nsd = i["nv"], i["ne"], i["nf"], i["ns"], i["n_sides"]
        vts = np.zeros((nv + 1, 3), np.float32)
        e = np.zeros((ne + 1, 2), np.int32)
        fo = np.zeros(nf + 1, np.int64)
        fv = np.zeros(nsd + 1, np.int32)
        fs = np.zeros((nf + 1, 2), np.int32)
        st = np.zeros((ns + 1, 3), np.float32)
        r = np.zeros(nv + 1, np.float32)
        val = np.zeros(nv + 1, np.uint8)
        self.ctx.check(self.ctx._lib.vc_vdiag_get(
            self.ctx.handle, self.handle, _p(vts), _p(e), _p(fo), _p(fv), _p(fs), _p(st),
            _p(r) if i["radii"] else None, _p(val)))
        return {"vts": vts[:nv], "edges": e[:ne], "face_off": fo, "face_vts": fv[:nsd],
                "face_sites": fs[:nf], "sites": st[:ns],
                "radii": r[:nv] if i["radii"] else None, "valid": val[:nv]}

    def extract_inside(self, all_valid=False, keep_complex=False):
        """(vts_idx, vm, edges (n, 2), em, tris (n, 3), from_fi, fm[, CellComplex
        handle]) of extractInsideWithMeasure(LAMBDA) on the device"""
        i = self.info()
        tcap = max(i["n_sides"] - 2 * i["nf"], 0)
        vcap, ecap = i["nv"], i["ne"] + 3 * tcap
        vi = np.zeros(vcap + 1, np.int32)
        vm = np.zeros(vcap + 1, np.float32)
        eo = np.zeros((ecap + 1, 2), np.int32)
        em = np.zeros(ecap + 1, np.float32)
        tr = np.zeros((tcap + 1, 3), np.int32)
        ff = np.zeros(tcap + 1, np.int32)
        fm = np.zeros(tcap + 1, np.float32)
        n1, n2, n3 = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        cc = ctypes.c_void_p()
        self.ctx.check(self.ctx._lib.vc_vdiag_extract_inside(
            self.ctx.handle, self.handle, int(bool(all_valid)), _p(vi), _p(vm), vcap,
            ctypes.byref(n1), _p(eo), _p(em), ecap, ctypes.byref(n2), _p(tr), _p(ff), _p(fm),
            tcap, ctypes.byref(n3), ctypes.byref(cc) if keep_complex else None))
        a, b, c = n1.value, n2.value, n3.value
        out = (vi[:a], vm[:a], eo[:b], em[:b], tr[:c], ff[:c], fm[:c])
        if keep_complex:
            h = CellComplex.__new__(CellComplex)
            h.ctx, h.nv, h.handle = self.ctx, a, cc
            h.face_off = np.arange(c + 1, dtype=np.int64) * 3
            h.face_vts = tr[:c].reshape(-1)
            out = out + (h,)
        return out

    def release(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.vc_vdiag_release(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class CellComplexThin:
    """CellComplexThinning over a CellComplex on the device (vc_ccthin)."""

    def __init__(self, cc):
        self.cc = cc
        self.ctx = cc.ctx
        h = ctypes.c_void_p()
        self.ctx.check(self.ctx._lib.vc_ccthin_create(self.ctx.handle, cc.handle,
                                                      ctypes.byref(h)))
        self.handle = h

    def assign(self, em, fm):
        em = np.ascontiguousarray(em, np.float32)
        fm = np.ascontiguousarray(fm, np.float32)
        self.ctx.check(self.ctx._lib.vc_ccthin_assign(self.ctx.handle, self.handle, _p(em),
                                                      _p(fm)))

    def prune(self, f_t, l_t):
        """Returns (status, queue generations); status VC_ERR_INVALID for a
        degenerate complex."""
        g = ctypes.c_int64()
        rc = self.ctx._lib.vc_ccthin_prune(self.ctx.handle, self.handle, float(f_t),
                                           float(l_t), ctypes.byref(g))
        if rc not in (VC_OK, VC_ERR_INVALID, VC_ERR_STATE):
            self.ctx.check(rc)
        return rc, g.value

    def removed(self):
        nv, ne, nf, _ = self.cc.sizes()
        rv, re, rf = (np.zeros(max(nv, 1), np.uint8), np.zeros(max(ne, 1), np.uint8),
                      np.zeros(max(nf, 1), np.uint8))
        ref = np.zeros(max(ne, 1), np.int32)
        self.ctx.check(self.ctx._lib.vc_ccthin_removed(self.ctx.handle, self.handle, _p(rv),
                                                       _p(re), _p(rf), _p(ref)))
        return rv[:nv], re[:ne], rf[:nf], ref[:ne]

    def remaining(self):
        """(kept vertex ids, edges (n, 2), triangles (n, 3), from_fi), sized by
        a count query first"""
        L = self.ctx._lib
        n = [ctypes.c_int64() for _ in range(3)]
        self.ctx.check(L.vc_ccthin_remaining(self.ctx.handle, self.handle, None, 0,
                                             ctypes.byref(n[0]), None, 0, ctypes.byref(n[1]),
                                             None, None, 0, ctypes.byref(n[2])))
        nv, ne, nt = (x.value for x in n)
        vi = np.zeros(max(nv, 1), np.int32)
        e = np.zeros((max(ne, 1), 2), np.int32)
        t = np.zeros((max(nt, 1), 3), np.int32)
        ff = np.zeros(max(nt, 1), np.int32)
        self.ctx.check(L.vc_ccthin_remaining(self.ctx.handle, self.handle, _p(vi), nv,
                                             ctypes.byref(n[0]), _p(e), ne, ctypes.byref(n[1]),
                                             _p(t), _p(ff), nt, ctypes.byref(n[2])))
        return vi[:nv], e[:ne], t[:nt], ff[:nt]

    def remained(self):
        """(getRemainedVts, getRemainedEdges, getRemainedFaces)"""
        L = self.ctx._lib
        n = [ctypes.c_int64() for _ in range(3)]
        self.ctx.check(L.vc_ccthin_remained(self.ctx.handle, self.handle, None, 0,
                                            ctypes.byref(n[0]), None, 0, ctypes.byref(n[1]),
                                            None, 0, ctypes.byref(n[2])))
        a = [np.zeros(max(x.value, 1), np.int32) for x in n]
        c = [x.value for x in n]
        self.ctx.check(L.vc_ccthin_remained(self.ctx.handle, self.handle, _p(a[0]), c[0],
                                            ctypes.byref(n[0]), _p(a[1]), c[1],
                                            ctypes.byref(n[1]), _p(a[2]), c[2],
                                            ctypes.byref(n[2])))
        return tuple(x[:k] for x, k in zip(a, c))

    def remaining_euler(self):
        t = CcTopo()
        self.ctx.check(self.ctx._lib.vc_ccthin_remaining_euler(self.ctx.handle, self.handle,
                                                               ctypes.byref(t)))
        return t.as_dict()

    def release(self):
        if getattr(self, "handle", None) and self.ctx.handle:
            self.ctx._lib.vc_ccthin_release(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def euler_char(ctx, nv, edges, tris):
    """util::computeEulerChar(vts, edges, tris) on the device (vc_euler_char)."""
    e = np.ascontiguousarray(np.asarray(edges, np.int32).reshape(-1, 2))
    t = np.ascontiguousarray(np.asarray(tris, np.int32).reshape(-1, 3))
    out = CcTopo()
    ctx.check(ctx._lib.vc_euler_char(ctx.handle, int(nv), _p(e), len(e), _p(t), len(t),
                                     ctypes.byref(out)))
    return out.as_dict()


# ---------------------------------------------------------------- Voronoi side
def _faces_csr(faces):
    """Face polygons as (offsets int64, vertex ids int32); a 2-D array of
    equal-size polygons is converted without a Python loop."""
    if isinstance(faces, np.ndarray) and faces.ndim == 2:
        nf, k = faces.shape
        foff = np.arange(nf + 1, dtype=np.int64) * k
        fv = np.ascontiguousarray(faces, np.int32).ravel()
    else:
        foff = np.zeros(len(faces) + 1, np.int64)
        foff[1:] = np.cumsum([len(f) for f in faces])
        fv = (np.concatenate([np.asarray(f, np.int32) for f in faces])
              if len(faces) else np.zeros(0, np.int32))
    return foff, np.ascontiguousarray(fv if fv.size else np.zeros(1, np.int32))


def _complex_arrays(nv, edges, faces_vts):
    e = np.ascontiguousarray(np.array(edges, np.int32).reshape(-1, 2))
    foff, fv = _faces_csr(faces_vts)
    return e, foff, fv


def voro_tag_points(vol, pts):
    """vc_voro_tag_points: inside flag of the voxel nearest to each point."""
    pts = np.ascontiguousarray(np.asarray(pts, np.float32).reshape(-1, 3))
    out = np.zeros(max(len(pts), 1), np.uint8)
    vol.ctx.check(library().vc_voro_tag_points(vol.ctx.handle, vol.handle, _p(pts), len(pts),
                                               _p(out)))
    return out[:len(pts)]


def voro_measures(ctx, nv, edges, face_sites, sites, face_valid, ve_off, ve, ef_off, ef):
    """vc_voro_measures: (lam_v, lam_e, lam_f) on the device."""
    e = np.array(edges, np.int32).reshape(-1, 2)
    fs = np.ascontiguousarray(np.array(face_sites, np.int32).reshape(-1, 2))
    st = np.ascontiguousarray(np.asarray(sites, np.float32).reshape(-1, 3))
    fvd = np.ascontiguousarray(np.asarray(face_valid, np.uint8))
    arrs = [np.ascontiguousarray(np.asarray(a, t)) for a, t in
            ((ve_off, np.int64), (ve, np.int32), (ef_off, np.int64), (ef, np.int32))]
    arrs = [a if a.size else np.zeros(1, a.dtype) for a in arrs]
    nf, ne = len(fs), len(e)
    lv, le, lf = (np.zeros(max(n, 1), np.float32) for n in (nv, ne, nf))
    ctx.check(library().vc_voro_measures(ctx.handle, nv, ne, nf, _p(st), len(st), _p(fs), _p(fvd),
                                         *[_p(a) for a in arrs], _p(lv), _p(le), _p(lf)))
    return lv[:nv], le[:ne], lf[:nf]


def voro_extract_inside(nv, edges, faces_vts, v_valid, lam_e, lam_f):
    """vc_voro_extract_inside: (vts_idx, edges, em, tris, from_fi, fm)."""
    e, foff, fv = _complex_arrays(nv, edges, faces_vts)
    vv = np.ascontiguousarray(np.asarray(v_valid, np.uint8))
    le = np.ascontiguousarray(np.asarray(lam_e, np.float32))
    lf = np.ascontiguousarray(np.asarray(lam_f, np.float32))
    le = le if le.size else np.zeros(1, np.float32)
    lf = lf if lf.size else np.zeros(1, np.float32)
    L = library()
    n1, n2, n3 = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    rc = L.vc_voro_extract_inside(nv, len(e), len(faces_vts), _p(e), _p(foff), _p(fv), _p(vv),
                                  _p(le), _p(lf), None, 0, ctypes.byref(n1), None, None, 0,
                                  ctypes.byref(n2), None, None, None, 0, ctypes.byref(n3))
    if rc not in (VC_OK, VC_ERR_CAPACITY):
        raise VcError(rc, "vc_voro_extract_inside")
    vi = np.zeros(max(n1.value, 1), np.int32)
    eo = np.zeros((max(n2.value, 1), 2), np.int32)
    em = np.zeros(max(n2.value, 1), np.float32)
    tr = np.zeros((max(n3.value, 1), 3), np.int32)
    ff = np.zeros(max(n3.value, 1), np.int32)
    fm = np.zeros(max(n3.value, 1), np.float32)
    rc = L.vc_voro_extract_inside(nv, len(e), len(faces_vts), _p(e), _p(foff), _p(fv), _p(vv),
                                  _p(le), _p(lf), _p(vi), n1.value, ctypes.byref(n1), _p(eo),
                                  _p(em), n2.value, ctypes.byref(n2), _p(tr), _p(ff), _p(fm),
                                  n3.value, ctypes.byref(n3))
    if rc != VC_OK:
        raise VcError(rc, "vc_voro_extract_inside")
    return (vi[:n1.value], eo[:n2.value], em[:n2.value], tr[:n3.value], ff[:n3.value],
            fm[:n3.value])


def voro_extract_inside_device(ctx, nv, edges, faces_vts, v_valid, lam_e, lam_f):
    """vc_voro_extract_inside_device: the same outputs as voro_extract_inside,
    computed on the device (one call: the capacities are the upper bounds)."""
    e, foff, fv = _complex_arrays(nv, edges, faces_vts)
    vv = np.ascontiguousarray(np.asarray(v_valid, np.uint8))
    vv = vv if vv.size else np.zeros(1, np.uint8)
    le = np.ascontiguousarray(np.asarray(lam_e, np.float32))
    lf = np.ascontiguousarray(np.asarray(lam_f, np.float32))
    le = le if le.size else np.zeros(1, np.float32)
    lf = lf if lf.size else np.zeros(1, np.float32)
    nt_max = int(np.maximum(np.diff(foff) - 2, 0).sum()) if len(foff) > 1 else 0
    ne_max = len(e) + 3 * nt_max
    vi = np.zeros(max(nv, 1), np.int32)
    eo = np.zeros((max(ne_max, 1), 2), np.int32)
    em = np.zeros(max(ne_max, 1), np.float32)
    tr = np.zeros((max(nt_max, 1), 3), np.int32)
    ff = np.zeros(max(nt_max, 1), np.int32)
    fm = np.zeros(max(nt_max, 1), np.float32)
    n1, n2, n3 = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    ctx.check(library().vc_voro_extract_inside_device(
        ctx.handle, nv, len(e), len(foff) - 1, _p(e), _p(foff), _p(fv), _p(vv), _p(le), _p(lf),
        _p(vi), len(vi), ctypes.byref(n1), _p(eo), _p(em), len(eo), ctypes.byref(n2), _p(tr),
        _p(ff), _p(fm), len(tr), ctypes.byref(n3)))
    return (vi[:n1.value], eo[:n2.value], em[:n2.value], tr[:n3.value], ff[:n3.value],
            fm[:n3.value])


def voro_merge_close(vts, edges, faces_vts, v_valid, eps, only_inside):
    """vc_voro_merge_close: (old_to_new, new_vts, new_faces, remove_face,
    new_edges, keep_edge)."""
    return _merge_close(vts, edges, faces_vts, v_valid, eps, only_inside, None)[:6]


def voro_merge_close_device(ctx, vts, edges, faces_vts, v_valid, eps, only_inside):
    """vc_voro_merge_close_device: the same six outputs plus whether the
    device path ran (nothing within eps along an edge)."""
    return _merge_close(vts, edges, faces_vts, v_valid, eps, only_inside, ctx)


def _merge_close(vts, edges, faces_vts, v_valid, eps, only_inside, ctx):
    call, out = merge_close_call(vts, edges, faces_vts, v_valid, eps, only_inside, ctx)
    rc = call()
    if rc != VC_OK:
        raise VcError(rc, "vc_voro_merge_close")
    return out()


def merge_close_call(vts, edges, faces_vts, v_valid, eps, only_inside, ctx):
    """(call, outputs): call() runs vc_voro_merge_close (ctx None) or
    vc_voro_merge_close_device on arrays prepared once and returns the
    status; outputs() converts the results (the wrapper's lists)."""
    vts = np.ascontiguousarray(np.asarray(vts, np.float32).reshape(-1, 3))
    nv = len(vts)
    e, foff, fv = _complex_arrays(nv, edges, faces_vts)
    # edges of each vertex in edge order (CSR)
    rows = e.ravel().astype(np.int64)
    cols = np.repeat(np.arange(len(e), dtype=np.int32), 2)
    order = np.lexsort((cols, rows))
    ve_off = np.zeros(nv + 1, np.int64)
    np.add.at(ve_off, rows + 1, 1)
    ve_off = np.cumsum(ve_off)
    ve = np.ascontiguousarray(cols[order] if len(cols) else np.zeros(1, np.int32), np.int32)
    vv = np.ascontiguousarray(np.asarray(v_valid, np.uint8)) if v_valid is not None else \
        np.ones(max(nv, 1), np.uint8)
    nf = len(faces_vts)
    o2n = np.zeros(max(nv, 1), np.int32)
    nvts = np.zeros(max(nv, 1), np.int32)
    nn = ctypes.c_int()
    nfo = np.zeros(nf + 1, np.int64)
    nfv = np.zeros(max(len(fv), 1), np.int32)
    rem = np.zeros(max(nf, 1), np.uint8)
    keep = np.zeros(max(len(e), 1), np.uint8)
    cap = len(e) + nf + 1
    ne_out = np.zeros((cap, 2), np.int32)
    nne = ctypes.c_int64()
    args = (nv, len(e), nf, _p(vts), _p(e), _p(foff), _p(fv), _p(ve_off), _p(ve), _p(vv),
            float(eps), 1 if only_inside else 0, _p(o2n), _p(nvts), ctypes.byref(nn), _p(nfo),
            _p(nfv), _p(rem), _p(ne_out), cap, ctypes.byref(nne), _p(keep))
    dev = ctypes.c_int(0)

    def call():
        if ctx is None:
            return library().vc_voro_merge_close(*args)
        return library().vc_voro_merge_close_device(ctx.handle, *args, ctypes.byref(dev))

    def outputs():
        kept = int((rem[:nf] == 0).sum())
        faces = [list(nfv[nfo[i]:nfo[i + 1]]) for i in range(kept)]
        return (list(o2n[:nv]), list(nvts[:nn.value]), faces, list(rem[:nf]),
                [tuple(x) for x in ne_out[:nne.value]], list(keep[:len(e)]), bool(dev.value))
    return call, outputs


def voro_radii(ctx, vts, faces_vts, face_sites, sites):
    """vc_voro_radii: per-vertex radius (last face holding the vertex decides)."""
    vts = np.ascontiguousarray(np.asarray(vts, np.float32).reshape(-1, 3))
    nv = len(vts)
    _, foff, fv = _complex_arrays(nv, [], faces_vts)
    fs = np.ascontiguousarray(np.array(face_sites, np.int32).reshape(-1, 2))
    st = np.ascontiguousarray(np.asarray(sites, np.float32).reshape(-1, 3))
    fs = fs if fs.size else np.zeros((1, 2), np.int32)
    st1 = st if st.size else np.zeros((1, 3), np.float32)
    r = np.zeros(max(nv, 1), np.float32)
    ctx.check(library().vc_voro_radii(ctx.handle, nv, len(faces_vts), _p(vts), _p(foff), _p(fv),
                                      _p(st1), len(st), _p(fs), _p(r)))
    return r[:nv]


def voro_from_ft(vol):
    """vc_voro_from_ft: the discrete Voronoi diagram of the feature transform
    (vol.edt() first): (vts (V, 3) float32, faces (F, 4) int32, face_sites
    (F, 2) int32, sites (S, 3) float32)."""
    ctx = vol.ctx
    L = library()
    nv, nf, ns = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    ctx.check(L.vc_voro_from_ft(ctx.handle, vol.handle, None, 0, ctypes.byref(nv), None,
                                None, 0, ctypes.byref(nf), None, 0, ctypes.byref(ns)))
    vts = np.zeros((max(nv.value, 1), 3), np.float32)
    faces = np.zeros((max(nf.value, 1), 4), np.int32)
    fs = np.zeros((max(nf.value, 1), 2), np.int32)
    sites = np.zeros((max(ns.value, 1), 3), np.float32)
    ctx.check(L.vc_voro_from_ft(ctx.handle, vol.handle, _p(vts), len(vts), ctypes.byref(nv),
                                _p(faces), _p(fs), len(faces), ctypes.byref(nf), _p(sites),
                                len(sites), ctypes.byref(ns)))
    return vts[:nv.value], faces[:nf.value], fs[:nf.value], sites[:ns.value]


def voro_part_inside(nv, faces_vts, v_valid):
    """vc_voro_part_inside: (vts_idx, edges, tris) of the valid part."""
    _, foff, fv = _complex_arrays(nv, [], faces_vts)
    vv = np.ascontiguousarray(np.asarray(v_valid, np.uint8))
    vv = vv if vv.size else np.zeros(1, np.uint8)
    L = library()
    n1, n2, n3 = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    rc = L.vc_voro_part_inside(nv, len(faces_vts), _p(foff), _p(fv), _p(vv), None, 0,
                               ctypes.byref(n1), None, 0, ctypes.byref(n2), None, 0,
                               ctypes.byref(n3))
    if rc not in (VC_OK, VC_ERR_CAPACITY):
        raise VcError(rc, "vc_voro_part_inside")
    vi = np.zeros(max(n1.value, 1), np.int32)
    eo = np.zeros((max(n2.value, 1), 2), np.int32)
    tr = np.zeros((max(n3.value, 1), 3), np.uint32)
    rc = L.vc_voro_part_inside(nv, len(faces_vts), _p(foff), _p(fv), _p(vv), _p(vi), n1.value,
                               ctypes.byref(n1), _p(eo), n2.value, ctypes.byref(n2), _p(tr),
                               n3.value, ctypes.byref(n3))
    if rc != VC_OK:
        raise VcError(rc, "vc_voro_part_inside")
    return vi[:n1.value], eo[:n2.value], tr[:n3.value]


# ---------------------------------------------------------------- output files
def _f32(a, k):
    a = np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, k))
    return a, (a if a.size else np.zeros((1, k), np.float32))


def _idx(a, k, dt):
    a = np.ascontiguousarray(np.asarray(a, dt).reshape(-1, k))
    return a, (a if a.size else np.zeros((1, k), dt))


def _mat(m):
    if m is None:
        return None, None
    m = np.ascontiguousarray(np.asarray(m, np.float64).reshape(16))
    return m, _p(m)


def _io(rc, what):
    if rc != VC_OK:
        raise VcError(rc, what)


def write_ply(path, vts, edges=(), tris=()):
    """vc_write_ply: ASCII PLY (vertex x y z, edge vertex1 vertex2, face
    vertex_indices) as the reference writes it."""
    v, v1 = _f32(vts, 3)
    e, e1 = _idx(edges, 2, np.int32)
    t, t1 = _idx(tris, 3, np.int32)
    _io(library().vc_write_ply(os.fsencode(path), _p(v1), len(v), _p(e1), len(e), _p(t1),
                               len(t)), "vc_write_ply")


def write_radii(path, vts, radii, vox_to_model=None):
    """vc_write_radii: .r file, vertices and radii through vox_to_model
    (16 doubles, column-major; None = identity)."""
    v, v1 = _f32(vts, 3)
    r, r1 = _f32(radii, 1)
    if len(r) != len(v):
        raise ValueError("one radius per vertex")
    m, mp = _mat(vox_to_model)
    _io(library().vc_write_radii(os.fsencode(path), _p(v1), _p(r1), len(v), mp),
        "vc_write_radii")


def write_dotma(path, vts, radii, edges=(), tris=()):
    """vc_write_dotma: QMAT .ma file."""
    v, v1 = _f32(vts, 3)
    r, r1 = _f32(radii, 1)
    if len(r) != len(v):
        raise ValueError("one radius per vertex")
    e, e1 = _idx(edges, 2, np.int32)
    t, t1 = _idx(tris, 3, np.uint32)
    _io(library().vc_write_dotma(os.fsencode(path), _p(v1), _p(r1), len(v), _p(e1), len(e),
                                 _p(t1), len(t)), "vc_write_dotma")


def write_off(path, vts, tris=(), vox_to_model=None):
    """vc_write_off: OFF as trimesh2 writes it (vertices through
    vox_to_model)."""
    v, v1 = _f32(vts, 3)
    t, t1 = _idx(tris, 3, np.uint32)
    m, mp = _mat(vox_to_model)
    _io(library().vc_write_off(os.fsencode(path), _p(v1), len(v), _p(t1), len(t), mp),
        "vc_write_off")


def write_tetnodes(path, vts):
    """vc_write_tetnodes: TetGen .node file."""
    v, v1 = _f32(vts, 3)
    _io(library().vc_write_tetnodes(os.fsencode(path), _p(v1), len(v)), "vc_write_tetnodes")


def xform_apply(m, pts):
    """vc_xform_apply: trimesh-style 4x4 transform of float points."""
    p, p1 = _f32(pts, 3)
    out = np.zeros_like(p1)
    mm, mp = _mat(m)
    _io(library().vc_xform_apply(mp, _p(p1), len(p), _p(out)), "vc_xform_apply")
    return out[:len(p)]


# ---------------------------------------------------------------- octree input
class Octree:
    """A parsed .sof / .sog octree (vc_octree)."""

    def __init__(self, data, sog=False):
        buf = (ctypes.c_uint8 * len(data)).from_buffer_copy(data)
        h = ctypes.c_void_p()
        self.info = OctreeInfo()
        rc = library().vc_octree_parse(buf, len(data), 1 if sog else 0, ctypes.byref(h),
                                       ctypes.byref(self.info))
        if rc != VC_OK:
            raise VcError(rc, "vc_octree_parse")
        self.handle = h

    @classmethod
    def read(cls, path):
        with open(path, "rb") as f:
            data = f.read()
        return cls(data, sog=path.lower().endswith(".sog"))

    def value(self, x, y, z):
        v = ctypes.c_double()
        rc = library().vc_octree_value(self.handle, x, y, z, ctypes.byref(v))
        if rc != VC_OK:
            raise VcError(rc, "vc_octree_value")
        return v.value

    def leaf_centers(self):
        n = ctypes.c_int64()
        library().vc_octree_leaf_centers(self.handle, None, 0, ctypes.byref(n))
        out = np.zeros((max(n.value, 1), 3), np.float32)
        rc = library().vc_octree_leaf_centers(self.handle, _p(out), n.value, ctypes.byref(n))
        if rc != VC_OK:
            raise VcError(rc, "vc_octree_leaf_centers")
        return out[:n.value]

    def volume(self, ctx):
        """Dense (R+1)^3 occupancy built on the device, as a Volume."""
        h = ctypes.c_void_p()
        ctx.check(library().vc_volume_from_octree(ctx.handle, self.handle, ctypes.byref(h)))
        n = self.info.resolution + 1
        v = Volume.__new__(Volume)
        v.ctx, v.shape, v.handle = ctx, (n, n, n), h
        return v

    def free(self):
        if getattr(self, "handle", None):
            library().vc_octree_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
