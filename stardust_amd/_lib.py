"""ctypes binding of libstardust_hip.so (the C ABI declared in include/sd_api.h).

The library is the only compute path: if it is missing, or no HIP device is
visible, every entry point raises. There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstardust_hip.so")

SD_OK = 0
SD_ERR_ARG = 1
SD_ERR_HIP = 2
SD_ERR_CAPACITY = 3
SD_ERR_STATE = 4
SD_ERR_NO_DEVICE = 5
SD_ERR_NONFINITE = 6


class SdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sd status %d: %s" % (code, msg))
        self.code = code


class Grid(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("cell", C.c_float), ("n", C.c_int * 3)]

    @staticmethod
    def make(lo, cell, n):
        g = Grid()
        for a in range(3):
            g.lo[a] = float(lo[a])
            g.n[a] = int(n[a])
        g.cell = float(cell)
        return g

    @property
    def n_cells(self):
        return self.n[0] * self.n[1] * self.n[2]


class ContactParams(C.Structure):
    _fields_ = [("kn", C.c_float), ("damping_ratio", C.c_float), ("kt_ratio", C.c_float), ("mu", C.c_float),
                ("dt", C.c_float), ("gravity", C.c_float * 3), ("walls", C.c_int), ("wall_lo", C.c_float * 3),
                ("wall_hi", C.c_float * 3), ("max_contacts", C.c_int)]

    @staticmethod
    def make(kn, damping_ratio, mu, dt, gravity=(0.0, 0.0, -9.81), walls=None, kt_ratio=2.0 / 7.0,
             max_contacts=16):
        c = ContactParams()
        c.kn, c.damping_ratio, c.kt_ratio, c.mu, c.dt = kn, damping_ratio, kt_ratio, mu, dt
        for a in range(3):
            c.gravity[a] = gravity[a]
        if walls is None:
            c.walls = 0
        else:
            c.walls = 1
            for a in range(3):
                c.wall_lo[a] = walls[0][a]
                c.wall_hi[a] = walls[1][a]
        c.max_contacts = max_contacts
        return c


class SphereParams(C.Structure):
    _fields_ = [("grid", Grid), ("contact", ContactParams), ("key64", C.c_int), ("use_graph", C.c_int),
                ("tile_items", C.c_int)]


class BatchParams(C.Structure):
    _fields_ = [("use_graph", C.c_int), ("tile_items", C.c_int), ("contact_block", C.c_int)]


class SceneDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("grid", Grid), ("contact", ContactParams)]


_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_SZ = C.c_size_t

# name -> argtypes (restype is always sd_status / int unless listed in _RESTYPE)
_SIGS = {
    "sd_last_error": [],
    "sd_version": [],
    "sd_device_count": [C.POINTER(C.c_int)],
    "sd_malloc": [C.POINTER(C.c_void_p), _SZ],
    "sd_free": [_P],
    "sd_memcpy_htod": [_P, _P, _SZ, _P],
    "sd_memcpy_dtoh": [_P, _P, _SZ, _P],
    "sd_memcpy_dtod": [_P, _P, _SZ, _P],
    "sd_memset_u32": [_P, C.c_uint32, _SZ, _P],
    "sd_stream_create": [C.POINTER(C.c_void_p)],
    "sd_stream_destroy": [_P],
    "sd_stream_synchronize": [_P],
    "sd_event_create": [C.POINTER(C.c_void_p)],
    "sd_event_destroy": [_P],
    "sd_event_record": [_P, _P],
    "sd_event_elapsed_ms": [_P, _P, C.POINTER(C.c_float)],
    "sd_hash_keys_u32": [_P, _P, _P, _I, C.POINTER(Grid), _P, _P, _P],
    "sd_hash_keys_u64": [_P, _P, _P, _I, C.POINTER(Grid), _P, _P],
    "sd_radix_sort_temp_bytes": [_I, C.POINTER(_SZ)],
    "sd_radix_sort_pairs_u32": [_P, _P, _P, _P, _I, _I, _I, _P, _SZ, _P],
    "sd_radix_sort_keys_u64": [_P, _P, _I, _I, _I, _P, _SZ, _P],
    "sd_radix_sort_pairs_u32_noalloc": [_P, _P, _P, _P, _I, _I, _I, _P],
    "sd_segmented_sort_temp_bytes": [_I, C.POINTER(_SZ)],
    "sd_segmented_sort_pairs_u32": [_P, _P, _P, _P, _I, _P, _I, _I, _P, _SZ, _P],
    "sd_body_reduce": [_P, _P, _P, _I, _P, _P, _P, _P, _P, _P],
    "sd_integrate_bodies": [_P, _P, _P, _P, _P, _P, _P, _I, C.POINTER(ContactParams), _P],
    "sd_mlp_forward": [_P, _I, _P, _P, _P],
    "sd_compat_pair_forces": [_P, _P, _P, _P, _I, _F, _F, _F, _P, _P, _P],
    "sd_compat_inverse3x3": [_P, _I, _P, _P],
    "sd_compat_angular_velocity": [_P, _P, _P, _I, _P, _P],
    "sd_cell_ranges_u32": [_P, _I, _I, _P, _P],
    "sd_cell_ranges_u64": [_P, _I, _I, _P, _P],
    "sd_neighbor_pairs": [_P, _I, C.POINTER(Grid), _P, _F, _P, _I, C.POINTER(C.c_int), _P, _SZ, _P],
    "sd_neighbor_pairs_scratch_bytes": [_I, C.POINTER(_SZ)],
    "sd_contact_forces": [_P, _P, _P, _P, _I, C.POINTER(Grid), _P, C.POINTER(ContactParams), _P, _P, _P, _P, _P,
                          _P, _P, _P, _P],
    "sd_integrate_spheres": [_P, _P, _P, _P, _P, _I, C.POINTER(ContactParams), _P],
    "sd_permute_spheres": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "sd_ctx_create": [C.POINTER(SphereParams), C.POINTER(C.c_void_p)],
    "sd_ctx_destroy": [_P],
    "sd_ctx_set_contact": [_P, C.POINTER(ContactParams)],
    "sd_ctx_upload_spheres": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "sd_ctx_step": [_P, _I],
    "sd_ctx_profile": [_P, _I, _P, _I],
    "sd_ctx_download_spheres": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "sd_ctx_contact_count": [_P, C.POINTER(C.c_longlong)],
    "sd_ctx_cell_changes": [_P, C.POINTER(_I), C.POINTER(_I)],
    "sd_ctx_diagnostics": [_P, _P],
    "sd_ctx_stream": [_P, C.POINTER(C.c_void_p)],
    "sd_ctx_state_bytes": [_P, C.POINTER(_SZ)],
    "sd_ctx_save": [_P, _P, _SZ],
    "sd_ctx_load": [_P, _P, _SZ],
    "sd_ctx_remove_spheres": [_P, _P, _I],
    "sd_ctx_add_spheres": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "sd_ctx_autotune": [_P, _I, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)],
    "sd_batch_create": [C.POINTER(BatchParams), C.POINTER(C.c_void_p)],
    "sd_batch_upload": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "sd_gen_spheres": [C.c_uint64, _I, _P, _P, _F, _F, _F, _F, _P, _P, _P, _P],
    "sd_gen_hulls": [C.c_uint64, _I, _P, _P, _I, _I, _F, _F, _F, _P, _P, _P, _P, _P, _P, _I, C.POINTER(C.c_int)],
    "sd_hull_aabbs": [_P, _P, _P, _P, _P, _I, _P, _P, _P],
    "sd_sap_scratch_bytes": [_I, C.POINTER(_SZ)],
    "sd_sap_pairs": [_P, _P, _I, _P, _I, C.POINTER(C.c_int), C.POINTER(C.c_int), _P, _SZ, _P],
    "sd_sap_batched_scratch_bytes": [_I, _I, C.POINTER(_SZ)],
    "sd_sap_pairs_batched": [_P, _P, _I, _P, _P, _I, _P, _I, C.POINTER(C.c_int), _P, _P, _SZ, _P],
    "sd_gjk_scratch_bytes": [_I, C.POINTER(_SZ)],
    "sd_gjk": [_P, _I, _P, _P, _P, _P, _P, _I, _I, _P, _I, _I, _F, _P, _P, _P, _P, _P, _SZ, _P],
    "sd_global_domain": [_P, _P, _I, _P, _P, _P],
    "sd_morton30": [_P, _P, _I, _P, _P, _P],
    "sd_lbvh_build": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "sd_lbvh_query_scratch_bytes": [_I, C.POINTER(_SZ)],
    "sd_lbvh_query": [_P, _I, _P, _P, _P, _P, _P, _P, _I, C.POINTER(C.c_int), _P, _P, _SZ, _P],
    "sd_clump_create": [_P, C.POINTER(C.c_void_p)],
    "sd_clump_destroy": [_P],
    "sd_clump_set_params": [_P, _P],
    "sd_clump_upload": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P],
    "sd_clump_step": [_P, _I],
    "sd_clump_profile": [_P, _I, _P, _I],
    "sd_clump_download": [_P, _P, _P, _P, _P],
    "sd_clump_pairs": [_P, _P, _I, C.POINTER(C.c_int)],
    "sd_clump_spheres": [_P, _P],
    "sd_hull_create": [_P, _P],
    "sd_batch_attach_hulls": [_P, _P],
    "sd_ctx_set_persistent": [_P, _I],
    "sd_hull_destroy": [_P],
    "sd_hull_set_params": [_P, _P],
    "sd_hull_upload": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _I],
    "sd_hull_step": [_P, _I],
    "sd_hull_set_warm": [_P, _P, _I],
    "sd_hull_narrowphase_stats": [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_longlong)],
    "sd_hull_download": [_P, _P, _P, _P, _P],
    "sd_hull_contacts": [_P, _P, _P, _P, _I, C.POINTER(C.c_int)],
    "sd_clump_state_bytes": [_P, C.POINTER(_SZ)],
    "sd_clump_save": [_P, _P, _SZ],
    "sd_clump_load": [_P, _P, _SZ],
    "sd_clump_remove_bodies": [_P, _P, _I],
    "sd_clump_add_bodies": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _I, _P],
    "sd_gen_clumps": [C.c_uint64, _I, _I, _F, _F, _I, _I, _F, _F, _F, _F, _F, _P, _P, _P, _P, _P, _P, _P, _P,
                      C.POINTER(C.c_int)],
    "sd_mlp_param_count": [],
    "sd_mlp_packed_bytes": [],
    "sd_hull_descriptors": [_P, _P, _P, _P, _I, _P, _P],
    "sd_mlp_features": [_P, _I, _P, _P, _P, _P, _P],
    "sd_mlp_targets": [_P, _I, _P, _P, _P, _P],
    "sd_mlp_init": [_P, C.c_uint64, _P],
    "sd_mlp_train_scratch_bytes": [_I, C.POINTER(_SZ)],
    "sd_mlp_forward_backward": [_P, _P, _P, _I, _P, _P, _P, _P, _SZ, _P],
    "sd_adam_step": [_P, _P, _P, _P, _I, _F, _F, _F, _F, _I, _P],
    "sd_mlp_pack_bf16": [_P, _P, _P],
    "sd_mlp_infer_bf16": [_P, _I, _P, _P, _P, _P, _P],
}
GJK_HIT, GJK_CONVERGED, GJK_BUDGET, GJK_EPA_LIMIT, GJK_DEGENERATE, GJK_WARM_USED = 1, 2, 4, 8, 16, 32
_RESTYPE = {"sd_last_error": C.c_char_p, "sd_version": C.c_char_p}

_lib = None


def load():
    """Loads the library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libstardust_hip.so is not built (%s); run `python -m stardust_amd.build`" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_SIGS)


def check(code):
    if code != SD_OK:
        raise SdError(code, load().sd_last_error().decode())
    return code


def call(name, *args):
    return check(getattr(load(), name)(*args))


def ptr(a):
    """Host pointer of a contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return C.c_void_p(a.ctypes.data)


def device_count():
    n = C.c_int(0)
    call("sd_device_count", C.byref(n))
    return n.value


def require_gpu():
    if device_count() < 1:
        raise RuntimeError("no HIP device visible: the stardust_amd path runs on the GPU only")


class DeviceBuffer:
    """A device allocation made through the C ABI (sd_malloc / sd_free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("sd_malloc", C.byref(p), max(self.nbytes, 16))
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        call("sd_memcpy_htod", self.ptr, ptr(a), a.nbytes, None)

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        call("sd_memcpy_dtoh", ptr(out), self.ptr, out.nbytes, None)
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            load().sd_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(stream=None):
    call("sd_stream_synchronize", stream)
