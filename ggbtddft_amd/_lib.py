"""ctypes binding of libggbtd.so (declarations in include/ggbtd.h).

The library is built in-tree by ``ggbtddft_amd.build.build()``.  Loading fails
loudly: there is no CPU fallback for the propagator.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libggbtd.so")

GGB_ABI_VERSION = 4
ERRORS = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "call out of order",
          -4: "device allocation failed"}

EXPORTS = ("ggb_abi_version", "ggb_create", "ggb_upload_static", "ggb_upload_orbitals",
           "ggb_set_field", "ggb_run", "ggb_download_orbitals", "ggb_download_potential",
           "ggb_debug_substep", "ggb_set_timing", "ggb_get_timing", "ggb_selftest_mfma",
           "ggb_field_ring", "ggb_set_field_slice", "ggb_upload_columns", "ggb_set_active",
           "ggb_rebuild_potential", "ggb_footprint", "ggb_selftest_copy", "ggb_set_graph",
           "ggb_spectrum", "ggb_set_kinetic_variant", "ggb_autotune", "ggb_ground_state",
           "ggb_set_channels", "ggb_get_channels", "ggb_spectrum_noalloc", "ggb_group_create",
           "ggb_group_run", "ggb_group_last_error", "ggb_group_destroy", "ggb_propagator_tables",
           "ggb_last_error", "ggb_destroy")
KIN_VARIANTS = ("default", "narrow", "tiled16", "tiled32", "split16")
NPHASE = 7
PHASES = ("half_potential_laser", "kinetic_qt", "kinetic_q", "laser_half_potential_obs",
          "density_moments", "hartree", "dv_phase")


class Problem(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_int32), ("n_r", ctypes.c_int32),
                ("n_l", ctypes.c_int32), ("n_theta", ctypes.c_int32),
                ("n_har", ctypes.c_int32), ("n_orb", ctypes.c_int32),
                ("n_inst", ctypes.c_int32), ("n_m", ctypes.c_int32),
                ("nk_max", ctypes.c_int32), ("interacting", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("device", ctypes.c_int32),
                ("dt", ctypes.c_double)]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class Static(ctypes.Structure):
    _fields_ = [("r", _dp), ("wr", _dp), ("v_gs", _dp), ("mask", _dp), ("Q", _dp),
                ("lam", _dp), ("nkeep", _ip), ("B", _dp), ("PL", _dp), ("chi", _dp),
                ("omega", _dp), ("m", _ip), ("occ", _dp), ("rho_frozen", _dp),
                ("phi", _dp), ("l0", _ip), ("eps", _dp)]


class GsProblem(ctypes.Structure):
    _fields_ = [("n_r", ctypes.c_int32), ("n_shell", ctypes.c_int32), ("sweeps", ctypes.c_int32),
                ("max_iter", ctypes.c_int32), ("interacting", ctypes.c_int32),
                ("device", ctypes.c_int32), ("Z", ctypes.c_double), ("mix", ctypes.c_double),
                ("gamma", ctypes.c_double), ("tol", ctypes.c_double)]


class GGBError(RuntimeError):
    pass


_lib = None


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("GGBTD_LIB", LIB_PATH)  # a differently built copy, for comparisons
    if not os.path.exists(path):
        raise GGBError("libggbtd.so not built (%s); run ggbtddft_amd.build.build()" % path)
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.ggb_abi_version.restype = ctypes.c_int
    lib.ggb_abi_version.argtypes = []
    lib.ggb_create.restype = ctypes.c_int
    lib.ggb_create.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(vp)]
    lib.ggb_upload_static.restype = ctypes.c_int
    lib.ggb_upload_static.argtypes = [vp, ctypes.POINTER(Static)]
    lib.ggb_upload_orbitals.restype = ctypes.c_int
    lib.ggb_upload_orbitals.argtypes = [vp, _dp]
    lib.ggb_set_field.restype = ctypes.c_int
    lib.ggb_set_field.argtypes = [vp, _dp, ctypes.c_int32]
    lib.ggb_run.restype = ctypes.c_int
    lib.ggb_run.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp]
    lib.ggb_download_orbitals.restype = ctypes.c_int
    lib.ggb_download_orbitals.argtypes = [vp, _dp]
    lib.ggb_download_potential.restype = ctypes.c_int
    lib.ggb_download_potential.argtypes = [vp, _dp]
    lib.ggb_debug_substep.restype = ctypes.c_int
    lib.ggb_debug_substep.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    lib.ggb_set_timing.restype = ctypes.c_int
    lib.ggb_set_timing.argtypes = [vp, ctypes.c_int32]
    lib.ggb_get_timing.restype = ctypes.c_int
    lib.ggb_get_timing.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_int64)]
    lib.ggb_field_ring.restype = ctypes.c_int
    lib.ggb_field_ring.argtypes = [vp, ctypes.c_int32]
    lib.ggb_set_field_slice.restype = ctypes.c_int
    lib.ggb_set_field_slice.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp]
    lib.ggb_upload_columns.restype = ctypes.c_int
    lib.ggb_upload_columns.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, _dp]
    lib.ggb_set_active.restype = ctypes.c_int
    lib.ggb_set_active.argtypes = [vp, _ip]
    lib.ggb_rebuild_potential.restype = ctypes.c_int
    lib.ggb_rebuild_potential.argtypes = [vp]
    lib.ggb_footprint.restype = ctypes.c_int
    lib.ggb_footprint.argtypes = [ctypes.POINTER(Problem), ctypes.POINTER(ctypes.c_uint64)]
    lib.ggb_selftest_copy.restype = ctypes.c_int
    lib.ggb_selftest_copy.argtypes = [ctypes.c_int32, ctypes.c_uint64, _dp]
    lib.ggb_set_kinetic_variant.restype = ctypes.c_int
    lib.ggb_set_kinetic_variant.argtypes = [vp, ctypes.c_int32]
    lib.ggb_autotune.restype = ctypes.c_int
    lib.ggb_autotune.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), _dp]
    lib.ggb_spectrum.restype = ctypes.c_int
    lib.ggb_spectrum.argtypes = [ctypes.c_int32, _dp, ctypes.c_int32, ctypes.c_double,
                                 ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                                 _dp, ctypes.c_int32, _dp, _dp]
    lib.ggb_spectrum_noalloc.restype = ctypes.c_int
    lib.ggb_spectrum_noalloc.argtypes = lib.ggb_spectrum.argtypes
    lib.ggb_ground_state.restype = ctypes.c_int
    lib.ggb_ground_state.argtypes = [ctypes.POINTER(GsProblem), _ip, _dp, _dp, _dp, _dp, _dp, _dp,
                                     _dp, _dp, _dp, _dp, _ip, _ip]
    lib.ggb_propagator_tables.restype = ctypes.c_int
    lib.ggb_propagator_tables.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp,
                                          ctypes.c_double, ctypes.c_int32, _dp, _dp, _ip, _ip]
    lib.ggb_set_channels.restype = ctypes.c_int
    lib.ggb_set_channels.argtypes = [vp, ctypes.c_int32]
    lib.ggb_get_channels.restype = ctypes.c_int
    lib.ggb_get_channels.argtypes = [vp, ctypes.c_int32, _dp]
    lib.ggb_set_graph.restype = ctypes.c_int
    lib.ggb_set_graph.argtypes = [vp, ctypes.c_int32]
    lib.ggb_selftest_mfma.restype = ctypes.c_int
    lib.ggb_selftest_mfma.argtypes = [ctypes.c_int32, _dp]
    _pp = ctypes.POINTER(_dp)
    lib.ggb_group_create.restype = ctypes.c_int
    lib.ggb_group_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.c_int32,
                                     ctypes.POINTER(vp)]
    lib.ggb_group_run.restype = ctypes.c_int
    lib.ggb_group_run.argtypes = [vp, _ip, ctypes.c_int32, _pp, _pp, _pp]
    lib.ggb_group_last_error.restype = ctypes.c_char_p
    lib.ggb_group_last_error.argtypes = [vp]
    lib.ggb_group_destroy.restype = None
    lib.ggb_group_destroy.argtypes = [vp]
    lib.ggb_last_error.restype = ctypes.c_char_p
    lib.ggb_last_error.argtypes = [vp]
    lib.ggb_destroy.restype = None
    lib.ggb_destroy.argtypes = [vp]
    if lib.ggb_abi_version() != GGB_ABI_VERSION:
        raise GGBError("libggbtd.so ABI %d, expected %d" % (lib.ggb_abi_version(), GGB_ABI_VERSION))
    _lib = lib
    return lib


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def check(rc, ctx=None, what=""):
    if rc != 0:
        lib = load()
        msg = lib.ggb_last_error(ctx)
        msg = msg.decode() if msg else ""
        raise GGBError("%s failed (%s): %s" % (what, ERRORS.get(rc, rc), msg))
