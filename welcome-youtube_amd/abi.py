"""ctypes mirror of include/oercel_mi.h and the loader of the HIP library.

The structs here must match the C header field for field; tests check the
sizes against the library (oc_abi_version) and against the header text.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "liboercel_mi.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

NTYPES = 7
MAXMLEN = 12
TYPE_NAMES = ("W", "H", "T", "N", "Ka", "Kb", "Z")
W, H, T, N, KA, KB, Z = range(7)
EV_CONV, EV_FORM, EV_BREAK = 0, 1, 2
EV_NAMES = ("conv", "form", "break")

OC_OK, OC_EINVAL, OC_ENOMEM, OC_EHIP, OC_EDEVICE, OC_ESTATE = range(6)


class oc_params(C.Structure):
    _fields_ = [
        ("L", C.c_double), ("rho", C.c_double), ("dt", C.c_double), ("lambda_", C.c_double),
        ("gamma", C.c_double), ("kT", C.c_double),
        ("k_bond", C.c_double), ("r0", C.c_double), ("k_angle", C.c_double),
        ("a", (C.c_double * 8) * 8),
        ("stir_A", C.c_double), ("kT_amp", C.c_double), ("kT_period", C.c_double),
        ("stir_n", C.c_int32), ("nsub", C.c_int32), ("lip_nH", C.c_int32), ("lip_nT", C.c_int32),
        ("warmup", C.c_int64), ("warm_end", C.c_int64), ("seed", C.c_uint64),
        ("f_lip", C.c_double), ("f_Ka", C.c_double), ("f_Kb", C.c_double), ("c_N", C.c_double),
        ("chem", C.c_int32), ("n_chem", C.c_int32),
        ("p_conv", C.c_double), ("r_conv", C.c_double), ("p_form", C.c_double), ("r_form", C.c_double),
        ("p_break", C.c_double),
        ("t_anchor", C.c_double), ("step_anchor", C.c_int64),
        ("lip_arch", C.c_int32), ("lip_tie", C.c_int32), ("r0_head", C.c_double), ("k_head", C.c_double),
        ("cos_head", C.c_double),
        ("rchem", C.c_int32), ("feed_mode", C.c_int32), ("p_ld", C.c_double), ("p_nd", C.c_double),
        ("p_feed", C.c_double), ("vent_shape", C.c_int32), ("pad2", C.c_int32), ("vent_x", C.c_double),
        ("vent_y", C.c_double), ("vent_r", C.c_double), ("vent_w", C.c_double),
        ("p_syn_a", C.c_double), ("p_syn_b", C.c_double), ("p_rep_a", C.c_double), ("p_rep_b", C.c_double),
        ("mu", C.c_double), ("p_kd", C.c_double), ("r_cat", C.c_double), ("feed_cap", C.c_double),
        ("syn_tries", C.c_int32), ("syn_relax", C.c_int32), ("syn_nN", C.c_int32), ("pad3", C.c_int32),
    ]


class oc_state(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("vx", C.POINTER(C.c_double)), ("vy", C.POINTER(C.c_double)),
        ("fx", C.POINTER(C.c_double)), ("fy", C.POINTER(C.c_double)),
        ("type", C.POINTER(C.c_uint8)), ("mpos", C.POINTER(C.c_int8)),
        ("nxt", C.POINTER(C.c_int32)), ("bond", C.POINTER(C.c_int32)),
    ]


class oc_event(C.Structure):
    _fields_ = [("step", C.c_int64), ("kind", C.c_int32), ("gid_a", C.c_uint32), ("gid_b", C.c_uint32),
                ("pad", C.c_int32), ("x", C.c_double), ("y", C.c_double)]


class oc_stats(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("time", C.c_double), ("T_kin", C.c_double), ("px", C.c_double), ("py", C.c_double),
        ("E_kin", C.c_double), ("E_pair", C.c_double), ("E_bond", C.c_double), ("max_bond", C.c_double),
        ("count", C.c_int64 * 8), ("n_lipids", C.c_int64), ("n_bonds", C.c_int64),
        ("virial", C.c_double), ("pressure", C.c_double),
    ]


class oc_rdf_spec(C.Structure):
    _fields_ = [("ta", C.c_int32), ("tb", C.c_int32), ("nbins", C.c_int32), ("pad", C.c_int32),
                ("rmax", C.c_double), ("width", C.c_double)]


def rdf_spec(ta=2, tb=2, nbins=64, rmax=3.0, width=None):
    """Smooth RDF of the type pair (ta, tb) on nbins bins of [0, rmax); the
    triweight kernel's half-width defaults to one bin."""
    return oc_rdf_spec(ta, tb, nbins, 0, rmax, width if width is not None else rmax / nbins)


EVENT_DTYPE = None  # numpy dtype, filled lazily


def event_dtype():
    import numpy as np
    global EVENT_DTYPE
    if EVENT_DTYPE is None:
        EVENT_DTYPE = np.dtype([("step", "<i8"), ("kind", "<i4"), ("gid_a", "<u4"), ("gid_b", "<u4"), ("pad", "<i4"),
                                ("x", "<f8"), ("y", "<f8")])
    return EVENT_DTYPE


def _ptr(arr, ctype):
    if arr is None:
        return C.POINTER(ctype)()
    return arr.ctypes.data_as(C.POINTER(ctype))


def make_state(arrs):
    """oc_state view of a dict of numpy arrays (x, y, vx, vy, fx, fy, type, mpos, nxt, bond); missing keys -> NULL."""
    g = arrs.get
    return oc_state(_ptr(g("x"), C.c_double), _ptr(g("y"), C.c_double), _ptr(g("vx"), C.c_double),
                    _ptr(g("vy"), C.c_double), _ptr(g("fx"), C.c_double), _ptr(g("fy"), C.c_double),
                    _ptr(g("type"), C.c_uint8), _ptr(g("mpos"), C.c_int8), _ptr(g("nxt"), C.c_int32),
                    _ptr(g("bond"), C.c_int32))


_LIB = None


def _declare(lib):
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "oc_last_error": (C.c_char_p, []),
        "oc_abi_version": (C.c_int, []),
        "oc_create": (C.c_int, [P(oc_params), P(C.c_int32), C.c_int32, P(vp)]),
        "oc_destroy": (None, [vp]),
        "oc_set_tuning": (C.c_int, [vp, C.c_int32, C.c_int32]),
        "oc_set_pair_variant": (C.c_int, [vp, C.c_int32]),
        "oc_soup_host": (C.c_int, [P(oc_params), C.c_int32, P(oc_state)]),
        "oc_start": (C.c_int, [vp, P(oc_state)]),
        "oc_set_state": (C.c_int, [vp, C.c_int32, P(oc_state), C.c_int64]),
        "oc_set_state_slots": (C.c_int, [vp, C.c_int32, P(oc_state), C.c_int64, P(C.c_uint32)]),
        "oc_get_state": (C.c_int, [vp, C.c_int32, P(oc_state)]),
        "oc_step": (C.c_int, [vp, C.c_int64, C.c_int32]),
        "oc_sort": (C.c_int, [vp]),
        "oc_forces": (C.c_int, [vp, C.c_int64, C.c_int32]),
        "oc_chem": (C.c_int, [vp, C.c_int64]),
        "oc_get_cells": (C.c_int, [vp, C.c_int32, P(C.c_int32), P(C.c_uint32), P(C.c_int32)]),
        "oc_get_events": (C.c_int, [vp, C.c_int32, P(oc_event), C.c_int64, P(C.c_int64)]),
        "oc_get_stats": (C.c_int, [vp, C.c_int32, P(oc_stats)]),
        "oc_get_step": (C.c_int, [vp, C.c_int32, P(C.c_int64)]),
        "oc_max_bond": (C.c_int, [vp, C.c_int32, P(C.c_double), P(C.c_double)]),
        "oc_stream": (vp, [vp]),
        "oc_launch_pair": (C.c_int, [vp, C.c_int32]),
        "oc_abi_sizes": (C.c_int, [P(C.c_int32), C.c_int32]),
        "oc_set_pair_params": (C.c_int, [vp, C.c_int32, P(C.c_double), C.c_double]),
        "oc_window_forward": (C.c_int, [vp, C.c_int32, C.c_int32, P(oc_rdf_spec), P(C.c_double)]),
        "oc_window_backward": (C.c_int, [vp, P(C.c_double), P(C.c_double)]),
        "oc_window_rewind": (C.c_int, [vp]),
        "oc_window_grad": (C.c_int, [vp, C.c_int32, C.c_int32, P(oc_rdf_spec), P(C.c_double), C.c_int32,
                                     P(C.c_double), P(C.c_double), P(C.c_double)]),
        "oc_forces_theta": (C.c_int, [vp, C.c_int64, C.c_int32, P(C.c_uint32), P(C.c_uint32), P(C.c_float),
                                      C.c_int64]),
        "oc_chem_uniforms": (C.c_int, [vp, C.c_int64, P(C.c_double), P(C.c_double), P(C.c_uint32), P(C.c_uint32),
                                       P(C.c_double), C.c_int64]),
        "oc_kinetic_dev_work": (C.c_size_t, [C.c_int64]),
        "oc_kinetic_dev": (C.c_int, [vp, vp, vp, vp, C.c_int64, C.c_double, vp, vp, C.c_size_t, vp]),
        "oc_kinetic_dev_noalloc": (C.c_int, [vp, vp, vp, vp, C.c_int64, C.c_double, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def load_library(path=None):
    """Load liboercel_mi.so (built in-tree by build.py).  Raises if it is missing:
    there is no CPU fallback for the step."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = _declare(C.CDLL(p, mode=C.RTLD_GLOBAL))
    if path is None:
        _LIB = lib
    return lib


class OcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oercel_mi error {code}: {msg}")
        self.code = code


def check(lib, rc):
    if rc != 0:
        raise OcError(rc, lib.oc_last_error().decode(errors="replace"))
    return rc
