"""ctypes binding of libdosehip.so (the C ABI declared in include/dosehip.h).

The library is loaded after ``torch`` so that it binds to the HIP runtime torch
already mapped (both use the soname libamdhip64.so.7): device pointers and
streams handed over from torch tensors are then valid inside the library.
There is no fallback: if the library is missing the import fails.
"""

import ctypes
import os

import torch  # noqa: F401  (must be loaded before libdosehip.so, see above)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DOSEHIP_LIB", os.path.join(_HERE, "libdosehip.so"))

DH_OK = 0
DH_ERR_ARG = -1
DH_ERR_HIP = -2


class dh_grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("spacing", ctypes.c_float)]


class dh_beam_geom(ctypes.Structure):
    _fields_ = [("iso", ctypes.c_float * 3), ("src", ctypes.c_float * 3),
                ("sin_g", ctypes.c_float), ("cos_g", ctypes.c_float),
                ("sin_c", ctypes.c_float), ("cos_c", ctypes.c_float),
                ("vsadx", ctypes.c_float), ("vsady", ctypes.c_float)]


class dh_spot(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("mu", ctypes.c_float),
                ("energy_id", ctypes.c_int32)]


class dh_layer(ctypes.Structure):
    _fields_ = [("spot_start", ctypes.c_int32), ("n_spots", ctypes.c_int32),
                ("energy_id", ctypes.c_int32), ("r80", ctypes.c_float), ("energy", ctypes.c_float)]


class dh_beam_model(ctypes.Structure):
    _fields_ = [("idd_sigma", ctypes.c_void_p), ("depths", ctypes.c_void_p), ("dvp", ctypes.c_void_p),
                ("n_energies", ctypes.c_int32), ("lut_len", ctypes.c_int32), ("dvp_len", ctypes.c_int32),
                ("depth_uniform", ctypes.c_int32), ("depth0", ctypes.c_float), ("ddepth", ctypes.c_float)]


class dh_spot_set(ctypes.Structure):
    _fields_ = [("spots", ctypes.c_void_p), ("layers", ctypes.c_void_p),
                ("n_spots", ctypes.c_int32), ("n_layers", ctypes.c_int32)]


class dh_batch_plan(ctypes.Structure):
    _fields_ = [("grid", dh_grid), ("rlsp_padded", ctypes.c_void_p), ("dose", ctypes.c_void_p),
                ("grad_dose", ctypes.c_void_p), ("beam_start", ctypes.c_int32), ("n_beams", ctypes.c_int32)]


class dh_batch_beam(ctypes.Structure):
    _fields_ = [("geom", dh_beam_geom), ("model", dh_beam_model), ("spots", dh_spot_set),
                ("wet", ctypes.c_void_p), ("wet_raw", ctypes.c_void_p), ("dose", ctypes.c_void_p),
                ("grad_mu", ctypes.c_void_p), ("grad_xy", ctypes.c_void_p)]


class dh_batch_tune(ctypes.Structure):
    _fields_ = [("adj_cs", ctypes.c_int32), ("adj_wg_target", ctypes.c_int32),
                ("adj_min_bricks", ctypes.c_int32), ("adj_max_slabs", ctypes.c_int32),
                ("fwd_blocks", ctypes.c_int32), ("ray_blocks", ctypes.c_int32),
                ("ray_shape", ctypes.c_int32)]


class dh_batch(ctypes.Structure):
    _fields_ = [("ws", ctypes.c_void_p), ("ws_bytes", ctypes.c_int64),
                ("n_plans", ctypes.c_int32), ("n_beams", ctypes.c_int32),
                ("cutoff_sigma", ctypes.c_float), ("magic", ctypes.c_int32),
                ("n_bricks", ctypes.c_int64), ("n_smooth_blocks", ctypes.c_int64),
                ("n_spots", ctypes.c_int64), ("n_layers", ctypes.c_int64),
                ("n_sum_blocks", ctypes.c_int64), ("n_partials", ctypes.c_int64),
                ("n_adj32", ctypes.c_int64), ("n_adj64", ctypes.c_int64), ("n_adj128", ctypes.c_int64),
                ("n_reduce_blocks", ctypes.c_int64), ("tab", ctypes.c_int64 * 16),
                ("fwd_blocks", ctypes.c_int32), ("ray_blocks", ctypes.c_int32),
                ("ray_shape", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("prepared", ctypes.c_int32), ("n_summary_rows", ctypes.c_int32),
                ("events", ctypes.c_void_p)]


class dh_mlc_pair(ctypes.Structure):
    _fields_ = [("x1", ctypes.c_float), ("x2", ctypes.c_float), ("y_offset", ctypes.c_float),
                ("y_width", ctypes.c_float)]


class dh_photon_model(ctypes.Structure):
    _fields_ = [("profile", ctypes.c_void_p), ("spectrum", ctypes.c_void_p),
                ("n_profile", ctypes.c_int32), ("n_spectrum", ctypes.c_int32),
                ("kernel", (ctypes.c_float * 6) * 6), ("mu_cal", ctypes.c_float),
                ("primary_src_dist", ctypes.c_float), ("scatter_src_dist", ctypes.c_float),
                ("primary_src_size", ctypes.c_float), ("scatter_src_size", ctypes.c_float),
                ("mlc_distance", ctypes.c_float), ("scatter_src_weight", ctypes.c_float),
                ("electron_attenuation", ctypes.c_float), ("electron_src_weight", ctypes.c_float),
                ("electron_fitted_dmax", ctypes.c_float)]


class dh_photon_beam(ctypes.Structure):
    _fields_ = [("geom", dh_beam_geom), ("sin_col", ctypes.c_float), ("cos_col", ctypes.c_float),
                ("mlc", ctypes.c_void_p), ("n_pairs", ctypes.c_int32), ("mu", ctypes.c_float)]


DH_TIME_RAYMARCH, DH_TIME_SMOOTH, DH_TIME_FORWARD, DH_TIME_ADJOINT, DH_TIME_NEED = range(5)
DH_TIME_N = 5
TIMED_KERNELS = {DH_TIME_RAYMARCH: "k_wet_raymarch_batch", DH_TIME_SMOOTH: "k_wet_smooth_tile_batch",
                 DH_TIME_FORWARD: "k_pb_forward_batch", DH_TIME_ADJOINT: "k_pb_adjoint_batch",
                 DH_TIME_NEED: "k_wet_need_batch+k_wet_dilate_batch"}


DH_BATCH_TAB_HEADER_END = 13   # dh_batch.tab index: end of the tables uploaded at setup
DH_BATCH_TAB_FLAGS = 14        # dh_batch.tab index: capability bits (not an offset)
DH_BATCH_FLAG_SMOOTH_TILE = 8  # ... bit: smoothing runs the tiled (LDS) kernel
DH_BATCH_REUSE_PREP = 1
DH_BATCH_WET_REACH = 2

# exported symbol -> (restype, argtypes); every declaration of include/dosehip.h
_P = ctypes.c_void_p
SIGNATURES = {
    "dh_version": (ctypes.c_int, []),
    "dh_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "dh_padded_len": (ctypes.c_int64, [dh_grid]),
    "dh_rlsp_from_hu": (ctypes.c_int, [_P, _P, dh_grid, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double), ctypes.c_int32, _P]),
    "dh_rlsp_pad": (ctypes.c_int, [_P, _P, dh_grid, _P]),
    "dh_wet_raymarch": (ctypes.c_int, [_P, _P, dh_grid, ctypes.POINTER(dh_beam_geom), _P]),
    "dh_wet_smooth": (ctypes.c_int, [_P, _P, dh_grid, ctypes.POINTER(dh_beam_geom), _P]),
    "dh_pb_workspace_size": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32, dh_grid]),
    "dh_pb_forward": (ctypes.c_int, [_P, _P, dh_grid, ctypes.POINTER(dh_beam_geom),
                                     ctypes.POINTER(dh_beam_model), ctypes.POINTER(dh_spot_set),
                                     ctypes.c_float, ctypes.c_int32, _P, ctypes.c_int64, _P]),
    "dh_pb_adjoint": (ctypes.c_int, [_P, _P, dh_grid, ctypes.POINTER(dh_beam_geom),
                                     ctypes.POINTER(dh_beam_model), ctypes.POINTER(dh_spot_set),
                                     ctypes.c_float, _P, _P, _P, ctypes.c_int64, _P]),
    "dh_pb_adjoint_noalloc": (ctypes.c_int, [_P, _P, dh_grid, ctypes.POINTER(dh_beam_geom),
                                             ctypes.POINTER(dh_beam_model), ctypes.POINTER(dh_spot_set),
                                             ctypes.c_float, _P, _P, _P]),
    "dh_pb_count_pairs": (ctypes.c_int, [_P, _P, dh_grid, ctypes.POINTER(dh_beam_geom),
                                         ctypes.POINTER(dh_beam_model), ctypes.POINTER(dh_spot_set),
                                         ctypes.c_float, _P, _P, ctypes.c_int64, _P]),
    "dh_spot_params": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int32, _P]),
    "dh_batch_workspace_size": (ctypes.c_int64, [ctypes.POINTER(dh_batch_plan), ctypes.c_int32,
                                                 ctypes.POINTER(dh_batch_beam), ctypes.c_int32,
                                                 ctypes.POINTER(dh_batch_tune)]),
    "dh_batch_setup": (ctypes.c_int, [ctypes.POINTER(dh_batch), ctypes.POINTER(dh_batch_plan), ctypes.c_int32,
                                      ctypes.POINTER(dh_batch_beam), ctypes.c_int32, ctypes.POINTER(dh_batch_tune),
                                      ctypes.c_float, _P, ctypes.c_int64, _P]),
    "dh_batch_wet": (ctypes.c_int, [ctypes.POINTER(dh_batch), ctypes.c_int32, _P]),
    "dh_batch_forward": (ctypes.c_int, [ctypes.POINTER(dh_batch), ctypes.c_int32, _P]),
    "dh_batch_adjoint": (ctypes.c_int, [ctypes.POINTER(dh_batch), ctypes.c_int32, _P]),
    "dh_batch_timing": (ctypes.c_int, [ctypes.POINTER(dh_batch), _P]),
    "dh_photon_record_len": (ctypes.c_int64, [dh_grid]),
    "dh_photon_density": (ctypes.c_int, [_P, _P, dh_grid, _P]),
    "dh_photon_density_from_hu": (ctypes.c_int, [_P, _P, dh_grid, ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double), ctypes.c_int32, _P]),
    "dh_photon_raytrace": (ctypes.c_int, [_P, _P, dh_grid, ctypes.POINTER(dh_photon_beam), _P]),
    "dh_photon_terma": (ctypes.c_int, [_P, _P, _P, dh_grid, ctypes.POINTER(dh_photon_model),
                                       ctypes.POINTER(dh_photon_beam), _P]),
    "dh_photon_terma_records": (ctypes.c_int, [_P, _P, dh_grid, _P]),
    "dh_photon_ccc": (ctypes.c_int, [_P, _P, _P, _P, dh_grid, ctypes.POINTER(dh_photon_model),
                                     ctypes.POINTER(dh_photon_beam), ctypes.c_float, ctypes.c_int32, _P]),
}

_lib = None


def load():
    """Load libdosehip.so once and declare its signatures."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libdosehip.so not found at %s; build it with "
                              "`make -C dosecuda-jax_amd/csrc` (or __graft_entry__.build())" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().dh_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc, what):
    """Map a library status to the reference extension's exceptions
    (dosemodule.cu:189-197): bad arguments -> ValueError, HIP failures ->
    RuntimeError with the message prefixed "CUDA error: " as the reference does."""
    if rc == DH_OK:
        return
    msg = "%s: %s" % (what, last_error())
    if rc == DH_ERR_ARG:
        raise ValueError(msg)
    raise RuntimeError("CUDA error: " + msg)


def ptr(t):
    """Device (or host) address of a tensor as a ctypes void pointer."""
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
