"""ctypes binding of libfsim.so (the C ABI declared in include/fsim.h).

The library is loaded from this package directory.  There is no fallback:
if libfsim.so is missing or fails to load, every entry point raises, so a GPU
run can never silently compute on the CPU.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# FSIM_LIBRARY selects another build of the same library (e.g. a variant
# compiled with different tile parameters); default: the in-tree build.
LIB_PATH = os.environ.get("FSIM_LIBRARY") or os.path.join(_HERE, "libfsim.so")

FSIM_OK = 0
FSIM_ERR_ARG = -1
FSIM_ERR_OOM = -2
FSIM_ERR_HIP = -3
FSIM_ERR_STATE = -4
FSIM_ERR_BUSY = -5


class FsimError(RuntimeError):
    """Mirror of the reference's RuntimeException("JNI-CUDA: ...")."""

    def __init__(self, code, msg):
        super().__init__(f"fsim error {code}: {msg}")
        self.code = code


class FsimBusy(FsimError):
    """Request refused by admission control (queue or memory full)."""


class RECON_PARAMS(C.Structure):
    """fsim_recon_params (include/fsim.h)."""
    _fields_ = [
        ("n", C.c_int),
        ("ndir", C.c_int),
        ("nphase", C.c_int),
        ("nband", C.c_int),
        ("px", C.c_double * 5),
        ("py", C.c_double * 5),
        ("phases", (C.c_double * 7) * 5),
        ("modul", (C.c_double * 3) * 5),
        ("vec_cpm", C.c_double),
        ("use_att", C.c_int),
        ("wiener", C.c_double),
        ("apo_bend", C.c_double),
        ("apo_cut", C.c_double),
        ("fade_px", C.c_int),
        ("mode", C.c_int),
        ("otf_band0", C.c_int),
    ]


class SIM_PARAMS(C.Structure):
    """fsim_sim_params (include/fsim.h)."""
    _fields_ = [("px", C.c_double * 5), ("py", C.c_double * 5), ("phases", (C.c_double * 7) * 5),
                ("modul", (C.c_double * 3) * 5)]


class ESTIMATE_PARAMS(C.Structure):
    """fsim_estimate_params (include/fsim.h)."""
    _fields_ = [("n", C.c_int), ("ndir", C.c_int), ("nphase", C.c_int), ("nband", C.c_int),
                ("fit_band", C.c_int), ("fit_exclude", C.c_double), ("vec_cpm", C.c_double),
                ("fade_px", C.c_int), ("live", C.c_int)]


class ESTIMATE_RESULT(C.Structure):
    """fsim_estimate_result (include/fsim.h)."""
    _fields_ = [("px", C.c_double), ("py", C.c_double), ("pha_off", C.c_double), ("mod1", C.c_double),
                ("mod2", C.c_double), ("peak_x", C.c_double), ("peak_y", C.c_double)]


class INGEST_CONFIG(C.Structure):
    """fsim_ingest_config (include/fsim.h)."""
    _fields_ = [("ndir", C.c_int), ("nphase", C.c_int), ("sync_mode", C.c_int), ("sync_delay", C.c_int64),
                ("sync_avr", C.c_int), ("seq_count", C.c_int), ("sort_capacity", C.c_int),
                ("queue_capacity", C.c_int), ("max_width", C.c_int), ("max_height", C.c_int)]


class INGEST_SEQ_INFO(C.Structure):
    """fsim_ingest_seq_info (include/fsim.h)."""
    _fields_ = [("channel", C.c_int), ("width", C.c_int), ("height", C.c_int), ("first_seq_nr", C.c_int64),
                ("time_camera", C.c_int64), ("pos2", C.c_int32)]


class INGEST_STATS(C.Structure):
    """fsim_ingest_stats (include/fsim.h)."""
    _fields_ = [("frames", C.c_int64), ("sequences", C.c_int64), ("missed_sequences", C.c_int64),
                ("dropped_frames", C.c_int64), ("sync_frames", C.c_int64), ("no_sync_windows", C.c_int64)]


class SERVER_CONFIG(C.Structure):
    """fsim_server_config (include/fsim.h)."""
    _fields_ = [("mem_budget", C.c_int64), ("max_batch", C.c_int), ("max_queue", C.c_int),
                ("nslots", C.c_int)]


class SERVER_STATS(C.Structure):
    """fsim_server_stats (include/fsim.h)."""
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("failed", C.c_int64), ("batches", C.c_int64), ("pending", C.c_int64),
                ("device_bytes", C.c_int64), ("mean_batch", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("max_ms", C.c_double)]


class TUNE_RESULT(C.Structure):
    """fsim_tune_result (include/fsim.h)."""
    _fields_ = [("sep", C.c_int), ("col", C.c_int), ("col_elems", C.c_int), ("col_width", C.c_int),
                ("candidates", C.c_int), ("ms_best", C.c_double), ("ms_default", C.c_double),
                ("chunk", C.c_int), ("streams", C.c_int), ("ms_whole", C.c_double),
                ("ms_chunk_default", C.c_double)]


FSIM_SUBMIT_DEVICE = 1

_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_F = C.c_float
_L = C.c_int64
_PD = C.POINTER(C.c_double)
_PF = C.POINTER(C.c_float)
_PI = C.POINTER(C.c_int)

# name: (argtypes); every function returns int status unless listed in _RET
_SIGS = {
    "fsim_init": [_I],
    "fsim_sync": [],
    "fsim_alloc_device": [_L, C.POINTER(_P)],
    "fsim_free_device": [_P],
    "fsim_alloc_host": [_L, C.POINTER(_P)],
    "fsim_free_host": [_P],
    "fsim_memcpy": [_P, _P, _L, _I],
    "fsim_vec_create": [_I, _I, _I, C.POINTER(_P)],
    "fsim_vec_destroy": [_P],
    "fsim_vec_shape": [_P, _PI, _PI, _PI],
    "fsim_vec_reshape": [_P, _I, _I],
    "fsim_vec_upload": [_P, _P],
    "fsim_vec_download": [_P, _P],
    "fsim_vec_zero": [_P],
    "fsim_vec_copy": [_P, _P],
    "fsim_vec_add": [_P, _P],
    "fsim_vec_add_const": [_P, _F, _F],
    "fsim_vec_axpy": [_P, _F, _F, _P],
    "fsim_vec_times": [_P, _P, _I],
    "fsim_vec_scal": [_P, _F, _F],
    "fsim_vec_norm2": [_P, _PD],
    "fsim_vec_sum": [_P, _PD, _PD],
    "fsim_vec_set_u16": [_P, _P],
    "fsim_vec_set_u16_device": [_P, _P],
    "fsim_vec_fft2d": [_P, _I],
    "fsim_vec_fourier_shift": [_P, _D, _D],
    "fsim_vec_paste_freq": [_P, _P, _I, _I],
    "fsim_vec_get": [_P, _I, _I, _PF, _PF],
    "fsim_vec_set": [_P, _I, _I, _F, _F],
    "fsim_otf_from_estimate": [_D, _D, _D, _I, _D, _D, _P, _P, _P, _PD, _PD],
    "fsim_otf_attenuation": [_I, _I, _D, _D, _D, _P, _P, _P],
    "fsim_otf_create": [_I, _I, _D, _D, _P, _P, _P, C.POINTER(_P)],
    "fsim_otf_destroy": [_P],
    "fsim_otf_to_vector": [_P, _P, _I, _D, _D, _D, _I, _I],
    "fsim_mask_otf": [_P, _P, _D, _D, _D],
    "fsim_write_apo": [_P, _P, _D, _D, _D],
    "fsim_write_attenuation": [_P, _D, _D, _D, _D, _D],
    "fsim_wiener_denominator": [_P, _P, _D, _I, _I, _PD, _PD, _I, _D],
    "fsim_fade_border": [_P, _I],
    "fsim_separation_matrix": [_PD, _I, _I, _PD, _PD],
    "fsim_separate_bands": [C.POINTER(_P), _I, C.POINTER(_P), _I, _PF],
    "fsim_locate_peak": [_P, _D, _PD],
    "fsim_common_region": [_P, _P, _I, _I, _P, _D, _D, _D, _D, _I, _D],
    "fsim_fit_peak": [_P, _P, _I, _I, _P, _D, _D, _D, _D, _D, C.POINTER(C.c_float), _PD],
    "fsim_get_peak": [_P, _P, _I, _I, _P, _D, _D, _D, _D, _PD],
    "fsim_estimate": [_P, _I, C.POINTER(ESTIMATE_PARAMS), _P, C.POINTER(ESTIMATE_RESULT)],
    "fsim_params_from_estimate": [C.POINTER(ESTIMATE_RESULT), _I, _I, _I, _I, C.POINTER(SIM_PARAMS),
                                  C.POINTER(SIM_PARAMS)],
    "fsim_channel_create": [C.POINTER(RECON_PARAMS), _P, C.POINTER(_P)],
    "fsim_channel_destroy": [_P],
    "fsim_reconstruct": [_P, _P, _P, _I, _P],
    "fsim_reconstruct_params": [_P, _P, _P, _I, C.POINTER(SIM_PARAMS), _P],
    "fsim_reconstruct_noalloc": [_P, _P, _P, _I, _P],
    "fsim_server_create": [C.POINTER(SERVER_CONFIG), C.POINTER(_P)],
    "fsim_server_destroy": [_P],
    "fsim_server_add_channel": [_P, _P, _PI],
    "fsim_server_submit": [_P, _I, _P, _P, _I, C.POINTER(C.c_uint64)],
    "fsim_server_wait": [_P, C.c_uint64, _I, _PD],
    "fsim_server_get_stats": [_P, C.POINTER(SERVER_STATS), _I],
    "fsim_server_submit_params": [_P, _I, _P, _P, _I, C.POINTER(SIM_PARAMS), _I, C.POINTER(C.c_uint64)],
    "fsim_batch_create": [_P, _I, _I, _P],
    "fsim_batch_run": [_P, _P, _P, _P],
    "fsim_batch_info": [_P, _P, _P, _P],
    "fsim_batch_destroy": [_P],
    "fsim_channel_class": [_P, _PI, _PI, _PI, _PI, _PI],
    "fsim_autotune": [_P, _I, _I, C.POINTER(TUNE_RESULT)],
    "fsim_autotune_set": [_I, _I, _I, _I, _I, _I, _I],
    "fsim_autotune_get": [_I, _I, _I, _I, _I, _PI, _PI],
    "fsim_autotune_set_chunks": [_I, _I, _I, _I, _I, _I, _I],
    "fsim_autotune_get_chunks": [_I, _I, _I, _I, _I, _I, _PI, _PI],
    "fsim_autotune_clear": [],
    "fsim_autotune_save": [C.c_char_p],
    "fsim_autotune_load": [C.c_char_p],
    "fsim_rl_deconvolve": [_P, _P, _I, _I],
    "fsim_clip_scale": [_P, _I, _I],
    "fsim_widefield": [_P, _P, _P, _I, _I, _P],
    "fsim_registration_create": [_P, _P, _I, _I, C.POINTER(_P)],
    "fsim_registration_destroy": [_P],
    "fsim_register_images": [_P, _I, _P, _P, _I, _I, _I, _P],
    "fsim_wire_parse_header": [_P, _I, _I, _P],
    "fsim_wire_write_header": [_P, _P],
    "fsim_ingest_create": [_P, C.POINTER(_I), _I, C.POINTER(_P)],
    "fsim_ingest_destroy": [_P],
    "fsim_ingest_push": [_P, _P, C.c_int64],
    "fsim_ingest_pop": [_P, _I, _P, C.c_int64, _P],
    "fsim_ingest_get_stats": [_P, _I, _P],
    "fsim_ingest_attach_server": [_P, _P, C.POINTER(_I), _I, _I],
    "fsim_ingest_result": [_P, _I, _P, _P],
    "fsim_stage_timing": [_I],
    "fsim_stage_times": [_PD, _PI, _I],
    "fsim_debug_fail_alloc": [_I],
    "fsim_debug_region_exact": [_I],
}
_RET = {
    "fsim_last_error": (C.c_char_p, []),
    "fsim_device_bytes": (C.c_int64, []),
    "fsim_vec_device_ptr": (_P, [_P]),
}

_lib = None


def lib():
    """Load libfsim.so once; raise if it is missing (no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FsimError(FSIM_ERR_STATE, f"native library not built: {LIB_PATH} "
                                        "(run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    for name, (ret, args) in _RET.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ret
    _lib = L
    return L


def declared_symbols():
    return list(_SIGS) + list(_RET)


def check(rc):
    """Raise the reference's exception type for a failing status."""
    if rc == FSIM_OK:
        return
    msg = lib().fsim_last_error().decode(errors="replace")
    if rc == FSIM_ERR_OOM:
        raise MemoryError(msg)
    if rc == FSIM_ERR_BUSY:
        raise FsimBusy(rc, msg)
    raise FsimError(rc, msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        call("fsim_init", device)
        _initialised = True
