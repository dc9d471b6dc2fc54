"""ctypes binding of libchs.so (the HIP/gfx950 library).  There is no CPU
fallback: if the library is missing or no GPU is present, loading fails with
an error."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchs.so")

CHS_OK, CHS_EINVAL, CHS_EHIP, CHS_ENOCONV, CHS_ENOMEM = 0, 1, 2, 3, 4
CHS_FP64, CHS_FP32, CHS_BF16, CHS_FP16 = 0, 1, 2, 3
CHS_LEARNED = 4   # chs_fermion's flexible CG with the learned preconditioner
PRECISIONS = {"fp64": CHS_FP64, "fp32": CHS_FP32, "bf16": CHS_BF16, "fp16": CHS_FP16}
BYTES_PER_COMPLEX = {CHS_FP64: 16, CHS_FP32: 8, CHS_BF16: 4, CHS_FP16: 4}
MAX_SHIFTS = 32

_lib = None


class ChsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libchs error %d: %s" % (code, msg))
        self.code = code
        self.msg = msg


class NoConvergence(ChsError):
    pass


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("reliable_updates", ctypes.c_int), ("restarts", ctypes.c_int),
                ("true_resid", ctypes.c_double), ("seconds", ctypes.c_double)]


class Rng(ctypes.Structure):
    _fields_ = [("s", ctypes.c_uint64), ("v", ctypes.c_double * 97), ("idum", ctypes.c_int),
                ("yran", ctypes.c_double), ("yrun", ctypes.c_double)]


class MsSystem(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("phi", ctypes.c_void_p), ("chi", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("hpin", ctypes.c_double), ("imass", ctypes.c_int), ("ndiag", ctypes.c_int),
                ("anum", ctypes.c_void_p), ("aden", ctypes.c_void_p), ("res", ctypes.c_double),
                ("itercg", ctypes.c_int), ("iflag", ctypes.c_int), ("dsdchi", ctypes.c_void_p)]


class BatchTune(ctypes.Structure):
    _fields_ = [("ept", ctypes.c_int), ("chunk", ctypes.c_int), ("graph", ctypes.c_int),
                ("us_per_iter", ctypes.c_double)]


class BatchStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("graph_replays", ctypes.c_int), ("launches", ctypes.c_int),
                ("seconds", ctypes.c_double), ("tune", BatchTune)]


def _declare(L):
    c, i, d, l, v, sz = ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_long, ctypes.c_void_p, ctypes.c_size_t
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "chs_ctx_create": ([i, i, i, i, d, i], v),
        "chs_ctx_destroy": ([v], None),
        "chs_last_error": ([v], c),
        "chs_create_error": ([], c),
        "chs_kferm": ([v], l),
        "chs_kvol": ([v], l),
        "chs_stream": ([v], v),
        "chs_synchronize": ([v], i),
        "chs_dslash": ([v, v, v, v, d, d, i], i),
        "chs_dslashd": ([v, v, v, v, d, d, i], i),
        "chs_qmrherm": ([v, v, d, ip, d, d, i, v, v, i, i, v, v, v], i),
        "chs_congrad": ([v, v, d, ip, d, d, i, v, v], i),
        "chs_force": ([v, v, d, d, d, i, v, d, v, v, v, v, i, v, v, ip], i),
        "chs_hamilton": ([v, v, v, v, v, v, d, i, d, d, i, v, v, d, v, v, i, v, ip], i),
        "chs_derivs": ([v, v, v, d, i, v], i),
        "chs_set_force_precision": ([v, i], i),
        "chs_set_lanczos_cap": ([v, i], i),
        "chs_set_shift_freeze": ([v, d], i),
        "chs_qmrherm_shifts_dev": ([v, v, d, ctypes.POINTER(i), d, i, v, i, v, v], i),
        "chs_mirror_register": ([v, v], i),
        "chs_mirror_touch": ([v, v], i),
        "chs_mirror_release": ([v, v], i),
        "chs_transfer_bytes": ([v, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_ulonglong)], i),
        "chs_fermion": ([v, v, d, d, i, i, i, v, v, ip], i),
        "chs_hmc_force": ([v, v, v, d, d, i, v, d, v, ip], i),
        "chs_hmc_hamilton": ([v, v, v, d, d, i, v, v, d, v, ip], i),
        "chs_alloc": ([v, sz], v),
        "chs_free": ([v, v], i),
        "chs_upload": ([v, v, v, sz], i),
        "chs_download": ([v, v, v, sz], i),
        "chs_convert_dev": ([v, v, i, v, i, l], i),
        "chs_layout_dev": ([v, v, v, i, i], i),
        "chs_dslash_dev": ([v, v, v, v, d, i, i, i], i),
        "chs_qmrherm_dev": ([v, v, d, ip, d, i, v, v, i, i, v, v, v], i),
        "chs_congrad_dev": ([v, v, d, ip, d, i, v, v], i),
        "chs_congrad_noalloc_dev": ([v, v, d, ip, d, i, v, v], i),
        "chs_solve_mdagm_dev": ([v, v, v, v, d, i, d, i, d, i, ctypes.POINTER(SolveStats)], i),
        "chs_time_dslash": ([v, i, i, i, i, d, i, ctypes.POINTER(ctypes.c_double)], i),
        "chs_time_cg_stencil": ([v, i, i, i, d, i, ctypes.POINTER(ctypes.c_double)], i),
        "chs_time_copy": ([v, i, i, i, ctypes.POINTER(ctypes.c_double)], i),
        "chs_time_copy_grid": ([v, i, i, i, ctypes.POINTER(ctypes.c_double)], i),
        "chs_nn_init": ([v, i, ctypes.c_ulonglong], i),
        "chs_nn_num_weights": ([v], l),
        "chs_nn_set_weights": ([v, v, l], i),
        "chs_nn_get_weights": ([v, v, l], i),
        "chs_nn_apply_dev": ([v, v, v], i),
        "chs_nn_loss_grad_dev": ([v, v, v, v, d, i, ctypes.POINTER(ctypes.c_double), v], i),
        "chs_nn_train_dev": ([v, v, v, i, i, ctypes.c_float, v, d, i, v], i),
        "chs_nn_collect": ([v, i], i),
        "chs_nn_num_collected": ([v], i),
        "chs_nn_pair_dev": ([v, i, ctypes.POINTER(v), ctypes.POINTER(v), ctypes.POINTER(v)], i),
        "chs_nn_train_collected": ([v, i, ctypes.c_float, d, i, v], i),
        "chs_solve_mdagm_fcg_dev": ([v, v, v, v, d, i, d, d, i, i, ctypes.POINTER(SolveStats)], i),
        "chs_batch_create": ([i], v),
        "chs_batch_destroy": ([v], None),
        "chs_batch_last_error": ([v], c),
        "chs_qmrherm_batch_dev": ([v, ctypes.POINTER(MsSystem), i, i, ctypes.POINTER(BatchTune),
                                   ctypes.POINTER(BatchStats)], i),
        "chs_qmrherm_prec_dev": ([v, v, d, ip, d, i, v, v, i, i, v, v, v, i], i),
        "chs_rng_set": ([ctypes.POINTER(Rng), d], None),
        "chs_rng_get": ([ctypes.POINTER(Rng)], d),
        "chs_rano": ([ctypes.POINTER(Rng), i], d),
        "chs_gauss0": ([ctypes.POINTER(Rng), v, i], None),
        "chs_gaussp": ([ctypes.POINTER(Rng), v, i], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


def lib():
    """The loaded libchs.so.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libchs.so not built: run `python -m chiralheisenberg_amd.build` "
                              "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def check(ctx_handle, rc):
    if rc == CHS_OK:
        return
    msg = lib().chs_last_error(ctx_handle)
    msg = msg.decode() if msg else ""
    if rc == CHS_ENOCONV:
        raise NoConvergence(rc, msg)
    raise ChsError(rc, msg)
