"""ctypes view of the native libraries (liberq.so: include/erq.h; liberq_builder.so).

Loading fails loudly: there is no Python or CPU fallback for the search path.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(_PKG, "lib")

c_size = ctypes.c_size_t
c_vp = ctypes.c_void_p
c_int = ctypes.c_int

ERQ_OK = 0
ERRORS = {1: "EINVAL", 2: "EIO", 3: "EFORMAT", 4: "EHIP", 5: "ENOMEM", 6: "EUNSUPPORTED"}

# name -> (restype, argtypes); the exported surface of include/erq.h
SIGNATURES = {
    "erq_last_error": (ctypes.c_char_p, []),
    "erq_index_load": (c_int, [ctypes.c_char_p, ctypes.POINTER(c_vp)]),
    "erq_index_create": (c_int, [c_size, c_size, c_size, c_size, c_vp, c_vp, c_vp, c_vp, c_vp,
                                 c_vp, c_vp, ctypes.POINTER(c_vp)]),
    "erq_index_create_nonha": (c_int, [c_size, c_size, c_size, c_size, c_vp, c_vp, c_vp, c_vp,
                                       c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)]),
    "erq_index_free": (None, [c_vp]),
    "erq_set_option": (c_int, [c_vp, c_int, ctypes.c_long]),
    "erq_get_option": (ctypes.c_long, [c_vp, c_int]),
    "erq_autotune": (c_int, [c_vp, c_vp, c_size, c_size, c_size, c_vp, c_vp]),
    "erq_padded_dim": (c_size, [c_vp]),
    "erq_num_clusters": (c_size, [c_vp]),
    "erq_dim": (c_size, [c_vp]),
    "erq_num_vectors": (c_size, [c_vp]),
    "erq_ex_bits": (c_size, [c_vp]),
    "erq_short_factors": (c_size, [c_vp]),
    "erq_rotate": (c_int, [c_vp, c_vp, c_size, c_vp]),
    "erq_search": (c_int, [c_vp, c_vp, c_size, c_size, c_vp, c_vp]),
    "erq_search_batch": (c_int, [c_vp, c_vp, c_size, c_size, c_size, c_vp, c_vp]),
    "erq_search_batch_nprobe": (c_int, [c_vp, c_vp, c_size, c_size, c_vp, c_vp, c_vp,
                                        ctypes.c_uint]),
    "erq_rotate_dev": (c_int, [c_vp, c_vp, c_size, c_vp, c_vp]),
    "erq_centroid_dist_dev": (c_int, [c_vp, c_vp, c_size, c_vp, c_vp]),
    "erq_probe_dev": (c_int, [c_vp, c_vp, c_size, c_size, c_vp, c_vp, c_vp]),
    "erq_scan_dev": (c_int, [c_vp, c_vp, c_size, c_size, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                             c_vp]),
    "erq_search_dev": (c_int, [c_vp, c_vp, c_size, c_size, c_size, c_vp, c_vp, c_vp, c_vp]),
    "erq_search_dev_noalloc": (c_int, [c_vp, c_vp, c_size, c_size, c_size, c_vp, c_vp, c_vp,
                                       c_vp]),
    "erq_search_noalloc_bytes": (c_size, [c_vp, c_size, c_size]),
    "erq_search_noalloc_capacity": (c_size, []),
    "erq_scan_batched_dev": (c_int, [c_vp, c_vp, c_size, c_size, c_vp, c_vp, c_vp, c_size, c_vp,
                                     c_vp, c_vp, c_vp]),
    "erq_search_batched_dev": (c_int, [c_vp, c_vp, c_size, c_size, c_size, c_vp, c_vp, c_vp,
                                       c_vp]),
    "erq_mlp_create": (c_int, [c_size, c_size, c_size, c_size, ctypes.c_uint64,
                               ctypes.POINTER(c_vp)]),
    "erq_mlp_free": (None, [c_vp]),
    "erq_mlp_num_params": (c_size, [c_vp]),
    "erq_mlp_in_dim": (c_size, [c_vp]),
    "erq_mlp_out_dim": (c_size, [c_vp]),
    "erq_mlp_set_params": (c_int, [c_vp, c_vp]),
    "erq_mlp_set_input_scale": (c_int, [c_vp, ctypes.c_float]),
    "erq_mlp_input_scale": (ctypes.c_float, [c_vp]),
    "erq_mlp_get_params": (c_int, [c_vp, c_vp, c_vp]),
    "erq_mlp_forward_dev": (c_int, [c_vp, c_vp, c_size, c_vp, c_vp]),
    "erq_mlp_loss_grad_dev": (c_int, [c_vp, c_vp, c_vp, c_size, ctypes.c_float, c_vp, c_vp]),
    "erq_mlp_adam_step_dev": (c_int, [c_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, c_vp]),
    "erq_probe_learned_dev": (c_int, [c_vp, c_vp, c_vp, c_size, ctypes.c_float, c_size, c_size,
                                      c_size, ctypes.c_float, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "erq_search_learned_dev": (c_int, [c_vp, c_vp, c_vp, c_size, c_size, ctypes.c_float, c_size,
                                       c_size, c_size, ctypes.c_float, c_vp, c_vp, c_vp, c_vp]),
    "erq_device_array": (c_vp, [c_vp, c_int]),
    "erq_timing_begin": (c_int, [c_vp, c_size]),
    "erq_timing_end": (c_int, [c_vp, c_vp, c_vp]),
    "erq_knn_dev": (c_int, [c_vp, ctypes.c_uint64, c_vp, ctypes.c_uint64, ctypes.c_uint32,
                            ctypes.c_uint32, c_vp, c_vp, c_vp]),
    "erq_kmeans_assign_dev": (c_int, [c_vp, ctypes.c_uint64, ctypes.c_uint32, c_vp,
                                      ctypes.c_uint32, c_vp, c_vp, c_vp]),
    "erq_kmeans_update_dev": (c_int, [c_vp, ctypes.c_uint64, ctypes.c_uint32, c_vp,
                                      ctypes.c_uint32, c_vp, c_vp, c_vp]),
    "erq_kmeans_dev": (c_int, [c_vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.c_uint32, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "erq_build_quantize_dev": (c_int, [c_vp, c_vp, ctypes.c_uint64, ctypes.c_uint32,
                                       ctypes.c_uint32, c_vp, c_vp, c_vp]),
    "erq_build_pack_dev": (c_int, [c_vp, c_vp, c_vp, ctypes.c_uint32, ctypes.c_uint32, c_vp,
                                   c_vp]),
    "erq_build_index_dev": (c_int, [c_vp, ctypes.c_uint64, ctypes.c_uint32, c_vp,
                                    ctypes.c_uint32, c_vp, c_vp, ctypes.c_uint32, c_vp, c_vp,
                                    c_vp, c_vp, c_vp, c_vp, c_vp]),
    "erq_build_index_dev_nonha": (c_int, [c_vp, ctypes.c_uint64, ctypes.c_uint32, c_vp,
                                    ctypes.c_uint32, c_vp, c_vp, ctypes.c_uint32, c_vp, c_vp,
                                    c_vp, c_vp, c_vp, c_vp, c_vp]),
}

BUILDER_SIGNATURES = {
    "erqb_quantize_cluster": (c_int, [c_vp, c_vp, c_size, c_size, c_int, c_vp, c_vp, c_vp]),
    "erqb_quantize_all": (c_int, [c_vp, c_vp, c_vp, c_size, c_size, c_int, c_vp, c_vp, c_vp,
                                  c_int]),
    "erqb_quantize_magnitudes": (ctypes.c_float, [c_vp, c_size, c_int, c_vp]),
    "erqb_quantize_all_nonha": (c_int, [c_vp, c_vp, c_vp, c_size, c_size, c_int, c_vp, c_vp, c_vp,
                                        c_int]),
}

_lib = None
_builder = None


def _load(name, sigs):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run extended-rabitq_amd/build.py "
                           "(no fallback exists for the native path)")
    lib = ctypes.CDLL(path)
    for fn, (res, args) in sigs.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    return lib


def lib():
    global _lib
    if _lib is None:
        _lib = _load("liberq.so", SIGNATURES)
    return _lib


def builder():
    global _builder
    if _builder is None:
        _builder = _load("liberq_builder.so", BUILDER_SIGNATURES)
    return _builder


class ErqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def check(rc):
    if rc != ERQ_OK:
        raise ErqError(rc, lib().erq_last_error().decode())
    return rc
