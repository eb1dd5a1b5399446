"""ctypes prototypes of the C ABI in include/minhashcuda.h.

Gives Python code (tests, benchmarks, users of the extension entry points)
typed access to every function libMHCUDA.so exports, with no torch types in
any signature.
"""
import ctypes

RESULT_NAMES = {
    0: "mhcudaSuccess",
    1: "mhcudaInvalidArguments",
    2: "mhcudaNoSuchDevice",
    3: "mhcudaMemoryAllocationFailure",
    4: "mhcudaRuntimeError",
    5: "mhcudaMemoryCopyError",
}

KERNEL_AUTO, KERNEL_PLAIN, KERNEL_PERSISTENT = 0, 1, 2

# Every function the public header declares, in header order.
EXPORTS = (
    "mhcuda_init", "mhcuda_get_parameters", "mhcuda_retrieve_random_vars",
    "mhcuda_assign_random_vars", "mhcuda_calc", "mhcuda_fini",
    "mhcuda_calc_device", "mhcuda_calc_device_noalloc", "mhcuda_calc_batch_device", "mhcuda_calc_batch",
    "mhcuda_set_tuning", "mhcuda_get_tuning", "mhcuda_set_stripes", "mhcuda_get_stripes",
    "mhcuda_set_skip", "mhcuda_get_skip", "mhcuda_set_seeds", "mhcuda_get_seeds",
    "mhcuda_autotune", "mhcuda_jaccard_pairs",
    "mhcuda_lsh_band_keys", "mhcuda_lsh_candidates", "mhcuda_save_random_vars", "mhcuda_load_random_vars",
    "mhcuda_params_file_info", "mhcuda_set_graph",
    "mhcuda_graph_captures",
    "mhcuda_set_timing", "mhcuda_kernel_time", "mhcuda_inject_fault",
    "mhcuda_serve_start", "mhcuda_serve_submit", "mhcuda_serve_wait", "mhcuda_serve_stats",
    "mhcuda_serve_stop",
)


class GeneratorParameters(ctypes.Structure):
    """MinhashCudaGeneratorParameters (minhashcuda.h)."""
    _fields_ = [("dim", ctypes.c_uint32), ("samples", ctypes.c_uint16),
                ("verbosity", ctypes.c_int)]


class Matrix(ctypes.Structure):
    """MHCUDAMatrix (minhashcuda.h)."""
    _fields_ = [("weights", ctypes.c_void_p), ("cols", ctypes.c_void_p),
                ("rows", ctypes.c_void_p), ("length", ctypes.c_uint32),
                ("output", ctypes.c_void_p)]


class TuneResult(ctypes.Structure):
    """MHCUDATuneResult (minhashcuda.h)."""
    _fields_ = [("key", ctypes.c_uint64), ("histogram_rows", ctypes.c_uint64),
                ("histogram_nnz", ctypes.c_uint64), ("kernel", ctypes.c_int),
                ("samples_per_lane", ctypes.c_int), ("segment", ctypes.c_int),
                ("waves_per_cu", ctypes.c_int), ("sample_ms", ctypes.c_float),
                ("sample_nnz", ctypes.c_uint64), ("cached", ctypes.c_int),
                ("stripe_mib", ctypes.c_int), ("seeds", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServeConfig(ctypes.Structure):
    """MHCUDAServeConfig (minhashcuda.h)."""
    _fields_ = [("memory_budget", ctypes.c_uint64), ("slots", ctypes.c_uint32),
                ("max_batch_requests", ctypes.c_uint32)]


class ServeStats(ctypes.Structure):
    """MHCUDAServeStats (minhashcuda.h)."""
    _fields_ = [("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("mean_requests_per_batch", ctypes.c_double),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("window_s", ctypes.c_double), ("nnz_samples_per_sec", ctypes.c_double),
                ("queue_ms", ctypes.c_double), ("stage_in_ms", ctypes.c_double),
                ("device_ms", ctypes.c_double), ("stage_out_ms", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def load(path):
    """Loads libMHCUDA.so and declares the prototypes of every export."""
    L = ctypes.CDLL(path)
    vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    L.mhcuda_init.restype = vp
    L.mhcuda_init.argtypes = [u32, ctypes.c_uint16, u32, i32, u32, i32, ctypes.POINTER(i32)]
    L.mhcuda_get_parameters.restype = GeneratorParameters
    L.mhcuda_get_parameters.argtypes = [vp]
    for name in ("mhcuda_retrieve_random_vars", "mhcuda_assign_random_vars"):
        getattr(L, name).restype = i32
        getattr(L, name).argtypes = [vp, vp, vp, vp]
    L.mhcuda_calc.restype = i32
    L.mhcuda_calc.argtypes = [vp, vp, vp, vp, u32, vp]
    L.mhcuda_fini.restype = i32
    L.mhcuda_fini.argtypes = [vp]
    L.mhcuda_calc_device.restype = i32
    L.mhcuda_calc_device.argtypes = [vp, vp, vp, vp, u32, vp, vp]
    L.mhcuda_calc_device_noalloc.restype = i32
    L.mhcuda_calc_device_noalloc.argtypes = [vp, vp, vp, vp, u32, vp, vp]
    L.mhcuda_calc_batch_device.restype = i32
    L.mhcuda_calc_batch_device.argtypes = [vp, u32, ctypes.POINTER(Matrix), i32, vp]
    L.mhcuda_calc_batch.restype = i32
    L.mhcuda_calc_batch.argtypes = [vp, u32, ctypes.POINTER(Matrix), i32]
    L.mhcuda_set_tuning.restype = i32
    L.mhcuda_set_tuning.argtypes = [vp, i32, i32, i32]
    L.mhcuda_get_tuning.restype = i32
    L.mhcuda_get_tuning.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                    ctypes.POINTER(i32)]
    L.mhcuda_set_stripes.restype = i32
    L.mhcuda_set_stripes.argtypes = [vp, i32]
    L.mhcuda_get_stripes.restype = i32
    L.mhcuda_get_stripes.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(u32),
                                     ctypes.POINTER(u32)]
    L.mhcuda_set_seeds.restype = i32
    L.mhcuda_set_seeds.argtypes = [vp, i32]
    L.mhcuda_get_seeds.restype = i32
    L.mhcuda_get_seeds.argtypes = [vp, ctypes.POINTER(i32)]
    L.mhcuda_set_skip.restype = i32
    L.mhcuda_set_skip.argtypes = [vp, i32]
    L.mhcuda_get_skip.restype = i32
    L.mhcuda_get_skip.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.mhcuda_autotune.restype = i32
    L.mhcuda_autotune.argtypes = [vp, u32, ctypes.POINTER(Matrix), vp, ctypes.POINTER(TuneResult)]
    L.mhcuda_jaccard_pairs.restype = i32
    L.mhcuda_jaccard_pairs.argtypes = [vp, vp, vp, vp, u32, vp, vp]
    L.mhcuda_lsh_band_keys.restype = i32
    L.mhcuda_lsh_band_keys.argtypes = [vp, vp, u32, u32, vp, vp]
    L.mhcuda_lsh_candidates.restype = i32
    L.mhcuda_lsh_candidates.argtypes = [vp, vp, u32, u32, u32, vp, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_uint64), vp]
    L.mhcuda_save_random_vars.restype = i32
    L.mhcuda_save_random_vars.argtypes = [vp, ctypes.c_char_p, i32]
    L.mhcuda_load_random_vars.restype = i32
    L.mhcuda_load_random_vars.argtypes = [vp, ctypes.c_char_p]
    L.mhcuda_params_file_info.restype = i32
    L.mhcuda_params_file_info.argtypes = [ctypes.c_char_p, ctypes.POINTER(u32),
                                          ctypes.POINTER(ctypes.c_uint16)]
    L.mhcuda_set_graph.restype = i32
    L.mhcuda_set_graph.argtypes = [vp, i32]
    L.mhcuda_graph_captures.restype = i32
    L.mhcuda_graph_captures.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.mhcuda_set_timing.restype = i32
    L.mhcuda_set_timing.argtypes = [vp, i32]
    L.mhcuda_kernel_time.restype = i32
    L.mhcuda_kernel_time.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.mhcuda_inject_fault.restype = i32
    L.mhcuda_inject_fault.argtypes = [vp, i32, u32]
    L.mhcuda_serve_start.restype = vp
    L.mhcuda_serve_start.argtypes = [vp, ctypes.POINTER(ServeConfig), ctypes.POINTER(i32)]
    L.mhcuda_serve_submit.restype = i32
    L.mhcuda_serve_submit.argtypes = [vp, vp, vp, vp, u32, vp, ctypes.POINTER(ctypes.c_uint64)]
    L.mhcuda_serve_wait.restype = i32
    L.mhcuda_serve_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    L.mhcuda_serve_stats.restype = i32
    L.mhcuda_serve_stats.argtypes = [vp, ctypes.POINTER(ServeStats), i32]
    L.mhcuda_serve_stop.restype = i32
    L.mhcuda_serve_stop.argtypes = [vp]
    return L


def check(code, what="call"):
    """Raises RuntimeError unless `code` is mhcudaSuccess."""
    if code != 0:
        raise RuntimeError("%s failed: %s" % (what, RESULT_NAMES.get(code, code)))
