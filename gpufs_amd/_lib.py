"""ctypes binding of libgpufs_amd.so (the C ABI in include/gpufs_amd.h).

The binding fails loudly when the native library is missing: there is no
fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpufs_amd.so")

O_RDONLY, O_WRONLY, O_RDWR, O_CREAT = os.O_RDONLY, os.O_WRONLY, os.O_RDWR, os.O_CREAT
O_GRDONLY, O_GWRONLY, O_GRDWR, O_GCREAT = O_RDONLY, O_WRONLY, O_RDWR, O_CREAT
O_GWRONCE = O_CREAT | O_WRONLY


class Config(ctypes.Structure):
    _fields_ = [
        ("cache_bytes", ctypes.c_uint64),
        ("page_size", ctypes.c_uint32),
        ("max_files", ctypes.c_uint32),
        ("ring_depth", ctypes.c_uint32),
        ("host_threads", ctypes.c_uint32),
        ("tree_nodes", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


STATS_FIELDS = [
    "page_hits", "page_misses", "page_waits", "evictions", "alloc_fail", "rpc_reads", "rpc_writes",
    "rpc_opens", "rpc_closes", "bytes_filled", "bytes_written", "nodes_used", "files_used",
    "host_requests", "host_bytes_read", "host_bytes_written", "frames", "frames_valid", "timeouts",
    "host_busy_ns", "vm_errors", "slots_lost", "files_reclaimed", "nodes_recycled",
    "spec_parts", "spec_rescans", "spec_changed",
]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in STATS_FIELDS]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n in STATS_FIELDS}


class ServeConfig(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("workgroups", "threads", "slots", "chunk_bytes", "n_words", "frame_budget", "parts_max",
                 "min_part_kib")]


# name -> (restype, argtypes) of every exported entry point
P, I, U32, U64, I32, I64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64,
                            ctypes.c_int32, ctypes.c_int64)
SIGNATURES = {
    "gpufs_config_default": (I, [ctypes.POINTER(Config)]),
    "gpufs_init": (I, [ctypes.POINTER(Config), ctypes.POINTER(P)]),
    "gpufs_destroy": (I, [P]),
    "gpufs_stream": (P, [P]),
    "gpufs_device_state": (P, [P]),
    "gpufs_get_config": (I, [P, ctypes.POINTER(Config)]),
    "gpufs_run_handler": (I, [P]),
    "gpufs_host_start": (I, [P]),
    "gpufs_host_stop": (I, [P]),
    "gpufs_host_kick": (I, [P]),
    "gpufs_cache_drop": (I, [P]),
    "gpufs_stats_get": (I, [P, ctypes.POINTER(Stats)]),
    "gpufs_stats_reset": (I, [P]),
    "gpufs_check_invariants": (I, [P]),
    "gpufs_strerror": (ctypes.c_char_p, [I]),
    "gpufs_launch_gread_ranges": (I, [P, P, U32, U64, P, P, U32, P]),
    "gpufs_launch_gwrite_ranges": (I, [P, P, I32, U32, U64, P, I32, P, U32, P]),
    "gpufs_launch_grep": (I, [P, P, U32, U32, P, U32, U32, P, P, U32, P]),
    "gpufs_launch_grep_split": (I, [P, P, U32, U32, P, U32, U32, P, P, U32, U64, U32, P]),
    "gpufs_launch_grep_noalloc": (I, [P, P, U32, U32, P, U32, U32, P, P, U32, U64, U32, P]),
    "gpufs_grep_batched_workspace_bytes": (U64, [U32, U32]),
    "gpufs_launch_grep_batched": (I, [P, P, U32, U32, P, U32, U32, P, P, P, U32, U32, U32, U64, P, P]),
    "gpufs_grep_batched_graph": (I, [P, P, U32, U32, P, U32, U32, P, P, P, U32, U32, U32, U64, P, ctypes.POINTER(P)]),
    "gpufs_graph_launch": (I, [P, P]),
    "gpufs_graph_destroy": (I, [P]),
    "gpufs_serve_start": (I, [P, ctypes.POINTER(ServeConfig), P, U32, U32, P, P, ctypes.POINTER(P)]),
    "gpufs_serve_run": (I, [P, P, P, U32, P, P, P, ctypes.POINTER(ctypes.c_double)]),
    "gpufs_serve_stop": (I, [P]),
    "gpufs_serve_pages": (I, [P, P, P]),
    "gpufs_serve_parts": (I, [P, P, P]),
    "gpufs_launch_mmap_warp_records": (I, [P, P, U64, U64, P, P, U32, U32, P]),
    "gpufs_grep_text": (I, [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, U32, U32, U32, P, P]),
    "gpufs_grep_text_records": (I, [P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, U32, U32, U32, P, P]),
    "gpufs_launch_vm_records": (I, [P, P, U64, P, U64, P, P, U32, U32, U32, P]),
    "gpufs_launch_gemv_mmap": (I, [P, P, U64, U64, P, P, U32, P, P, U32, P]),
    "gpufs_fnv1a64_ranges": (I, [P, U32, U64, P, P]),
    "gpufs_launch_read": (I, [P, P, I32, U64, U64, P, P, U32, P]),
    "gpufs_launch_write": (I, [P, P, I32, U64, U64, P, I32, P, U32, P]),
    "gpufs_launch_api_script": (I, [P, P, P, U32, P, P, U32, P]),
}

_LIB = None


def _adopt_torch_hip_runtime():
    """Load PyTorch's HIP runtime first when PyTorch is installed.

    libgpufs_amd.so needs libamdhip64.so.7.  PyTorch ships its own copy with the
    same soname; if that copy is already mapped the dynamic linker reuses it,
    so the process has one HIP runtime shared by torch tensors and this
    library.  Loaded the other way round, torch would map a second runtime and
    lose the device."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    global _LIB
    if _LIB is None:
        _adopt_torch_hip_runtime()
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m gpufs_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class GpufsError(RuntimeError):
    def __init__(self, where, code):
        super().__init__(f"{where}: {code} ({lib().gpufs_strerror(code).decode()})")
        self.code = code


def check(where, rc):
    if rc < 0:
        raise GpufsError(where, rc)
    return rc


def default_config(**kw):
    c = Config()
    check("gpufs_config_default", lib().gpufs_config_default(ctypes.byref(c)))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Context:
    """One gpufs context (page cache + file table + RPC ring) on one GPU."""

    def __init__(self, **cfg):
        self.cfg = default_config(**cfg)
        h = ctypes.c_void_p()
        check("gpufs_init", lib().gpufs_init(ctypes.byref(self.cfg), ctypes.byref(h)))
        self.h = h

    @property
    def stream(self):
        return lib().gpufs_stream(self.h)

    def host_start(self):
        check("gpufs_host_start", lib().gpufs_host_start(self.h))

    def host_stop(self):
        check("gpufs_host_stop", lib().gpufs_host_stop(self.h))

    def run_handler(self):
        check("gpufs_run_handler", lib().gpufs_run_handler(self.h))

    def cache_drop(self):
        check("gpufs_cache_drop", lib().gpufs_cache_drop(self.h))

    def stats(self):
        s = Stats()
        check("gpufs_stats_get", lib().gpufs_stats_get(self.h, ctypes.byref(s)))
        return s.as_dict()

    def stats_reset(self):
        check("gpufs_stats_reset", lib().gpufs_stats_reset(self.h))

    def check_invariants(self):
        return lib().gpufs_check_invariants(self.h)

    def close(self):
        if self.h:
            lib().gpufs_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
