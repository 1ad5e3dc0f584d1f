"""hipCUB (rocPRIM) over unpacked u64 elements: the GPU library baselines the
reference times beside its own kernels, run on this GPU.

  src/benchs/bench_compressstore_baseline_cub.cu:29-35  cub::DeviceSelect::Flagged
  src/benchs/bench_hop_baseline_cub.cu:32-38            cub::DeviceReduce (sum)

Measurement only: bench.py times these next to the library's operators and a
GPU test checks that they compute the same results. Nothing in amethyst_amd
imports this module.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhipcub_baseline.so")
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (make -C amethyst_amd/baseline)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64 = ctypes.c_void_p, ctypes.c_uint64
        L.hb_select_flagged_u64.argtypes = [vp, vp, vp, vp, u64, vp, ctypes.POINTER(ctypes.c_size_t), vp]
        L.hb_select_flagged_u64.restype = ctypes.c_int
        L.hb_select_flagged_u64_bytes.argtypes = [vp, vp, vp, vp, u64, vp, ctypes.POINTER(ctypes.c_size_t), vp]
        L.hb_select_flagged_u64_bytes.restype = ctypes.c_int
        L.hb_reduce_sum_u64.argtypes = [vp, vp, u64, vp, ctypes.POINTER(ctypes.c_size_t), vp]
        L.hb_reduce_sum_u64.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class SelectFlagged:
    """hipcub::DeviceSelect::Flagged of n u64 elements (an int64 device tensor)
    by an LSB-first u64 bitmask, or with bytes=True by one flag byte per
    element (a uint8 device tensor); run() writes out[:count] and count[0]."""

    def __init__(self, vals, mask, n, bytes=False):
        import torch
        self.vals, self.mask, self.n = vals, mask, int(n)
        self.fn = lib().hb_select_flagged_u64_bytes if bytes else lib().hb_select_flagged_u64
        self.out = torch.empty(max(1, self.n), dtype=torch.int64, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int64, device="cuda")
        tb = ctypes.c_size_t(0)
        if self.fn(_ptr(vals), _ptr(mask), _ptr(self.out), _ptr(self.count), self.n, None, ctypes.byref(tb),
                   _stream()):
            raise RuntimeError("hipcub::DeviceSelect::Flagged (size query) failed")
        self.tb = tb
        self.temp = torch.empty(max(1, (tb.value + 7) // 8), dtype=torch.int64, device="cuda")

    def run(self):
        if self.fn(_ptr(self.vals), _ptr(self.mask), _ptr(self.out), _ptr(self.count), self.n, _ptr(self.temp),
                   ctypes.byref(self.tb), _stream()):
            raise RuntimeError("hipcub::DeviceSelect::Flagged failed")


class ReduceSum:
    """hipcub::DeviceReduce::Sum of n u64 elements (wrapping); run() writes out[0]."""

    def __init__(self, vals, n):
        import torch
        self.vals, self.n = vals, int(n)
        self.out = torch.zeros(1, dtype=torch.int64, device="cuda")
        tb = ctypes.c_size_t(0)
        if lib().hb_reduce_sum_u64(_ptr(vals), _ptr(self.out), self.n, None, ctypes.byref(tb), _stream()):
            raise RuntimeError("hipcub::DeviceReduce::Sum (size query) failed")
        self.tb = tb
        self.temp = torch.empty(max(1, (tb.value + 7) // 8), dtype=torch.int64, device="cuda")

    def run(self):
        if lib().hb_reduce_sum_u64(_ptr(self.vals), _ptr(self.out), self.n, _ptr(self.temp), ctypes.byref(self.tb),
                                   _stream()):
            raise RuntimeError("hipcub::DeviceReduce::Sum failed")
