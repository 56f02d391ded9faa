"""Thin Python binding over the C ABI in include/swift_spmm.h.

The C driver (host/swift_driver.cpp) is the primary host of the library; this
module exists for the test harness, the benchmark and the autograd wrapper.
It uses ctypes to call ``libswift_amd.so`` (built in place by
``csrc/Makefile``) and torch only for device memory and streams.  Every
compute call goes to the HIP library: there is no CPU fallback, and importing
the module without the built library raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswift_amd.so")

F32, F64, BF16, F16 = 0, 1, 2, 3
ROW_MAJOR, COL_MAJOR = 0, 1
KERNEL_DEFAULT, KERNEL_BALANCED, KERNEL_UNBALANCED, KERNEL_KCHUNK, KERNEL_ROWSORT, KERNEL_ROWVIEW, \
    KERNEL_FLAT, KERNEL_WINDOW = range(8)
B_SORTED, ACCUMULATE = 0x1, 0x2
BATCH_WIDTH_TUNED, BATCH_WIDTH_FULL, BATCH_WIDTH_HALF = 0, 1, 2

STATUS = {0: "OK", -1: "ERR_OPEN", -2: "ERR_BANNER", -3: "ERR_DATA", -4: "ERR_SIZE",
          -10: "ERR_ARG", -11: "ERR_HIP", -12: "ERR_ALLOC", -13: "ERR_UNSUPPORTED",
          -14: "ERR_VERIFY"}

_P = C.POINTER
_vp = C.c_void_p


class SwiftError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__("%s -> %s (%d): %s" % (fn, STATUS.get(code, "?"), code, msg))
        self.code = code


class Info(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("m", "n", "nnz", "vtype", "tilenum", "regular_nnz",
                                       "reside_col", "reside_nnz", "segsum", "numblk",
                                       "nsteps", "has_rowsort")] + [
        (n, _vp) for n in ("perm", "sortnnz", "sortrowidx_tmp", "sortval_tmp", "tile_ptr",
                           "tile_len", "tile_colidx", "tile_tall", "tile_format",
                           "sortrowidx", "sortval", "countrow", "segmentoffset",
                           "sortrowindex", "reside_cscptr", "reside_cscrowidx", "reside_val",
                           "colorblk", "blkstart", "blkstop", "blkptr", "val_position",
                           "val_offset", "rowsorted", "sortsrc", "reside_src")] + [
        ("build_ms", C.c_float * 4), ("rowview_nnz", C.c_int), ("segment_items", C.c_int),
        ("tuned_K", C.c_int), ("tuned_slots", C.c_int), ("tuned_kernel", C.c_int),
        ("tuned_cols", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("sparsity", C.c_double), ("empty_block_ratio", C.c_double),
                ("blocks", C.c_longlong), ("nonempty_blocks", C.c_longlong),
                ("regular_nnz", C.c_int), ("reside_nnz", C.c_int),
                ("regular_ratio", C.c_double), ("tilenum", C.c_int),
                ("tile_format_count", C.c_int * 3), ("max_row_nnz", C.c_int),
                ("max_col_nnz", C.c_int), ("mean_row_nnz", C.c_double),
                ("dense_blocks", C.c_longlong), ("dense_block_nnz", C.c_longlong),
                ("max_block_nnz", C.c_int)]


# every symbol include/swift_spmm.h declares
EXPORTS = (
    "swift_last_error", "swift_version", "swift_mmio_allinone", "swift_mmio_allinone_f32",
    "swift_mmio_allinone_ex", "swift_host_free", "swift_generate", "swift_generate_dense", "swift_csr_to_csc",
    "swift_col_sort", "swift_build", "swift_build_host", "swift_destroy",
    "swift_row_idx_sort", "swift_set_segment_items", "swift_get_info", "swift_matrix_stats",
    "swift_permute_dense", "swift_spmm", "swift_spmm_noalloc",
    "swift_spmm_timed",
    "swift_spmm_host", "swift_prepare_values", "swift_convert_dense", "swift_spmm_scaled",
    "swift_build_transpose", "swift_sddmm", "swift_backward_fused", "swift_update_values",
    "swift_batch_create", "swift_batch_spmm", "swift_batch_graph_launch",
    "swift_batch_params", "swift_batch_destroy", "swift_autotune", "swift_batch_set_lane_width",
)

_lib = None


def lib():
    """Load libswift_amd.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libswift_amd.so is not built (run __graft_entry__.build())")
        try:  # load torch's HIP runtime first so that one runtime serves both
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.swift_last_error.restype = C.c_char_p
        L.swift_version.restype = C.c_char_p
        L.swift_generate.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint64, _P(C.c_int),
                                     _P(C.c_int), _P(C.c_int), _P(_P(C.c_int)),
                                     _P(_P(C.c_int)), _P(_P(C.c_double)), _P(C.c_int)]
        L.swift_generate_dense.argtypes = [C.c_size_t, C.c_int, C.c_double, C.c_double,
                                           C.c_uint64, _vp]
        L.swift_host_free.argtypes = [_vp]
        L.swift_build.argtypes = [_P(_vp), C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                  _vp]
        L.swift_build_host.argtypes = [_P(_vp), C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp,
                                       _vp]
        L.swift_destroy.argtypes = [_vp]
        L.swift_destroy.restype = None
        L.swift_get_info.argtypes = [_vp, _P(Info)]
        L.swift_matrix_stats.argtypes = [_vp, _P(Stats), _vp]
        L.swift_row_idx_sort.argtypes = [_vp, _vp]
        L.swift_set_segment_items.argtypes = [_vp, C.c_int, _vp]
        L.swift_spmm.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, _vp, _vp,
                                 _vp]
        L.swift_spmm_noalloc.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint,
                                         _vp, _vp, _vp]
        L.swift_spmm_timed.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, _vp,
                                       _vp, _vp, _P(C.c_float)]
        L.swift_spmm_scaled.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_uint, _vp, _vp,
                                        _vp, _vp]
        L.swift_spmm_host.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                      C.c_int, C.c_int, _P(C.c_float), _P(C.c_double)]
        L.swift_permute_dense.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]
        L.swift_prepare_values.argtypes = [_vp, C.c_int, _vp]
        L.swift_convert_dense.argtypes = [_vp, _vp, C.c_size_t, C.c_int, C.c_int, C.c_int,
                                          _vp, _vp]
        L.swift_csr_to_csc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp,
                                       _vp, _vp, _vp, _vp]
        L.swift_col_sort.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp,
                                     _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        L.swift_build_transpose.argtypes = [_vp, _P(_vp), _vp]
        L.swift_sddmm.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp]
        L.swift_backward_fused.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]
        L.swift_update_values.argtypes = [_vp, _vp, _vp]
        L.swift_mmio_allinone.argtypes = [_P(C.c_int)] * 4 + [_P(_P(C.c_int)), _P(_P(C.c_int)),
                                                                _P(_P(C.c_double)), C.c_char_p]
        L.swift_mmio_allinone_ex.argtypes = [_P(C.c_int)] * 4 + [_P(_P(C.c_int)),
                                                                   _P(_P(C.c_int)), _P(_vp),
                                                                   C.c_int, C.c_char_p, C.c_int]
        L.swift_batch_create.argtypes = [_P(_vp), _vp, _vp, C.c_int, _vp]
        L.swift_batch_spmm.argtypes = [_vp, _vp, _vp, _vp]
        L.swift_batch_graph_launch.argtypes = [_vp, _vp, _vp, _vp]
        L.swift_batch_params.argtypes = [_vp, _P(C.c_int), _P(C.c_int), _P(C.c_int)]
        L.swift_batch_destroy.argtypes = [_vp]
        L.swift_batch_destroy.restype = None
        L.swift_autotune.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                     _P(C.c_int)]
        L.swift_batch_destroy.restype = None
        _lib = L
    return _lib


def check(fn, rc):
    if rc != 0:
        raise SwiftError(fn, rc, lib().swift_last_error().decode(errors="replace"))


def version():
    return lib().swift_version().decode()


# ------------------------------------------------------------------ host --

def _take(ptr, n, dtype):
    if n <= 0:
        out = np.zeros(0, dtype=dtype)
    else:
        out = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)
    lib().swift_host_free(C.cast(ptr, _vp))
    return out


def generate(which, scale=1.0, seed=0, index=0):
    """Seeded configuration matrix -> dict(m, n, nnz, rowptr, colidx, val, k)."""
    m, n, nnz, k = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rp, ci, vv = _P(C.c_int)(), _P(C.c_int)(), _P(C.c_double)()
    check("swift_generate", lib().swift_generate(which, index, scale, seed, C.byref(m),
                                                 C.byref(n), C.byref(nnz), C.byref(rp),
                                                 C.byref(ci), C.byref(vv), C.byref(k)))
    return dict(m=m.value, n=n.value, nnz=nnz.value, k=k.value,
                rowptr=_take(rp, m.value + 1, np.int32),
                colidx=_take(ci, nnz.value, np.int32),
                val=_take(vv, nnz.value, np.float64))


def generate_dense(count, dist=0, lo=-1.0, hi=1.0, seed=0):
    out = np.empty(count, np.float64)
    check("swift_generate_dense", lib().swift_generate_dense(count, dist, lo, hi, seed,
                                                             out.ctypes.data))
    return out


def mmio_load(path, threads=None, dtype=F64):
    """swift_mmio_allinone (threads None) or swift_mmio_allinone_ex with that
    parser thread count (0 automatic, 1 sequential) -> (status, dict or None)."""
    m, n, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rp, ci = _P(C.c_int)(), _P(C.c_int)()
    if threads is None:
        vv = _P(C.c_double)()
        rc = lib().swift_mmio_allinone(C.byref(m), C.byref(n), C.byref(nnz), C.byref(sym),
                                       C.byref(rp), C.byref(ci), C.byref(vv), path.encode())
    else:
        raw = _vp()
        rc = lib().swift_mmio_allinone_ex(C.byref(m), C.byref(n), C.byref(nnz), C.byref(sym),
                                          C.byref(rp), C.byref(ci), C.byref(raw), dtype,
                                          path.encode(), threads)
        if rc == 0:
            vv = C.cast(raw, _P(C.c_double if dtype == F64 else C.c_float))
    if rc != 0:
        return rc, None
    if threads is not None and dtype == F32:
        return 0, dict(m=m.value, n=n.value, nnz=nnz.value, sym=sym.value,
                       rowptr=_take(rp, m.value + 1, np.int32),
                       colidx=_take(ci, nnz.value, np.int32),
                       val=_take(vv, nnz.value, np.float32))
    return 0, dict(m=m.value, n=n.value, nnz=nnz.value, sym=sym.value,
                   rowptr=_take(rp, m.value + 1, np.int32),
                   colidx=_take(ci, nnz.value, np.int32),
                   val=_take(vv, nnz.value, np.float64))


# ---------------------------------------------------------------- device --

def _torch():
    import torch
    return torch


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _dtype_code(dt):
    torch = _torch()
    return {torch.float32: F32, torch.float64: F64, torch.bfloat16: BF16,
            torch.float16: F16}[dt]


class SwiftMatrix:
    """A device-resident Swift format (swift_build) and its SpMM entry points."""

    def __init__(self, m, n, rowptr, colidx, val, dtype=None, device="cuda"):
        torch = _torch()
        if dtype is None:
            dtype = torch.float32
        self.m, self.n = int(m), int(n)
        rp = torch.as_tensor(np.asarray(rowptr, np.int32)).to(device)
        ci = torch.as_tensor(np.asarray(colidx, np.int32)).to(device)
        vv = torch.as_tensor(np.asarray(val)).to(device=device, dtype=dtype)
        self.nnz = int(ci.numel())
        self.dtype = dtype
        self.vtype = _dtype_code(dtype)
        h = _vp()
        check("swift_build", lib().swift_build(C.byref(h), self.m, self.n, self.nnz, self.vtype,
                                               _ptr(rp), _ptr(ci), _ptr(vv), _stream()))
        torch.cuda.synchronize()
        self.h = h
        self._t = None
        self.lowp = None

    @classmethod
    def from_handle(cls, h, like):
        obj = cls.__new__(cls)
        obj.h = h
        info = Info()
        check("swift_get_info", lib().swift_get_info(h, C.byref(info)))
        obj.m, obj.n, obj.nnz = info.m, info.n, info.nnz
        obj.dtype, obj.vtype = like.dtype, like.vtype
        obj._t = None
        obj.lowp = None
        return obj

    def close(self):
        if getattr(self, "h", None):
            if self._t is not None:
                self._t.close()
                self._t = None
            lib().swift_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- format dump ------------------------------------------------------
    def info(self):
        inf = Info()
        check("swift_get_info", lib().swift_get_info(self.h, C.byref(inf)))
        return inf

    def stats(self):
        """swift_matrix_stats: sparsity, empty 32x32 block share, part sizes."""
        st = Stats()
        check("swift_matrix_stats", lib().swift_matrix_stats(self.h, C.byref(st), _stream()))
        return st

    def arrays(self, names=None):
        """Copy format arrays to host numpy (for bit-exact comparisons)."""
        torch = _torch()
        inf = self.info()
        vnp = np.float64 if self.vtype == F64 else np.float32
        sizes = {
            "perm": (inf.n, np.int32), "sortnnz": (inf.n, np.int32),
            "sortrowidx_tmp": (inf.nnz, np.int32), "sortval_tmp": (inf.nnz, vnp),
            "tile_ptr": (inf.tilenum + 1, np.int32), "tile_len": (inf.tilenum, np.int32),
            "tile_colidx": (inf.tilenum, np.int32), "tile_tall": (inf.tilenum, np.int32),
            "tile_format": (inf.tilenum, np.int32), "sortrowidx": (inf.regular_nnz, np.int32),
            "sortval": (inf.regular_nnz, vnp), "countrow": (inf.tilenum + 1, np.int32),
            "segmentoffset": (inf.segsum + 1, np.int32),
            "sortrowindex": (inf.regular_nnz, np.int32),
            "reside_cscptr": (inf.reside_col + 1, np.int32),
            "reside_cscrowidx": (inf.reside_nnz, np.int32), "reside_val": (inf.reside_nnz, vnp),
            "colorblk": (inf.numblk, np.uint32), "blkstart": (inf.numblk, np.int32),
            "blkstop": (inf.numblk, np.int32), "blkptr": (inf.numblk, np.int32),
            "sortsrc": (inf.regular_nnz, np.int32), "reside_src": (inf.reside_nnz, np.int32),
        }
        if inf.has_rowsort:
            for k in ("val_position", "val_offset", "rowsorted"):
                sizes[k] = (inf.regular_nnz, np.int32)
        out = {k: getattr(inf, k) for k in ("m", "n", "nnz", "tilenum", "regular_nnz",
                                            "reside_col", "reside_nnz", "segsum", "numblk",
                                            "nsteps")}
        torch.cuda.synchronize()
        for k, (cnt, dt) in sizes.items():
            if names is not None and k not in names:
                continue
            arr = np.zeros(cnt, dtype=dt)
            ptr = getattr(inf, k)
            if cnt > 0 and ptr:
                hip = _hip()
                rc = hip.hipMemcpy(C.c_void_p(arr.ctypes.data), C.c_void_p(ptr),
                                   C.c_size_t(arr.nbytes), 2)  # device to host
                if rc != 0:
                    raise RuntimeError("hipMemcpy failed: %d" % rc)
            out[k] = arr
        return out

    def autotune(self, B, out):
        """swift_autotune on these operands (out is overwritten); returns the
        chosen segment size."""
        B = B.contiguous()
        best = C.c_int(0)
        check("swift_autotune", lib().swift_autotune(self.h, B.shape[1], _dtype_code(B.dtype),
                                                     ROW_MAJOR, _ptr(B), _ptr(out), _stream(),
                                                     C.byref(best)))
        return best.value

    def set_segment_items(self, items):
        check("swift_set_segment_items", lib().swift_set_segment_items(self.h, items, _stream()))
        return self

    def row_idx_sort(self):
        check("swift_row_idx_sort", lib().swift_row_idx_sort(self.h, _stream()))
        return self

    # -- compute ----------------------------------------------------------
    def spmm(self, B, out=None, layout=ROW_MAJOR, kernel=KERNEL_DEFAULT, flags=0,
             dtype=None, stream=None):
        """C = A*B.  B is (n, K) for ROW_MAJOR or (K, n) for COL_MAJOR."""
        torch = _torch()
        B = B.contiguous()
        K = B.shape[1] if layout == ROW_MAJOR else B.shape[0]
        code = _dtype_code(B.dtype) if dtype is None else dtype
        cdt = torch.float64 if code == F64 else torch.float32
        if out is None:
            shape = (self.m, K) if layout == ROW_MAJOR else (K, self.m)
            out = torch.empty(shape, dtype=cdt, device=B.device)
        check("swift_spmm", lib().swift_spmm(self.h, K, layout, code, kernel, flags, _ptr(B),
                                             _ptr(out), _stream(stream)))
        return out

    def spmm_timed(self, B, out, layout=ROW_MAJOR, kernel=KERNEL_DEFAULT, flags=0,
                   stream=None):
        """swift_spmm_timed: returns [zero, regular, reside] kernel ms."""
        B = B.contiguous()
        K = B.shape[1] if layout == ROW_MAJOR else B.shape[0]
        ms = (C.c_float * 3)()
        check("swift_spmm_timed", lib().swift_spmm_timed(self.h, K, layout,
                                                         _dtype_code(B.dtype), kernel, flags,
                                                         _ptr(B), _ptr(out), _stream(stream),
                                                         ms))
        return [ms[0], ms[1], ms[2]]

    def spmm_scaled(self, Bh, scale, out=None, layout=ROW_MAJOR, kernel=KERNEL_DEFAULT,
                    flags=0):
        torch = _torch()
        K = Bh.shape[1] if layout == ROW_MAJOR else Bh.shape[0]
        if out is None:
            shape = (self.m, K) if layout == ROW_MAJOR else (K, self.m)
            out = torch.empty(shape, dtype=torch.float32, device=Bh.device)
        check("swift_spmm_scaled", lib().swift_spmm_scaled(self.h, K, layout, kernel, flags,
                                                           _ptr(Bh), _ptr(scale), _ptr(out),
                                                           _stream()))
        return out

    def prepare_values(self, code):
        check("swift_prepare_values", lib().swift_prepare_values(self.h, code, _stream()))
        self.lowp = code

    def permute_dense(self, B, layout=ROW_MAJOR):
        torch = _torch()
        B = B.contiguous()
        K = B.shape[1] if layout == ROW_MAJOR else B.shape[0]
        out = torch.empty_like(B)
        check("swift_permute_dense", lib().swift_permute_dense(self.h, K, layout,
                                                               _dtype_code(B.dtype), _ptr(B),
                                                               _ptr(out), _stream()))
        return out

    def transpose(self):
        """Swift layout of A^T (built once, cached)."""
        if self._t is None:
            h = _vp()
            check("swift_build_transpose", lib().swift_build_transpose(self.h, C.byref(h),
                                                                       _stream()))
            self._t = SwiftMatrix.from_handle(h, self)
        return self._t

    def sddmm(self, dC, B):
        torch = _torch()
        odt = torch.float32 if dC.dtype == torch.bfloat16 else dC.dtype
        dvals = torch.empty(self.nnz, dtype=odt, device=dC.device)
        check("swift_sddmm", lib().swift_sddmm(self.h, dC.shape[1], _dtype_code(dC.dtype),
                                               _ptr(dC.contiguous()), _ptr(B.contiguous()),
                                               _ptr(dvals), _stream()))
        return dvals

    def backward_fused(self, dC, B, want_dB=True, want_dvals=True):
        """dB = A^T dC and dvals = SDDMM(dC, B).  BF16 dC and B (FP32 build):
        BF16 values of A^T (prepared on first use), FP32 dB and dvals."""
        torch = _torch()
        K = dC.shape[1]
        odt = torch.float32 if dC.dtype == torch.bfloat16 else dC.dtype
        dB = torch.empty((self.n, K), dtype=odt, device=dC.device) if want_dB else None
        dv = torch.empty(self.nnz, dtype=odt, device=dC.device) if want_dvals else None
        T = self.transpose()
        if dC.dtype == torch.bfloat16 and getattr(T, "lowp", None) != BF16:
            T.prepare_values(BF16)
        check("swift_backward_fused", lib().swift_backward_fused(
            T.h, K, _dtype_code(dC.dtype), _ptr(dC.contiguous()),
            _ptr(B.contiguous()) if B is not None else None, _ptr(dB), _ptr(dv), _stream()))
        return dB, dv

    def update_values(self, csr_val):
        check("swift_update_values", lib().swift_update_values(self.h, _ptr(csr_val.contiguous()),
                                                               _stream()))
        if self._t is not None:
            check("swift_update_values",
                  lib().swift_update_values(self._t.h, _ptr(csr_val.contiguous()), _stream()))


def convert_dense(B, code, layout=ROW_MAJOR, col_scale=False):
    """FP32 B -> BF16/F16 (and per-column scales for F16)."""
    torch = _torch()
    B = B.contiguous()
    K = B.shape[1] if layout == ROW_MAJOR else B.shape[0]
    rows = B.numel() // K
    tdt = torch.bfloat16 if code == BF16 else torch.float16
    out = torch.empty(B.shape, dtype=tdt, device=B.device)
    sc = torch.empty(K, dtype=torch.float32, device=B.device) if col_scale else None
    check("swift_convert_dense", lib().swift_convert_dense(_ptr(B), _ptr(out), rows, K, layout,
                                                           code, _ptr(sc), _stream()))
    return (out, sc) if col_scale else out


_hip_lib = None


def _hip():
    global _hip_lib
    if _hip_lib is None:
        lib()  # the runtime already in the process (same soname)
        _hip_lib = C.CDLL("libamdhip64.so.7")
        _hip_lib.hipMemcpy.argtypes = [_vp, _vp, C.c_size_t, C.c_int]
    return _hip_lib


# -------------------------------------------------------------- batched --

def batch_set_lane_width(mode):
    """Lane widths the batch autotuner times for later batches (process-wide)."""
    check("swift_batch_set_lane_width", lib().swift_batch_set_lane_width(mode))


class Batch:
    """swift_batch_*: many FP32 row-major products C_i = A_i * B_i in one
    persistent launch (optionally replayed from a captured HIP graph)."""

    def __init__(self, mats, Ks):
        self.mats = list(mats)
        self.Ks = [int(k) for k in Ks]
        n = len(self.mats)
        handles = (_vp * n)(*[m.h for m in self.mats])
        ks = (C.c_int * n)(*self.Ks)
        h = _vp()
        check("swift_batch_create", lib().swift_batch_create(C.byref(h), handles, ks, n,
                                                             _stream()))
        self.h = h

    def _ptrs(self, Bs, Cs):
        n = len(self.mats)
        if len(Bs) != n or len(Cs) != n:
            raise ValueError("need %d B and %d C tensors" % (n, n))
        pb = (_vp * n)(*[_ptr(b) for b in Bs])
        pc = (_vp * n)(*[_ptr(c) for c in Cs])
        return pb, pc

    def spmm(self, Bs, Cs, graph=False):
        pb, pc = self._ptrs(Bs, Cs)
        fn = "swift_batch_graph_launch" if graph else "swift_batch_spmm"
        check(fn, getattr(lib(), fn)(self.h, pb, pc, _stream()))
        return Cs

    def params(self):
        u, d, t = C.c_int(), C.c_int(), C.c_int()
        check("swift_batch_params", lib().swift_batch_params(self.h, C.byref(u), C.byref(d),
                                                             C.byref(t)))
        return {"work_units": u.value, "depth": d.value, "block_threads": t.value}

    def close(self):
        if getattr(self, "h", None):
            lib().swift_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- autograd --

def spmm_function():
    """torch.autograd.Function: C = A(vals) * B with gradients for B and vals."""
    torch = _torch()

    class SwiftSpMM(torch.autograd.Function):
        """BF16 B (FP32 matrix): BF16 storage of A's values and B with FP32
        C; the fused backward takes dC rounded to BF16 and returns dB rounded
        to BF16 and FP32 dvals."""

        @staticmethod
        def forward(ctx, vals, B, mat, fused=True):
            mat.update_values(vals.detach())
            ctx.mat, ctx.fused = mat, fused
            ctx.save_for_backward(B)
            if B.dtype == torch.bfloat16:
                if mat.lowp != BF16:
                    mat.prepare_values(BF16)
                return mat.spmm(B.detach(), dtype=BF16)
            return mat.spmm(B.detach())

        @staticmethod
        def backward(ctx, dC):
            (B,) = ctx.saved_tensors
            mat = ctx.mat
            dC = dC.contiguous()
            need_v, need_b = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
            dB = dv = None
            if B.dtype == torch.bfloat16:
                dB, dv = mat.backward_fused(dC.to(torch.bfloat16), B.detach(),
                                            want_dB=need_b, want_dvals=need_v)
                return dv, (dB.to(torch.bfloat16) if dB is not None else None), None, None
            if ctx.fused:
                dB, dv = mat.backward_fused(dC, B.detach(), want_dB=need_b, want_dvals=need_v)
            else:
                if need_b:
                    dB = mat.transpose().spmm(dC)
                if need_v:
                    dv = mat.sddmm(dC, B.detach())
            return dv, dB, None, None

    return SwiftSpMM
