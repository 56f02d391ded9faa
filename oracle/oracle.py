"""ctypes front end of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (the plain-C restatement in swift_oracle.c) and,
when present, ``oracle/_ref/libswiftref_{f64,f32}.so`` (the reference's own CPU
headers compiled in place by oracle/Makefile).  Only tests, the smoke entry
point and the benchmark's CPU-baseline leg import this module, always to
compute expected results; the product path never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = C.POINTER
_i32p = _P(C.c_int)
_f64p = _P(C.c_double)
_f32p = _P(C.c_float)
_u32p = _P(C.c_uint)


class _Swift(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("m", "n", "nnz")] + [
        ("perm", _i32p), ("sortnnz", _i32p), ("sortrowidx_tmp", _i32p),
        ("sortval_tmp", _f64p),
        ("tilenum", C.c_int), ("regular_nnz", C.c_int),
        ("reside_col", C.c_int), ("reside_nnz", C.c_int),
        ("tile_ptr", _i32p), ("tile_len", _i32p), ("tile_colidx", _i32p),
        ("tile_tall", _i32p), ("sortrowidx", _i32p), ("sortval", _f64p),
        ("reside_cscptr", _i32p), ("reside_cscrowidx", _i32p),
        ("reside_val", _f64p), ("tile_format", _i32p), ("segsum", C.c_int),
        ("countrow", _i32p), ("segmentoffset", _i32p), ("sortrowindex", _i32p),
        ("numblk", C.c_int), ("colorblk", _u32p), ("blkstart", _i32p),
        ("blkstop", _i32p), ("blkptr", _i32p), ("val_position", _i32p),
        ("val_offset", _i32p), ("rowsorted", _i32p)]


def _arr(ptr, n, dtype):
    if n <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _ip(a):
    return a.ctypes.data_as(_i32p)


def _dp(a):
    return a.ctypes.data_as(_f64p)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so not built (make -C oracle)")
        L = C.CDLL(path)
        L.oracle_build.restype = _P(_Swift)
        L.oracle_build_from_sorted.restype = _P(_Swift)
        L.oracle_swift_cpu_f32_timed.restype = C.c_double
        _lib = L
    return _lib


# the names of the format fields and their lengths, as (name, len-attr, dtype)
_FIELDS = [
    ("perm", "n", np.int32), ("sortnnz", "n", np.int32),
    ("sortrowidx_tmp", "nnz", np.int32), ("sortval_tmp", "nnz", np.float64),
    ("tile_ptr", "tilenum+1", np.int32), ("tile_len", "tilenum", np.int32),
    ("tile_colidx", "tilenum", np.int32), ("tile_tall", "tilenum", np.int32),
    ("sortrowidx", "regular_nnz", np.int32), ("sortval", "regular_nnz", np.float64),
    ("reside_cscptr", "reside_col+1", np.int32),
    ("reside_cscrowidx", "reside_nnz", np.int32),
    ("reside_val", "reside_nnz", np.float64),
    ("tile_format", "tilenum", np.int32), ("countrow", "tilenum+1", np.int32),
    ("segmentoffset", "segsum+1", np.int32),
    ("sortrowindex", "regular_nnz", np.int32),
    ("colorblk", "numblk", np.uint32), ("blkstart", "numblk", np.int32),
    ("blkstop", "numblk", np.int32), ("blkptr", "numblk", np.int32),
]
_ROWSORT = [("val_position", np.int32), ("val_offset", np.int32),
            ("rowsorted", np.int32)]
SCALARS = ("m", "n", "nnz", "tilenum", "regular_nnz", "reside_col",
           "reside_nnz", "segsum", "numblk")


def _length(s, spec):
    if spec.endswith("+1"):
        return getattr(s, spec[:-2]) + 1
    return getattr(s, spec)


class SwiftFormat:
    """Owns one oracle_swift; ``.a`` holds numpy copies of every field."""

    def __init__(self, ptr):
        self._p = ptr
        s = ptr.contents
        self.a = {k: getattr(s, k) for k in SCALARS}
        for name, spec, dt in _FIELDS:
            self.a[name] = _arr(getattr(s, name), _length(s, spec), dt)

    def __getitem__(self, k):
        return self.a[k]

    def row_idx_sort(self):
        lib().oracle_row_idx_sort(self._p)
        s = self._p.contents
        for name, dt in _ROWSORT:
            self.a[name] = _arr(getattr(s, name), s.regular_nnz, dt)
        return self

    def swift_cpu(self, K, sortB, dtype=np.float64):
        """Swift_cpu (src/Swift_cpu.h:31-65): row-major sorted B -> C."""
        m = self.a["m"]
        if dtype == np.float32:
            B = np.ascontiguousarray(sortB, dtype=np.float32)
            out = np.zeros((m, K), dtype=np.float32)
            lib().oracle_swift_cpu_f32(self._p, K, B.ctypes.data_as(_f32p),
                                       out.ctypes.data_as(_f32p))
        else:
            B = np.ascontiguousarray(sortB, dtype=np.float64)
            out = np.zeros((m, K), dtype=np.float64)
            lib().oracle_swift_cpu_f64(self._p, K, _dp(B), _dp(out))
        return out

    def swift_cpu_f32_timed(self, K, sortB):
        B = np.ascontiguousarray(sortB, dtype=np.float32)
        out = np.zeros((self.a["m"], K), dtype=np.float32)
        ms = lib().oracle_swift_cpu_f32_timed(self._p, K, B.ctypes.data_as(_f32p),
                                              out.ctypes.data_as(_f32p))
        return out, ms

    def __del__(self):
        if getattr(self, "_p", None) is not None and _lib is not None:
            _lib.oracle_swift_free(self._p)
            self._p = None


def build(m, n, rowptr, colidx, val):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colidx = np.ascontiguousarray(colidx, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    return SwiftFormat(lib().oracle_build(m, n, int(rowptr[-1]), _ip(rowptr),
                                          _ip(colidx), _dp(val)))


def build_from_sorted(m, n, perm, sortnnz, sortrowidx_tmp, sortval_tmp):
    sortnnz = np.ascontiguousarray(sortnnz, dtype=np.int32)
    sr = np.ascontiguousarray(sortrowidx_tmp, dtype=np.int32)
    sv = np.ascontiguousarray(sortval_tmp, dtype=np.float64)
    perm = None if perm is None else np.ascontiguousarray(perm, dtype=np.int32)
    return SwiftFormat(lib().oracle_build_from_sorted(
        m, n, len(sr), None if perm is None else _ip(perm), _ip(sortnnz),
        _ip(sr), _dp(sv)))


def csr2csc(m, n, rowptr, colidx, val):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colidx = np.ascontiguousarray(colidx, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    nnz = int(rowptr[-1])
    ptr = np.zeros(n + 1, np.int32)
    row = np.zeros(nnz, np.int32)
    cv = np.zeros(nnz, np.float64)
    npc = np.zeros(n, np.int32)
    lib().oracle_csr2csc(m, n, nnz, _ip(rowptr), _ip(colidx), _dp(val), _ip(ptr),
                         _ip(row), _dp(cv), _ip(npc))
    return ptr, row, cv, npc


def col_sort(n, npc, cscptr, cscrow, cscval):
    nnz = len(cscrow)
    perm = np.zeros(n, np.int32)
    sn = np.zeros(n, np.int32)
    sr = np.zeros(nnz, np.int32)
    sv = np.zeros(nnz, np.float64)
    lib().oracle_col_sort(n, _ip(np.ascontiguousarray(npc, np.int32)),
                          _ip(np.ascontiguousarray(cscptr, np.int32)),
                          _ip(np.ascontiguousarray(cscrow, np.int32)),
                          _dp(np.ascontiguousarray(cscval, np.float64)),
                          _ip(perm), _ip(sn), _ip(sr), _dp(sv))
    return perm, sn, sr, sv


def mmio_load(path):
    m, n, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rp, ci, vv = _i32p(), _i32p(), _f64p()
    rc = lib().oracle_mmio_load(path.encode(), C.byref(m), C.byref(n),
                                C.byref(nnz), C.byref(sym), C.byref(rp),
                                C.byref(ci), C.byref(vv))
    if rc != 0:
        return rc, None
    out = dict(m=m.value, n=n.value, nnz=nnz.value, sym=sym.value,
               rowptr=_arr(rp, m.value + 1, np.int32),
               colidx=_arr(ci, nnz.value, np.int32),
               val=_arr(vv, nnz.value, np.float64))
    for p in (rp, ci, vv):
        lib().oracle_free_ptr(p)
    return 0, out


def spmm_f64(m, n, rowptr, colidx, val, B, want_abs=True):
    """FP64 C = A*B (row-major) and |A|*|B| for error normalisation."""
    K = B.shape[1]
    B = np.ascontiguousarray(B, dtype=np.float64)
    Cm = np.zeros((m, K), np.float64)
    A = np.zeros((m, K), np.float64) if want_abs else None
    lib().oracle_spmm_csr_f64(
        m, n, _ip(np.ascontiguousarray(rowptr, np.int32)),
        _ip(np.ascontiguousarray(colidx, np.int32)),
        _dp(np.ascontiguousarray(val, np.float64)), K, _dp(B), _dp(Cm),
        _dp(A) if want_abs else None)
    return Cm, A


def spmm_t_f64(m, n, rowptr, colidx, val, dC):
    K = dC.shape[1]
    dC = np.ascontiguousarray(dC, dtype=np.float64)
    dB = np.zeros((n, K), np.float64)
    lib().oracle_spmm_t_csr_f64(
        m, n, _ip(np.ascontiguousarray(rowptr, np.int32)),
        _ip(np.ascontiguousarray(colidx, np.int32)),
        _dp(np.ascontiguousarray(val, np.float64)), K, _dp(dC), _dp(dB))
    return dB


def sddmm_f64(m, n, rowptr, colidx, dC, B):
    K = dC.shape[1]
    out = np.zeros(int(rowptr[-1]), np.float64)
    lib().oracle_sddmm_csr_f64(
        m, n, _ip(np.ascontiguousarray(rowptr, np.int32)),
        _ip(np.ascontiguousarray(colidx, np.int32)), K,
        _dp(np.ascontiguousarray(dC, np.float64)),
        _dp(np.ascontiguousarray(B, np.float64)), _dp(out))
    return out


def block_stats(m, n, rowptr, colidx):
    """(blocks, empty 32x32 blocks, sparsity nnz/(m*n)) of a CSR pattern
    (test/MTX_info/main.cu:143-176, :484)."""
    blocks, empty = C.c_longlong(), C.c_longlong()
    lib().oracle_block_stats(m, n, _ip(np.ascontiguousarray(rowptr, np.int32)),
                             _ip(np.ascontiguousarray(colidx, np.int32)),
                             C.byref(blocks), C.byref(empty))
    nnz = int(rowptr[-1]) if len(rowptr) else 0
    return blocks.value, empty.value, (nnz / (m * n) if m and n else 0.0)


def block_density(m, n, rowptr, colidx, thr=512):
    """(32x32 blocks with >= thr entries, their entries, largest block count)."""
    d, dz, mx = C.c_longlong(), C.c_longlong(), C.c_int()
    lib().oracle_block_density(m, n, _ip(np.ascontiguousarray(rowptr, np.int32)),
                               _ip(np.ascontiguousarray(colidx, np.int32)), thr,
                               C.byref(d), C.byref(dz), C.byref(mx))
    return d.value, dz.value, mx.value


# ------------------------------------------------------------------ _ref --

class _RefView(C.Structure):
    pass


def _ref_view_fields(vp):
    return [(n, C.c_int) for n in ("tilenum", "regular_nnz", "reside_col",
                                   "reside_nnz", "segsum", "numblk")] + [
        ("sortval", vp), ("sortrowidx", _i32p), ("tile_ptr", _i32p),
        ("tile_len", _i32p), ("tile_colidx", _i32p), ("tile_tall", _i32p),
        ("tile_format", _i32p), ("segmentoffset", _i32p),
        ("sortrowindex", _i32p), ("reside_cscptr", _i32p),
        ("reside_cscrowidx", _i32p), ("reside_val", vp), ("colorblk", _u32p),
        ("blkstart", _i32p), ("blkstop", _i32p), ("blkptr", _i32p),
        ("val_position", _i32p), ("val_offset", _i32p)]


class Ref:
    """The reference's CPU headers compiled in place (oracle/_ref)."""

    def __init__(self, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        tag = "f32" if self.dtype == np.float32 else "f64"
        path = os.path.join(HERE, "_ref", "libswiftref_%s.so" % tag)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = C.CDLL(path)
        self.vp = _f32p if tag == "f32" else _f64p
        self.ctype = C.c_float if tag == "f32" else C.c_double

        class View(C.Structure):
            _fields_ = _ref_view_fields(self.vp)
        self.View = View
        self.L.ref_format.restype = C.c_void_p
        self.L.ref_swift_cpu.restype = C.c_double
        self.L.ref_export.argtypes = [C.c_void_p, C.POINTER(View)]
        self.L.ref_row_idx_sort.argtypes = [C.c_void_p, self.vp, C.c_int,
                                            C.c_int, C.c_int]
        self.L.ref_free.argtypes = [C.c_void_p]
        self.L.ref_swift_cpu.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int,
                                         C.c_int, self.vp, self.vp]

    @staticmethod
    def available(dtype=np.float64):
        tag = "f32" if np.dtype(dtype) == np.float32 else "f64"
        return os.path.exists(os.path.join(HERE, "_ref", "libswiftref_%s.so" % tag))

    def _v(self, a):
        return np.ascontiguousarray(a, dtype=self.dtype)

    def mmio_load(self, path):
        m, n, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        rp, ci, vv = _i32p(), _i32p(), _f64p()
        rc = self.L.ref_mmio_load(path.encode(), C.byref(m), C.byref(n),
                                  C.byref(nnz), C.byref(sym), C.byref(rp),
                                  C.byref(ci), C.byref(vv))
        if rc != 0:
            return rc, None
        out = dict(m=m.value, n=n.value, nnz=nnz.value, sym=sym.value,
                   rowptr=_arr(rp, m.value + 1, np.int32),
                   colidx=_arr(ci, nnz.value, np.int32),
                   val=_arr(vv, nnz.value, np.float64))
        for p in (rp, ci, vv):
            self.L.ref_free_ptr(p)
        return 0, out

    def csr2csc(self, m, n, rowptr, colidx, val):
        nnz = int(rowptr[-1])
        rp = np.ascontiguousarray(rowptr, np.int32)
        ci = np.ascontiguousarray(colidx, np.int32)
        vv = self._v(val)
        row, ptr, npc, cv = _i32p(), _i32p(), _i32p(), self.vp()
        self.L.ref_csr2csc(nnz, m, n, _ip(ci), _ip(rp), vv.ctypes.data_as(self.vp),
                           C.byref(row), C.byref(ptr), C.byref(npc), C.byref(cv))
        out = (_arr(ptr, n + 1, np.int32), _arr(row, nnz, np.int32),
               _arr(cv, nnz, self.dtype), _arr(npc, n, np.int32))
        for p in (row, ptr, npc, cv):
            self.L.ref_free_ptr(p)
        return out

    def col_sort(self, n, npc, cscptr, cscrow, cscval, B):
        """Reference col_sort (std::sort, unstable).  B is n x K row-major."""
        nnz = len(cscrow)
        K = B.shape[1]
        sr = np.zeros(nnz, np.int32)
        sv = np.zeros(nnz, self.dtype)
        sn = np.zeros(n + 1, np.int32)
        Bc = self._v(B)
        sB = np.zeros_like(Bc)
        self.L.ref_col_sort(n, K, _ip(np.ascontiguousarray(npc, np.int32)),
                            _ip(np.ascontiguousarray(cscptr, np.int32)),
                            _ip(np.ascontiguousarray(cscrow, np.int32)),
                            self._v(cscval).ctypes.data_as(self.vp), _ip(sr),
                            sv.ctypes.data_as(self.vp), _ip(sn),
                            Bc.ctypes.data_as(self.vp), sB.ctypes.data_as(self.vp))
        return sr, sv, sn[:n], sB

    def format(self, m, n, sortrowidx_tmp, sortval_tmp, sortnnz):
        """formattransation + block_catgory + balance; returns (handle, dict)."""
        sr = np.ascontiguousarray(sortrowidx_tmp, np.int32).copy()
        sv = self._v(sortval_tmp).copy()
        sn = np.ascontiguousarray(sortnnz, np.int32).copy()
        h = self.L.ref_format(_ip(sr), sv.ctypes.data_as(self.vp), _ip(sn),
                              len(sr), m, n)
        return h, self.export(h)

    def export(self, h):
        v = self.View()
        self.L.ref_export(h, C.byref(v))
        sizes = {"sortval": v.regular_nnz, "sortrowidx": v.regular_nnz,
                 "tile_ptr": v.tilenum + 1, "tile_len": v.tilenum,
                 "tile_colidx": v.tilenum, "tile_tall": v.tilenum,
                 "tile_format": v.tilenum, "segmentoffset": v.segsum + 1,
                 "sortrowindex": v.regular_nnz,
                 "reside_cscptr": v.reside_col + 1,
                 "reside_cscrowidx": v.reside_nnz, "reside_val": v.reside_nnz,
                 "colorblk": v.numblk, "blkstart": v.numblk,
                 "blkstop": v.numblk, "blkptr": v.numblk,
                 "val_position": v.regular_nnz, "val_offset": v.regular_nnz}
        out = {k: getattr(v, k) for k in ("tilenum", "regular_nnz", "reside_col",
                                          "reside_nnz", "segsum", "numblk")}
        for k, n in sizes.items():
            dt = (self.dtype if k in ("sortval", "reside_val") else
                  np.uint32 if k == "colorblk" else np.int32)
            out[k] = _arr(getattr(v, k), n, dt)
        return out

    def row_idx_sort(self, h, m, n, K=32):
        dummy = np.zeros(max(1, n * K), self.dtype)
        self.L.ref_row_idx_sort(h, dummy.ctypes.data_as(self.vp), m, n, K)
        return self.export(h)

    def swift_cpu(self, h, m, n, nnz, sortB):
        K = sortB.shape[1]
        B = self._v(sortB)
        out = np.zeros((m, K), self.dtype)
        ms = self.L.ref_swift_cpu(h, nnz, m, n, K, B.ctypes.data_as(self.vp),
                                  out.ctypes.data_as(self.vp))
        return out, ms

    def free(self, h):
        self.L.ref_free(h)
