"""TEST INFRASTRUCTURE -- ctypes binding of the CPU oracle (liboracle_cnv.so).

The oracle restates CubismNova's lab assembly, BCs and stencil operators on
the CPU (oracle/cnv_oracle.cpp).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it, and only to check results or as the
timed CPU baseline; the product never imports it.

BCs are plain tuples (kind, dir, side, value) with kind 0 Base, 1 Absorbing,
2 Dirichlet, 3 Symmetry.  Slabs are 1-D numpy arrays in the grid's block
layout (blocks x-fastest, `block_stride` elements apart).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_cnv.so")

BASE, ABSORBING, DIRICHLET, SYMMETRY = 0, 1, 2, 3


class _bc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("dir", ctypes.c_int), ("side", ctypes.c_int),
                ("reserved", ctypes.c_int), ("value", ctypes.c_double)]


class _field(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dtype", ctypes.c_int), ("nblocks", ctypes.c_int * 3),
                ("cells", ctypes.c_int * 3), ("plus", ctypes.c_int), ("block_stride", ctypes.c_longlong),
                ("bc_offsets", ctypes.c_void_p), ("bcs", ctypes.c_void_p)]


class _stencil(ctypes.Structure):
    _fields_ = [("begin", ctypes.c_int * 3), ("end", ctypes.c_int * 3), ("tensorial", ctypes.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        if os.path.exists("/proc/cpuinfo") and " fma" not in open("/proc/cpuinfo").read():
            raise RuntimeError("the oracle is built for x86-64 hosts with FMA3 instructions")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, f64, ll = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
        L.orc_lab_load.argtypes = [vp, vp, vp, i32, vp, i32, f64, vp]
        L.orc_lab_load.restype = None
        L.orc_diffusion.argtypes = [vp, vp, ll, vp, vp, f64, i32, i32]
        L.orc_diffusion.restype = f64
        L.orc_weno5.argtypes = [vp, vp, ll, i32, i32, vp, i32]
        L.orc_weno5.restype = i32
        L.orc_weno5_point.argtypes = [vp, vp, vp]
        L.orc_weno5_point.restype = None
        L.orc_weno5_point_textbook.argtypes = [vp, vp, vp]
        L.orc_weno5_point_textbook.restype = None
        L.orc_jacobi.argtypes = [vp, vp, ll, vp, ll, vp, vp, vp, f64, i32, i32]
        L.orc_jacobi.restype = f64
        L.orc_max_threads.restype = i32
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _s3(v):
    return [int(v)] * 3 if np.isscalar(v) else [int(x) for x in v]


class Field:
    """One scalar component in block layout plus per-block BC lists.

    bcs: dict block-flat-index -> list of (kind, dir, side, value).
    """

    def __init__(self, slab, nblocks, cells, block_stride=None, bcs=None, plus=(0, 0, 0)):
        """plus[d] = 1: blocks at the upper end of dim d hold cells[d] + 1
        elements (node fields in every dim, face fields in their direction)."""
        self.slab = np.ascontiguousarray(slab)
        assert self.slab.dtype in (np.float32, np.float64)
        self.nblocks, self.cells = _s3(nblocks), _s3(cells)
        self.plus = [int(bool(x)) for x in plus]
        if block_stride is None:
            block_stride = int(np.prod([self.cells[d] + self.plus[d] for d in range(3)]))
        self.block_stride = int(block_stride)
        nbt = int(np.prod(self.nblocks))
        offs = np.zeros(nbt + 1, dtype=np.int32)
        flat = []
        bcs = bcs or {}
        for b in range(nbt):
            lst = bcs.get(b, [])
            offs[b + 1] = offs[b] + len(lst)
            flat.extend(lst)
        self._offs = offs
        self._bcs = (_bc * max(len(flat), 1))()
        for i, (k, d, s, v) in enumerate(flat):
            self._bcs[i] = _bc(k, d, s, 0, float(v))
        self._has = len(flat) > 0
        f = _field()
        f.data = self.slab.ctypes.data
        f.dtype = 1 if self.slab.dtype == np.float64 else 0
        for d in range(3):
            f.nblocks[d], f.cells[d] = self.nblocks[d], self.cells[d]
        f.block_stride = self.block_stride
        f.plus = self.plus[0] | (self.plus[1] << 1) | (self.plus[2] << 2)
        f.bc_offsets = self._offs.ctypes.data if self._has else None
        f.bcs = ctypes.cast(self._bcs, ctypes.c_void_p).value if self._has else None
        self._f = f

    def ref(self):
        return ctypes.byref(self._f)


def stencil(begin, end, tensorial=False):
    s = _stencil()
    b, e = _s3(begin), _s3(end)
    for d in range(3):
        s.begin[d], s.end[d] = b[d], e[d]
    s.tensorial = 1 if tensorial else 0
    return s


def lab_extent(cells, begin, end):
    c, b, e = _s3(cells), _s3(begin), _s3(end)
    return [c[d] + e[d] - 1 - b[d] for d in range(3)]


def block_extent(field, bidx):
    """Element extent of block bidx (one more at the upper end of plus dims)."""
    return [field.cells[d] + (1 if field.plus[d] and bidx[d] == field.nblocks[d] - 1 else 0) for d in range(3)]


def lab_load(field, bidx, begin, end, tensorial=False, apply_bc=True, extern=None, poison=np.nan):
    """Lab of block `bidx` over the active lab range, as a (Z, Y, X) array."""
    L = lab_extent(block_extent(field, bidx), begin, end)
    out = np.empty(L[2] * L[1] * L[0], dtype=field.slab.dtype)
    st = stencil(begin, end, tensorial)
    bi = (ctypes.c_int * 3)(*bidx)
    if extern is None:
        lib().orc_lab_load(field.ref(), bi, ctypes.byref(st), 1 if apply_bc else 0, None, -1, float(poison), _p(out))
    else:
        arr = (_bc * max(len(extern), 1))()
        for i, (k, d, s, v) in enumerate(extern):
            arr[i] = _bc(k, d, s, 0, float(v))
        lib().orc_lab_load(field.ref(), bi, ctypes.byref(st), 1 if apply_bc else 0, arr, len(extern), float(poison),
                           _p(out))
    return out.reshape(L[2], L[1], L[0])


def laplacian(field, h, nthreads=1):
    out = np.zeros_like(field.slab)
    lib().orc_diffusion(field.ref(), _p(out), field.block_stride, None,
                        (ctypes.c_double * 3)(*[float(x) for x in h]), 0.0, 0, int(nthreads))
    return out


def diffusion_step(field, h, nu_dt, nthreads=1, want_lap=False):
    """(u_new, sum of u_new^2[, laplacian]) for one forward-Euler step."""
    out = np.zeros_like(field.slab)
    lap = np.zeros_like(field.slab) if want_lap else None
    s = lib().orc_diffusion(field.ref(), _p(out), field.block_stride, _p(lap) if want_lap else None,
                            (ctypes.c_double * 3)(*[float(x) for x in h]), float(nu_dt), 1, int(nthreads))
    return (out, s, lap) if want_lap else (out, s)


def weno5(field, prec=1, nthreads=1, blocks=None):
    """Six state slabs (u- x, u+ x, u- y, u+ y, u- z, u+ z) of one component
    in float64 (prec=1, the only precision the oracle evaluates).  With
    `blocks` (flat block indices) only those blocks are computed and returned
    consecutively, one block_stride each."""
    if prec != 1:
        raise ValueError("the WENO5 oracle evaluates in float64 only (prec=1)")
    n = field.slab.shape[0] if blocks is None else len(blocks) * field.block_stride
    outs = [np.zeros(n, dtype=np.float64) for _ in range(6)]
    ptrs = (ctypes.c_void_p * 6)(*[o.ctypes.data for o in outs])
    if blocks is None:
        rc = lib().orc_weno5(field.ref(), ptrs, field.block_stride, 1, int(nthreads), None, 0)
    else:
        sel = np.ascontiguousarray(blocks, dtype=np.int64)
        rc = lib().orc_weno5(field.ref(), ptrs, field.block_stride, 1, int(nthreads), _p(sel), len(sel))
    assert rc == 0
    return outs


def weno5_point(v, textbook=False):
    """(u- at i+1/2, u+ at i-1/2) of one 5-point stencil in FP64: the
    difference form the kernels evaluate, or the textbook statement."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    um, up = ctypes.c_double(), ctypes.c_double()
    fn = lib().orc_weno5_point_textbook if textbook else lib().orc_weno5_point
    fn(_p(v), ctypes.byref(um), ctypes.byref(up))
    return um.value, up.value


def jacobi(field, f_slab, czm, czp, h, omega, mode, nthreads=1):
    """mode 0: one weighted-Jacobi sweep (returns new slab); mode 1: defect
    f - A u (returns (defect slab, sum of squares))."""
    out = np.zeros_like(field.slab)
    f_slab = np.ascontiguousarray(f_slab, dtype=field.slab.dtype)
    czm = np.ascontiguousarray(czm, dtype=np.float64)
    czp = np.ascontiguousarray(czp, dtype=np.float64)
    s = lib().orc_jacobi(field.ref(), _p(f_slab), field.block_stride, _p(out), field.block_stride, _p(czm), _p(czp),
                         (ctypes.c_double * 3)(*[float(x) for x in h]), float(omega), int(mode), int(nthreads))
    return out if mode == 0 else (out, s)


def max_threads():
    return lib().orc_max_threads()
