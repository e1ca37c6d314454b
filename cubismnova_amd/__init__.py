"""cubismnova_amd -- MI355X-native block-structured stencil path.

Python mirror of CubismNova's host API for the stencil path (the names follow
include/Cubism/{Core/Stencil.h, BC/*.h, Grid/Cartesian.h, Block/FieldLab.h});
every compute call goes through the C ABI of libcubismnova_amd.so
(include/cubismnova_amd.h).  Storage lives on the GPU in the reference's
block layout; numpy arrays are only the host side of uploads and downloads.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import CnvError, call, lib
from .layout import blocks_to_global, global_to_blocks

__all__ = [
    "Stencil", "Base", "Absorbing", "Dirichlet", "Symmetry", "BoundaryInfo", "Cartesian", "FieldLab",
    "DeviceBuffer", "Stream", "Graph", "laplacian", "diffusion_step", "diffusion_run", "grid_sums", "CnvError", "lib",
    "global_to_blocks", "blocks_to_global",
]

# ------------------------------------------------------------------ stencil


class Stencil:
    """Core::Stencil<3> (include/Cubism/Core/Stencil.h:20-108)."""

    def __init__(self, begin, end, tensorial=False):
        b = [int(begin)] * 3 if np.isscalar(begin) else [int(v) for v in begin]
        e = [int(end)] * 3 if np.isscalar(end) else [int(v) for v in end]
        # Vector comparisons are all-component (Core/Vector.h:392-420)
        if all(v > 0 for v in b):
            raise RuntimeError("Stencil: begin_ must be <= 0")
        if all(v < 1 for v in e):
            raise RuntimeError("Stencil: end_ must be > 0")
        self._b, self._e, self._t = b, e, bool(tensorial)

    def getBegin(self):
        return list(self._b)

    def getEnd(self):
        return list(self._e)

    def isTensorial(self):
        return self._t

    def _c(self):
        s = _abi.cnv_stencil()
        for d in range(3):
            s.begin[d] = self._b[d]
            s.end[d] = self._e[d]
        s.tensorial = 1 if self._t else 0
        return s


# ------------------------------------------------------------------ BCs


class BoundaryInfo:
    """BC::BoundaryInfo (include/Cubism/BC/Base.h:21-28)."""

    def __init__(self, dir, side, is_periodic):
        self.dir, self.side, self.is_periodic = dir, side, is_periodic


class Base:
    """BC::Base (include/Cubism/BC/Base.h:39-79): periodic, no-op apply."""
    _kind = _abi.CNV_BC_BASE
    _name = "Base"
    _periodic = True

    def __init__(self, dir, side):
        self._info = BoundaryInfo(int(dir), int(side), self._periodic)
        self._value = 0.0

    def getBoundaryInfo(self):
        return self._info

    def name(self):
        return self._name

    def _c(self):
        return _abi.cnv_bc(self._kind, self._info.dir, self._info.side, 0, float(self._value))


class Absorbing(Base):
    """BC::Absorbing, zeroth-order absorbing (include/Cubism/BC/Absorbing.h:24-105)."""
    _kind = _abi.CNV_BC_ABSORBING
    _name = "Zeroth-Order Absorbing"
    _periodic = False


class Dirichlet(Base):
    """BC::Dirichlet, constant ghost value (include/Cubism/BC/Dirichlet.h:24-114)."""
    _kind = _abi.CNV_BC_DIRICHLET
    _name = "Dirichlet"
    _periodic = False

    def __init__(self, dir, side, value=0.0):
        super().__init__(dir, side)
        self._value = value

    def getValue(self):
        return self._value

    def setValue(self, v):
        self._value = v


class Symmetry(Base):
    """BC::Symmetry, ghost = sign * mirror (include/Cubism/BC/Symmetry.h:24-115)."""
    _kind = _abi.CNV_BC_SYMMETRY
    _name = "Symmetry"
    _periodic = False

    def __init__(self, dir, side, sign=1.0):
        super().__init__(dir, side)
        self._value = sign


def _bc_array(bcs):
    n = len(bcs)
    arr = (_abi.cnv_bc * max(n, 1))()
    for i, b in enumerate(bcs):
        arr[i] = b._c()
    return arr, n


# ------------------------------------------------------------------ device memory


class Stream:
    """A HIP stream owned by the library (hipStreamNonBlocking)."""

    def __init__(self):
        h = ctypes.c_void_p()
        call("cnv_stream_create", ctypes.byref(h))
        self.handle = h

    def synchronize(self):
        call("cnv_stream_synchronize", self.handle)

    def __del__(self):
        try:
            if self.handle:
                lib().cnv_stream_destroy(self.handle)
        except Exception:
            pass


class Graph:
    """Executable HIP graph recorded from the calls issued on `stream`
    between capture() and end()."""

    def __init__(self, stream):
        self.stream = stream
        self.handle = None

    def capture(self):
        call("cnv_graph_begin", self.stream.handle)
        return self

    def end(self):
        h = ctypes.c_void_p()
        call("cnv_graph_end", self.stream.handle, ctypes.byref(h))
        self.handle = h
        return self

    def launch(self, stream=None):
        call("cnv_graph_launch", self.handle, (stream or self.stream).handle)

    def __del__(self):
        try:
            if self.handle:
                lib().cnv_graph_destroy(self.handle)
        except Exception:
            pass


def _st(stream):
    return stream.handle if stream is not None else None


class DeviceBuffer:
    """Raw device allocation for lab dumps, norms and other outputs."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        call("cnv_malloc", ctypes.byref(p), ctypes.c_size_t(self.nbytes))
        self.ptr = p

    def upload(self, arr, stream=None):
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than buffer")
        call("cnv_memcpy_h2d", self.ptr, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes),
             _st(stream))

    def download(self, dtype, count=None, stream=None):
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dt)
        call("cnv_memcpy_d2h", out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes),
             _st(stream))
        return out

    def fill(self, dtype, value, count=None, stream=None):
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else int(count)
        call("cnv_fill", self.ptr, _abi.CNV_F64 if dt == np.float64 else _abi.CNV_F32, float(value),
             ctypes.c_size_t(n), _st(stream))

    def at(self, byte_offset):
        return ctypes.c_void_p(self.ptr.value + int(byte_offset))

    def free(self):
        if self.ptr:
            lib().cnv_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ------------------------------------------------------------------ grid


class _Mesh:
    """Mesh::StructuredUniform queries used by the path (StructuredUniform.h:56-64,361-365)."""

    def __init__(self, begin, end, cells):
        self.begin = np.asarray(begin, dtype=np.float64)
        self.end = np.asarray(end, dtype=np.float64)
        self.cells = np.asarray(cells)
        self.h = (self.end - self.begin) / self.cells

    def getCellSize(self, i=0):
        return self.h.copy()

    def getCoordsCell(self, p):
        return self.begin + (np.asarray(p, dtype=np.float64) + 0.5) * self.h


ENTITIES = {"cell": _abi.CNV_CELL, "node": _abi.CNV_NODE, "face": _abi.CNV_FACE}


class Cartesian:
    """Grid::Cartesian for scalar components on the GPU.

    Mirrors include/Cubism/Grid/Cartesian.h:44-550: one allocation, component
    slabs of block_bytes-aligned blocks, blocks numbered x-fastest.  entity
    "node" / "face" (Cubism::EntityType) gives (N+1)^3-element blocks whose
    upper-end blocks hold N+1 along their plus dimensions; a face grid has
    3 * ncomp slabs, slab d * ncomp + c (BlockFieldAssembler.h:111-185), and
    every `comp` argument is a slab index.  The stencil operators take cell
    grids only.
    """

    def __init__(self, nblocks, block_cells, begin=0.0, end=1.0, dtype=np.float64, ncomp=1, entity="cell",
                 device_data=None, device_bytes=0):
        """device_data (an address, with device_bytes): build the grid over
        that device memory instead of a new allocation (cnv_grid_wrap); the
        memory must stay alive as long as the grid."""
        nb = [int(nblocks)] * 3 if np.isscalar(nblocks) else [int(v) for v in nblocks]
        bc = [int(block_cells)] * 3 if np.isscalar(block_cells) else [int(v) for v in block_cells]
        b = [float(begin)] * 3 if np.isscalar(begin) else [float(v) for v in begin]
        e = [float(end)] * 3 if np.isscalar(end) else [float(v) for v in end]
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float32, np.float64):
            raise ValueError("dtype must be float32 or float64")
        d = _abi.cnv_grid_desc()
        for k in range(3):
            d.nblocks[k], d.block_cells[k], d.begin[k], d.end[k] = nb[k], bc[k], b[k], e[k]
        d.ncomp = int(ncomp)
        d.dtype = _abi.CNV_F64 if self.dtype == np.float64 else _abi.CNV_F32
        if entity not in ENTITIES:
            raise ValueError("entity must be 'cell', 'node' or 'face'")
        d.entity = ENTITIES[entity]
        self.entity = entity
        self.nslabs = 3 * int(ncomp) if entity == "face" else int(ncomp)
        h = ctypes.c_void_p()
        if device_data is None:
            call("cnv_grid_create", ctypes.byref(d), ctypes.byref(h))
        else:
            addr = device_data.value if isinstance(device_data, ctypes.c_void_p) else int(device_data)
            call("cnv_grid_wrap", ctypes.byref(d), ctypes.c_void_p(addr), ctypes.c_size_t(int(device_bytes)),
                 ctypes.byref(h))
        self.handle = h
        self._owner = None
        bb, cb = ctypes.c_size_t(), ctypes.c_size_t()
        call("cnv_grid_info", h, None, ctypes.byref(bb), ctypes.byref(cb))
        self.block_bytes, self.component_bytes = bb.value, cb.value
        self.block_stride = self.block_bytes // self.dtype.itemsize
        self.nblocks, self.block_cells, self.ncomp = nb, bc, int(ncomp)
        self.nblocks_total = nb[0] * nb[1] * nb[2]
        self._mesh = _Mesh(b, e, [nb[k] * bc[k] for k in range(3)])
        self._bcs = {}

    @classmethod
    def from_tensor(cls, t, nblocks, block_cells, begin=0.0, end=1.0, ncomp=1, entity="cell"):
        """Grid over the storage of a contiguous torch tensor on the GPU
        (float32 or float64), no copy.  The tensor holds the reference
        layout (component slabs of 32-byte rounded blocks, x-fastest) and is
        kept alive by the grid.  torch brings its own HIP runtime, which sees
        the device only if it initialises before this library's first GPU
        call in the process: touch the GPU with torch first."""
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError("need a contiguous tensor on the GPU")
        import torch
        dt = {torch.float32: np.float32, torch.float64: np.float64}.get(t.dtype)
        if dt is None:
            raise ValueError("tensor dtype must be float32 or float64")
        g = cls(nblocks, block_cells, begin, end, dtype=dt, ncomp=ncomp, entity=entity, device_data=t.data_ptr(),
                device_bytes=t.numel() * t.element_size())
        g._owner = t
        return g

    # -- reference queries
    def size(self):
        return self.nblocks_total

    def getSize(self):
        return list(self.nblocks)

    def getBlockCells(self):
        return list(self.block_cells)

    def getMesh(self):
        return self._mesh

    def block_flat(self, bi):
        return bi[0] + self.nblocks[0] * (bi[1] + self.nblocks[1] * bi[2])

    def plus(self, comp=0):
        """Dimensions in which the upper-end blocks of slab `comp` hold one
        more element (BlockFieldAssembler.h:123-131)."""
        if self.entity == "node":
            return (1, 1, 1)
        if self.entity == "face":
            d = comp // self.ncomp
            return tuple(1 if e == d else 0 for e in range(3))
        return (0, 0, 0)

    def block_extent(self, bi, comp=0):
        """Elements of block bi of slab `comp` per dimension."""
        p = self.plus(comp)
        return [self.block_cells[d] + (1 if p[d] and bi[d] == self.nblocks[d] - 1 else 0) for d in range(3)]

    def lab_geometry(self, stencil, comp=0, block_index=None):
        """(pitch, active): the slab's largest lab extent and block_index's own."""
        pitch, active = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        bi = None if block_index is None else (ctypes.c_int * 3)(*block_index)
        sc = stencil._c()
        call("cnv_lab_geometry", self.handle, comp, bi, ctypes.byref(sc), pitch, active)
        return list(pitch), list(active)

    # -- data movement (reference byte layout)
    def upload(self, comp, slab, stream=None):
        a = np.ascontiguousarray(slab, dtype=self.dtype)
        if a.nbytes != self.component_bytes:
            raise ValueError("slab must hold component_bytes = %d bytes" % self.component_bytes)
        call("cnv_grid_upload", self.handle, comp, a.ctypes.data_as(ctypes.c_void_p),
             ctypes.c_size_t(a.nbytes), _st(stream))

    def download(self, comp, stream=None):
        out = np.empty(self.component_bytes // self.dtype.itemsize, dtype=self.dtype)
        call("cnv_grid_download", self.handle, comp, out.ctypes.data_as(ctypes.c_void_p),
             ctypes.c_size_t(out.nbytes), _st(stream))
        return out

    def download_blocks(self, comp, blocks):
        """Host copy of selected blocks (flat indices) of one component, one
        block_stride each, in the order given."""
        base = self.data_ptr(comp).value
        out = np.empty((len(blocks), self.block_stride), dtype=self.dtype)
        for j, b in enumerate(blocks):
            call("cnv_memcpy_d2h", out[j].ctypes.data_as(ctypes.c_void_p),
                 ctypes.c_void_p(base + int(b) * self.block_bytes), ctypes.c_size_t(self.block_bytes), None)
        return out.reshape(-1)

    def upload_global(self, comp, u, stream=None):
        """Upload a global (Z, Y, X) array into the block layout."""
        self.upload(comp, global_to_blocks(np.asarray(u, dtype=self.dtype), self.nblocks, self.block_cells,
                                           self.block_stride), stream)

    def download_global(self, comp, stream=None):
        return blocks_to_global(self.download(comp, stream), self.nblocks, self.block_cells, self.block_stride)

    def data_ptr(self, comp=0):
        p = ctypes.c_void_p()
        call("cnv_grid_data", self.handle, comp, ctypes.byref(p))
        return p

    # -- BCs (Block::Field::getBC, Block/Field.h:437-442)
    def setBC(self, block_index, bcs, comp=0):
        arr, n = _bc_array(bcs)
        bi = (ctypes.c_int * 3)(*block_index)
        call("cnv_grid_set_block_bcs", self.handle, comp, bi, arr, n)
        self._bcs[(comp, tuple(block_index))] = list(bcs)

    def getBC(self, block_index, comp=0):
        return list(self._bcs.get((comp, tuple(block_index)), []))

    def addFaceBC(self, bc, comp=-1):
        c = bc._c()
        call("cnv_grid_add_face_bc", self.handle, comp, ctypes.byref(c))
        d, s = bc.getBoundaryInfo().dir, bc.getBoundaryInfo().side
        comps = range(self.nslabs) if comp == -1 else [comp]
        for cc in comps:
            for bz in range(self.nblocks[2]):
                for by in range(self.nblocks[1]):
                    for bx in range(self.nblocks[0]):
                        bi = (bx, by, bz)
                        if bi[d] == (0 if s == 0 else self.nblocks[d] - 1):
                            self._bcs.setdefault((cc, bi), []).append(bc)

    def clearBC(self, comp=-1):
        call("cnv_grid_clear_bcs", self.handle, comp)
        self._bcs = {k: v for k, v in self._bcs.items() if comp != -1 and k[0] != comp}

    def prepare(self):
        """Build device BC tables (and z metric) now, e.g. before graph capture."""
        call("cnv_grid_prepare", self.handle)

    def loadLab(self, block_index, lab, comp=0):
        """Grid::Cartesian::loadLab (Grid/Cartesian.h:411-427)."""
        lab.loadData(block_index, self, comp=comp)

    def lab_all(self, stencil, comp=0, out=None, stream=None):
        """Labs of every block (device buffer, blocks consecutive, one slot of
        the largest block's lab extent each, a smaller block's lab in the low
        corner); returns (buffer, slot extent)."""
        sc = stencil._c()
        ext, _ = self.lab_geometry(stencil, comp)
        n = ext[0] * ext[1] * ext[2] * self.nblocks_total
        if out is None:
            out = DeviceBuffer(n * self.dtype.itemsize)
        call("cnv_lab_load_all", self.handle, comp, ctypes.byref(sc), out.ptr, _st(stream))
        return out, list(ext)

    def __del__(self):
        try:
            if self.handle:
                lib().cnv_grid_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


# ------------------------------------------------------------------ lab


class FieldLab:
    """Block::FieldLab (include/Cubism/Block/FieldLab.h:30-492), assembled on the GPU.

    loadData() runs the device lab kernel and keeps a host copy of the active
    lab range so that lab[(ix, iy, iz)] (ghost indices allowed) reads it.
    """

    def __init__(self):
        self._stencil = None
        self._alloc = False
        self._data = None
        self._field = None

    def allocate(self, stencil, max_extent):
        self._stencil = stencil
        self._max = [int(v) for v in max_extent]
        self._alloc = True

    def isAllocated(self):
        return self._alloc

    def loadData(self, block_index, grid, comp=0, apply_bc=True, extern_bc=None, poison=None):
        if not self._alloc:
            raise RuntimeError("FieldLab: can not load lab data when not allocated first")
        if any(grid.block_cells[d] > self._max[d] + 2 for d in range(3)):
            raise RuntimeError("FieldLab: block larger than the allocated maximum range")
        s = self._stencil
        sc = s._c()
        _, ext = grid.lab_geometry(s, comp, block_index)
        n = ext[0] * ext[1] * ext[2]
        buf = DeviceBuffer(n * grid.dtype.itemsize)
        buf.fill(grid.dtype, np.nan if poison is None else poison)
        bi = (ctypes.c_int * 3)(*block_index)
        if extern_bc is None:
            call("cnv_lab_load", grid.handle, comp, bi, ctypes.byref(sc), 1 if apply_bc else 0, None, -1, buf.ptr,
                 None)
        else:
            arr, nb = _bc_array(extern_bc)
            call("cnv_lab_load", grid.handle, comp, bi, ctypes.byref(sc), 1 if apply_bc else 0, arr, nb, buf.ptr,
                 None)
        self._data = buf.download(grid.dtype).reshape(ext[2], ext[1], ext[0])
        self._ext = list(ext)
        self._N = grid.block_extent(block_index, comp)
        self._field = (grid, tuple(block_index), comp)
        return self._data

    def __getitem__(self, p):
        b = self._stencil.getBegin()
        return self._data[p[2] - b[2], p[1] - b[1], p[0] - b[0]]

    def __call__(self, ix, iy=0, iz=0):
        return self[(ix, iy, iz)]

    def array(self):
        """Active lab range as a (Z, Y, X) array starting at the stencil begin."""
        return self._data

    def getActiveStencil(self):
        return self._stencil

    def getActiveRange(self):
        return [0, 0, 0], list(self._N)

    def getActiveLabRange(self):
        b = self._stencil.getBegin()
        e = self._stencil.getEnd()
        return list(b), [self._N[d] + e[d] - 1 for d in range(3)]

    def getActiveField(self):
        if self._field is None:
            raise RuntimeError("FieldLab: no field loaded")
        return self._field


# ------------------------------------------------------------------ operators


def laplacian(u, lap, comp_in=0, comp_out=0, stream=None):
    """7-point Laplacian of u into lap (Laplacian.cpp:110-131 semantics)."""
    call("cnv_laplacian", u.handle, comp_in, lap.handle, comp_out, _st(stream))


def diffusion_step(u, unew, nu_dt, lap=None, sumsq=None, stream=None):
    """unew = u + nu_dt * lap(u) for every component; optional FP64 sum of unew^2
    into the device double at `sumsq` (a DeviceBuffer or ctypes.c_void_p)."""
    p = sumsq.ptr if isinstance(sumsq, DeviceBuffer) else sumsq
    call("cnv_diffusion_step", u.handle, unew.handle, ctypes.c_double(nu_dt), lap.handle if lap is not None else None,
         p, _st(stream))


def diffusion_run(u, tmp, nsteps, nu_dt, norms=None, stream=None):
    """nsteps Euler steps alternating u -> tmp -> u ...; returns the grid that
    holds the result (u for even nsteps, tmp for odd).  norms: optional
    DeviceBuffer of nsteps doubles receiving the FP64 sum of squares after
    every step."""
    p = norms.ptr if isinstance(norms, DeviceBuffer) else norms
    call("cnv_diffusion_run", u.handle, tmp.handle, int(nsteps), ctypes.c_double(nu_dt), p, _st(stream))
    return u if nsteps % 2 == 0 else tmp


FIELD_OPS = {"add": _abi.CNV_FIELD_ADD, "sub": _abi.CNV_FIELD_SUB, "mul": _abi.CNV_FIELD_MUL,
             "div": _abi.CNV_FIELD_DIV}


def _dt(dtype):
    return _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32


def field_op(op, a, b, dst, n, dtype=np.float64, stream=None):
    """dst[i] = a[i] OP b[i] (b a device pointer) or a[i] OP b (b a number),
    op in add/sub/mul/div, on n elements of device memory
    (FieldOperator.h:11-141)."""
    if isinstance(b, (int, float, np.floating)):
        call("cnv_field_op1", FIELD_OPS[op], _dt(dtype), a, ctypes.c_double(float(b)), dst, int(n), _st(stream))
    else:
        call("cnv_field_op2", FIELD_OPS[op], _dt(dtype), a, b, dst, int(n), _st(stream))


def field_rcp(a, scalar, dst, n, dtype=np.float64, stream=None):
    """dst[i] = scalar * (1 / a[i]) on device memory (FieldOperator.h:25-30)."""
    call("cnv_field_rcp", _dt(dtype), a, ctypes.c_double(float(scalar)), dst, int(n), _st(stream))


GRID_OPS = dict(FIELD_OPS, rcp=_abi.CNV_FIELD_RCP, neg=_abi.CNV_FIELD_NEG)


def grid_field_op(op, dst, a, b=None, scalar=0.0, comps=None, stream=None):
    """Field / FieldContainer arithmetic over whole grids (Field.h:444-556,
    1039-1154), every block over its own elements:
    op add/sub/mul/div: dst = a OP b (b a Cartesian) or a OP scalar (b None);
    rcp: dst = scalar * (1 / a); neg: dst = -a.  comps = (cd, ca, cb) slab
    indices, or None for every slab.  dst may be a or b."""
    cd, ca, cb = (-1, -1, -1) if comps is None else (tuple(comps) + (0,))[:3]
    call("cnv_grid_field_op", GRID_OPS[op], dst.handle, int(cd), a.handle, int(ca),
         b.handle if b is not None else None, int(cb), ctypes.c_double(float(scalar)), _st(stream))


def _box(v):
    return None if v is None else (ctypes.c_int * 3)(*[int(x) for x in v])


def write_hdf(grid, fname, aname="data", time=0.0, file_dtype=None, face_dir=0, create_xdmf=True, begin=None,
              end=None, stream=None):
    """<fname>.h5 (dataset "data", Z x Y x X x ncomp) and <fname>.xmf of the
    grid's entity span, or of the global index box [begin, end)
    (IO/CartesianHDF.h:46-136).  file_dtype: numpy dtype of the file
    elements (default: the grid's)."""
    ft = _dt(grid.dtype if file_dtype is None else file_dtype)
    call("cnv_grid_write_hdf", grid.handle, str(fname).encode(), str(aname).encode(), ft, ctypes.c_double(time),
         int(face_dir), 1 if create_xdmf else 0, _box(begin), _box(end), _st(stream))


def read_hdf(grid, fname, file_dtype=None, face_dir=0, begin=None, end=None, stream=None):
    """Fill the grid (or the box [begin, end) of it) from <fname>.h5
    (IO/CartesianHDF.h:138-215)."""
    ft = _dt(grid.dtype if file_dtype is None else file_dtype)
    call("cnv_grid_read_hdf", grid.handle, str(fname).encode(), ft, int(face_dir), _box(begin), _box(end),
         _st(stream))


def CartesianWriteHDF(fname, aname, grid, time, face_dir=0, create_xdmf=True, file_dtype=None):
    """IO::CartesianWriteHDF<FileDataType>(fname, aname, grid, time, face_dir,
    create_xdmf) (IO/CartesianHDF.h:120-136)."""
    write_hdf(grid, fname, aname, time, file_dtype, face_dir, create_xdmf)


def CartesianReadHDF(fname, grid, face_dir=0, file_dtype=None):
    """IO::CartesianReadHDF<FileDataType>(fname, grid, face_dir) (IO/CartesianHDF.h:203-210)."""
    read_hdf(grid, fname, file_dtype, face_dir)


def grid_sums(g, comp=0, stream=None):
    """(sum, sum of squares) of one component, FP64, deterministic."""
    buf = DeviceBuffer(16)
    call("cnv_grid_sums", g.handle, comp, buf.ptr, _st(stream))
    r = buf.download(np.float64, 2, stream)
    return float(r[0]), float(r[1])


def autotune_stencil7(u, out, iters=20):
    """Pick (workgroup size, z-planes per tile) of the diffusion kernel for u's
    (dtype, block edge) by timing every candidate; returns (wg, zchunk, table)."""
    wg, zc, nr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    res = np.zeros(3 * 64, dtype=np.float64)
    call("cnv_autotune_stencil7", u.handle, out.handle, int(iters), ctypes.byref(wg), ctypes.byref(zc),
         res.ctypes.data_as(ctypes.c_void_p), 64, ctypes.byref(nr))
    table = [(int(res[3 * i]), int(res[3 * i + 1]), float(res[3 * i + 2])) for i in range(nr.value)]
    return wg.value, zc.value, table


def autotune_diffusion_run(u, tmp, iters=20):
    """Pick (workgroup size, z-planes per tile) of the two-step kernel that
    diffusion_run uses on periodic grids of 16^3 / 32^3 blocks; returns
    (wg, zchunk, table).  tmp is overwritten."""
    wg, zc, nr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    res = np.zeros(3 * 64, dtype=np.float64)
    call("cnv_autotune_diffusion_run", u.handle, tmp.handle, int(iters), ctypes.byref(wg), ctypes.byref(zc),
         res.ctypes.data_as(ctypes.c_void_p), 64, ctypes.byref(nr))
    table = [(int(res[3 * i]), int(res[3 * i + 1]), float(res[3 * i + 2])) for i in range(nr.value)]
    return wg.value, zc.value, table


def set_tune_diffusion_run(dtype, edge, wg, zchunk):
    call("cnv_tune_set_diffusion_run", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         int(wg), int(zchunk))


def tune_diffusion_run(dtype, edge):
    wg, zc = ctypes.c_int(), ctypes.c_int()
    call("cnv_tune_get_diffusion_run", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         ctypes.byref(wg), ctypes.byref(zc))
    return wg.value, zc.value


def tune_stencil7(dtype, edge):
    wg, zc = ctypes.c_int(), ctypes.c_int()
    call("cnv_tune_get_stencil7", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         ctypes.byref(wg), ctypes.byref(zc))
    return wg.value, zc.value


def set_tune_stencil7(dtype, edge, wg, zchunk):
    call("cnv_tune_set_stencil7", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         int(wg), int(zchunk))


def weno5(u, out, stream=None):
    """WENO5 face states of every component of u into out (6 components each:
    6c+2d = u^-_{i+1/2}, 6c+2d+1 = u^+_{i-1/2} in direction d)."""
    call("cnv_weno5", u.handle, out.handle, _st(stream))


def autotune_weno5(u, out, iters=5):
    wg, zc, nr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    res = np.zeros(3 * 64, dtype=np.float64)
    call("cnv_autotune_weno5", u.handle, out.handle, int(iters), ctypes.byref(wg), ctypes.byref(zc),
         res.ctypes.data_as(ctypes.c_void_p), 64, ctypes.byref(nr))
    table = [(int(res[3 * i]), int(res[3 * i + 1]), float(res[3 * i + 2])) for i in range(nr.value)]
    return wg.value, zc.value, table


class Batch:
    """Independent FP32 grids processed together: one persistent launch over
    a device list of work units (or, when the autotuner finds it faster, one
    launch per (operator, edge) class of cubic 16^3 / 32^3 blocks plus one
    persistent launch for the rest).

    items: list of (in_grid, out_grid, op, ghost, nu_dt) with op "diffusion"
    (out = in + nu_dt lap(in)) or "weno5" (6 states per component)."""

    OPS = {"diffusion": _abi.CNV_OP_DIFFUSION, "weno5": _abi.CNV_OP_WENO5}

    def __init__(self, items):
        arr = (_abi.cnv_batch_item * len(items))()
        for i, (gi, go, op, ghost, nudt) in enumerate(items):
            arr[i] = _abi.cnv_batch_item(gi.handle.value, go.handle.value, self.OPS[op], int(ghost), float(nudt))
        self._keep = [(it[0], it[1]) for it in items]
        h = ctypes.c_void_p()
        call("cnv_batch_create", arr, len(items), ctypes.byref(h))
        self.handle = h

    def autotune(self, iters=5):
        """Returns the trials as (op, edge, wg, depth, microseconds): op >= 0
        a class launch, -1 the class form's persistent launch, -2 the single
        launch, -3 each form as a whole (wg: the form, as in form())."""
        res = np.zeros(5 * 512, dtype=np.float64)
        nr = ctypes.c_int()
        call("cnv_batch_autotune", self.handle, int(iters), res.ctypes.data_as(ctypes.c_void_p), 512,
             ctypes.byref(nr))
        return [tuple(res[5 * i:5 * i + 5]) for i in range(min(nr.value, 512))]

    # the forms of a step: class launches overlapping on the batch's streams,
    # one persistent launch, class launches one after another
    CLASSES, SINGLE, CLASSES_SERIAL = 0, 1, 2
    FORM_NAMES = {0: "class_launches", 1: "single_launch", 2: "class_launches_serial"}

    def set_form(self, form):
        """Pin the form (Batch.CLASSES, SINGLE or CLASSES_SERIAL) until the
        next autotune."""
        call("cnv_batch_set_form", self.handle, int(form))

    def form(self):
        """(form, units): the form that runs and the single launch's work units."""
        f, nu = ctypes.c_int(), ctypes.c_int()
        call("cnv_batch_form", self.handle, ctypes.byref(f), ctypes.byref(nu))
        return f.value, nu.value

    def classes(self):
        """[(op, edge, wg, z depth, items)] of the per-class launches."""
        out = np.zeros(5 * 16, dtype=np.int32)
        n = ctypes.c_int()
        call("cnv_batch_class_info", self.handle, 16, out.ctypes.data_as(_abi.c_int_p), ctypes.byref(n))
        return [tuple(int(x) for x in out[5 * i:5 * i + 5]) for i in range(min(n.value, 16))]

    def run(self, stream=None):
        call("cnv_batch_run", self.handle, _st(stream))

    def info(self):
        n, wg, lds = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        call("cnv_batch_info", self.handle, ctypes.byref(n), ctypes.byref(wg), ctypes.byref(lds))
        return {"items": n.value, "wg": wg.value, "lds_bytes": lds.value}

    def __del__(self):
        try:
            if self.handle:
                lib().cnv_batch_destroy(self.handle)
        except Exception:
            pass


def set_zfaces(g, zfaces):
    """Stretched z mesh of grid g: nz_total + 1 increasing face positions."""
    zf = np.ascontiguousarray(zfaces, dtype=np.float64)
    call("cnv_grid_set_zfaces", g.handle, zf.ctypes.data_as(ctypes.c_void_p), int(zf.size))


def jacobi_sweep(u, f, out, omega=2.0 / 3.0, cu=0, cf=0, co=0, stream=None):
    call("cnv_jacobi_sweep", u.handle, cu, f.handle, cf, out.handle, co, ctypes.c_double(omega), _st(stream))


def jacobi_sweep2(u, f, out, omega=2.0 / 3.0, cu=0, cf=0, co=0, stream=None):
    """Two sweeps u -> out in one launch, equal to two jacobi_sweep calls
    (each domain face one rule on all its blocks: periodic, Dirichlet,
    Symmetry or Absorbing)."""
    call("cnv_jacobi_sweep2", u.handle, cu, f.handle, cf, out.handle, co, ctypes.c_double(omega), _st(stream))


def set_tune_jacobi2(dtype, edge, wg, zchunk, ysplit=1):
    """Pins the two-sweep tile: workgroup size, z planes per tile, and tiles
    per block plane (1 whole, 2 or 4 parts in y)."""
    call("cnv_tune_set_jacobi2", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         int(wg), int(zchunk), int(ysplit))


def get_tune_jacobi2(dtype, edge):
    """(wg, zchunk, ysplit) of the two-sweep tile for (dtype, block edge)."""
    wg, zc, ys = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("cnv_tune_get_jacobi2", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         ctypes.byref(wg), ctypes.byref(zc), ctypes.byref(ys))
    return wg.value, zc.value, ys.value


def jacobi_defect(u, f, r, sumsq=None, cu=0, cf=0, cr=0, stream=None):
    p = sumsq.ptr if isinstance(sumsq, DeviceBuffer) else sumsq
    call("cnv_jacobi_defect", u.handle, cu, f.handle, cf, r.handle, cr, p, _st(stream))


def add_correction(u, e, cu=0, ce=0, stream=None):
    call("cnv_add_correction", u.handle, cu, e.handle, ce, _st(stream))


def jacobi_solve(u, f, work, refine=True, sweeps_per_cycle=50, max_cycles=1000, omega=2.0 / 3.0, tol=1e-12,
                 history=True, stream=None):
    """Weighted-Jacobi solve of A u = f in place; returns (cycles, sweeps,
    relative residual, residual history per cycle)."""
    o = _abi.cnv_solve_opts(1 if refine else 0, int(sweeps_per_cycle), int(max_cycles), float(omega), float(tol))
    info = _abi.cnv_solve_info()
    hist = np.zeros(max_cycles + 1, dtype=np.float64)
    call("cnv_jacobi_solve", u.handle, f.handle, work.handle, ctypes.byref(o), ctypes.byref(info),
         hist.ctypes.data_as(ctypes.c_void_p) if history else None, int(hist.size), _st(stream))
    return info.cycles, info.sweeps, info.rel_residual, hist[:info.cycles + 1]


def grid_compare(a, b, comp_a=0, comp_b=0, stream=None):
    """(sum (a-b)^2, sum b^2, max|a-b|) over one component, FP64 on the device."""
    buf = DeviceBuffer(24)
    call("cnv_grid_compare", a.handle, comp_a, b.handle, comp_b, buf.ptr, _st(stream))
    r = buf.download(np.float64, 3, stream)
    return float(r[0]), float(r[1]), float(r[2])


def set_tune_weno5(dtype, edge, wg, zchunk):
    call("cnv_tune_set_weno5", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         int(wg), int(zchunk))


def autotune_jacobi(u, f, out, iters=5):
    """Times the Jacobi sweep tiles (workgroup size, z-planes per tile, f
    prefetch) on u -> out (component 0; out is overwritten) and keeps the
    fastest for (dtype, x/y block edge); returns (wg, zchunk, prefetch,
    [(wg, zchunk, prefetch, us per sweep), ...])."""
    wg, zc, pf, nr = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    res = np.zeros(4 * 64, dtype=np.float64)
    call("cnv_autotune_jacobi", u.handle, f.handle, out.handle, int(iters), ctypes.byref(wg), ctypes.byref(zc),
         ctypes.byref(pf), res.ctypes.data_as(ctypes.c_void_p), 64, ctypes.byref(nr))
    table = [(int(res[4 * i]), int(res[4 * i + 1]), int(res[4 * i + 2]), float(res[4 * i + 3]))
             for i in range(nr.value)]
    return wg.value, zc.value, pf.value, table


def set_tune_jacobi(dtype, edge, wg, zchunk, prefetch):
    call("cnv_tune_set_jacobi", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         int(wg), int(zchunk), int(prefetch))


def tune_jacobi(dtype, edge):
    wg, zc, pf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    call("cnv_tune_get_jacobi", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         ctypes.byref(wg), ctypes.byref(zc), ctypes.byref(pf))
    return wg.value, zc.value, pf.value


def tune_weno5(dtype, edge):
    wg, zc = ctypes.c_int(), ctypes.c_int()
    call("cnv_tune_get_weno5", _abi.CNV_F64 if np.dtype(dtype) == np.float64 else _abi.CNV_F32, int(edge),
         ctypes.byref(wg), ctypes.byref(zc))
    return wg.value, zc.value
