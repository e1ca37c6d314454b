"""GPU parity of the element-wise field operations, through both the
stream-ordered entry points and the reference's own hook names
(FieldOperator.h:149-227), against numpy's IEEE arithmetic in the same
precision: bit-identical for every operation, size and alignment."""
import ctypes

import numpy as np
import pytest

import cubismnova_amd as cnv
from cubismnova_amd import _abi

pytestmark = pytest.mark.gpu

REF = {"add": np.add, "sub": np.subtract, "mul": np.multiply, "div": np.divide}


def values(seed, n, dtype):
    rng = np.random.Generator(np.random.PCG64(seed))
    v = (rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)).astype(dtype)
    v[::97] = 0.0                      # zeros: division by zero gives +-inf
    if n > 10:
        v[5] = np.inf
        v[7] = -2.5
    return v


def dev(arr):
    b = cnv.DeviceBuffer(arr.nbytes)
    b.upload(arr)
    return b


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("n,off", [(1, 0), (1000, 0), (1 << 20, 3), ((1 << 18) + 5, 1)])
def test_field_ops_bitexact(gpu, dtype, n, off):
    isz = np.dtype(dtype).itemsize
    a = values(1, n + off, dtype)
    b = values(2, n + off, dtype)
    da, db, dd = dev(a), dev(b), cnv.DeviceBuffer((n + off) * isz)
    pa, pb, pd = da.at(off * isz), db.at(off * isz), dd.at(off * isz)
    s = dtype(-0.37)
    with np.errstate(all="ignore"):
        for op, f in REF.items():
            cnv.field_op(op, pa, pb, pd, n, dtype)
            got = dd.download(dtype, n + off)[off:]
            np.testing.assert_array_equal(got, f(a[off:], b[off:]), err_msg=op)
            cnv.field_op(op, pa, float(s), pd, n, dtype)
            np.testing.assert_array_equal(dd.download(dtype, n + off)[off:], f(a[off:], s), err_msg=op + "1")
        cnv.field_rcp(pa, float(s), pd, n, dtype)
        want = s * (dtype(1) / a[off:])
        np.testing.assert_array_equal(dd.download(dtype, n + off)[off:], want)


@pytest.mark.parametrize("dtype,tname", [(np.float32, "float"), (np.float64, "double")])
def test_reference_hook_names(gpu, dtype, tname):
    n = 4099
    a, b = values(3, n, dtype), values(4, n, dtype)
    da, db, dd = dev(a), dev(b), cnv.DeviceBuffer(a.nbytes)
    L = _abi.lib()
    cs = ctypes.c_float if dtype == np.float32 else ctypes.c_double
    with np.errstate(all="ignore"):
        for op, f in REF.items():
            getattr(L, "field%s2_%s" % (op.capitalize(), tname))(da.ptr, db.ptr, dd.ptr, n)
            np.testing.assert_array_equal(dd.download(dtype, n), f(a, b))
            getattr(L, "field%s1_%s" % (op.capitalize(), tname))(da.ptr, cs(1.75), dd.ptr, n)
            np.testing.assert_array_equal(dd.download(dtype, n), f(a, dtype(1.75)))
        getattr(L, "fieldRcp_%s" % tname)(da.ptr, cs(3.0), dd.ptr, n)
        np.testing.assert_array_equal(dd.download(dtype, n), dtype(3.0) * (dtype(1) / a))


def test_field_op_errors(gpu):
    with pytest.raises(cnv.CnvError):
        cnv.field_op("add", None, None, None, 10)
    with pytest.raises(cnv.CnvError):
        _abi.call("cnv_field_op2", 7, _abi.CNV_F64, None, None, None, 0, None)
    cnv.field_op("add", None, None, None, 0)      # empty: no-op


# ------------------------------------------------ whole grids (Field.h:444-556, 1039-1154)

GRID_SHAPES = [("cell", (2, 3, 2), (8, 8, 8), 1), ("cell", (3, 1, 2), (5, 7, 3), 2),
               ("node", (2, 2, 3), (4, 5, 3), 1), ("face", (2, 2, 2), (3, 4, 5), 2)]


def grid_expected(op, g, ain, bin_, s, dst_before, comps):
    """numpy over each block's own extent; padding keeps dst's old values."""
    out = {c: dst_before[c].copy() for c in dst_before}
    nb = g.nblocks
    todo = [(c, c, c) for c in range(g.nslabs)] if comps is None else [comps]
    T = g.dtype.type
    for (cd, ca, cb) in todo:
        o = out[cd].reshape(g.nblocks_total, g.block_stride)
        x = ain[ca].reshape(g.nblocks_total, g.block_stride)
        y = None if bin_ is None else bin_[cb].reshape(g.nblocks_total, g.block_stride)
        for blk in range(g.nblocks_total):
            bi = (blk % nb[0], (blk // nb[0]) % nb[1], blk // (nb[0] * nb[1]))
            ne = int(np.prod(g.block_extent(bi, ca)))
            xa = x[blk, :ne]
            with np.errstate(divide="ignore", invalid="ignore"):  # IEEE inf/nan are part of the contract
                if op == "neg":
                    o[blk, :ne] = -xa
                elif op == "rcp":
                    o[blk, :ne] = T(s) * (T(1) / xa)
                else:
                    o[blk, :ne] = REF[op](xa, y[blk, :ne] if y is not None else T(s))
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", GRID_SHAPES, ids=[s[0] + str(i) for i, s in enumerate(GRID_SHAPES)])
@pytest.mark.parametrize("op", ["add", "sub", "mul", "div", "rcp", "neg"])
def test_grid_field_op_every_slab(gpu, dtype, shape, op):
    ent, nb, n, nc = shape
    mk = lambda: cnv.Cartesian(nb, n, dtype=dtype, ncomp=nc, entity=ent)  # noqa: E731
    a, b, d = mk(), mk(), mk()
    cnt = a.component_bytes // a.dtype.itemsize
    A = {c: values(10 + c, cnt, dtype) for c in range(a.nslabs)}
    Bv = {c: values(20 + c, cnt, dtype) for c in range(a.nslabs)}
    D = {c: np.full(cnt, 7.0, dtype=dtype) for c in range(a.nslabs)}
    for c in range(a.nslabs):
        a.upload(c, A[c])
        b.upload(c, Bv[c])
        d.upload(c, D[c])
    two = op in REF
    s = -1.75
    cnv.grid_field_op(op, d, a, b if two else None, s)
    want = grid_expected(op, a, A, Bv if two else None, s, D, None)
    for c in range(a.nslabs):
        assert np.array_equal(d.download(c), want[c], equal_nan=True), (op, c)
    if two:  # grid OP scalar
        cnv.grid_field_op(op, d, a, None, s)
        want = grid_expected(op, a, A, None, s, want, None)
        for c in range(a.nslabs):
            assert np.array_equal(d.download(c), want[c], equal_nan=True), (op, "scalar", c)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_grid_field_op_one_slab_in_place(gpu, dtype):
    # a += b on slab 3 from slab 5 of a face grid (both faces normal to z, ncomp = 2)
    nb, n = (2, 3, 2), (4, 4, 6)
    a = cnv.Cartesian(nb, n, dtype=dtype, ncomp=2, entity="face")
    b = cnv.Cartesian(nb, n, dtype=dtype, ncomp=2, entity="face")
    cnt = a.component_bytes // a.dtype.itemsize
    A = {c: values(30 + c, cnt, dtype) for c in range(6)}
    Bv = {c: values(40 + c, cnt, dtype) for c in range(6)}
    for c in range(6):
        a.upload(c, A[c])
        b.upload(c, Bv[c])
    cnv.grid_field_op("add", a, a, b, comps=(5, 5, 4))
    want = grid_expected("add", a, A, Bv, 0.0, A, (5, 5, 4))
    for c in range(6):
        assert np.array_equal(a.download(c), want[c], equal_nan=True), c
    with pytest.raises(cnv.CnvError):  # x faces and z faces differ in shape
        cnv.grid_field_op("add", a, a, b, comps=(0, 0, 5))
    c = cnv.Cartesian(nb, n, dtype=dtype, ncomp=2, entity="node")
    with pytest.raises(cnv.CnvError):
        cnv.grid_field_op("add", a, a, c)
    with pytest.raises(cnv.CnvError):
        cnv.grid_field_op("neg", a, a, b)
