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
