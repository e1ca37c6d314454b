"""The allocation-free Euler step (cnv_diffusion_step_noalloc) against
cnv_diffusion_step: the field bit-identical and the FP64 norm identical, on
two calls in a row, the second on a smaller grid of another dtype and more
components; a grid that needs more norm partials than the static scratch
array holds is refused before any launch."""
import ctypes

import numpy as np
import pytest

import cubismnova_amd as cnv
from cubismnova_amd import _abi
from tests.helpers import random_slab

pytestmark = pytest.mark.gpu


def desc_of(g):
    d = _abi.cnv_grid_desc()
    cnv.call("cnv_grid_info", g.handle, ctypes.byref(d), None, None)
    return d


def step_both(nb, n, dtype, ncomp, seed):
    """One step through each entry point on the same bytes; returns the two
    fields and the two norms."""
    u = cnv.Cartesian(nb, n, dtype=dtype, ncomp=ncomp)
    out = cnv.Cartesian(nb, n, dtype=dtype, ncomp=ncomp)
    slabs = [random_slab(seed + c, nb, n, dtype, u.block_stride) for c in range(ncomp)]
    for c in range(ncomp):
        u.upload(c, slabs[c])
    h = 1.0 / (nb[0] * n[0])
    nudt = 0.1 * h * h
    s_ref = cnv.DeviceBuffer(8)
    cnv.diffusion_step(u, out, nudt, sumsq=s_ref)
    want = [out.download(c) for c in range(ncomp)]

    nbytes = u.component_bytes * ncomp
    a, b, s = cnv.DeviceBuffer(nbytes), cnv.DeviceBuffer(nbytes), cnv.DeviceBuffer(8)
    a.upload(np.concatenate(slabs))
    b.fill(dtype, -3.0)
    s.fill(np.float64, np.nan)
    cnv.call("cnv_diffusion_step_noalloc", ctypes.byref(desc_of(u)), a.ptr, b.ptr, ctypes.c_double(nudt), s.ptr,
             None)
    cnv.call("cnv_stream_synchronize", None)
    got = b.download(dtype).reshape(ncomp, -1)
    return want, [got[c] for c in range(ncomp)], s_ref.download(np.float64, 1)[0], s.download(np.float64, 1)[0]


@pytest.mark.parametrize("order", ["large-then-small"])
def test_noalloc_step_matches_diffusion_step(gpu, order):
    # first call: 256^3 cells in 32^3 blocks, FP32, one component
    want, got, nref, ngot = step_both((8, 8, 8), (32, 32, 32), np.float32, 1, 11)
    assert np.array_equal(got[0], want[0])
    assert nref == ngot and nref > 0
    # second call, smaller: 2x3x2 blocks of 16^3, FP64, three components (the
    # array keeps the first call's partials; the second writes fewer)
    want, got, nref, ngot = step_both((2, 3, 2), (16, 16, 16), np.float64, 3, 23)
    for c in range(3):
        assert np.array_equal(got[c], want[c]), c
    assert nref == ngot and nref > 0


def test_noalloc_rejects_grid_beyond_scratch(gpu):
    # 41^3 blocks of 16^3 cells in FP32 with 8-plane tiles: 2 tiles per block,
    # 137842 partials > 2^17; the buffers are full size (1.13 GB each), so a
    # launch could not run out of bounds even if the check were missing
    base = cnv.tune_stencil7(np.float32, 16)
    cnv.set_tune_stencil7(np.float32, 16, 64, 8)
    try:
        d = _abi.cnv_grid_desc()
        for k in range(3):
            d.nblocks[k], d.block_cells[k], d.begin[k], d.end[k] = 41, 16, 0.0, 1.0
        d.ncomp, d.dtype = 1, _abi.CNV_F32
        nbytes = 41 ** 3 * 16 ** 3 * 4
        a, b, s = cnv.DeviceBuffer(nbytes), cnv.DeviceBuffer(nbytes), cnv.DeviceBuffer(8)
        a.fill(np.float32, 1.0)
        with pytest.raises(cnv.CnvError) as e:
            cnv.call("cnv_diffusion_step_noalloc", ctypes.byref(d), a.ptr, b.ptr, ctypes.c_double(1e-6), s.ptr,
                     None)
        assert e.value.status == _abi.CNV_ERR_SIZE
        # without a norm there are no partials and the same grid steps
        cnv.call("cnv_diffusion_step_noalloc", ctypes.byref(d), a.ptr, b.ptr, ctypes.c_double(1e-6), None, None)
        cnv.call("cnv_stream_synchronize", None)
        assert np.all(b.download(np.float32, 1 << 20) == 1.0)
    finally:
        cnv.set_tune_stencil7(np.float32, 16, *base)
