"""Every tile shape of the one-step diffusion kernel and of the WENO5 kernel
on 32^3 blocks at 256^3 (enough tiles for several rounds of workgroups): the
field is bit-identical to the built-in tile's, and the FP64 norm of the
diffusion step agrees to 1e-13 (each tile shape sums its own partials in its
own fixed tree).  The autotuner may pick any of these shapes."""
import numpy as np
import pytest

import cubismnova_amd as cnv

pytestmark = pytest.mark.gpu

NB, N = (8, 8, 8), (32, 32, 32)


def grid(dtype, seed):
    g = cnv.Cartesian(NB, N, dtype=dtype)
    rng = np.random.Generator(np.random.PCG64(seed))
    g.upload(0, rng.standard_normal(g.component_bytes // np.dtype(dtype).itemsize).astype(dtype))
    return g


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_one_step_every_tile_full_size(gpu, dtype):
    g, out = grid(dtype, 21), cnv.Cartesian(NB, N, dtype=dtype)
    s = cnv.DeviceBuffer(8)
    base = cnv.tune_stencil7(dtype, 32)
    cnv.diffusion_step(g, out, 1e-6, sumsq=s)
    ref, rs = out.download(0), s.download(np.float64, 1)[0]
    tried = 0
    try:
        for wg in (64, 128, 256, 512):
            for z in (1, 2, 4, 8, 16, 32):
                try:
                    cnv.set_tune_stencil7(dtype, 32, wg, z)
                except cnv.CnvError:
                    continue
                tried += 1
                cnv.diffusion_step(g, out, 1e-6, sumsq=s)
                assert np.array_equal(out.download(0), ref), (wg, z)
                assert abs(s.download(np.float64, 1)[0] - rs) <= 1e-13 * rs, (wg, z)
    finally:
        cnv.set_tune_stencil7(dtype, 32, *base)
    assert tried >= 8


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_weno5_every_tile_full_size(gpu, dtype):
    g, out = grid(dtype, 22), cnv.Cartesian(NB, N, dtype=dtype, ncomp=6)
    base = cnv.tune_weno5(dtype, 32)
    cnv.weno5(g, out)
    ref = [out.download(k) for k in range(6)]
    tried = 0
    try:
        for wg in (64, 128, 256, 512):
            for z in (4, 8, 16, 32):
                try:
                    cnv.set_tune_weno5(dtype, 32, wg, z)
                except cnv.CnvError:
                    continue
                tried += 1
                cnv.weno5(g, out)
                for k in range(6):
                    assert np.array_equal(out.download(k), ref[k]), (wg, z, k)
    finally:
        cnv.set_tune_weno5(dtype, 32, *base)
    assert tried >= 8
