"""Grids over caller-owned device memory (cnv_grid_wrap): the Alloc seam of
Grid::Cartesian (Cartesian.h:481,527,537).  A wrapped allocation in the
reference layout gives the same operator results as the library's own grid,
bit for bit, and stays owned by the caller (a torch tensor or a raw
buffer)."""
import numpy as np
import pytest

import cubismnova_amd as cnv
from tests.helpers import random_slab

pytestmark = pytest.mark.gpu


# torch loads its own HIP runtime; it sees the device only when it initialises
# before this library's runtime does, so the torch cases run in a fresh
# interpreter that imports torch first
TORCH_CASE = r"""
import sys
sys.path.insert(0, sys.argv[1])
import torch
assert torch.cuda.is_available()
torch.ones(1, device="cuda").sum().item()
import numpy as np
import cubismnova_amd as cnv
from tests.helpers import random_slab
for dtype in (np.float64, np.float32):
    nb, n = (2, 3, 2), (16, 16, 16)
    own = cnv.Cartesian(nb, n, dtype=dtype)
    slab = random_slab(77, nb, n, dtype, block_stride=own.block_stride)
    own.upload(0, slab)
    t = torch.from_numpy(slab.copy()).to("cuda")
    torch.cuda.synchronize()
    w = cnv.Cartesian.from_tensor(t, nb, n)
    assert w.block_stride == own.block_stride and w.component_bytes == own.component_bytes
    a, b = cnv.Cartesian(nb, n, dtype=dtype), cnv.Cartesian(nb, n, dtype=dtype)
    cnv.laplacian(own, a)
    cnv.laplacian(w, b)
    assert np.array_equal(a.download(0), b.download(0))
    # no copy: a change made through torch is seen by the grid, and results
    # the library writes into the grid are seen by torch
    t.mul_(2.0)
    torch.cuda.synchronize()
    own.upload(0, 2 * slab)
    cnv.laplacian(own, a)
    cnv.laplacian(w, b)
    assert np.array_equal(a.download(0), b.download(0))
    cnv.diffusion_step(own, w, 1e-6)
    out = cnv.Cartesian(nb, n, dtype=dtype)
    cnv.diffusion_step(own, out, 1e-6)
    assert np.array_equal(t.cpu().numpy(), out.download(0))
    try:
        cnv.Cartesian.from_tensor(t[:-1], nb, n)
        raise SystemExit("short tensor accepted")
    except cnv.CnvError:
        pass
    try:
        cnv.Cartesian.from_tensor(t.cpu(), nb, n)
        raise SystemExit("host tensor accepted")
    except ValueError:
        pass
print("TORCH CASES OK")
"""


def test_tensor_grid_matches_own_grid(gpu):
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", TORCH_CASE, root], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "TORCH CASES OK" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]


def test_wrap_checks(gpu):
    nb, n = (2, 2, 2), (8, 8, 8)
    buf = cnv.DeviceBuffer(8 * 512 * 8)
    cnv.Cartesian(nb, n, device_data=buf.ptr, device_bytes=buf.nbytes)
    with pytest.raises(cnv.CnvError):  # one element short
        cnv.Cartesian(nb, n, device_data=buf.ptr, device_bytes=buf.nbytes - 8)
    with pytest.raises(cnv.CnvError):  # not 32-byte aligned
        cnv.Cartesian((1, 1, 1), (4, 4, 4), device_data=buf.ptr.value + 8, device_bytes=1024)
    with pytest.raises(cnv.CnvError):  # host memory
        host = np.zeros(8 * 512)
        cnv.Cartesian(nb, n, device_data=host.ctypes.data, device_bytes=host.nbytes)
    # a node grid over a buffer: (N+1)^3 elements per block
    buf = cnv.DeviceBuffer(8 * 732 * 8)
    g = cnv.Cartesian(nb, n, entity="node", device_data=buf.ptr, device_bytes=buf.nbytes)
    assert g.block_stride == 732  # 9^3 = 729 doubles rounded up to 32 bytes
    with pytest.raises(cnv.CnvError):
        cnv.Cartesian(nb, n, entity="node", device_data=buf.ptr, device_bytes=8 * 729 * 8)


def test_wrapped_memory_outlives_grid(gpu):
    nb, n = (1, 1, 2), (8, 8, 8)
    buf = cnv.DeviceBuffer(2 * 512 * 4)
    buf.fill(np.float32, 3.5)
    g = cnv.Cartesian(nb, n, dtype=np.float32, device_data=buf.ptr, device_bytes=buf.nbytes)
    assert (g.download(0) == 3.5).all()
    del g
    assert (buf.download(np.float32) == 3.5).all()


def test_handles_over_shared_memory_count_as_one_grid(gpu):
    """Two handles wrapped over the same (or overlapping) memory are one grid
    for the operators' input/output checks and for a batch's items."""
    nb, n = (2, 2, 2), (16, 16, 16)
    own = cnv.Cartesian(nb, n, dtype=np.float32)
    alias = cnv.Cartesian(nb, n, dtype=np.float32, device_data=own.data_ptr(0), device_bytes=own.component_bytes)
    with pytest.raises(cnv.CnvError, match="share memory"):
        cnv.diffusion_step(own, alias, 1e-4)
    with pytest.raises(cnv.CnvError, match="share memory"):
        cnv.laplacian(own, alias)
    with pytest.raises(cnv.CnvError, match="share memory"):
        cnv.Batch([(own, alias, "diffusion", 1, 1e-4)])
    # a handle over the second half of another grid's memory overlaps it too
    big = cnv.Cartesian((2, 2, 4), n, dtype=np.float32)
    half = cnv.Cartesian(nb, n, dtype=np.float32, device_data=big.data_ptr(0).value + big.component_bytes // 2,
                         device_bytes=big.component_bytes // 2)
    with pytest.raises(cnv.CnvError, match="share memory"):
        cnv.Batch([(half, big, "diffusion", 1, 1e-4)])
    # apart: fine
    other = cnv.Cartesian(nb, n, dtype=np.float32)
    own.upload(0, random_slab(3, nb, n, np.float32))
    cnv.diffusion_step(own, other, 1e-4)
