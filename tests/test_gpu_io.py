"""The on-disk format (IO/CartesianHDF.h:46-215, src/IO/HDFDriver.cpp:26-148,
IO/XDMFDriver.h:29-160): the file written from the GPU grid holds the
dataset "data" of shape (Z, Y, X, ncomp) in the reference's interleaved
order, checked against a numpy assembly of the block layout and read by an
independent HDF5 reader (h5py of the image's conda Python) when present;
reading it back restores the grid bit for bit.  Restates
CartesianHDFTest.cpp:221-252 (write, read back, every block equal to its
index)."""
import os
import subprocess

import numpy as np
import pytest

import cubismnova_amd as cnv
from tests.helpers import random_slab

pytestmark = pytest.mark.gpu

H5PY = "/opt/conda/bin/python3.9"


def h5_read(path):
    """The dataset "data" of an HDF5 file, read by h5py in another interpreter
    (None where the image has no such reader; the in-library read-back
    checks still run)."""
    if not os.path.exists(H5PY):
        return None
    out = path + ".npy"
    code = "import sys, h5py, numpy; numpy.save(sys.argv[2], h5py.File(sys.argv[1], 'r')['data'][...])"
    r = subprocess.run([H5PY, "-c", code, path, out], capture_output=True, text=True, timeout=120)
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        return None
    assert r.returncode == 0, r.stderr[-2000:]
    return np.load(out)


def plus_of(g, slab):
    return g.plus(slab)


def to_global(g, slabs, first_slab, nc):
    """(Z, Y, X, nc) array of the entity's global span from downloaded slabs,
    each block at its own extent (FieldAOS.h AOSDriver)."""
    p = plus_of(g, first_slab)
    n, nb = g.block_cells, g.nblocks
    G = [nb[d] * n[d] + p[d] for d in range(3)]
    out = np.zeros((G[2], G[1], G[0], nc), dtype=slabs[0].dtype)
    for b in range(g.nblocks_total):
        bi = (b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1]))
        e = g.block_extent(bi, first_slab)
        cnt = e[0] * e[1] * e[2]
        for c in range(nc):
            blk = slabs[c][b * g.block_stride:b * g.block_stride + cnt].reshape(e[2], e[1], e[0])
            out[bi[2] * n[2]:bi[2] * n[2] + e[2], bi[1] * n[1]:bi[1] * n[1] + e[1],
                bi[0] * n[0]:bi[0] * n[0] + e[0], c] = blk
    return out


def filled(entity, nb, n, dtype, ncomp, seed):
    g = cnv.Cartesian(nb, n, dtype=dtype, ncomp=ncomp, entity=entity)
    el = tuple(x + (0 if entity == "cell" else 1) for x in n)
    slabs = [random_slab(seed + s, nb, el, dtype, block_stride=g.block_stride) for s in range(g.nslabs)]
    for s in range(g.nslabs):
        g.upload(s, slabs[s])
    return g, slabs


def test_write_read_back_block_index(gpu, tmp_path):
    # CartesianHDFTest.cpp:221-252: every block filled with its index
    nb, n = (3, 3, 3), (8, 8, 8)
    src = cnv.Cartesian(nb, n)
    slab = np.zeros(src.component_bytes // 8)
    for k in range(27):
        slab[k * src.block_stride:(k + 1) * src.block_stride] = k
    src.upload(0, slab)
    fn = str(tmp_path / "framen")
    cnv.CartesianWriteHDF(fn, "ramen", src, 0.0)
    dst = cnv.Cartesian(nb, n)
    cnv.CartesianReadHDF(fn, dst)
    got = dst.download(0).reshape(27, -1)[:, :512]
    assert (got == np.arange(27)[:, None]).all()
    with pytest.raises(cnv.CnvError, match="does not exist"):
        cnv.CartesianReadHDF(str(tmp_path / "nothing"), dst)


CASES = [("cell", (2, 3, 2), (8, 4, 6), 1, 0), ("cell", (2, 2, 1), (4, 4, 4), 3, 0),
         ("node", (2, 2, 3), (4, 5, 3), 2, 0), ("face", (2, 1, 2), (4, 3, 5), 2, 0),
         ("face", (2, 1, 2), (4, 3, 5), 2, 1), ("face", (2, 1, 2), (4, 3, 5), 2, 2)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", CASES, ids=["%s%d_c%d_f%d" % (c[0], i, c[3], c[4]) for i, c in enumerate(CASES)])
def test_file_layout_and_round_trip(gpu, tmp_path, dtype, case):
    ent, nb, n, nc, fd = case
    g, slabs = filled(ent, nb, n, dtype, nc, 900)
    first = fd * nc if ent == "face" else 0
    fn = str(tmp_path / "f")
    cnv.write_hdf(g, fn, "u", time=0.5, face_dir=fd)
    want = to_global(g, slabs[first:first + nc], first, nc)
    d = h5_read(fn + ".h5")
    if d is not None:
        assert d.dtype == np.dtype(dtype) and d.shape == want.shape
        assert np.array_equal(d, want)
    back, _ = filled(ent, nb, n, dtype, nc, 5000)
    cnv.read_hdf(back, fn, face_dir=fd)
    for c in range(nc):
        s = first + c
        a, b = g.download(s).reshape(g.nblocks_total, -1), back.download(s).reshape(g.nblocks_total, -1)
        for blk in range(g.nblocks_total):
            bi = (blk % nb[0], (blk // nb[0]) % nb[1], blk // (nb[0] * nb[1]))
            cnt = int(np.prod(g.block_extent(bi, s)))
            assert np.array_equal(a[blk, :cnt], b[blk, :cnt]), (c, bi)
    xmf = fn + ".xmf"
    if ent == "face":
        assert not os.path.exists(xmf)
    else:
        t = open(xmf).read()
        dz, dy, dx = want.shape[:3]
        assert 'DataItem Dimensions="%d %d %d %d"' % (dz, dy, dx, nc) in t
        m = (dz, dy, dx) if ent == "node" else (dz + 1, dy + 1, dx + 1)
        assert 'TopologyType="3DCoRectMesh" Dimensions="%d %d %d"' % m in t
        assert 'Center="%s"' % ("Node" if ent == "node" else "Cell") in t
        assert "./f.h5:/data" in t and 'Attribute Name="u"' in t
        assert 'Precision="%d"' % np.dtype(dtype).itemsize in t


def test_box_and_file_dtype(gpu, tmp_path):
    nb, n = (2, 2, 2), (8, 8, 8)
    g, slabs = filled("cell", nb, n, np.float64, 1, 31)
    want = to_global(g, slabs, 0, 1)
    fn = str(tmp_path / "box")
    lo, hi = (3, 0, 5), (12, 7, 16)
    cnv.write_hdf(g, fn, file_dtype=np.float32, begin=lo, end=hi)
    box = want[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
    d = h5_read(fn + ".h5")
    if d is not None:
        assert d.dtype == np.float32 and d.shape == box.shape
        assert np.array_equal(d, box.astype(np.float32))
    t = open(fn + ".xmf").read()
    assert "%.16g %.16g %.16g" % (lo[2] / 16, lo[1] / 16, lo[0] / 16) in t  # origin of the box, ZYX
    # read the box back into a zero grid: only the box changes
    z = cnv.Cartesian(nb, n)
    z.upload(0, np.zeros(z.component_bytes // 8))
    cnv.read_hdf(z, fn, file_dtype=np.float32, begin=lo, end=hi)
    got = to_global(z, [z.download(0)], 0, 1)
    exp = np.zeros_like(want)
    exp[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = box.astype(np.float32).astype(np.float64)
    assert np.array_equal(got, exp)
    # a file smaller than the span asked for is an error
    with pytest.raises(cnv.CnvError):
        cnv.read_hdf(z, fn, file_dtype=np.float32)
