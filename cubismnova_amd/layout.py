"""Conversions between a global (Z, Y, X) array and the block layout of a
Grid::Cartesian component slab (include/Cubism/Grid/Cartesian.h:490-529,
include/Cubism/Grid/BlockFieldAssembler.h:188-199): blocks numbered x-fastest,
each block x-fastest inside, consecutive blocks `block_stride` elements apart
(block_bytes rounded up to 32 bytes)."""
import numpy as np


def global_to_blocks(u, nblocks, cells, block_stride=None):
    nbx, nby, nbz = nblocks
    nx, ny, nz = cells
    u = np.asarray(u)
    assert u.shape == (nbz * nz, nby * ny, nbx * nx), u.shape
    blk = u.reshape(nbz, nz, nby, ny, nbx, nx).transpose(0, 2, 4, 1, 3, 5)
    nbt, n = nbx * nby * nbz, nx * ny * nz
    bs = n if block_stride is None else int(block_stride)
    out = np.zeros((nbt, bs), dtype=u.dtype)
    out[:, :n] = blk.reshape(nbt, n)
    return out.reshape(-1)


def blocks_to_global(slab, nblocks, cells, block_stride=None):
    nbx, nby, nbz = nblocks
    nx, ny, nz = cells
    nbt, n = nbx * nby * nbz, nx * ny * nz
    bs = n if block_stride is None else int(block_stride)
    b = np.asarray(slab).reshape(nbt, bs)[:, :n].reshape(nbz, nby, nbx, nz, ny, nx)
    return np.ascontiguousarray(b.transpose(0, 3, 1, 4, 2, 5).reshape(nbz * nz, nby * ny, nbx * nx))
