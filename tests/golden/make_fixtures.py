"""Writes tests/golden/lab_fixtures.npz: known-answer lab dumps and
Laplacian / Euler outputs on tiny grids.

The values come from a deliberately naive evaluator written here from the
reference's definitions, cell by cell on a dictionary keyed by lab
multi-index (no padded buffers, no row copies, no shared code with the C++
oracle or the HIP kernels):

  loadInner   Block/FieldLabLoader.h:129-173  (copy of the block)
  loadGhosts  Block/FieldLabLoader.h:175-287  (26 neighbours, periodic wrap,
              skip / periodic from the last BC per direction, FieldLab.h:206-214,
              faces only for non-tensorial stencils)
  BCs         BC/Absorbing.h:62-104, BC/Dirichlet.h:75-113,
              BC/Symmetry.h:69-114, applied in list order (FieldLab.h:219-223),
              skipped for zero stencil width (BC/Base.h:73-78)
  Laplacian   Laplacian.cpp:123-130: ih2 * ((l[+] - 2 c) + l[-]), x + y + z
  Euler       Field.h:457-512: lap *= nu_dt; u += lap (two roundings)

Node / face cases (plus flags) store blocks at the upper end of a plus
dimension with one more element and copy ghosts below a block from the
neighbour's own end (FieldLabLoader.h:217-219); their labs are laid out in
the largest block's geometry, cells outside a block's active range NaN.

Inputs are multiples of 1/64 in [-1, 1] drawn from a fixed splitmix64 stream,
exact in FP32 and FP64, so one input serves both precisions.  Cells the
reference never writes, and BC values read from such cells (tensorial edges
filled from a cell no load has written yet), are NaN.  Run from the repository root:
    python tests/golden/make_fixtures.py
"""
import itertools
import json
import os

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lab_fixtures.npz")
M64 = (1 << 64) - 1


def splitmix(seed, n):
    s, out = seed & M64, []
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out.append(z ^ (z >> 31))
    return out


def field_values(seed, nblocks, n, plus=(0, 0, 0)):
    cnt = nblocks * (n[0] + plus[0]) * (n[1] + plus[1]) * (n[2] + plus[2])
    return np.array([(v % 129) - 64 for v in splitmix(seed, cnt)], dtype=np.float64) / 64.0


class Grid:
    """Blocks of n elements per dim; with plus[d] the blocks at the upper end
    of dim d hold n[d] + 1 (node / face fields, BlockFieldAssembler.h:121-131).
    Block b occupies [b*stride, ...) x-fastest in its own extent."""

    def __init__(self, nb, n, values, plus=(0, 0, 0)):
        self.nb, self.n, self.plus = tuple(nb), tuple(n), tuple(plus)
        self.stride = (n[0] + plus[0]) * (n[1] + plus[1]) * (n[2] + plus[2])
        self.v = values

    def block_id(self, B):
        return B[0] + self.nb[0] * (B[1] + self.nb[1] * B[2])

    def extent(self, B):
        return tuple(self.n[d] + (1 if self.plus[d] and B[d] == self.nb[d] - 1 else 0) for d in range(3))

    def cell(self, B, p):
        e = self.extent(B)
        assert all(0 <= p[d] < e[d] for d in range(3))
        return self.v[self.block_id(B) * self.stride + p[0] + e[0] * (p[1] + e[1] * p[2])]


def lab_of(g, B, sb, se, tensorial, bcs):
    """dict lab[(x, y, z)] for block B (tuple), stencil [sb, se)."""
    n = g.extent(B)
    lab = {}
    # inner
    for p in itertools.product(*[range(n[d]) for d in range(3)]):
        lab[p] = g.cell(B, p)
    # skip / periodic
    periodic, skip = [True] * 3, [1] * 3
    for (kind, d, side, val) in bcs:
        periodic[d] = kind == "base"
        skip[d] = -1 if side == 0 else 1
    for bi in itertools.product((-1, 0, 1), repeat=3):
        if bi == (0, 0, 0):
            continue
        if any((not periodic[d]) and bi[d] == skip[d] for d in range(3)):
            continue
        if not tensorial and sum(abs(x) for x in bi) > 1:
            continue
        rng = []
        for d in range(3):
            if bi[d] < 0:
                rng.append(range(sb[d], 0))
            elif bi[d] == 0:
                rng.append(range(0, n[d]))
            else:
                rng.append(range(n[d], n[d] + se[d] - 1))
        NB = tuple((B[d] + bi[d]) % g.nb[d] for d in range(3))
        ne = g.extent(NB)
        # below the block the source counts back from the neighbour's own end
        # (FieldLabLoader.h:217-219)
        shift = [ne[d] - n[d] if bi[d] < 0 else 0 for d in range(3)]
        for q in itertools.product(*rng):
            lab[q] = g.cell(NB, tuple(q[d] - bi[d] * n[d] + shift[d] for d in range(3)))
    # BCs in list order
    for (kind, d, side, val) in bcs:
        if kind == "base" or sb[d] == 0 or se[d] == 1:
            continue
        other = []
        for e in range(3):
            if e == d:
                other.append(None)
            elif tensorial:
                other.append(range(sb[e], n[e] + se[e] - 1))
            else:
                other.append(range(0, n[e]))
        layers = range(-1, sb[d] - 1, -1) if side == 0 else range(n[d], n[d] + se[d] - 1)
        for o in itertools.product(*[r if r is not None else [0] for r in other]):
            for k in layers:
                q = list(o)
                q[d] = k
                q = tuple(q)
                if kind == "dirichlet":
                    lab[q] = val
                elif kind == "absorbing":
                    r = list(q)
                    r[d] = 0 if side == 0 else n[d] - 1
                    lab[q] = lab.get(tuple(r), np.nan)   # unwritten stays undefined
                elif kind == "symmetry":
                    r = list(q)
                    r[d] = -1 - k if side == 0 else 2 * n[d] - 1 - k
                    lab[q] = val * lab.get(tuple(r), np.nan)
    return lab


def lab_array(lab, n, sb, se):
    ext = [n[d] + se[d] - 1 - sb[d] for d in range(3)]
    a = np.full((ext[2], ext[1], ext[0]), np.nan)
    for (x, y, z), v in lab.items():
        a[z - sb[2], y - sb[1], x - sb[0]] = v
    return a


def face_lists(nb, faces):
    """BC list per block id from (kind, dir, side, value) face BCs, in order."""
    out = {}
    for B in itertools.product(*[range(nb[d]) for d in range(3)]):
        b = B[0] + nb[0] * (B[1] + nb[1] * B[2])
        lst = [f for f in faces if B[f[1]] == (0 if f[2] == 0 else nb[f[1]] - 1)]
        if lst:
            out[b] = lst
    return out


def diffusion(g, bcl, h, nu_dt, T):
    """Laplacian and Euler update of every block in precision T."""
    n = g.n
    ih = [T(1.0 / (h[d] * h[d])) for d in range(3)]
    s = T(nu_dt)
    nbt = g.nb[0] * g.nb[1] * g.nb[2]
    lap = np.zeros(nbt * g.stride, dtype=T)
    un = np.zeros(nbt * g.stride, dtype=T)
    two = T(2)
    for B in itertools.product(*[range(g.nb[d]) for d in range(3)]):
        b = g.block_id(B)
        L = lab_of(g, B, (-1, -1, -1), (2, 2, 2), False, bcl.get(b, []))
        L = {k: T(v) for k, v in L.items()}
        for z, y, x in itertools.product(range(n[2]), range(n[1]), range(n[0])):
            c = L[(x, y, z)]
            ddx = ih[0] * ((L[(x + 1, y, z)] - two * c) + L[(x - 1, y, z)])
            ddy = ih[1] * ((L[(x, y + 1, z)] - two * c) + L[(x, y - 1, z)])
            ddz = ih[2] * ((L[(x, y, z + 1)] - two * c) + L[(x, y, z - 1)])
            i = b * g.stride + x + n[0] * (y + n[1] * z)
            lp = (ddx + ddy) + ddz
            lap[i] = lp
            un[i] = c + lp * s
    return lap, un


CASES = [
    # name, nblocks, cells, seed, face BCs, labs [(sb, se, tensorial)], diffusion?
    ("periodic", (2, 2, 2), (4, 4, 4), 1, [],
     [((-1, -1, -1), (2, 2, 2), False), ((-1, -1, -1), (2, 2, 2), True), ((-3, -3, -3), (4, 4, 4), False)], True),
    ("mixed", (2, 2, 2), (4, 4, 4), 2,
     [("absorbing", 1, 0, 0.0), ("dirichlet", 1, 1, 0.25), ("symmetry", 2, 0, -1.0), ("dirichlet", 2, 1, -0.5)],
     [((-1, -1, -1), (2, 2, 2), False), ((-1, -1, -1), (2, 2, 2), True), ((-3, -3, -3), (4, 4, 4), False),
      ((-3, -3, -3), (4, 4, 4), True)], True),
    ("order", (2, 1, 2), (4, 3, 5), 3,
     [("dirichlet", 0, 0, 1.0), ("absorbing", 2, 1, 0.0), ("symmetry", 0, 1, 1.0), ("dirichlet", 2, 1, -2.0),
      ("absorbing", 1, 0, 0.0)],
     [((-2, -1, -3), (3, 2, 4), True), ((-1, -2, 0), (2, 1, 3), False), ((-2, -2, -2), (3, 3, 3), False)], True),
    ("one_block", (1, 1, 1), (4, 4, 4), 4,
     [("dirichlet", 0, 0, -1.0), ("dirichlet", 0, 1, 2.0), ("symmetry", 2, 0, 1.0), ("absorbing", 2, 1, 0.0),
      ("base", 1, 0, 0.0)],
     [((-2, -2, -2), (3, 3, 3), True), ((-2, -2, -2), (3, 3, 3), False)], True),
    # node and face fields: upper-end blocks one element larger, shifted copies
    ("node_periodic", (2, 2, 2), (4, 4, 4), 5, [],
     [((-1, -1, -1), (2, 2, 2), False), ((-2, -2, -2), (3, 3, 3), True)], False, (1, 1, 1)),
    ("face_y_mixed", (2, 3, 1), (4, 3, 4), 6,
     [("absorbing", 0, 0, 0.0), ("dirichlet", 1, 1, 0.5), ("symmetry", 2, 0, -1.0)],
     [((-2, -2, -2), (3, 3, 3), True), ((-1, -1, -1), (2, 2, 2), False)], False, (0, 1, 0)),
    ("node_one_block", (1, 1, 1), (3, 3, 3), 7,
     [("dirichlet", 0, 1, 1.5), ("absorbing", 2, 0, 0.0)],
     [((-2, -2, -2), (3, 3, 3), True)], False, (1, 1, 1)),
]


def main():
    data, meta = {}, []
    for case in CASES:
        name, nb, n, seed, faces, labs, diff = case[:7]
        plus = case[7] if len(case) > 7 else (0, 0, 0)
        nbt = nb[0] * nb[1] * nb[2]
        vals = field_values(0x5EED0100 + seed, nbt, n, plus)
        g = Grid(nb, n, vals, plus)
        bcl = face_lists(nb, faces)
        data[name + "/input"] = vals
        entry = {"name": name, "nblocks": nb, "cells": n, "plus": plus, "faces": faces,
                 "bc_lists": {str(k): v for k, v in bcl.items()}, "labs": []}
        nmax = tuple(n[d] + plus[d] for d in range(3))  # labs of all blocks in the largest block's geometry
        for li, (sb, se, tens) in enumerate(labs):
            arr = np.stack([lab_array(lab_of(g, B, sb, se, tens, bcl.get(g.block_id(B), [])), nmax, sb, se)
                            for B in [(b % nb[0], (b // nb[0]) % nb[1], b // (nb[0] * nb[1])) for b in range(nbt)]])
            data["%s/lab%d" % (name, li)] = arr
            entry["labs"].append({"begin": sb, "end": se, "tensorial": tens})
        if diff:
            h = [1.0 / (nb[d] * n[d]) for d in range(3)]
            nu_dt = 0.1 * min(h) ** 2
            for T, tn in ((np.float64, "f64"), (np.float32, "f32")):
                lap, un = diffusion(g, bcl, h, nu_dt, T)
                data["%s/lap_%s" % (name, tn)] = lap
                data["%s/euler_%s" % (name, tn)] = un
            entry["h"], entry["nu_dt"] = h, nu_dt
        meta.append(entry)
    data["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **data)
    print("wrote %s (%d arrays)" % (OUT, len(data)))


if __name__ == "__main__":
    main()
