"""Host-side evaluation of the tile index arithmetic of the z-march kernels,
lane by lane, for every tile shape the kernels compile and for the FP32 WENO5
tile of 8-byte vectors that is not compiled.

The shape rules and the index expressions below restate those of
`X2Shape` / `k_stencil7x2` (cubismnova_amd/csrc/cnv_stencil7x2.hip) and of
`W5Shape` / `w5_march` (cubismnova_amd/csrc/cnv_march.h).  For every thread
of a tile they check that:
  * every LDS store falls inside the tile's LDS planes and no two threads
    store to the same element of one plane;
  * every LDS load reads an element that a store of the same plane wrote;
  * every global load and store stays inside its block plane (the two-step
    kernel: on a small periodic grid, every staged u^n value is the periodic
    neighbour's cell at that position);
  * the lanes whose x neighbours come from DPP row shifts have their source
    lane in the same 16-lane DPP row and in the same tile row.
CPU only: nothing here runs on a GPU."""
import itertools

import numpy as np
import pytest


def x2_shape(ts, N, WG):
    """X2Shape<T, N, WG> restated (cnv_stencil7x2.hip)."""
    VB = 16 if N * N * ts // 16 >= WG else 8
    VEC = VB // ts
    TX = N // VEC
    ROWS = WG // TX if TX else 0
    R = N // ROWS if ROWS else 0
    WPE = ts // 4
    PA0 = N + 2 * VEC
    if (PA0 * WPE) % 8 == 4:
        PA = PA0
    elif ((PA0 + VEC) * WPE) % 8 == 4:
        PA = PA0 + VEC
    else:
        PA = PA0 + 2 * VEC
    s = dict(VEC=VEC, TX=TX, ROWS=ROWS, R=R, PA=PA, PLA=(N + 4) * PA, PLB=(N + 2) * PA, NHX=4 * N,
             NHY=4 * TX, NR=4 * N)
    s["NH"] = s["NHX"] + s["NHY"] + 4
    s["HS"] = -(-s["NH"] // WG)
    s["RS"] = -(-s["NR"] // WG)
    s["valid"] = (N % VEC == 0 and ROWS > 0 and WG % TX == 0 and ROWS <= N and N % ROWS == 0 and VEC >= 2
                  and R <= 2 and s["HS"] <= 2 and TX <= 16 and 16 % TX == 0)
    return s


X2_SHAPES = [(ts, N, WG) for ts in (4, 8) for N in (16, 32) for WG in (64, 128, 256, 512)
             if x2_shape(ts, N, WG)["valid"]]


def test_x2_shape_list():
    """The shapes the kernel compiles, including both 8-byte-vector tiles
    (FP32 16^3 x 128 and 32^3 x 512)."""
    assert (4, 16, 128) in X2_SHAPES and (4, 32, 512) in X2_SHAPES
    assert x2_shape(4, 16, 128)["VEC"] == 2 and x2_shape(4, 32, 512)["VEC"] == 2
    assert x2_shape(4, 32, 256)["VEC"] == 4 and x2_shape(8, 32, 256)["VEC"] == 2
    assert len(X2_SHAPES) >= 8


@pytest.mark.parametrize("ts,N,WG", X2_SHAPES)
def test_x2_tile_indices(ts, N, WG):
    S = x2_shape(ts, N, WG)
    VEC, TX, ROWS, R, PA, PLA, PLB = (S[k] for k in ("VEC", "TX", "ROWS", "R", "PA", "PLA", "PLB"))
    NHX, NHY, NH, HS, NR, RS = (S[k] for k in ("NHX", "NHY", "NH", "HS", "NR", "RS"))
    NN = N * N
    nbx, nby = 3, 3
    bstride = N ** 3          # N^3 * sizeof(T) is a multiple of 32 bytes for these edges
    layer = nbx * nby * bstride

    def wrap(i, n):
        return i + n if i < 0 else (i - n if i >= n else i)

    # a field whose value is its global (x, y) position, for one z-layer of
    # blocks and one plane p of every block
    p = 5
    field = np.full(layer, -1, dtype=np.int64)
    for by, bx, y, x in itertools.product(range(nby), range(nbx), range(N), range(N)):
        field[(bx + nbx * by) * bstride + p * NN + x + N * y] = (by * N + y) * 1000 + (bx * N + x)

    for bx, by in ((0, 0), (1, 2), (2, 1)):
        xyself = (bx + nbx * by) * bstride
        lb = p * NN
        A = {}
        # registers -> LDS plane A (store_u0), from the global loads (load_u0)
        for tid in range(WG):
            tx, ty = tid % TX, tid // TX
            x0 = tx * VEC
            own = x0 + N * ty
            for r in range(R):
                g0 = lb + xyself + own + N * r * ROWS
                assert xyself + lb <= g0 and g0 + VEC <= xyself + lb + NN
                l0 = (ty + r * ROWS + 2) * PA + VEC + x0
                for v in range(VEC):
                    assert 0 <= l0 + v < PLA and (l0 + v) not in A
                    A[l0 + v] = field[g0 + v]
            for h in range(HS):
                s = tid + h * WG
                kind = 0
                if s < NHX:
                    side, lyr, y = s // (2 * N), (s // N) & 1, s % N
                    nb = (wrap(bx + (1 if side else -1), nbx) + nbx * by) * bstride
                    src = nb + (lyr if side else N - 1 - lyr) + N * y
                    dst = (y + 2) * PA + VEC + (N + lyr if side else -1 - lyr)
                    kind = 1
                elif s < NHX + NHY:
                    s2 = s - NHX
                    side, lyr, xs = s2 // (2 * TX), (s2 // TX) & 1, (s2 % TX) * VEC
                    nb = (bx + nbx * wrap(by + (1 if side else -1), nby)) * bstride
                    src = nb + xs + N * (lyr if side else N - 1 - lyr)
                    dst = ((N + lyr if side else -1 - lyr) + 2) * PA + VEC + xs
                    kind = 2
                elif s < NH:
                    d = s - NHX - NHY
                    sx, sy = d & 1, d >> 1
                    nb = (wrap(bx + (1 if sx else -1), nbx) + nbx * wrap(by + (1 if sy else -1), nby)) * bstride
                    src = nb + (0 if sx else N - 1) + N * (0 if sy else N - 1)
                    dst = ((N if sy else -1) + 2) * PA + VEC + (N if sx else -1)
                    kind = 1
                if kind:
                    n = VEC if kind == 2 else 1
                    assert 0 <= src and src + n <= layer
                    for v in range(n):
                        assert 0 <= dst + v < PLA and (dst + v) not in A
                        A[dst + v] = field[lb + src + v]

        def a_at(x, y):
            """the staged u^n value at tile position (x, y), checked against
            the periodic field"""
            i = (y + 2) * PA + VEC + x
            assert i in A, (x, y)
            gx, gy = (bx * N + x) % (nbx * N), (by * N + y) % (nby * N)
            assert A[i] == gy * 1000 + gx, (x, y, A[i])
            return i

        B = set()
        for tid in range(WG):
            tx, ty = tid % TX, tid // TX
            x0 = tx * VEC
            if tx:   # DPP row shift source: same 16-lane row, same tile row
                assert (tid - 1) // 16 == tid // 16 and (tid - 1) // TX == ty
            if tx < TX - 1:
                assert (tid + 1) // 16 == tid // 16 and (tid + 1) // TX == ty
            for r in range(R):
                y = ty + r * ROWS
                # step 1 reads: rows y -+ 1 (vectors), the row's halo cells
                for v in range(VEC):
                    a_at(x0 + v, y - 1), a_at(x0 + v, y + 1), a_at(x0 + v, y)
                a_at(-1, y), a_at(N, y)
                o = (y + 2) * PA + VEC + x0
                assert o - x0 - 1 == a_at(-1, y) and o - x0 + N == a_at(N, y)
                for v in range(VEC):
                    b = (y + 1) * PA + VEC + x0 + v
                    assert 0 <= b < PLB and b not in B
                    B.add(b)
        for tid in range(WG):
            for h in range(RS):
                s = tid + h * WG
                if s < NR:
                    if s < 2 * N:
                        x, y = (N if s // N else -1), s % N
                    else:
                        y, x = (N if (s - 2 * N) // N else -1), (s - 2 * N) % N
                    for dx, dy in ((0, 0), (-1, 0), (1, 0), (0, -1), (0, 1)):
                        a_at(x + dx, y + dy)
                    b = (y + 1) * PA + VEC + x
                    assert 0 <= b < PLB and b not in B
                    B.add(b)
        # step 2 reads of u^{n+1}: rows y -+ 1 and the row ends, all written
        for tid in range(WG):
            tx, ty = tid % TX, tid // TX
            x0 = tx * VEC
            for r in range(R):
                y = ty + r * ROWS
                o = (y + 1) * PA + VEC + x0
                for v in range(VEC):
                    assert o - PA + v in B and o + PA + v in B
                assert o - x0 - 1 in B and o - x0 + N in B
                # the output store stays in the block plane
                assert 0 <= x0 + N * y and x0 + N * y + VEC <= NN


def w5_shape(ts, NX, WG, vb=16):
    """W5Shape<T, NX, NX, WG> restated (cnv_march.h); vb = 8 gives the FP32
    tile of 8-byte vectors that is not compiled."""
    G = 2
    VEC = vb // ts
    TX = NX // VEC
    ROWS = WG // TX
    R = NX // ROWS if ROWS else 0
    P = NX + 2 * VEC
    s = dict(G=G, VEC=VEC, TX=TX, ROWS=ROWS, R=R, P=P, PLANE=(NX + 2 * G) * P, NHX=2 * G * NX, NHY=2 * G * TX)
    s["HX"] = -(-s["NHX"] // WG)
    s["HY"] = -(-s["NHY"] // WG)
    s["valid"] = NX % VEC == 0 and ROWS > 0 and WG % TX == 0 and ROWS <= NX and NX % ROWS == 0 and VEC >= G
    return s


W5_SHAPES = [(ts, NX, WG, 16) for ts in (4, 8) for NX in (16, 32) for WG in (64, 128, 256, 512)
             if w5_shape(ts, NX, WG)["valid"]] + [(4, 32, 512, 8), (4, 16, 128, 8)]


@pytest.mark.parametrize("ts,NX,WG,vb", W5_SHAPES)
def test_w5_tile_indices(ts, NX, WG, vb):
    """Every LDS store of a WENO5 plane in bounds and unique, every LDS load
    written, every global access inside its plane.  The last two shapes are
    the FP32 tiles of 8-byte vectors (512 threads on 32^3, 128 on 16^3): the
    512-thread one took an illegal-address fault in its first launch in an
    uncommitted build, and its index arithmetic as restated here stays in
    bounds, so the fault is not explained by these expressions."""
    S = w5_shape(ts, NX, WG, vb)
    assert S["valid"]
    G, VEC, TX, ROWS, R, P, PLANE = (S[k] for k in ("G", "VEC", "TX", "ROWS", "R", "P", "PLANE"))
    NHX, NHY, HX, HY = S["NHX"], S["NHY"], S["HX"], S["HY"]
    NY = NX
    L = {}

    def put(i, what):
        assert 0 <= i < PLANE and i not in L, (i, what)
        L[i] = what

    for tid in range(WG):
        tx, ty = tid % TX, tid // TX
        x0 = tx * VEC
        for r in range(R):
            y = ty + r * ROWS
            assert x0 + NX * y + VEC <= NX * NY          # centre load and the six state stores
            base = (y + G) * P + VEC + x0
            assert base % VEC == 0                       # vector-aligned LDS address
            for v in range(VEC):
                put(base + v, ("c", x0 + v, y))
        for h in range(HX):
            s = tid + h * WG
            if s < NHX:
                side, layer, y = s // (G * NY), 1 + (s // NY) % G, s % NY
                assert 0 <= NX * y < NX * NY
                x = NX - 1 + layer if side else -layer
                put((y + G) * P + (VEC + NX - 1 + layer if side else VEC - layer), ("hx", x, y))
        for h in range(HY):
            s = tid + h * WG
            if s < NHY:
                side, layer, x = s // (G * TX), 1 + (s // TX) % G, (s % TX) * VEC
                assert x + VEC <= NX
                row = NY - 1 + layer + G if side else G - layer
                assert (row * P + VEC + x) % VEC == 0          # vector-aligned LDS address
                for v in range(VEC):
                    put(row * P + VEC + x + v, ("hy", x + v, row - G))

    def read(x, y):
        i = (y + G) * P + VEC + x
        assert i in L, (x, y)
        assert L[i][1:] == (x, y), (x, y, L[i])

    for tid in range(WG):
        tx, ty = tid % TX, tid // TX
        x0 = tx * VEC
        for r in range(R):
            y = ty + r * ROWS
            for v in range(VEC):
                for dy in (-2, -1, 1, 2):
                    read(x0 + v, y + dy)
            for dx in (-2, -1, VEC, VEC + 1):
                read(x0 + dx, y)
