// cnv_jacobi2.hip -- two weighted-Jacobi sweeps per launch, u -> u' -> u'',
// on the stretched-z operator of cnv_jacobi.hip, for grids whose BC lists
// give every domain face one rule on all its blocks (periodic on both sides
// of an axis, or per side a Dirichlet constant, a Symmetry mirror or an
// Absorbing edge copy) and neighbour copies inside the domain.  On such a
// grid every value the two sweeps read is a grid cell (wrapped or mirrored,
// times the mirror's sign) or a face constant, so a tile reads u and f once,
// with a two-cell x/y halo of u, and writes u'' once: about half the HBM
// traffic of two single sweeps.
//
// Per z step (plane p, one barrier): u'(p) on the block's cells and, for
// planes of the tile, on the one-cell x/y ring around the block (the ring
// cells are the neighbour blocks' edge cells, which the second sweep reads);
// then u''(p - 1) on the block's cells from u'(p - 2), u'(p - 1), u'(p).
// Every value is the single sweep's expression on the same operands
// (cnv_jacobi.hip), so u'' is bit-identical to two calls of the single sweep.
#include "cnv_device.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>

namespace cnv
{

constexpr int kJ2ZMax = 64; // planes per tile at most

// Rule of a domain face as the two-sweep kernel applies it to the ghost
// cells the sweeps read past it: the periodic neighbour, a constant
// (Dirichlet, BC/Dirichlet.h:75-113), sign times the mirror cell (Symmetry,
// BC/Symmetry.h:69-114: ghost layer l reads interior cell l - 1 from the
// face), or the face's own cell (Absorbing, BC/Absorbing.h:66-105: every
// layer reads interior cell 0 from the face).
enum { FM_PERIODIC = 0, FM_CONST = 1, FM_MIRROR = 2, FM_EDGE = 3 };

struct J2Args {
    const char *u;  // grid base of u (component cu)
    const char *f;  // grid base of f (component cf)
    char *out;      // grid base of u'' (component co)
    Geom G;         // geometry (the three grids share it)
    long long cs_u, cs_f, cs_o; // component strides (elements)
    int cu, cf, co;
    int nzc, zchunk;
    int fm[3][2];    // domain face rule: FM_PERIODIC, FM_CONST, FM_MIRROR, FM_EDGE
    double fv[3][2]; // its constant (FM_CONST) or sign (FM_MIRROR)
    const double *czm, *czp;
    double ihx2, ihy2, omega;
};

// A tile is NX x TY cells of a block's plane (TY = NX: the whole plane; TY =
// NX / 2 or NX / 4: a half or a quarter of it, for less LDS per workgroup)
// times a chunk of z planes.
template <typename T, int NX, int TY, int WG>
struct J2Shape {
    static constexpr int VEC = 16 / (int)sizeof(T);
    static constexpr int TX = NX / VEC;
    static constexpr int ROWS = WG / TX;
    static constexpr int R = ROWS > 0 ? TY / ROWS : 1;
    static constexpr int PA = NX + 2 * VEC;       // row pitch of both LDS plane kinds
    static constexpr int PLA = (TY + 4) * PA;     // u: two halo layers in x and y
    static constexpr int PLB = (TY + 2) * PA;     // u': the one-cell ring
    static constexpr int NHT = 2 * TY + 4 * TX + 4; // threads holding halo cells by role
    // blocks of 16 cells per edge: one scalar halo cell per thread and slot
    // instead (one or two waves: the roles would diverge within a wave)
    static constexpr bool SCALAR = NX == 16;
    static constexpr int NR1 = 2 * NX + 2 * TY;        // ring cells per plane
    static constexpr int NH1 = 2 * NR1 + 4;            // halo cells of a u plane
    static constexpr int HS1 = (NH1 + WG - 1) / WG;    // scalar halo cells per thread
    static constexpr int HN = SCALAR ? HS1 : VEC;      // held halo values per thread
    static constexpr size_t lds = sizeof(T) * (size_t)(2 * PLA + 2 * PLB);
    // a halo role per thread at most; FP64 whole-plane tiles of more than 512
    // threads spill registers to scratch memory; LDS within one compute unit's
    static constexpr bool valid = (TY == NX || 2 * TY == NX || 4 * TY == NX) && NX % VEC == 0 && WG % TX == 0 && ROWS >= 1 &&
                                  ROWS <= TY && TY % ROWS == 0 &&
                                  R <= 8 && WG >= (SCALAR ? NR1 : NHT) && (sizeof(T) == 4 || WG <= 512 || TY < NX) &&
                                  lds + 2 * (kJ2ZMax + 2) * sizeof(double) <= 160 * 1024;
};

// MIR: some domain face is a mirror or edge face (the code for those rules
// is compiled into this variant only)
template <typename T, int NX, int TY, int WG, bool MIR>
__global__ __launch_bounds__(WG) void k_jacobi2(const J2Args a)
{
    using S = J2Shape<T, NX, TY, WG>;
    constexpr int NSPLIT = NX / TY; // tiles per block plane
    constexpr int VEC = S::VEC, TX = S::TX, ROWS = S::ROWS, R = S::R, PA = S::PA, PLA = S::PLA, PLB = S::PLB;
    static_assert(S::valid, "tile shape");
    using V = VecT<T, VEC>;
    constexpr bool SCALAR = S::SCALAR;
    constexpr int NR1 = S::NR1, NH1 = S::NH1, HS1 = S::HS1;
    using HV = VecT<T, S::HN>; // the held halo values of one plane
    constexpr long long SXY = (long long)NX * NX;

    __shared__ alignas(16) T sA[2 * PLA]; // u planes p (two slots)
    __shared__ alignas(16) T sB[2 * PLB]; // u' planes (two slots)
    __shared__ double czl[2][kJ2ZMax + 2];

    const int tid = threadIdx.x;
    // tiles in the order (block, z chunk, y part): the parts of one chunk
    // are neighbours in the launch and share their halo rows in L2
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const long long blk = tile / (a.nzc * NSPLIT);
    const int rem = tile - (int)(blk * (a.nzc * NSPLIT));
    const int zc = rem / NSPLIT;
    const int y0 = (rem - zc * NSPLIT) * TY; // first block row of the tile
    const int NZ = a.G.n[2];
    const int nbx = a.G.nb[0], nby = a.G.nb[1], nbz = a.G.nb[2];
    const int bx = (int)(blk % nbx), by = (int)((blk / nbx) % nby), bz = (int)(blk / ((long long)nbx * nby));
    const int z0 = zc * a.zchunk;
    const int z1 = min(NZ, z0 + a.zchunk);
    const int nzt = z1 - z0;
    const int NGX = nbx * NX, NGY = nby * NX, NGZ = nbz * NZ;
    const int per2 = a.fm[2][0] == FM_PERIODIC;
    const T *ubase = reinterpret_cast<const T *>(a.u) + (long long)a.cu * a.cs_u;
    const T *fbase = reinterpret_cast<const T *>(a.f) + (long long)a.cf * a.cs_f;
    T *obase = reinterpret_cast<T *>(a.out) + (long long)a.co * a.cs_o;
    const long long bstride = a.G.bstride;
    // global index g on axis d (at most two cells past the grid of n cells)
    // to the cell it reads: periodic wrap, mirror or edge cell (m = the
    // factor it is read with), or none (in = false; the face's constant cv)
    auto face_map = [&](int d, int g, int n, bool &in, T &cv, T &m) -> int {
        if (g >= 0 && g < n)
            return g;
        const int side = g < 0 ? 0 : 1;
        const int rule = a.fm[d][side];
        if (rule == FM_PERIODIC)
            return g < 0 ? g + n : g - n;
        if (rule == FM_CONST) {
            if (in)
                cv = (T)a.fv[d][side];
            in = false;
            return 0;
        }
        if constexpr (MIR) {
            if (rule == FM_MIRROR) {
                m = m * (T)a.fv[d][side];
                return g < 0 ? -1 - g : 2 * n - 1 - g;
            }
            return g < 0 ? 0 : n - 1;
        }
        return 0;
    };

    // global plane p of this block's column (at most two planes past the
    // grid) as the offset of its cells from a grid base, by the z faces'
    // rules: in false past a constant face (every cell of the plane reads
    // c), m the factor of a mirror face (uniform over the workgroup: one
    // division per plane, in scalar registers)
    struct ZPart {
        long long off;
        int g;  // bz * NZ + p (before the face rule)
        int lz; // plane within its block
        bool in;
        T c;
        T m;
    };
    const long long zblk = (long long)nbx * nby * bstride;
    auto zpart = [&](int p) -> ZPart {
        int gz = bz * NZ + p;
        ZPart r;
        r.off = 0;
        r.g = gz;
        r.lz = 0;
        r.in = true;
        r.c = (T)0;
        r.m = (T)1;
        gz = face_map(2, gz, NGZ, r.in, r.c, r.m);
        if (!r.in)
            return r;
        const int qz = gz / NZ, lz = gz - qz * NZ;
        r.off = qz * zblk + SXY * lz;
        r.lz = lz;
        return r;
    };
    // the plane after z, without a division where it stays inside the grid
    auto zadvance = [&](const ZPart &z) -> ZPart {
        const int g = z.g + 1;
        if (!z.in || g <= 0 || g >= NGZ)
            return zpart(g - bz * NZ);
        ZPart r = z;
        r.g = g;
        r.lz = z.lz + 1;
        r.off = z.off + SXY;
        if (r.lz == NZ) {
            r.lz = 0;
            r.off += zblk - NZ * SXY;
        }
        return r;
    };
    const long long colxy = ((long long)bx + (long long)nbx * by) * bstride;
    // base of block-local plane p of this block's column (the plane may lie
    // in the block below or above); nullptr past a Dirichlet z face
    auto plane_of = [&](const T *base, const ZPart &z) -> const T * {
        return z.in ? base + colxy + z.off : nullptr;
    };
    const int tx = tid % TX, ty = tid / TX;
    const int x0 = tx * VEC;
    auto load_cells = [&](const T *base, const ZPart &z, V (&Q)[R]) {
        const T *pl = plane_of(base, z);
        if (pl) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                Q[r] = *reinterpret_cast<const V *>(pl + x0 + NX * (y0 + ty + r * ROWS));
            if (MIR && z.m != (T)1) {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int v = 0; v < VEC; ++v)
                        Q[r].v[v] = Q[r].v[v] * z.m;
            }
        } else {
            const T c = z.c;
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                    Q[r].v[v] = c;
        }
    };
    auto load_f = [&](const ZPart &z, V (&Q)[R]) {
        const T *pl = plane_of(fbase, z);
        if (pl) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                Q[r] = load_nt<T, VEC>(pl + x0 + NX * (y0 + ty + r * ROWS));
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                    Q[r].v[v] = (T)0;
        }
    };

    // The halo of a u plane by thread: threads tid < 2 TY hold a pair of
    // cells of an x side (x = -2, -1, or NX, NX + 1, at row tid % TY), the
    // next 4 TX threads VEC cells of a row y = -2, -1, TY or TY + 1, the next
    // four a corner cell (x = -1 or NX, y = -1 or TY); coordinates from the
    // tile's first cell.  The cells next to the tile (x = -1, NX; y = -1,
    // TY) are the ring: their holders compute u' there.  Per thread: the
    // first cell (hx, hy), the x/y part of its offset in the grid (hoff:
    // periodic axes wrapped; the held cells are contiguous in one block) and,
    // for cells past a Dirichlet x or y face, that face's value (hin false,
    // hcv).
    constexpr int NXS = 2 * TY, NYS = 4 * TX;
    int role = 0; // 0 none, 1 x-side pair, 2 row vector, 3 corner
    int hx = 0, hy = 0;
    int nring = 0, ringv = 0; // ring cells held and the first one's index
    if (tid < NXS) {
        role = 1;
        const int k = tid / TY;
        hx = k == 0 ? -2 : NX;
        hy = tid - k * TY;
        nring = 1;
        ringv = k == 0 ? 1 : 0;
    } else if (tid < NXS + NYS) {
        role = 2;
        const int i = tid - NXS, k = i / TX;
        hx = (i - k * TX) * VEC;
        hy = k == 0 ? -2 : k == 1 ? -1 : k == 2 ? TY : TY + 1;
        nring = (k == 1 || k == 2) ? VEC : 0;
    } else if (tid < NXS + NYS + 4) {
        role = 3;
        const int d = tid - NXS - NYS;
        hx = (d & 1) ? NX : -1;
        hy = (d & 2) ? TY : -1;
    }
    // scalar halo (blocks of 16 cells per edge): slot s of the plane's halo,
    // layer 1 (the ring positions, s < NR1: the x sides, then the y sides),
    // layer 2 (s < 2 NR1), the four corner cells (s < NH1)
    int shx[HS1], shy[HS1];
    long long shoff[HS1];
    bool shin[HS1];
    T shcv[HS1], shm[HS1];
    // x/y part of the grid offset of the cell that tile cell (x, y) reads by
    // the x and y face rules, with its factor m; past a constant face, in =
    // false and cv = the face's value
    auto xy_offset = [&](int x, int y, bool &in, T &cv, T &m) -> long long {
        in = true;
        cv = (T)0;
        m = (T)1;
        const int gx = face_map(0, bx * NX + x, NGX, in, cv, m);
        const int gy = face_map(1, by * NX + y0 + y, NGY, in, cv, m);
        const int qx = gx / NX, lx = gx - qx * NX;
        const int qy = gy / NX, ly = gy - qy * NX;
        return ((long long)qx + (long long)nbx * qy) * bstride + lx + (long long)NX * ly;
    };
    if constexpr (SCALAR) {
#pragma unroll
        for (int h = 0; h < HS1; ++h) {
            const int s = tid + h * WG;
            int x = 0, y = 0;
            if (s < 2 * NR1) {
                const int layer = s < NR1 ? 1 : 2;
                const int i = s < NR1 ? s : s - NR1;
                if (i < 2 * TY) {
                    const int k = i / TY;
                    x = k == 0 ? -layer : NX - 1 + layer;
                    y = i - k * TY;
                } else {
                    const int k = (i - 2 * TY) / NX;
                    x = (i - 2 * TY) - k * NX;
                    y = k == 0 ? -layer : TY - 1 + layer;
                }
            } else {
                const int d = s - 2 * NR1;
                x = (d & 1) ? NX : -1;
                y = (d & 2) ? TY : -1;
            }
            shx[h] = x;
            shy[h] = y;
            shoff[h] = xy_offset(x, y, shin[h], shcv[h], shm[h]);
        }
    }
    bool hin;
    T hcv, hm;
    const long long hoff = xy_offset(hx, hy, hin, hcv, hm);
    // an x-side pair is one vector load unless a mirror face reverses it
    long long hoff2 = hoff + 1;
    if (role == 1) {
        bool in2;
        T cv2, m2;
        hoff2 = xy_offset(hx + 1, hy, in2, cv2, m2);
    }
    const bool pair_vec = !MIR || hoff2 == hoff + 1;
    // tile cell (x, y) past a mirror or edge face: its u' is written by the
    // thread that computes the cell it reads
    auto mirrored = [&](int x, int y) -> bool {
        const int gx = bx * NX + x, gy = by * NX + y0 + y;
        const bool ox = gx < 0 || gx >= NGX, oy = gy < 0 || gy >= NGY;
        return (ox && a.fm[0][gx < 0 ? 0 : 1] >= FM_MIRROR) || (oy && a.fm[1][gy < 0 ? 0 : 1] >= FM_MIRROR);
    };
    // the held halo cells of plane p; the loads read valid addresses for
    // every role, also past a Dirichlet face
    auto gather = [&](const ZPart &z, HV &H) {
        if constexpr (SCALAR) {
#pragma unroll
            for (int h = 0; h < HS1; ++h) {
                const T v = ubase[shoff[h] + z.off];
                const T vm = MIR ? v * shm[h] * z.m : v;
                H.v[h] = tid + h * WG < NH1 ? (shin[h] ? (z.in ? vm : z.c) : shcv[h]) : (T)0;
            }
        } else {
            const T *q = ubase + hoff + z.off;
            if (role == 1) {
                if (pair_vec) {
                    const VecT<T, 2> t = load_vec<T, 2>(q);
                    H.v[0] = t.v[0];
                    H.v[1] = t.v[1];
                } else {
                    H.v[0] = q[0];
                    H.v[1] = ubase[hoff2 + z.off];
                }
            } else if (role == 2) {
                H = load_vec<T, VEC>(q);
            } else if (role == 3) {
                H.v[0] = *q;
            }
#pragma unroll
            for (int v = 0; v < VEC; ++v)
                H.v[v] = hin ? (z.in ? (MIR ? H.v[v] * hm * z.m : H.v[v]) : z.c) : hcv;
        }
    };
    auto put_halo = [&](T *A, const HV &H) {
        if constexpr (SCALAR) {
#pragma unroll
            for (int h = 0; h < HS1; ++h)
                if (tid + h * WG < NH1)
                    A[(shy[h] + 2) * PA + VEC + shx[h]] = H.v[h];
        } else {
            T *o = A + (hy + 2) * PA + VEC + hx;
            if (role == 1) {
                o[0] = H.v[0];
                o[1] = H.v[1];
            } else if (role == 2) {
                *reinterpret_cast<V *>(o) = H;
            } else if (role == 3) {
                o[0] = H.v[0];
            }
        }
    };
    // past a constant face the ring's u' is that face's constant; past a
    // mirror or edge face the thread that computes the cell it reads
    // writes it (ring_skip)
    const bool ring_out = SCALAR ? !shin[0] : !hin;
    const bool ring_skip =
        MIR && (SCALAR ? mirrored(shx[0], shy[0]) : (role == 1 ? mirrored(hx + ringv, hy) : mirrored(hx, hy)));
    // the ghost u' this thread writes for a mirror or edge face next to
    // the tile: the factor (0: none) for the cell left of its first cell
    // (x0 = 0), right of its last (x0 = NX - VEC), and for the row below
    // row 0 / above row TY - 1 of its cells
    auto ghost_factor = [&](int d, int side, bool at_edge) -> T {
        if (!MIR || !at_edge || a.fm[d][side] < FM_MIRROR)
            return (T)0;
        return a.fm[d][side] == FM_MIRROR ? (T)a.fv[d][side] : (T)1;
    };
    const T gxl = ghost_factor(0, 0, bx == 0 && x0 == 0);
    const T gxr = ghost_factor(0, 1, bx == nbx - 1 && x0 + VEC == NX);
    const T gyl = ghost_factor(1, 0, by == 0 && y0 == 0);
    const T gyh = ghost_factor(1, 1, by == nby - 1 && y0 + TY == NX);
    // sweep two at the first / last plane of a z mirror or edge face reads
    // its ghost u' as this factor times u' of that plane
    const T gzl = ghost_factor(2, 0, true), gzh = ghost_factor(2, 1, true);
    // f at the held ring cells (planes of the tile only: z inside the grid)
    auto fetch_fring = [&](const ZPart &z, V &Fr) {
        if constexpr (SCALAR) {
            const T v = fbase[shoff[0] + z.off];
            Fr.v[0] = tid < NR1 ? v : (T)0;
        } else {
            const T *q = fbase + hoff + z.off;
            if (role == 1) {
                Fr.v[0] = q[ringv];
            } else if (nring > 0) {
                Fr = load_vec<T, VEC>(q);
            }
        }
    };

    // z metric of planes z0 - 1 .. z1 (step j holds plane z0 - 1 + j)
    for (int k = tid; k < 2 * (nzt + 2); k += WG) {
        const int j = k >> 1;
        int gz = bz * NZ + z0 - 1 + j;
        double v = 0.0;
        if (gz >= 0 && gz < NGZ) {
            v = (k & 1) ? a.czp[gz] : a.czm[gz];
        } else if (per2) {
            gz = gz < 0 ? gz + NGZ : gz - NGZ;
            v = (k & 1) ? a.czp[gz] : a.czm[gz];
        }
        czl[k & 1][j] = v;
    }

    const T ihx2 = (T)a.ihx2, ihy2 = (T)a.ihy2, w = (T)a.omega;
    // single-sweep update of one cell (cnv_jacobi.hip, MODE 0)
    auto upd = [&](T c, T xm, T xp, T ym, T yp, T zlo, T zhi, T fv, T cm, T cp, T wd) -> T {
        const T ax = ihx2 * (fms2(c, xp) + xm);
        const T ay = ihy2 * (fms2(c, yp) + ym);
        const T az = cp * (zhi - c) - cm * (c - zlo);
        const T au = ax + ay + az;
        const T rr = fv - au;
        return c + rr * wd;
    };

    V U0[R], U1[R], U2[R], U3[R]; // u at planes p - 1, p, p + 1, p + 2
    V P0[R], P1[R], P2[R];        // u' at planes p - 2, p - 1, p
    V F1[R], FN[R], FQ[R];        // f at planes p, p + 1, p - 1
    HV H0, H1, H2;                // held halo cells of u at planes p, p + 1, p + 2
    HV RZ;                        // held halo cells of u at plane p - 1
    V FR, FRN;                    // f at the held ring cells, planes p, p + 1
#pragma unroll
    for (int v = 0; v < S::HN; ++v)
        H0.v[v] = H1.v[v] = H2.v[v] = RZ.v[v] = (T)0;
#pragma unroll
    for (int v = 0; v < VEC; ++v)
        FR.v[v] = FRN.v[v] = (T)0;
    {
        const ZPart zm2 = zpart(z0 - 2), zm1 = zpart(z0 - 1), zc0 = zpart(z0);
        load_cells(ubase, zm2, U0);
        load_cells(ubase, zm1, U1);
        load_cells(ubase, zc0, U2);
        gather(zm1, H0);
        gather(zc0, H1);
        load_f(zm1, F1);
    }
    // planes p + 1 and p + 2 of the step
    ZPart za = zpart(z0), zb = zpart(z0 + 1);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        P0[r] = U0[r];
        P1[r] = U0[r];
        FQ[r] = F1[r];
    }

    for (int j = 0; j < nzt + 2; ++j) {
        const int p = z0 - 1 + j;
        prio_ramp(j, nzt + 2);
        T *A = sA + (j & 1) * PLA;
        T *B = sB + (j & 1) * PLB;
        const T *Bq = sB + ((j + 1) & 1) * PLB; // plane p - 1, written in the step before
        // u at plane p into LDS: the block's cells and the halo
#pragma unroll
        for (int r = 0; r < R; ++r)
            *reinterpret_cast<V *>(A + (ty + r * ROWS + 2) * PA + VEC + x0) = U1[r];
        put_halo(A, H0);
        // the next planes' loads (f first: the nearest use); planes past
        // those the tile reads (z0 - 2 .. z1 + 1) are not loaded
        if (j <= nzt) {
            load_f(za, FN);
            if (j < nzt) {
                fetch_fring(za, FRN);
                gather(zb, H2);
            }
            load_cells(ubase, zb, U3);
        }
        __syncthreads();

        // first sweep at plane p
        const int gzp = bz * NZ + p;
        const bool ghost = !per2 && (gzp < 0 || gzp >= NGZ);
        if (ghost) {
            // (past a mirror or edge face the second sweep does not read it)
            const int side = gzp < 0 ? 0 : 1;
            const T c = a.fm[2][side] == FM_CONST ? (T)a.fv[2][side] : (T)0;
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                    P2[r].v[v] = c;
        } else {
            const T cm = (T)czl[0][j], cp = (T)czl[1][j];
            const T diag = -(2 * ihx2 + 2 * ihy2 + cm + cp);
            const T wd = w / diag;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int y = ty + r * ROWS;
                const T *row = A + (y + 2) * PA + VEC;
                const V ym = *reinterpret_cast<const V *>(row - PA + x0);
                const V yp = *reinterpret_cast<const V *>(row + PA + x0);
                const T xl = row[x0 - 1];
                const T xr = row[x0 + VEC];
                V res;
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    const T c = U1[r].v[v];
                    const T xm = v == 0 ? xl : U1[r].v[v - 1];
                    const T xp = v == VEC - 1 ? xr : U1[r].v[v + 1];
                    res.v[v] = upd(c, xm, xp, ym.v[v], yp.v[v], U0[r].v[v], U2[r].v[v], F1[r].v[v], cm, cp, wd);
                }
                P2[r] = res;
                *reinterpret_cast<V *>(B + (y + 1) * PA + VEC + x0) = res;
                // ghost u' past a mirror or edge face, for the tile's planes
                if (MIR && j >= 1 && j <= nzt) {
                    if (gxl != (T)0)
                        B[(y + 1) * PA + VEC - 1] = res.v[0] * gxl;
                    if (gxr != (T)0)
                        B[(y + 1) * PA + VEC + NX] = res.v[VEC - 1] * gxr;
                    if ((y == 0 && gyl != (T)0) || (y == TY - 1 && gyh != (T)0)) {
                        const T gf = y == 0 ? gyl : gyh;
                        V gv;
#pragma unroll
                        for (int v = 0; v < VEC; ++v)
                            gv.v[v] = res.v[v] * gf;
                        *reinterpret_cast<V *>(B + (y == 0 ? 0 : TY + 1) * PA + VEC + x0) = gv;
                    }
                }
            }
            // the ring, for planes of the tile
            if constexpr (SCALAR) {
                // the ring cell of thread tid < NR1 is its halo slot 0
                if (j >= 1 && j <= nzt && tid < NR1 && !ring_skip) {
                    const int rx = shx[0], ry = shy[0];
                    const T *o = A + (ry + 2) * PA + VEC + rx;
                    B[(ry + 1) * PA + VEC + rx] =
                        ring_out ? shcv[0]
                                 : upd(H0.v[0], o[-1], o[1], o[-PA], o[PA], RZ.v[0], H1.v[0], FR.v[0], cm, cp, wd);
                }
            } else if (j >= 1 && j <= nzt && nring > 0 && !ring_skip) {
                const T *o = A + (hy + 2) * PA + VEC + hx;
                T *ob = B + (hy + 1) * PA + VEC + hx;
                if (role == 1) {
                    // (element ringv of the pair, selected by value: a
                    // runtime index into registers would go through scratch)
                    const int v = ringv;
                    const T c = v ? H0.v[1] : H0.v[0];
                    const T zlo = v ? RZ.v[1] : RZ.v[0];
                    const T zhi = v ? H1.v[1] : H1.v[0];
                    ob[v] = ring_out ? hcv
                                     : upd(c, o[v - 1], o[v + 1], o[v - PA], o[v + PA], zlo, zhi, FR.v[0], cm, cp, wd);
                } else {
                    V res;
#pragma unroll
                    for (int v = 0; v < VEC; ++v)
                        res.v[v] = ring_out ? hcv
                                            : upd(H0.v[v], o[v - 1], o[v + 1], o[v - PA], o[v + PA], RZ.v[v],
                                                  H1.v[v], FR.v[v], cm, cp, wd);
                    *reinterpret_cast<V *>(ob) = res;
                }
            }
        }

        // second sweep at plane p - 1 (a plane of the tile)
        if (j >= 2) {
            const T cm = (T)czl[0][j - 1], cp = (T)czl[1][j - 1];
            const T diag = -(2 * ihx2 + 2 * ihy2 + cm + cp);
            const T wd = w / diag;
            const int q = p - 1;
            T *oq = obase + blk * bstride + SXY * q;
            const int gq = bz * NZ + q;
            const T fzl = gq == 0 ? gzl : (T)0, fzh = gq == NGZ - 1 ? gzh : (T)0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int y = ty + r * ROWS;
                const T *row = Bq + (y + 1) * PA + VEC;
                const V ym = *reinterpret_cast<const V *>(row - PA + x0);
                const V yp = *reinterpret_cast<const V *>(row + PA + x0);
                const T xl = row[x0 - 1];
                const T xr = row[x0 + VEC];
                V res;
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    const T c = P1[r].v[v];
                    const T xm = v == 0 ? xl : P1[r].v[v - 1];
                    const T xp = v == VEC - 1 ? xr : P1[r].v[v + 1];
                    const T zlo = (MIR && fzl != (T)0) ? c * fzl : P0[r].v[v];
                    const T zhi = (MIR && fzh != (T)0) ? c * fzh : P2[r].v[v];
                    res.v[v] = upd(c, xm, xp, ym.v[v], yp.v[v], zlo, zhi, FQ[r].v[v], cm, cp, wd);
                }
                store_nt<T, VEC>(oq + x0 + NX * (y0 + y), res);
            }
        }

        // shift the z windows by one plane
        za = zb;
        zb = zadvance(zb);
        RZ = H0;
        H0 = H1;
        H1 = H2;
        FR = FRN;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            U0[r] = U1[r];
            U1[r] = U2[r];
            U2[r] = U3[r];
            P0[r] = P1[r];
            P1[r] = P2[r];
            FQ[r] = F1[r];
            F1[r] = FN[r];
        }
    }
}

// ---------------------------------------------------------------- host side

struct J2Tune {
    int wg;
    int zchunk;
    int split; // tiles per block plane: 1 (whole planes), 2 or 4 (parts in y)
};

template <typename T, int NX, int TY, int WG>
constexpr bool j2_valid()
{
    return J2Shape<T, NX, TY, WG>::valid;
}

template <typename T, int NX, int TY, int WG>
static bool j2_launch(const J2Args &a, int ntiles, hipStream_t st)
{
    if constexpr (j2_valid<T, NX, TY, WG>()) {
        bool mir = false;
        for (int d = 0; d < 3; ++d)
            for (int s = 0; s < 2; ++s)
                mir = mir || a.fm[d][s] >= FM_MIRROR;
        if (mir)
            hipLaunchKernelGGL((k_jacobi2<T, NX, TY, WG, true>), dim3(ntiles), dim3(WG), 0, st, a);
        else
            hipLaunchKernelGGL((k_jacobi2<T, NX, TY, WG, false>), dim3(ntiles), dim3(WG), 0, st, a);
        return true;
    }
    return false;
}

// Calls f(integral_constant<NX>, integral_constant<TY>, integral_constant<WG>)
// for the tile (edge, split, wg); false when there is none.
template <typename F>
static bool j2_with_tile(int edge, int split, int wg, F &&f)
{
    auto by_wg = [&](auto nx, auto ty) -> bool {
        switch (wg) {
        case 64:
            return f(nx, ty, std::integral_constant<int, 64>{});
        case 128:
            return f(nx, ty, std::integral_constant<int, 128>{});
        case 256:
            return f(nx, ty, std::integral_constant<int, 256>{});
        case 512:
            return f(nx, ty, std::integral_constant<int, 512>{});
        case 1024:
            return f(nx, ty, std::integral_constant<int, 1024>{});
        default:
            return false;
        }
    };
    auto by_split = [&](auto nx) -> bool {
        constexpr int NX = decltype(nx)::value;
        if (split == 1)
            return by_wg(nx, std::integral_constant<int, NX>{});
        if (split == 2)
            return by_wg(nx, std::integral_constant<int, NX / 2>{});
        if constexpr (NX >= 32) {
            if (split == 4)
                return by_wg(nx, std::integral_constant<int, NX / 4>{});
        }
        return false;
    };
    switch (edge) {
    case 16:
        return by_split(std::integral_constant<int, 16>{});
    case 32:
        return by_split(std::integral_constant<int, 32>{});
    case 64:
        return by_split(std::integral_constant<int, 64>{});
    default:
        return false;
    }
}

template <typename T>
static bool j2_dispatch(const J2Args &a, int edge, int split, int wg, int ntiles, hipStream_t st)
{
    return j2_with_tile(edge, split, wg, [&](auto nx, auto ty, auto w) {
        return j2_launch<T, decltype(nx)::value, decltype(ty)::value, decltype(w)::value>(a, ntiles, st);
    });
}

static bool j2_tile_ok(int dtype, int edge, int split, int wg)
{
    return j2_with_tile(edge, split, wg, [&](auto nx, auto ty, auto w) {
        constexpr int NX = decltype(nx)::value, TY = decltype(ty)::value, WG = decltype(w)::value;
        return dtype == CNV_F64 ? j2_valid<double, NX, TY, WG>() : j2_valid<float, NX, TY, WG>();
    });
}

static J2Tune j2_default(int dtype, int edge)
{
    if (edge >= 64)
        return dtype == CNV_F64 ? J2Tune{256, 64, 2} : J2Tune{256, 64, 1};
    if (edge >= 32)
        return {256, 32, 1};
    return {dtype == CNV_F64 ? 128 : 64, 16, 1};
}

static std::mutex g_j2_mtx;
static std::map<std::pair<int, int>, J2Tune> g_j2_tune;

static J2Tune j2_tune(int dtype, int edge)
{
    std::lock_guard<std::mutex> lk(g_j2_mtx);
    auto it = g_j2_tune.find({dtype, edge});
    return it == g_j2_tune.end() ? j2_default(dtype, edge) : it->second;
}

int jacobi2_set_tune(int dtype, int edge, int wg, int zchunk, int split)
{
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    if (zchunk < 1)
        return fail(CNV_ERR_INVALID, "zchunk must be >= 1");
    if (!j2_tile_ok(dtype, edge, split, wg))
        return fail(CNV_ERR_UNSUPPORTED,
                    "no two-sweep Jacobi tile of this workgroup size and split for this dtype and block edge");
    std::lock_guard<std::mutex> lk(g_j2_mtx);
    g_j2_tune[{dtype, edge}] = J2Tune{wg, zchunk, split};
    return CNV_OK;
}

int jacobi2_get_tune(int dtype, int edge, int *wg, int *zchunk, int *split)
{
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    const J2Tune t = j2_tune(dtype, edge);
    if (wg)
        *wg = t.wg;
    if (zchunk)
        *zchunk = t.zchunk;
    if (split)
        *split = t.split;
    return CNV_OK;
}

// The grid's domain faces as the two-sweep kernel needs them (component
// comp): every block's BC list resolved for the sweep's stencil (-1, 2)
// gives neighbour copies on every face inside the domain and, on each domain
// face, the same rule on every block of that face: a periodic neighbour on
// both sides of an axis, or on each side a Dirichlet constant, a Symmetry
// mirror of nonzero sign, or an Absorbing edge copy.  Fills fm and fv;
// false otherwise.
bool jacobi2_faces(cnv_grid_s *g, int comp, int fm[3][2], double fv[3][2])
{
    const int *nb = g->desc.nblocks;
    std::lock_guard<std::mutex> lk(g->mtx);
    const auto &lists = g->bcs[comp];
    const int sb[3] = {-1, -1, -1}, se[3] = {2, 2, 2};
    int kind0[3][2];
    double val0[3][2];
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s) {
            kind0[d][s] = -1;
            val0[d][s] = 0.0;
        }
    for (long long b = 0; b < g->nblocks_total; ++b) {
        const int bi[3] = {(int)(b % nb[0]), (int)((b / nb[0]) % nb[1]), (int)(b / ((long long)nb[0] * nb[1]))};
        std::vector<DevBC> tmp;
        for (const auto &x : lists[(size_t)b])
            tmp.push_back({x.kind, x.dir, x.side, 0, x.value});
        int kind[6];
        double val[6];
        resolve_faces(tmp.data(), (int)tmp.size(), sb, se, kind, val);
        for (int f = 0; f < 6; ++f) {
            const int d = f >> 1, s = f & 1;
            const bool edge = bi[d] == (s == 0 ? 0 : nb[d] - 1);
            if (!edge) {
                if (kind[f] != FACE_NBR)
                    return false;
                continue;
            }
            if (kind[f] == FACE_UNSET || (kind[f] == FACE_SYMMETRY && val[f] == 0.0))
                return false;
            const double v = (kind[f] == FACE_DIRICHLET || kind[f] == FACE_SYMMETRY) ? val[f] : 0.0;
            if (kind0[d][s] < 0) {
                kind0[d][s] = kind[f];
                val0[d][s] = v;
            } else if (kind[f] != kind0[d][s] || v != val0[d][s]) {
                return false;
            }
        }
    }
    for (int d = 0; d < 3; ++d) {
        const bool p0 = kind0[d][0] == FACE_NBR, p1 = kind0[d][1] == FACE_NBR;
        if (p0 != p1)
            return false;
        for (int s = 0; s < 2; ++s) {
            const int k = kind0[d][s];
            fm[d][s] = k == FACE_NBR ? FM_PERIODIC : k == FACE_DIRICHLET ? FM_CONST : k == FACE_SYMMETRY ? FM_MIRROR
                                                                                                        : FM_EDGE;
            fv[d][s] = val0[d][s];
        }
    }
    // a periodic or mirror z reads planes two cells past the grid from
    // inside it with one addition
    const int ngz = nb[2] * g->desc.block_cells[2];
    if (ngz < 2 && (fm[2][0] == FM_PERIODIC || fm[2][0] == FM_MIRROR || fm[2][1] == FM_MIRROR))
        return false;
    return true;
}

// Whether the two-sweep kernel runs u -> out on these grids (u, f, out of
// one dtype and shape, x/y block edge 16, 32 or 64, faces as above).
bool jacobi2_applies(cnv_grid_s *u, int cu, cnv_grid_s *f, cnv_grid_s *out)
{
    const int *n = u->desc.block_cells;
    if (!(n[0] == n[1] && (n[0] == 16 || n[0] == 32 || n[0] == 64)))
        return false;
    if (f->desc.dtype != u->desc.dtype || out->desc.dtype != u->desc.dtype)
        return false;
    if (u->desc.entity != CNV_CELL)
        return false;
    int fm[3][2];
    double fv[3][2];
    return jacobi2_faces(u, cu, fm, fv);
}

int launch_jacobi2(cnv_grid_s *u, int cu, cnv_grid_s *f, int cf, cnv_grid_s *out, int co, double omega,
                   hipStream_t st, const int *tile)
{
    const int *n = u->desc.block_cells;
    if (!(n[0] == n[1] && (n[0] == 16 || n[0] == 32 || n[0] == 64)))
        return fail(CNV_ERR_UNSUPPORTED, "two-sweep Jacobi needs x/y block edges of 16, 32 or 64");
    if (f->desc.dtype != u->desc.dtype || out->desc.dtype != u->desc.dtype)
        return fail(CNV_ERR_SIZE, "u, f and out must share a dtype");
    J2Args a;
    if (!jacobi2_faces(u, cu, a.fm, a.fv))
        return fail(CNV_ERR_UNSUPPORTED, "two-sweep Jacobi needs every domain face periodic, Dirichlet, Symmetry "
                                         "(nonzero sign) or Absorbing, the same on every block of the face");
    {
        std::lock_guard<std::mutex> lk(u->mtx);
        int rc = ensure_zmetric(u);
        if (rc != CNV_OK)
            return rc;
    }
    a.u = static_cast<const char *>(u->data);
    a.f = static_cast<const char *>(f->data);
    a.out = static_cast<char *>(out->data);
    a.G = geom_of(u);
    a.cs_u = (long long)(u->component_bytes / u->elem);
    a.cs_f = (long long)(f->component_bytes / f->elem);
    a.cs_o = (long long)(out->component_bytes / out->elem);
    a.cu = cu;
    a.cf = cf;
    a.co = co;
    a.czm = u->czm_dev;
    a.czp = u->czp_dev;
    a.ihx2 = 1. / (u->h[0] * u->h[0]);
    a.ihy2 = 1. / (u->h[1] * u->h[1]);
    a.omega = omega;
    const J2Tune tu = tile ? J2Tune{tile[0], tile[1], tile[2]} : j2_tune(u->desc.dtype, n[0]);
    a.zchunk = std::min(std::min(tu.zchunk, n[2]), kJ2ZMax);
    a.nzc = (n[2] + a.zchunk - 1) / a.zchunk;
    const long long ntiles = u->nblocks_total * a.nzc * tu.split;
    if (ntiles >= (1LL << 31))
        return fail(CNV_ERR_UNSUPPORTED, "grid too large for the two-sweep Jacobi kernel");
    const bool ok = u->desc.dtype == CNV_F64 ? j2_dispatch<double>(a, n[0], tu.split, tu.wg, (int)ntiles, st)
                                             : j2_dispatch<float>(a, n[0], tu.split, tu.wg, (int)ntiles, st);
    if (!ok)
        return fail(CNV_ERR_UNSUPPORTED, "no two-sweep Jacobi tile for this block edge and workgroup size");
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

} // namespace cnv
