// cnv_stencil7x2.hip -- two forward-Euler diffusion steps per launch on
// periodic grids (every face of every block a neighbour copy, no BC list).
//
// A workgroup owns the x-y planes [z0, z1) of one block of one component.
// It stages the planes of u^n it needs, each with a two-cell x/y halo from
// the neighbour blocks (periodic wrap, the four diagonal cells included), in
// a rolling set of three LDS planes; evaluates u^{n+1} on the block's cells
// and on the one-cell ring around them (the face-neighbour cells that the
// second step reads), and u^{n+2} on the block's cells.  Each cell of
// u^{n+1} and u^{n+2} is computed with the operation order of the one-step
// kernel (Laplacian.cpp:125-129, then the two Field.h:457-512 roundings), so
// the result is bit-identical to two launches of it, at one HBM read and one
// write for two steps.  The FP64 sums of squares of both steps are produced
// as per-tile partials.
#include "cnv_march.h"

#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

namespace cnv
{

struct X2Args {
    const char *u; // input grid base
    char *out;     // output grid base
    Geom G;
    int nzc, tiles_per_comp; // z-chunks per block, tiles per component
    double ih2[3];
    double nudt;
    double *part1, *part2; // per-tile partials of |u^{n+1}|^2, |u^{n+2}|^2 (or null)
    // deferred sums of the previous launch of a run (as in the one-step kernel)
    const double *prev_p[2];
    int prev_n[2];
    double *prev_out[2];
};

template <typename T, int N, int WG>
struct X2Shape {
    // 16-byte vectors per thread, or 8-byte ones when a plane has fewer
    // 16-byte vectors than the workgroup has threads (more waves per tile)
    static constexpr int VB = N * N * (int)sizeof(T) / 16 >= WG ? 16 : 8;
    static constexpr int VEC = VB / (int)sizeof(T);
    static constexpr int TX = N / VEC;
    static constexpr int ROWS = WG / TX;
    static constexpr int R = N / ROWS;
    // row pitch: x = -VEC .. N+VEC-1 plus padding so that a column of 32
    // rows spreads over 8 of the 32 LDS banks (pitch in 4-byte words = 4 mod 8)
    static constexpr int WPE = (int)sizeof(T) / 4;
    static constexpr int PA0 = N + 2 * VEC;
    static constexpr int PA = (PA0 * WPE) % 8 == 4 ? PA0 : ((PA0 + VEC) * WPE) % 8 == 4 ? PA0 + VEC : PA0 + 2 * VEC;
    static constexpr int PLA = (N + 4) * PA; // u^n plane, y = -2 .. N+1
    static constexpr int PLB = (N + 2) * PA; // u^{n+1} plane, y = -1 .. N
    static constexpr int NHX = 4 * N;        // x halo: 2 layers x 2 sides x N rows (scalars)
    static constexpr int NHY = 4 * TX;       // y halo: 2 layers x 2 sides x TX vectors
    static constexpr int NH = NHX + NHY + 4; // + the four diagonal cells
    static constexpr int HS = (NH + WG - 1) / WG;
    static constexpr int NR = 4 * N; // ring cells of u^{n+1} per plane
    static constexpr int RS = (NR + WG - 1) / WG;
    // shapes with more than two rows per thread run out of registers
    // rows of TX lanes must not straddle the 16-lane rows of DPP lane shifts
    static constexpr bool valid = N % VEC == 0 && WG % TX == 0 && ROWS <= N && N % ROWS == 0 && VEC >= 2 &&
                                  R <= 2 && HS <= 2 && TX <= 16 && 16 % TX == 0;
};

// Value held by lane i-1 (CTRL = row_shr:1) or lane i+1 (row_shl:1) within a
// 16-lane DPP row; the lane with no source in the row (lane 0 for row_shr,
// lane 15 for row_shl) gets `old`, so every lane's result is defined.
template <typename T, int CTRL>
__device__ inline T dpp_shift(T old, T v)
{
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
    } else {
        const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
        const int hi =
            __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
        return __builtin_bit_cast(T, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    }
}
constexpr int kDppRowShr1 = 0x111; // lane i <- lane i-1
constexpr int kDppRowShl1 = 0x101; // lane i <- lane i+1

// ZC planes per tile: the ZC + 2 planes of u^{n+1} run as a static rotation
// of the three register slots, so every slot index is a constant.
template <typename T, int N, int WG, int ZC>
__global__ __launch_bounds__(WG) void k_stencil7x2(const X2Args a)
{
    using S = X2Shape<T, N, WG>;
    static_assert(N % ZC == 0, "z chunk");
    constexpr int VEC = S::VEC, TX = S::TX, ROWS = S::ROWS, R = S::R, PA = S::PA, PLA = S::PLA, PLB = S::PLB;
    constexpr int NHX = S::NHX, NHY = S::NHY, NH = S::NH, HS = S::HS, NR = S::NR, RS = S::RS;
    static_assert(S::valid, "tile shape");
    using V = VecT<T, VEC>;
    constexpr long long NN = (long long)N * N;

    __shared__ alignas(16) T sA[3 * PLA];
    __shared__ alignas(16) T sB[3 * PLB];
    __shared__ double red[WG / 64];

    const int tid = threadIdx.x;
    if (blockIdx.x == 0)
        for (int k = 0; k < 2; ++k)
            if (a.prev_p[k])
                wg_sum_partials<WG>(a.prev_p[k], a.prev_n[k], a.prev_out[k], red);

    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int comp = tile / a.tiles_per_comp;
    const int tt = tile - comp * a.tiles_per_comp;
    const long long blk = tt / a.nzc;
    const int zc = tt - (int)blk * a.nzc;
    const int z0 = zc * ZC;
    const int z1 = z0 + ZC;
    const int nbx = a.G.nb[0], nby = a.G.nb[1], nbz = a.G.nb[2];
    const int bx = (int)(blk % nbx), by = (int)((blk / nbx) % nby), bz = (int)(blk / ((long long)nbx * nby));
    const T *cbase = reinterpret_cast<const T *>(a.u) + (long long)comp * a.G.cstride;
    T *obase = reinterpret_cast<T *>(a.out) + (long long)comp * a.G.cstride;
    const long long bstride = a.G.bstride;
    // buffer resources over the input and output component slabs (fewer
    // than 2^31 bytes, host-checked): scalar bases, 32-bit lane offsets
    constexpr unsigned TS = sizeof(T);
    const auto rin = buf_rsrc(cbase, (unsigned)(a.G.cstride * TS));
    const auto rout = buf_rsrc(obase, (unsigned)(a.G.cstride * TS));
    const unsigned oblk = (unsigned)(blk * bstride * TS); // this block in the output slab

    auto wrap = [](int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); };
    const long long layer = (long long)nbx * nby * bstride; // elements per z-layer of blocks
    const long long xyself = (bx + (long long)nbx * by) * bstride;

    const int tx = tid % TX, ty = tid / TX;
    const int x0 = tx * VEC;

    // Fixed per-thread roles for the halo of a u^n plane, resolved once: the
    // element offset of the source within its z-layer of blocks and plane
    // (periodic neighbour blocks in x and y), the LDS offset, and the kind
    // (0 none, 1 scalar, 2 vector).  A plane p then only adds its layer and
    // plane offset.
    unsigned hsrc[HS]; // byte offset within the source plane's z-layer of blocks
    int hlds[HS], hkind[HS];
#pragma unroll
    for (int h = 0; h < HS; ++h) {
        const int s = tid + h * WG;
        hkind[h] = 0;
        hsrc[h] = 0;
        hlds[h] = 0;
        if (s < NHX) {
            const int side = s / (2 * N), lyr = (s / N) & 1, y = s % N;
            const long long nb = (wrap(bx + (side ? 1 : -1), nbx) + (long long)nbx * by) * bstride;
            hsrc[h] = (unsigned)((nb + (side ? lyr : N - 1 - lyr) + N * y) * TS);
            hlds[h] = (y + 2) * PA + VEC + (side ? N + lyr : -1 - lyr);
            hkind[h] = 1;
        } else if (s < NHX + NHY) {
            const int s2 = s - NHX;
            const int side = s2 / (2 * TX), lyr = (s2 / TX) & 1, xs = (s2 % TX) * VEC;
            const long long nb = (bx + (long long)nbx * wrap(by + (side ? 1 : -1), nby)) * bstride;
            hsrc[h] = (unsigned)((nb + xs + N * (side ? lyr : N - 1 - lyr)) * TS);
            hlds[h] = ((side ? N + lyr : -1 - lyr) + 2) * PA + VEC + xs;
            hkind[h] = 2;
        } else if (s < NH) {
            const int d = s - NHX - NHY, sx = d & 1, sy = d >> 1;
            const long long nb =
                (wrap(bx + (sx ? 1 : -1), nbx) + (long long)nbx * wrap(by + (sy ? 1 : -1), nby)) * bstride;
            hsrc[h] = (unsigned)((nb + (sx ? 0 : N - 1) + N * (sy ? 0 : N - 1)) * TS);
            hlds[h] = ((sy ? N : -1) + 2) * PA + VEC + (sx ? N : -1);
            hkind[h] = 1;
        }
    }
    const int own = x0 + N * ty; // this thread's first row offset within a plane

    struct U0 {
        V c[R];
        V h[HS];
    };
    // plane p (z0-2 <= p <= z1+1, possibly in the z-neighbour block) into registers
    auto load_u0 = [&](int p) -> U0 {
        U0 Q; // scalar halo entries use element 0 only
        const int dz = p < 0 ? -1 : (p >= N ? 1 : 0);
        // byte offset of plane p of this block row's z-layer (uniform)
        const unsigned lb = (unsigned)(((long long)wrap(bz + dz, nbz) * layer + (long long)(p - dz * N) * NN) * TS);
        const unsigned sb = lb + (unsigned)(xyself * TS);
#pragma unroll
        for (int r = 0; r < R; ++r)
            Q.c[r] = buf_load<T, VEC>(rin, (unsigned)((own + N * r * ROWS) * TS), sb);
#pragma unroll
        for (int h = 0; h < HS; ++h) {
            if (hkind[h] == 2) {
                Q.h[h] = buf_load<T, VEC>(rin, hsrc[h], lb);
            } else if (hkind[h] == 1) {
                Q.h[h].v[0] = buf_load1<T>(rin, hsrc[h], lb);
            }
        }
        return Q;
    };
    // registers -> LDS plane A, A(x, y) = A[(y + 2) * PA + VEC + x]
    auto store_u0 = [&](const U0 &Q, T *A) {
#pragma unroll
        for (int r = 0; r < R; ++r)
            *reinterpret_cast<V *>(A + (ty + r * ROWS + 2) * PA + VEC + x0) = Q.c[r];
#pragma unroll
        for (int h = 0; h < HS; ++h) {
            if (hkind[h] == 2)
                *reinterpret_cast<V *>(A + hlds[h]) = Q.h[h];
            else if (hkind[h] == 1)
                A[hlds[h]] = Q.h[h].v[0];
        }
    };

    const T ihx2 = (T)a.ih2[0], ihy2 = (T)a.ih2[1], ihz2 = (T)a.ih2[2];
    const T nudt = (T)a.nudt;
    // one Euler step of one cell, in the order of the one-step kernel
    auto euler = [&](T c, T xm, T xp, T ym, T yp, T zm, T zp) -> T {
        const T ddx = ihx2 * (fms2<T>(c, xp) + xm);
        const T ddy = ihy2 * (fms2<T>(c, yp) + ym);
        const T ddz = ihz2 * (fms2<T>(c, zp) + zm);
        const T lap = ddx + ddy + ddz;
        const T l = lap * nudt;
        return c + l;
    };
    // LDS: three u^n planes (q and q + 1 read while plane q is evaluated,
    // q + 2 stored at the next plane) and three u^{n+1} planes, in slots
    // (p - z0 + 1) % 3; one barrier per plane orders every write before its
    // reads and every read before the write that reuses its slot
    auto slotA = [&](int k) { return sA + k * PLA; };
    auto slotB = [&](int k) { return sB + k * PLB; };

    // Register rings of three planes, unrolled by three so that every slot
    // index is a constant (plane p in slot (p - z0 + 2) % 3): U0r holds u^n on
    // this thread's cells, U1r u^{n+1}, rz u^n on this thread's ring cells;
    // Ph the halo of the next u^n plane.
    V U0r[3][R], U1r[3][R];
    T rz[3][RS];
    V Ph[HS];
    {
        const U0 Q0 = load_u0(z0 - 2);
        const U0 Q1 = load_u0(z0 - 1);
        store_u0(Q1, slotA(0));
        const U0 Q2 = load_u0(z0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            U0r[0][r] = Q0.c[r];
            U0r[1][r] = Q1.c[r];
            U0r[2][r] = Q2.c[r];
        }
#pragma unroll
        for (int h = 0; h < HS; ++h)
            Ph[h] = Q2.h[h];
    }

    double acc1 = 0.0, acc2 = 0.0;
    // plane q of the tile, register slots in the static rotation u = (q - z0 + 1) % 3
    auto plane = [&](auto U, const int q) __attribute__((always_inline)) {
        constexpr int u = decltype(U)::value;
        V(&c0m)[R] = U0r[u % 3];
        V(&c0c)[R] = U0r[(u + 1) % 3];
        V(&c0p)[R] = U0r[(u + 2) % 3];
        V(&m1)[R] = U1r[u % 3];
        V(&c1)[R] = U1r[(u + 1) % 3];
        V(&n1)[R] = U1r[(u + 2) % 3];
        {
            U0 P;
#pragma unroll
            for (int r = 0; r < R; ++r)
                P.c[r] = c0p[r];
#pragma unroll
            for (int h = 0; h < HS; ++h)
                P.h[h] = Ph[h];
            store_u0(P, slotA((u + 1) % 3)); // plane q + 1
        }
        __syncthreads();
        const T *Ac = slotA(u), *Ap = slotA((u + 1) % 3);
        const bool inner = q >= z0 && q < z1;
        T *Bq = slotB(u);
        // u^{n+1} on the block's cells of plane q (z neighbours from registers)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int y = ty + r * ROWS;
            const int o = (y + 2) * PA + VEC + x0;
            const V ym = *reinterpret_cast<const V *>(Ac + o - PA);
            const V yp = *reinterpret_cast<const V *>(Ac + o + PA);
            // x neighbours of the end cells from the adjacent lanes; every
            // lane of a row also reads the row's two halo values (one LDS
            // address per row), and the first / last lane of the row selects
            // them
            const T hl = Ac[o - x0 - 1], hr = Ac[o - x0 + N];
            const T sl = dpp_shift<T, kDppRowShr1>(hl, c0c[r].v[VEC - 1]);
            const T sr = dpp_shift<T, kDppRowShl1>(hr, c0c[r].v[0]);
            const T xl = tx == 0 ? hl : sl;
            const T xr = tx == TX - 1 ? hr : sr;
            V res;
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
                const T xm = v == 0 ? xl : c0c[r].v[v - 1];
                const T xp = v == VEC - 1 ? xr : c0c[r].v[v + 1];
                res.v[v] = euler(c0c[r].v[v], xm, xp, ym.v[v], yp.v[v], c0m[r].v[v], c0p[r].v[v]);
            }
            n1[r] = res;
            if (inner)
                *reinterpret_cast<V *>(Bq + (y + 1) * PA + VEC + x0) = res;
        }
        // plane q - 1 has served: the next u^n plane goes into its slot
        if (q + 2 <= z1 + 1) {
            const U0 Q = load_u0(q + 2);
#pragma unroll
            for (int r = 0; r < R; ++r)
                c0m[r] = Q.c[r];
#pragma unroll
            for (int h = 0; h < HS; ++h)
                Ph[h] = Q.h[h];
        }
        // u^{n+1} on the ring of face-neighbour cells (planes the second step
        // computes); their z neighbours come from the register ring rz
        if (q < z1) {
#pragma unroll
            for (int h = 0; h < RS; ++h) {
                const int s = tid + h * WG;
                if (s < NR) {
                    int x, y;
                    if (s < 2 * N) {
                        x = (s / N) ? N : -1;
                        y = s % N;
                    } else {
                        y = ((s - 2 * N) / N) ? N : -1;
                        x = (s - 2 * N) % N;
                    }
                    const int o = (y + 2) * PA + VEC + x;
                    rz[(u + 2) % 3][h] = Ap[o];
                    if (q == z0 - 1)
                        rz[(u + 1) % 3][h] = Ac[o];
                    if (inner)
                        Bq[(y + 1) * PA + VEC + x] = euler(rz[(u + 1) % 3][h], Ac[o - 1], Ac[o + 1], Ac[o - PA],
                                                           Ac[o + PA], rz[u % 3][h], rz[(u + 2) % 3][h]);
                }
            }
        }
        // u^{n+2} on the block's cells of plane q - 1 (its u^{n+1} plane was
        // completed before this plane's barrier); the sums of squares of
        // u^{n+1} and u^{n+2} over the tile's cells
        if (q >= z0 + 1) {
            const T *Br = slotB((u + 2) % 3);
            const unsigned ops = oblk + (unsigned)(NN * (q - 1) * TS); // plane q - 1 of the block
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int y = ty + r * ROWS;
                const int o = (y + 1) * PA + VEC + x0;
                const V ym = *reinterpret_cast<const V *>(Br + o - PA);
                const V yp = *reinterpret_cast<const V *>(Br + o + PA);
                const T hl = Br[o - x0 - 1], hr = Br[o - x0 + N];
                const T sl = dpp_shift<T, kDppRowShr1>(hl, c1[r].v[VEC - 1]);
                const T sr = dpp_shift<T, kDppRowShl1>(hr, c1[r].v[0]);
                const T xl = tx == 0 ? hl : sl;
                const T xr = tx == TX - 1 ? hr : sr;
                V res;
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    const T xm = v == 0 ? xl : c1[r].v[v - 1];
                    const T xp = v == VEC - 1 ? xr : c1[r].v[v + 1];
                    res.v[v] = euler(c1[r].v[v], xm, xp, ym.v[v], yp.v[v], m1[r].v[v], n1[r].v[v]);
                    acc1 = __builtin_fma((double)c1[r].v[v], (double)c1[r].v[v], acc1);
                    acc2 = __builtin_fma((double)res.v[v], (double)res.v[v], acc2);
                }
                buf_store<T, VEC>(rout, (unsigned)((x0 + N * y) * TS), ops, res);
            }
        }
    };
    constexpr int kIters = (ZC + 2) / 3;
    for (int it = 0; it < kIters; ++it) {
        const int qb = z0 - 1 + 3 * it;
        prio_ramp(it, kIters);
        plane(std::integral_constant<int, 0>{}, qb);
        plane(std::integral_constant<int, 1>{}, qb + 1);
        plane(std::integral_constant<int, 2>{}, qb + 2);
    }
    // the last (ZC + 2) % 3 planes continue the rotation
    constexpr int kTail = (ZC + 2) % 3;
    const int qt = z0 - 1 + 3 * ((ZC + 2) / 3);
    if constexpr (kTail >= 1)
        plane(std::integral_constant<int, 0>{}, qt);
    if constexpr (kTail >= 2)
        plane(std::integral_constant<int, 1>{}, qt + 1);

    // per-tile partials in a fixed wave tree and wave order
    auto tile_sum = [&](double v, double *part) {
        for (int o = 32; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        __syncthreads();
        if ((tid & 63) == 0)
            red[tid >> 6] = v;
        __syncthreads();
        if (tid == 0 && part) {
            double t = 0.0;
            for (int w = 0; w < WG / 64; ++w)
                t += red[w];
            part[tile] = t;
        }
    };
    tile_sum(acc1, a.part1);
    tile_sum(acc2, a.part2);
}

// ---------------------------------------------------------------- host side

namespace
{

struct X2Tune {
    int wg, zchunk;
};
std::mutex g_x2_mtx;
std::map<std::pair<int, int>, X2Tune> g_x2_tune; // (dtype, edge) -> choice

X2Tune x2_default(int dtype, int edge)
{
    if (edge == 32)
        return dtype == CNV_F64 ? X2Tune{256, 32} : X2Tune{256, 16};
    return dtype == CNV_F64 ? X2Tune{128, 16} : X2Tune{64, 16};
}

X2Tune x2_tune(int dtype, int edge)
{
    std::lock_guard<std::mutex> lk(g_x2_mtx);
    auto it = g_x2_tune.find({dtype, edge});
    return it == g_x2_tune.end() ? x2_default(dtype, edge) : it->second;
}

// z-chunks compiled: 8, 16 and 32 planes (up to the block edge)
inline bool x2_zchunk_ok(int zc) { return zc == 8 || zc == 16 || zc == 32; }

template <typename T, int N, int WG>
bool x2_launch1(const X2Args &a, int zc, int ntiles, hipStream_t st)
{
    if constexpr (X2Shape<T, N, WG>::valid) {
        if (zc == 16)
            hipLaunchKernelGGL((k_stencil7x2<T, N, WG, 16>), dim3(ntiles), dim3(WG), 0, st, a);
        else if (zc == 8)
            hipLaunchKernelGGL((k_stencil7x2<T, N, WG, 8>), dim3(ntiles), dim3(WG), 0, st, a);
        else if constexpr (N == 32) {
            if (zc != 32)
                return false;
            hipLaunchKernelGGL((k_stencil7x2<T, N, WG, 32>), dim3(ntiles), dim3(WG), 0, st, a);
        } else
            return false;
        return true;
    }
    return false;
}

template <typename T, int N>
bool x2_dispatch_wg(const X2Args &a, int wg, int zc, int ntiles, hipStream_t st)
{
    switch (wg) {
    case 64:
        return x2_launch1<T, N, 64>(a, zc, ntiles, st);
    case 128:
        return x2_launch1<T, N, 128>(a, zc, ntiles, st);
    case 256:
        return x2_launch1<T, N, 256>(a, zc, ntiles, st);
    case 512:
        return x2_launch1<T, N, 512>(a, zc, ntiles, st);
    }
    return false;
}

template <typename T, int N>
bool x2_valid_wg(int wg)
{
    switch (wg) {
    case 64:
        return X2Shape<T, N, 64>::valid;
    case 128:
        return X2Shape<T, N, 128>::valid;
    case 256:
        return X2Shape<T, N, 256>::valid;
    case 512:
        return X2Shape<T, N, 512>::valid;
    }
    return false;
}

bool x2_valid(int dtype, int edge, int wg)
{
    if (dtype == CNV_F64)
        return edge == 16 ? x2_valid_wg<double, 16>(wg) : x2_valid_wg<double, 32>(wg);
    return edge == 16 ? x2_valid_wg<float, 16>(wg) : x2_valid_wg<float, 32>(wg);
}

bool x2_dispatch(int dtype, int edge, int wg, int zc, const X2Args &a, int ntiles, hipStream_t st)
{
    if (dtype == CNV_F64)
        return edge == 16 ? x2_dispatch_wg<double, 16>(a, wg, zc, ntiles, st)
                          : x2_dispatch_wg<double, 32>(a, wg, zc, ntiles, st);
    return edge == 16 ? x2_dispatch_wg<float, 16>(a, wg, zc, ntiles, st)
                      : x2_dispatch_wg<float, 32>(a, wg, zc, ntiles, st);
}

} // namespace

bool x2_applicable(cnv_grid_s *u)
{
    const int *n = u->desc.block_cells;
    if (!(n[0] == n[1] && n[1] == n[2] && (n[0] == 16 || n[0] == 32)))
        return false;
    // the kernel addresses a component slab with 32-bit byte offsets
    if (u->component_bytes >= ((size_t)1 << 31))
        return false;
    std::lock_guard<std::mutex> lk(u->mtx);
    for (int c = 0; c < u->desc.ncomp; ++c)
        for (const auto &l : u->bcs[c])
            if (!l.empty())
                return false;
    return true;
}

long long x2_tiles(const cnv_grid_s *u, int zchunk)
{
    const int edge = u->desc.block_cells[0];
    const int zc = zchunk < edge ? zchunk : edge;
    return u->nblocks_total * u->desc.ncomp * ((edge + zc - 1) / zc);
}

int launch_diffusion_x2(cnv_grid_s *u, cnv_grid_s *out, double nu_dt, double *part1, double *part2,
                        const double *const prev_p[2], const int prev_n[2], double *const prev_out[2], int wg,
                        int zchunk, int *nparts, hipStream_t st)
{
    const int edge = u->desc.block_cells[0];
    if (wg <= 0 || zchunk <= 0) {
        const X2Tune t = x2_tune(u->desc.dtype, edge);
        wg = t.wg;
        zchunk = t.zchunk;
    }
    X2Args a;
    a.u = static_cast<const char *>(u->data);
    a.out = static_cast<char *>(out->data);
    a.G = geom_of(u);
    if (!x2_zchunk_ok(zchunk))
        return fail(CNV_ERR_UNSUPPORTED, "two-step kernel: zchunk must be 8, 16 or 32");
    if (zchunk > edge)
        return fail(CNV_ERR_UNSUPPORTED, "two-step kernel: zchunk larger than the block edge");
    a.nzc = edge / zchunk;
    a.tiles_per_comp = (int)(u->nblocks_total * a.nzc);
    for (int d = 0; d < 3; ++d)
        a.ih2[d] = 1. / (u->h[d] * u->h[d]);
    a.nudt = nu_dt;
    a.part1 = part1;
    a.part2 = part2;
    for (int k = 0; k < 2; ++k) {
        a.prev_p[k] = prev_p ? prev_p[k] : nullptr;
        a.prev_n[k] = prev_n ? prev_n[k] : 0;
        a.prev_out[k] = prev_out ? prev_out[k] : nullptr;
    }
    const long long ntiles = (long long)a.tiles_per_comp * u->desc.ncomp;
    if (ntiles >= (1LL << 31))
        return fail(CNV_ERR_UNSUPPORTED, "grid too large for the two-step kernel");
    if (!x2_dispatch(u->desc.dtype, edge, wg, zchunk, a, (int)ntiles, st))
        return fail(CNV_ERR_UNSUPPORTED, "two-step tile shape not available for this block edge and dtype");
    CNV_HIP(hipGetLastError());
    if (nparts)
        *nparts = (int)ntiles;
    return CNV_OK;
}

int x2_set_tune(int dtype, int edge, int wg, int zchunk)
{
    if (edge != 16 && edge != 32)
        return fail(CNV_ERR_INVALID, "two-step kernel: block edge must be 16 or 32");
    if (!x2_zchunk_ok(zchunk) || zchunk > edge)
        return fail(CNV_ERR_UNSUPPORTED, "two-step kernel: zchunk must be 8, 16 or 32, at most the edge");
    if (!x2_valid(dtype, edge, wg))
        return fail(CNV_ERR_UNSUPPORTED, "two-step tile shape not available for this block edge and dtype");
    std::lock_guard<std::mutex> lk(g_x2_mtx);
    g_x2_tune[{dtype, edge}] = X2Tune{wg, zchunk};
    return CNV_OK;
}

int x2_get_tune(int dtype, int edge, int *wg, int *zchunk)
{
    const X2Tune t = x2_tune(dtype, edge);
    if (wg)
        *wg = t.wg;
    if (zchunk)
        *zchunk = t.zchunk;
    return CNV_OK;
}

// Times every valid (workgroup size, z-planes per tile) pair of the two-step
// kernel on u -> out and keeps the fastest for (dtype, edge).
int autotune_x2(cnv_grid_s *u, cnv_grid_s *out, int iters, int *best_wg, int *best_z, double *results,
                int max_results, int *nresults)
{
    const int edge = u->desc.block_cells[0];
    hipStream_t st;
    CNV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    CNV_HIP(hipEventCreate(&e0));
    CNV_HIP(hipEventCreate(&e1));
    const int wgs[4] = {64, 128, 256, 512};
    float best = 1e30f;
    int bw = 0, bz = 0, cnt = 0, rc = CNV_OK;
    for (int wi = 0; wi < 4 && rc == CNV_OK; ++wi)
        for (int z = 8; z <= edge && rc == CNV_OK; z *= 2) {
            if (x2_set_tune(u->desc.dtype, edge, wgs[wi], z) != CNV_OK)
                continue;
            rc = launch_diffusion_x2(u, out, 1e-9, nullptr, nullptr, nullptr, nullptr, nullptr, wgs[wi], z, nullptr,
                                     st);
            if (rc != CNV_OK)
                break;
            (void)hipEventRecord(e0, st);
            for (int i = 0; i < iters && rc == CNV_OK; ++i)
                rc = launch_diffusion_x2(u, out, 1e-9, nullptr, nullptr, nullptr, nullptr, nullptr, wgs[wi], z,
                                         nullptr, st);
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(CNV_ERR_HIP, "autotune: two-step kernel failed");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const float us = ms * 1000.f / iters;
            if (results && cnt < max_results) {
                results[3 * cnt] = wgs[wi];
                results[3 * cnt + 1] = z;
                results[3 * cnt + 2] = us;
            }
            ++cnt;
            if (us < best) {
                best = us;
                bw = wgs[wi];
                bz = z;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (rc != CNV_OK)
        return rc;
    if (bw == 0)
        return fail(CNV_ERR_UNSUPPORTED, "no two-step tile shape for this grid");
    x2_set_tune(u->desc.dtype, edge, bw, bz);
    if (best_wg)
        *best_wg = bw;
    if (best_z)
        *best_z = bz;
    if (nresults)
        *nresults = cnt < max_results ? cnt : max_results;
    return CNV_OK;
}

} // namespace cnv
