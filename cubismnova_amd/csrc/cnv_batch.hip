// cnv_batch.hip -- many independent grids of uneven block counts, block
// sizes, component counts, ghost widths and BCs processed together.
//
// A work item is (grid, component, block, z-range).  The batch runs in one
// of three forms, chosen by cnv_batch_autotune from their measured times
// (the single launch until then):
//  * one persistent launch (k_batch_one) over a device list of work units,
//    largest first: a unit is one item of a cubic 32^3 block run by the
//    whole workgroup on the single-grid tile body of its operator (s7_march,
//    w5_march), or up to WG/64 items of cubic 16^3 blocks, one per 64-thread
//    slice on the same tile bodies, or one item of any other block shape,
//    staged in LDS with its face halo (one layer for the 7-point Euler step,
//    two for WENO5) by the generic body;
//  * one launch per (operator, edge) class of 16^3/32^3 blocks (k_bmarch,
//    one workgroup per item on the same tile bodies) plus the persistent
//    generic launch (k_batch) for the other blocks, forked onto streams of
//    the batch when the items are independent, or one after another on the
//    caller's stream.
// Operators: op 0: u_new = u + nu_dt * lap(u) in the Laplacian.cpp:125-129
// order; op 1: WENO5 u^-_{i+1/2}, u^+_{i-1/2} in x, y, z (outputs 6c+2d,
// 6c+2d+1).  Ghosts follow the grid's Stencil(-g, g+1) and BC lists.
// FP32 grids only.  Either form is capturable as one unit in a HIP graph on
// the caller's stream.
#include "cnv_march.h"

#include <algorithm>
#include <chrono>
#include <map>

namespace cnv
{

constexpr int kBMaxComp = 3;

// device descriptor of one grid of the batch
struct BGrid {
    const float *in; // component 0
    float *out;      // component 0 of the output grid
    Geom G;
    long long ocstride;
    int op, g, ncomp, pad;
    int has_bc[kBMaxComp];
    const DevBCHead *heads[kBMaxComp];
    const DevBC *entries[kBMaxComp];
    float ih2[3];
    float nudt;
};

struct BItem {
    int grid;
    int comp;
    int blk;
    short z0, nz;
};

// (ix, iy, iz) of a flat index over a box of nx x ny x ... cells, x fastest,
// advanced by a fixed step without divisions (the step's digits are found
// once): the generic item visits its cells in strides of the workgroup size.
struct Walk3 {
    int ix, iy, iz, dx, dy, dz, nx, ny;
    __device__ Walk3(int start, int step, int nx_, int ny_) : nx(nx_), ny(ny_)
    {
        ix = start % nx;
        iy = (start / nx) % ny;
        iz = start / (nx * ny);
        dx = step % nx;
        dy = (step / nx) % ny;
        dz = step / (nx * ny);
    }
    __device__ void next()
    {
        ix += dx;
        iy += dy;
        iz += dz;
        if (ix >= nx) { // dx < nx: one carry at most
            ix -= nx;
            ++iy;
        }
        if (iy >= ny) { // dy + 1 <= ny: one carry at most
            iy -= ny;
            ++iz;
        }
    }
};

// One item of any block shape: its cells plus the face halo its operator
// reads staged in LDS (every ghost by the lab rules), then the operator.
template <int WG>
__device__ __attribute__((always_inline)) inline void gen_item(const BGrid *__restrict__ grids, const BItem I, float *lds)
{
    const int tid = threadIdx.x;
    const BGrid &Gd = grids[I.grid];
    const int c = I.comp;
    const int nx = Gd.G.n[0], ny = Gd.G.n[1], nzb = Gd.G.n[2];
    const int gl = Gd.op == 0 ? 1 : 2; // layers read by the operator
    const int Lx = nx + 2 * gl, Ly = ny + 2 * gl, Lz = I.nz + 2 * gl;
    const float *cbase = Gd.in + (long long)c * Gd.G.cstride;
    const float *ub = cbase + (long long)I.blk * Gd.G.bstride;
    FaceL<float> F[6];
    const int sb[3] = {-Gd.g, -Gd.g, -Gd.g}, se[3] = {Gd.g + 1, Gd.g + 1, Gd.g + 1};
    make_faces<float>(Gd.G, Gd.heads[c], Gd.entries[c], Gd.has_bc[c], I.blk, cbase, sb, se, F);
    const long long sx = nx, sxy = (long long)nx * ny;
    // stage the tile and its face halo, U cells per thread and pass: their
    // loads are all issued before the first LDS store, so they are in
    // flight together
    constexpr int U = 4;
    const int ltot = Lx * Ly * Lz;
    Walk3 ws(tid, WG, Lx, Ly);
    for (int l0 = tid; l0 < ltot; l0 += U * WG) {
        float v[U];
        bool w[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int l = l0 + k * WG;
            const int x = ws.ix - gl, y = ws.iy - gl, z = I.z0 + ws.iz - gl;
            ws.next();
            const bool ox = x < 0 || x >= nx, oy = y < 0 || y >= ny, oz = z < 0 || z >= nzb;
            // edges and corners are never read
            w[k] = l < ltot && (int)ox + (int)oy + (int)oz <= 1;
            v[k] = 0.f;
            if (!w[k])
                continue;
            if (ox)
                v[k] = x < 0 ? face_layer(F[0], -x, 1, sx * y + sxy * z)
                             : face_layer(F[1], x - nx + 1, 1, sx * y + sxy * z);
            else if (oy)
                v[k] = y < 0 ? face_layer(F[2], -y, sx, x + sxy * z) : face_layer(F[3], y - ny + 1, sx, x + sxy * z);
            else if (oz)
                v[k] = z < 0 ? face_layer(F[4], -z, sxy, x + sx * y)
                             : face_layer(F[5], z - nzb + 1, sxy, x + sx * y);
            else
                v[k] = ub[x + sx * y + sxy * z];
        }
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (w[k])
                lds[l0 + k * WG] = v[k];
    }
    __syncthreads();
    const int ctot = nx * ny * I.nz;
    float *ob = Gd.out + (long long)I.blk * Gd.G.bstride;
    const int LXY = Lx * Ly;
    Walk3 wc(tid, WG, nx, ny);
    for (int l = tid; l < ctot; l += WG, wc.next()) {
        const int x = wc.ix, y = wc.iy, zt = wc.iz;
        const int z = I.z0 + zt;
        const int p = (x + gl) + Lx * (y + gl) + LXY * (zt + gl);
        const long long o = x + sx * y + sxy * z;
        const float cc = lds[p];
        if (Gd.op == 0) {
            const float ddx = Gd.ih2[0] * (fms2(cc, lds[p + 1]) + lds[p - 1]);
            const float ddy = Gd.ih2[1] * (fms2(cc, lds[p + Lx]) + lds[p - Lx]);
            const float ddz = Gd.ih2[2] * (fms2(cc, lds[p + LXY]) + lds[p - LXY]);
            const float lap = ddx + ddy + ddz;
            const float lt = lap * Gd.nudt;
            ob[(long long)c * Gd.ocstride + o] = cc + lt;
        } else {
            float um, up;
            float *oc = ob + (long long)(6 * c) * Gd.ocstride + o;
            weno5<float>(lds[p - 2], lds[p - 1], cc, lds[p + 1], lds[p + 2], um, up);
            oc[0] = um;
            oc[Gd.ocstride] = up;
            weno5<float>(lds[p - 2 * Lx], lds[p - Lx], cc, lds[p + Lx], lds[p + 2 * Lx], um, up);
            oc[2 * Gd.ocstride] = um;
            oc[3 * Gd.ocstride] = up;
            weno5<float>(lds[p - 2 * LXY], lds[p - LXY], cc, lds[p + LXY], lds[p + 2 * LXY], um, up);
            oc[4 * Gd.ocstride] = um;
            oc[5 * Gd.ocstride] = up;
        }
    }
}

template <int WG>
__global__ __launch_bounds__(WG) void k_batch(const BGrid *__restrict__ grids, const BItem *__restrict__ items,
                                              int nitems)
{
    extern __shared__ float lds[];
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        gen_item<WG>(grids, items[it], lds);
        __syncthreads(); // the next item overwrites the tile
    }
}

// The 7-point Euler tile body of one item on cubic blocks of edge NX (the
// single-grid tile), run by WG threads or by one WG-thread slice (SUB).
template <int NX, int WG, int SUB = 0>
__device__ __attribute__((always_inline)) inline void s7_item(const BGrid *__restrict__ grids, const BItem I, float *lds)
{
    const BGrid &Gd = grids[I.grid];
    const int c = I.comp;
    const float *cbase = Gd.in + (long long)c * Gd.G.cstride;
    const float *ub = cbase + (long long)I.blk * Gd.G.bstride;
    float *ob = Gd.out + (long long)I.blk * Gd.G.bstride;
    FaceSrc<float> F[6];
    block_faces<float>(Gd.G, Gd.heads[c], Gd.entries[c], Gd.has_bc[c], I.blk, cbase, F);
    (void)s7_march<float, NX, NX, WG, 1, false, SUB>(ub, ob + (long long)c * Gd.ocstride, nullptr, F, NX, I.z0,
                                                      I.z0 + I.nz, Gd.ih2[0], Gd.ih2[1], Gd.ih2[2], Gd.nudt, lds);
}

// The WENO5 tile body of one item on cubic blocks of edge NX.
template <int NX, int WG, int SUB = 0>
__device__ __attribute__((always_inline)) inline void w5_item(const BGrid *__restrict__ grids, const BItem I, float *lds)
{
    const BGrid &Gd = grids[I.grid];
    const int c = I.comp;
    const float *cbase = Gd.in + (long long)c * Gd.G.cstride;
    const float *ub = cbase + (long long)I.blk * Gd.G.bstride;
    float *ob = Gd.out + (long long)I.blk * Gd.G.bstride;
    FaceL<float> F[6];
    const int sb[3] = {-Gd.g, -Gd.g, -Gd.g}, se[3] = {Gd.g + 1, Gd.g + 1, Gd.g + 1};
    make_faces<float>(Gd.G, Gd.heads[c], Gd.entries[c], Gd.has_bc[c], I.blk, cbase, sb, se, F);
    w5_march<float, NX, NX, WG, SUB>(ub, ob + (long long)(6 * c) * Gd.ocstride, Gd.ocstride, F, NX, I.z0,
                                     I.z0 + I.nz, lds);
}

// A unit of the single persistent launch.
enum BKind { BK_S7_32 = 0, BK_W5_32 = 1, BK_S7_16 = 2, BK_W5_16 = 3, BK_GEN = 4 };
struct BUnit {
    int kind;
    int n;     // items of the unit: one, or one per 64-thread slice (16^3 kinds)
    int item0; // first item
    int pad;
};

// LDS floats of the tile bodies of k_batch_one<WG> (the generic body's tile
// comes on top as dynamic size): a 16^3 slice gets 2 planes of its tile
template <int WG>
struct OneLds {
    static constexpr int S16 = 2 * S7Shape<float, 16, 16, 64>::PLANE;
    static constexpr int W16 = 2 * W5Shape<float, 16, 16, 64>::PLANE;
    static constexpr int S32 = 2 * S7Shape<float, 32, 32, WG>::PLANE;
    static constexpr int W32 = 2 * W5Shape<float, 32, 32, WG>::PLANE;
    static constexpr int a = S32 > W32 ? S32 : W32;
    static constexpr int b = (WG / 64) * (S16 > W16 ? S16 : W16);
    static constexpr int floats = a > b ? a : b;
};

// The whole batch in one persistent launch over a list of units sorted by
// cost, in rounds of gridDim.x units: in round k workgroup w runs unit
// k * gridDim.x + xcd_remap(w), so that each XCD takes a contiguous run of
// the round's units (neighbouring blocks of one grid, whose halos then meet
// in that XCD's L2).  In a 16^3 unit every slice runs an item (slices past
// the unit's items repeat its last item, writing the same values again), so
// the tile bodies' barriers match.
template <int WG>
__global__ __launch_bounds__(WG) void k_batch_one(const BGrid *__restrict__ grids, const BItem *__restrict__ items,
                                                  const BUnit *__restrict__ units, int nunits)
{
    extern __shared__ __align__(16) float lds[];
    const int mine = xcd_remap(blockIdx.x, gridDim.x);
    for (int u0 = 0; u0 < nunits; u0 += gridDim.x) {
        const int u = u0 + mine;
        if (u >= nunits)
            break;
        const BUnit U = units[u];
        if (U.kind == BK_GEN) {
            gen_item<WG>(grids, items[U.item0], lds);
        } else if (U.kind == BK_S7_32) {
            s7_item<32, WG>(grids, items[U.item0], lds);
        } else if (U.kind == BK_W5_32) {
            w5_item<32, WG>(grids, items[U.item0], lds);
        } else {
            // the slice index is uniform in a wave: scalar loads of the
            // item and its grid descriptor
            const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
            const BItem I = items[U.item0 + (w < U.n ? w : U.n - 1)];
            if (U.kind == BK_S7_16)
                s7_item<16, 64, 64>(grids, I, lds + w * OneLds<WG>::S16);
            else
                w5_item<16, 64, 64>(grids, I, lds + w * OneLds<WG>::W16);
        }
        __syncthreads(); // the next unit overwrites the tiles
    }
}

template <int OP, int NX, int WG>
constexpr bool bm_valid()
{
    return OP == 0 ? S7Shape<float, NX, NX, WG>::valid : W5Shape<float, NX, NX, WG>::valid;
}

template <int OP, int NX, int WG>
constexpr int bm_plane()
{
    return OP == 0 ? S7Shape<float, NX, NX, WG>::PLANE : W5Shape<float, NX, NX, WG>::PLANE;
}

// One item of an (operator, edge) class per workgroup: the single-grid tile
// bodies on the item's grid, component, block and z-range.
template <int OP, int NX, int WG>
__global__ __launch_bounds__(WG) void k_bmarch(const BGrid *__restrict__ grids, const BItem *__restrict__ items,
                                               int nitems)
{
    __shared__ alignas(16) float lds[2 * bm_plane<OP, NX, WG>()];
    const BItem I = items[xcd_remap(blockIdx.x, gridDim.x)];
    if constexpr (OP == 0)
        s7_item<NX, WG>(grids, I, lds);
    else
        w5_item<NX, WG>(grids, I, lds);
}

} // namespace cnv

// ---------------------------------------------------------------- host side

using namespace cnv;

// An (operator, block edge) class run by k_bmarch.
struct BClass {
    int op = 0, edge = 0;
    int wg = 64;
    int zdepth = 8;
    std::vector<int> grids; // indices into cnv_batch_s::src
    std::vector<BItem> items;
    BItem *d_items = nullptr;
    size_t d_cap = 0;
};

struct cnv_batch_s {
    std::vector<cnv_batch_item> src;
    std::vector<BGrid> grids;
    std::vector<std::vector<uint64_t>> versions; // BC versions the descriptors were built from
    BGrid *d_grids = nullptr;
    // persistent kernel over the items of no class
    BItem *d_items = nullptr;
    int nitems = 0;
    int wg = 256;
    int nwg = 0;
    size_t lds_bytes = 0;
    std::map<std::pair<int, int>, int> zdepth; // (edge, ghost) -> planes per item
    std::vector<BItem> items;
    std::vector<BClass> classes;
    // launches of the classes and the persistent launch run concurrently
    // when no item writes what another item reads or writes: forked from
    // the caller's stream onto `side` and joined back (capturable)
    bool concurrent = false;
    bool serial = false; // class launches one after another even when they could overlap
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    // the single persistent launch (k_batch_one): its items, grouped in
    // units and sorted by cost, planes per item per (block edge, ghost
    // width), workgroup size
    bool one = true;
    int one_wg = 256;
    std::map<std::pair<int, int>, int> zmap;
    std::vector<BItem> one_items;
    std::vector<BUnit> units;
    BItem *d_one_items = nullptr;
    size_t one_items_cap = 0;
    BUnit *d_units = nullptr;
    size_t units_cap = 0;
    size_t one_lds = 0;
    int one_per_cu = 0; // resident workgroups per CU for (one_wg, one_lds)
};

static int class_key_edge(const cnv_grid_s *g) { return g->desc.block_cells[0]; }

// Items of this grid go to an (operator, edge) class: cubic blocks of edge 16 or 32.
static bool marchable(const cnv_batch_item &it)
{
    const int *n = it.in->desc.block_cells;
    return n[0] == n[1] && n[1] == n[2] && (n[0] == 16 || n[0] == 32);
}

static const int kWgCand[3] = {64, 128, 256};

template <int OP, int NX, int WG>
static bool bm_launch1(const BClass &c, const BGrid *g, hipStream_t st)
{
    if constexpr (bm_valid<OP, NX, WG>()) {
        hipLaunchKernelGGL((k_bmarch<OP, NX, WG>), dim3((unsigned)c.items.size()), dim3(WG), 0, st, g, c.d_items,
                           (int)c.items.size());
        return true;
    }
    return false;
}

template <int OP, int NX>
static bool bm_launch_wg(const BClass &c, const BGrid *g, hipStream_t st)
{
    switch (c.wg) {
    case 64:
        return bm_launch1<OP, NX, 64>(c, g, st);
    case 128:
        return bm_launch1<OP, NX, 128>(c, g, st);
    case 256:
        return bm_launch1<OP, NX, 256>(c, g, st);
    }
    return false;
}

static bool bm_launch(const BClass &c, const BGrid *g, hipStream_t st)
{
    if (c.op == 0)
        return c.edge == 16 ? bm_launch_wg<0, 16>(c, g, st) : bm_launch_wg<0, 32>(c, g, st);
    return c.edge == 16 ? bm_launch_wg<1, 16>(c, g, st) : bm_launch_wg<1, 32>(c, g, st);
}

static bool bm_wg_valid(int op, int edge, int wg)
{
    if (op == 0)
        return edge == 16 ? (wg == 64 ? bm_valid<0, 16, 64>() : wg == 128 ? bm_valid<0, 16, 128>() : bm_valid<0, 16, 256>())
                          : (wg == 64 ? bm_valid<0, 32, 64>() : wg == 128 ? bm_valid<0, 32, 128>() : bm_valid<0, 32, 256>());
    return edge == 16 ? (wg == 64 ? bm_valid<1, 16, 64>() : wg == 128 ? bm_valid<1, 16, 128>() : bm_valid<1, 16, 256>())
                      : (wg == 64 ? bm_valid<1, 32, 64>() : wg == 128 ? bm_valid<1, 32, 128>() : bm_valid<1, 32, 256>());
}

// Default (workgroup, z depth) of a class: the single-grid kernels' choices.
static void class_defaults(BClass &c)
{
    if (c.op == 0) {
        c.wg = 64;
        c.zdepth = 8;
    } else {
        c.wg = c.edge == 32 ? 256 : 64;
        c.zdepth = c.edge;
    }
}

static int upload_items(BItem **d, size_t *cap, const std::vector<BItem> &v)
{
    if (*d) // a rebuild: launches still queued on any stream may read the old list
        CNV_HIP(hipDeviceSynchronize());
    if (v.size() > *cap) {
        if (*d)
            CNV_HIP(hipFree(*d));
        *d = nullptr;
        *cap = 0;
        CNV_HIP(hipMalloc((void **)d, sizeof(BItem) * v.size()));
        *cap = v.size();
    }
    if (!v.empty())
        CNV_HIP(hipMemcpy(*d, v.data(), sizeof(BItem) * v.size(), hipMemcpyHostToDevice));
    return CNV_OK;
}

static int build_class_items(cnv_batch_s *b, BClass &c)
{
    c.items.clear();
    const int nz = std::max(1, std::min(c.zdepth, c.edge));
    for (int gi : c.grids) {
        const cnv_grid_s *g = b->src[gi].in;
        for (int comp = 0; comp < g->desc.ncomp; ++comp)
            for (long long blk = 0; blk < g->nblocks_total; ++blk)
                for (int z0 = 0; z0 < c.edge; z0 += nz) {
                    BItem I;
                    I.grid = gi;
                    I.comp = comp;
                    I.blk = (int)blk;
                    I.z0 = (short)z0;
                    I.nz = (short)std::min(nz, c.edge - z0);
                    c.items.push_back(I);
                }
    }
    return upload_items(&c.d_items, &c.d_cap, c.items);
}

static size_t tile_bytes(const cnv_grid_s *g, int op, int nz)
{
    const int gl = op == 0 ? 1 : 2;
    return sizeof(float) * (size_t)(g->desc.block_cells[0] + 2 * gl) * (g->desc.block_cells[1] + 2 * gl) *
           (nz + 2 * gl);
}

static constexpr size_t kMaxLds = 64 * 1024;

// Items of the persistent kernel for the current depth table, sorted by cost
// (cells x op weight).  only: restrict to these (edge, ghost) keys.
static int build_items(cnv_batch_s *b, const std::map<std::pair<int, int>, int> *only)
{
    b->items.clear();
    b->lds_bytes = 0;
    for (size_t gi = 0; gi < b->src.size(); ++gi) {
        const cnv_batch_item &it = b->src[gi];
        if (marchable(it))
            continue;
        const cnv_grid_s *g = it.in;
        const std::pair<int, int> key(class_key_edge(g), it.ghost);
        if (only && !only->count(key))
            continue;
        int nz = b->zdepth.count(key) ? b->zdepth.at(key) : g->desc.block_cells[2];
        nz = std::max(1, std::min(nz, g->desc.block_cells[2]));
        while (nz > 1 && tile_bytes(g, it.op, nz) > kMaxLds)
            nz /= 2;
        if (tile_bytes(g, it.op, nz) > kMaxLds)
            return fail(CNV_ERR_UNSUPPORTED, "block too large for the batched kernel");
        b->lds_bytes = std::max(b->lds_bytes, tile_bytes(g, it.op, nz));
        for (int c = 0; c < g->desc.ncomp; ++c)
            for (long long blk = 0; blk < g->nblocks_total; ++blk)
                for (int z0 = 0; z0 < g->desc.block_cells[2]; z0 += nz) {
                    BItem I;
                    I.grid = (int)gi;
                    I.comp = c;
                    I.blk = (int)blk;
                    I.z0 = (short)z0;
                    I.nz = (short)std::min(nz, g->desc.block_cells[2] - z0);
                    b->items.push_back(I);
                }
    }
    auto cost = [&](const BItem &I) {
        const cnv_grid_s *g = b->src[I.grid].in;
        return (long long)g->desc.block_cells[0] * g->desc.block_cells[1] * I.nz * (b->src[I.grid].op == 1 ? 8 : 1);
    };
    std::stable_sort(b->items.begin(), b->items.end(),
                     [&](const BItem &x, const BItem &y) { return cost(x) > cost(y); });
    b->nitems = (int)b->items.size();
    if (b->d_items) { // a rebuild: launches still queued on any stream may read the old list
        CNV_HIP(hipDeviceSynchronize());
        CNV_HIP(hipFree(b->d_items));
    }
    b->d_items = nullptr;
    if (b->nitems > 0) {
        CNV_HIP(hipMalloc((void **)&b->d_items, sizeof(BItem) * b->items.size()));
        CNV_HIP(hipMemcpy(b->d_items, b->items.data(), sizeof(BItem) * b->items.size(), hipMemcpyHostToDevice));
    }
    return CNV_OK;
}

static int build_classes(cnv_batch_s *b)
{
    b->classes.clear();
    std::map<std::pair<int, int>, size_t> at;
    for (size_t gi = 0; gi < b->src.size(); ++gi) {
        const cnv_batch_item &it = b->src[gi];
        if (!marchable(it))
            continue;
        const std::pair<int, int> key(it.op, it.in->desc.block_cells[0]);
        if (!at.count(key)) {
            at[key] = b->classes.size();
            BClass c;
            c.op = key.first;
            c.edge = key.second;
            class_defaults(c);
            b->classes.push_back(c);
        }
        b->classes[at[key]].grids.push_back((int)gi);
    }
    for (auto &c : b->classes) {
        const int rc = build_class_items(b, c);
        if (rc != CNV_OK)
            return rc;
    }
    return CNV_OK;
}

static int build_grids(cnv_batch_s *b)
{
    b->grids.clear();
    b->versions.clear();
    for (const auto &it : b->src) {
        cnv_grid_s *g = it.in;
        BGrid d{};
        d.in = static_cast<const float *>(g->data);
        d.out = static_cast<float *>(it.out->data);
        d.G = geom_of(g);
        d.ocstride = (long long)(it.out->component_bytes / 4);
        d.op = it.op;
        d.g = it.ghost;
        d.ncomp = g->desc.ncomp;
        std::vector<uint64_t> ver;
        std::lock_guard<std::mutex> lk(g->mtx);
        for (int c = 0; c < g->desc.ncomp; ++c) {
            int rc = ensure_bc_table(g, c);
            if (rc != CNV_OK)
                return rc;
            const BCTable &t = g->tables[c];
            d.has_bc[c] = (t.any && t.heads && t.entries) ? 1 : 0;
            d.heads[c] = t.heads;
            d.entries[c] = t.entries;
            // every face the operator reads must be written by the lab rules
            const int sb[3] = {-it.ghost, -it.ghost, -it.ghost}, se[3] = {it.ghost + 1, it.ghost + 1, it.ghost + 1};
            for (const auto &l : g->bcs[c]) {
                if (l.empty())
                    continue;
                std::vector<DevBC> tmp;
                for (const auto &x : l)
                    tmp.push_back({x.kind, x.dir, x.side, 0, x.value});
                int kind[6];
                double val[6];
                resolve_faces(tmp.data(), (int)tmp.size(), sb, se, kind, val);
                for (int f = 0; f < 6; ++f)
                    if (kind[f] == FACE_UNSET)
                        return fail(CNV_ERR_UNSUPPORTED, "a BC list leaves ghost cells unwritten");
            }
        }
        ver.push_back(g->bc_version);
        b->versions.push_back(ver);
        for (int d2 = 0; d2 < 3; ++d2)
            d.ih2[d2] = (float)(1. / (g->h[d2] * g->h[d2]));
        d.nudt = (float)it.nu_dt;
        b->grids.push_back(d);
    }
    if (!b->d_grids) {
        CNV_HIP(hipMalloc((void **)&b->d_grids, sizeof(BGrid) * b->grids.size()));
    } else {
        // a rebuild (BC lists changed): launches still queued on the caller's
        // or the batch's streams may read the old table
        CNV_HIP(hipDeviceSynchronize());
    }
    CNV_HIP(hipMemcpy(b->d_grids, b->grids.data(), sizeof(BGrid) * b->grids.size(), hipMemcpyHostToDevice));
    return CNV_OK;
}

static int launch_generic(cnv_batch_s *b, hipStream_t st)
{
    if (b->nitems == 0)
        return CNV_OK;
    int dev = 0, ncu = 0;
    CNV_HIP(hipGetDevice(&dev));
    CNV_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int per_cu = b->wg >= 256 ? 4 : 8;
    const int nwg = std::min(b->nitems, ncu * per_cu);
    b->nwg = nwg;
    if (b->wg == 128)
        hipLaunchKernelGGL(k_batch<128>, dim3(nwg), dim3(128), b->lds_bytes, st, b->d_grids, b->d_items, b->nitems);
    else if (b->wg == 512)
        hipLaunchKernelGGL(k_batch<512>, dim3(nwg), dim3(512), b->lds_bytes, st, b->d_grids, b->d_items, b->nitems);
    else
        hipLaunchKernelGGL(k_batch<256>, dim3(nwg), dim3(256), b->lds_bytes, st, b->d_grids, b->d_items, b->nitems);
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

// planes per item of the single launch for a grid (per block edge and ghost
// width; the single-grid kernels' depths by default)
static int one_zdepth(const cnv_batch_s *b, const cnv_batch_item &it)
{
    const cnv_grid_s *g = it.in;
    const std::pair<int, int> key(class_key_edge(g), it.ghost);
    int nz = b->zmap.count(key) ? b->zmap.at(key) : (key.first >= 32 ? 32 : 8);
    return std::max(1, std::min(nz, g->desc.block_cells[2]));
}

static int upload_units(BUnit **d, size_t *cap, const std::vector<BUnit> &v)
{
    if (*d) // a rebuild: launches still queued on any stream may read the old list
        CNV_HIP(hipDeviceSynchronize());
    if (v.size() > *cap) {
        if (*d)
            CNV_HIP(hipFree(*d));
        *d = nullptr;
        *cap = 0;
        CNV_HIP(hipMalloc((void **)d, sizeof(BUnit) * v.size()));
        *cap = v.size();
    }
    if (!v.empty())
        CNV_HIP(hipMemcpy(*d, v.data(), sizeof(BUnit) * v.size(), hipMemcpyHostToDevice));
    return CNV_OK;
}

// Items and units of the single launch for the current depths and workgroup
// size: a 32^3 item or an item of another shape is a unit of its own; the
// items of cubic 16^3 blocks of one kind and depth go WG/64 to a unit.
// Units are sorted by cost (cells x operator weight), largest first.
static int build_one(cnv_batch_s *b)
{
    b->one_items.clear();
    b->units.clear();
    size_t gen_lds = 0;
    struct Tmp {
        int kind, nz;
        BItem item;
    };
    std::vector<Tmp> all;
    for (size_t gi = 0; gi < b->src.size(); ++gi) {
        const cnv_batch_item &it = b->src[gi];
        const cnv_grid_s *g = it.in;
        const int edge = g->desc.block_cells[0];
        int kind = BK_GEN;
        if (marchable(it))
            kind = edge == 32 ? (it.op == 0 ? BK_S7_32 : BK_W5_32) : (it.op == 0 ? BK_S7_16 : BK_W5_16);
        int nz = one_zdepth(b, it);
        if (kind == BK_GEN) {
            while (nz > 1 && tile_bytes(g, it.op, nz) > kMaxLds)
                nz /= 2;
            if (tile_bytes(g, it.op, nz) > kMaxLds)
                return fail(CNV_ERR_UNSUPPORTED, "block too large for the batched kernel");
            gen_lds = std::max(gen_lds, tile_bytes(g, it.op, nz));
        }
        for (int c = 0; c < g->desc.ncomp; ++c)
            for (long long blk = 0; blk < g->nblocks_total; ++blk)
                for (int z0 = 0; z0 < g->desc.block_cells[2]; z0 += nz) {
                    BItem I;
                    I.grid = (int)gi;
                    I.comp = c;
                    I.blk = (int)blk;
                    I.z0 = (short)z0;
                    I.nz = (short)std::min(nz, g->desc.block_cells[2] - z0);
                    all.push_back(Tmp{kind, I.nz, I});
                }
    }
    // 16^3 items: grouped by (kind, depth) so that the slices of a unit march
    // through the same number of planes
    std::stable_sort(all.begin(), all.end(), [](const Tmp &x, const Tmp &y) {
        return x.kind != y.kind ? x.kind < y.kind : x.nz < y.nz;
    });
    const int slices = b->one_wg / 64;
    std::vector<std::pair<long long, BUnit>> units;
    for (size_t i = 0; i < all.size();) {
        const Tmp &t = all[i];
        const bool sliced = t.kind == BK_S7_16 || t.kind == BK_W5_16;
        size_t n = 1;
        if (sliced)
            while (n < (size_t)slices && i + n < all.size() && all[i + n].kind == t.kind && all[i + n].nz == t.nz)
                ++n;
        BUnit u;
        u.kind = t.kind;
        u.n = (int)n;
        u.item0 = (int)b->one_items.size();
        u.pad = 0;
        const cnv_grid_s *g = b->src[t.item.grid].in;
        const long long cost = (long long)g->desc.block_cells[0] * g->desc.block_cells[1] * t.nz *
                               (b->src[t.item.grid].op == 1 ? 8 : 1);
        for (size_t k = 0; k < n; ++k)
            b->one_items.push_back(all[i + k].item);
        units.push_back({cost, u});
        i += n;
    }
    std::stable_sort(units.begin(), units.end(),
                     [](const std::pair<long long, BUnit> &x, const std::pair<long long, BUnit> &y) {
                         return x.first > y.first;
                     });
    for (const auto &u : units)
        b->units.push_back(u.second);
    const size_t fixed = sizeof(float) * (size_t)(b->one_wg == 128 ? OneLds<128>::floats : OneLds<256>::floats);
    b->one_lds = std::max(fixed, gen_lds);
    b->one_per_cu = 0;
    int rc = upload_items(&b->d_one_items, &b->one_items_cap, b->one_items);
    if (rc != CNV_OK)
        return rc;
    return upload_units(&b->d_units, &b->units_cap, b->units);
}

static int launch_one(cnv_batch_s *b, hipStream_t st)
{
    if (b->units.empty())
        return CNV_OK;
    int dev = 0, ncu = 0;
    CNV_HIP(hipGetDevice(&dev));
    CNV_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (b->one_per_cu <= 0) {
        int n = 0;
        if (b->one_wg == 128)
            CNV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_batch_one<128>, 128, b->one_lds));
        else
            CNV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_batch_one<256>, 256, b->one_lds));
        b->one_per_cu = std::max(1, n);
    }
    const long long want = (long long)ncu * b->one_per_cu;
    const int nwg = (int)std::min<long long>((long long)b->units.size(), want);
    if (b->one_wg == 128)
        hipLaunchKernelGGL(k_batch_one<128>, dim3(nwg), dim3(128), b->one_lds, st, b->d_grids, b->d_one_items,
                           b->d_units, (int)b->units.size());
    else
        hipLaunchKernelGGL(k_batch_one<256>, dim3(nwg), dim3(256), b->one_lds, st, b->d_grids, b->d_one_items,
                           b->d_units, (int)b->units.size());
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

static int launch_class(cnv_batch_s *b, const BClass &c, hipStream_t st)
{
    if (c.items.empty())
        return CNV_OK;
    if (!bm_launch(c, b->d_grids, st))
        return fail(CNV_ERR_UNSUPPORTED, "no batched kernel for this class and workgroup size");
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

static int launch_batch(cnv_batch_s *b, hipStream_t st)
{
    if (b->one)
        return launch_one(b, st);
    // the non-empty launches: classes in order, then the persistent one (-1)
    std::vector<int> parts;
    for (size_t k = 0; k < b->classes.size(); ++k)
        if (!b->classes[k].items.empty())
            parts.push_back((int)k);
    if (b->nitems > 0)
        parts.push_back(-1);
    auto run = [&](int k, hipStream_t s) { return k < 0 ? launch_generic(b, s) : launch_class(b, b->classes[k], s); };
    if (!b->concurrent || b->serial || parts.size() < 2 || b->side.size() + 1 < parts.size()) {
        for (int k : parts) {
            const int rc = run(k, st);
            if (rc != CNV_OK)
                return rc;
        }
        return CNV_OK;
    }
    // fork: every launch after the first waits for the work before it on
    // st, runs on its own stream and is joined back into st
    CNV_HIP(hipEventRecord(b->fork, st));
    int rc = CNV_OK;
    size_t forked = 0;
    for (size_t i = 1; i < parts.size() && rc == CNV_OK; ++i) {
        hipStream_t s = b->side[i - 1];
        CNV_HIP(hipStreamWaitEvent(s, b->fork, 0));
        rc = run(parts[i], s);
        CNV_HIP(hipEventRecord(b->joins[i - 1], s));
        ++forked;
    }
    if (rc == CNV_OK)
        rc = run(parts[0], st);
    for (size_t i = 0; i < forked; ++i)
        CNV_HIP(hipStreamWaitEvent(st, b->joins[i], 0));
    return rc;
}

// Concurrent launches need the memory of every output grid apart from every
// other item's output and from every input (handles wrapped over the same
// memory count as one grid).
static bool independent_items(const cnv_batch_s *b)
{
    for (size_t i = 0; i < b->src.size(); ++i)
        for (size_t j = 0; j < b->src.size(); ++j) {
            if (i != j && grids_overlap(b->src[i].out, b->src[j].out))
                return false;
            if (grids_overlap(b->src[i].out, b->src[j].in))
                return false;
        }
    return true;
}

static int make_side_streams(cnv_batch_s *b)
{
    b->concurrent = independent_items(b);
    if (!b->concurrent)
        return CNV_OK;
    const size_t n = b->classes.size(); // launches beyond the first: at most #classes
    CNV_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (size_t i = 0; i < n; ++i) {
        hipStream_t s = nullptr;
        CNV_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        b->side.push_back(s);
        hipEvent_t e = nullptr;
        CNV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->joins.push_back(e);
    }
    return CNV_OK;
}

static bool stale(const cnv_batch_s *b)
{
    for (size_t i = 0; i < b->src.size(); ++i)
        if (b->versions[i][0] != b->src[i].in->bc_version)
            return true;
    return false;
}

extern "C" {

int cnv_batch_create(const cnv_batch_item *items, int n, cnv_batch *out)
{
    try {
        if (!items || n < 1 || !out)
            return fail(CNV_ERR_INVALID, "need at least one batch item and an output handle");
        *out = nullptr;
        for (int i = 0; i < n; ++i) {
            const cnv_batch_item &it = items[i];
            if (!it.in || !it.out || grids_overlap(it.in, it.out))
                return fail(CNV_ERR_INVALID, "batch item needs input and output grids that do not share memory");
            if (it.in->desc.dtype != CNV_F32 || it.out->desc.dtype != CNV_F32)
                return fail(CNV_ERR_UNSUPPORTED, "the batched kernel runs FP32 grids");
            if (it.in->desc.entity != CNV_CELL || it.out->desc.entity != CNV_CELL)
                return fail(CNV_ERR_UNSUPPORTED, "operator takes cell-centred grids only");
            if (it.op != CNV_OP_DIFFUSION && it.op != CNV_OP_WENO5)
                return fail(CNV_ERR_INVALID, "unknown batch operator");
            if (it.ghost < (it.op == CNV_OP_WENO5 ? 2 : 1) || it.ghost > 8)
                return fail(CNV_ERR_INVALID, "ghost width too small for the operator");
            if (it.in->desc.ncomp > kBMaxComp)
                return fail(CNV_ERR_UNSUPPORTED, "at most 3 components per batched grid");
            for (int d = 0; d < 3; ++d) {
                if (it.in->desc.nblocks[d] != it.out->desc.nblocks[d] ||
                    it.in->desc.block_cells[d] != it.out->desc.block_cells[d])
                    return fail(CNV_ERR_SIZE, "input and output grids differ in shape");
                if (it.in->desc.block_cells[d] < it.ghost && it.op == CNV_OP_WENO5)
                    return fail(CNV_ERR_UNSUPPORTED, "block edge smaller than the ghost width");
            }
            const int need = it.op == CNV_OP_WENO5 ? 6 * it.in->desc.ncomp : it.in->desc.ncomp;
            if (it.out->desc.ncomp < need)
                return fail(CNV_ERR_SIZE, "output grid has too few components");
            if (it.in->nblocks_total > 0x7fffffff)
                return fail(CNV_ERR_UNSUPPORTED, "grid too large");
        }
        cnv_batch_s *b = new cnv_batch_s();
        b->src.assign(items, items + n);
        int rc = build_grids(b);
        if (rc == CNV_OK)
            rc = build_classes(b);
        if (rc == CNV_OK)
            rc = build_items(b, nullptr);
        if (rc == CNV_OK)
            rc = build_one(b);
        if (rc == CNV_OK)
            rc = make_side_streams(b);
        if (rc != CNV_OK) {
            cnv_batch_destroy(b);
            return rc;
        }
        *out = b;
        return CNV_OK;
    } catch (...) {
        return fail(CNV_ERR_ALLOC, "batch creation failed");
    }
}

int cnv_batch_destroy(cnv_batch b)
{
    if (!b)
        return CNV_OK;
    if (b->d_grids)
        (void)hipFree(b->d_grids);
    if (b->d_items)
        (void)hipFree(b->d_items);
    if (b->d_one_items)
        (void)hipFree(b->d_one_items);
    if (b->d_units)
        (void)hipFree(b->d_units);
    for (auto &c : b->classes)
        if (c.d_items)
            (void)hipFree(c.d_items);
    for (hipStream_t s : b->side)
        (void)hipStreamDestroy(s);
    for (hipEvent_t e : b->joins)
        (void)hipEventDestroy(e);
    if (b->fork)
        (void)hipEventDestroy(b->fork);
    delete b;
    return CNV_OK;
}

int cnv_batch_run(cnv_batch b, cnv_stream s)
{
    if (!b)
        return fail(CNV_ERR_INVALID, "batch is NULL");
    if (stale(b)) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (s)
            (void)hipStreamIsCapturing((hipStream_t)s, &cs);
        if (cs != hipStreamCaptureStatusNone)
            return fail(CNV_ERR_INVALID, "BC lists changed since the batch was built; run it once before capture");
        int rc = build_grids(b);
        if (rc != CNV_OK)
            return rc;
    }
    return launch_batch(b, (hipStream_t)s);
}

int cnv_batch_info(cnv_batch b, int *nitems, int *wg, size_t *lds_bytes)
{
    if (!b)
        return fail(CNV_ERR_INVALID, "batch is NULL");
    if (nitems) {
        long long n = b->nitems;
        for (const auto &c : b->classes)
            n += (long long)c.items.size();
        *nitems = (int)(b->one ? (long long)b->one_items.size() : n);
    }
    if (wg)
        *wg = b->one ? b->one_wg : b->wg;
    if (lds_bytes)
        *lds_bytes = b->one ? b->one_lds : b->lds_bytes;
    return CNV_OK;
}

int cnv_batch_set_form(cnv_batch b, int form)
{
    if (!b)
        return fail(CNV_ERR_INVALID, "batch is NULL");
    if (form < 0 || form > 2)
        return fail(CNV_ERR_INVALID, "form must be 0 (class launches), 1 (single launch) or 2 (class launches "
                                     "one after another)");
    b->one = form == 1;
    b->serial = form == 2;
    return CNV_OK;
}

int cnv_batch_form(cnv_batch b, int *form, int *nunits)
{
    if (!b)
        return fail(CNV_ERR_INVALID, "batch is NULL");
    if (form)
        *form = b->one ? 1 : (b->serial ? 2 : 0);
    if (nunits)
        *nunits = (int)b->units.size();
    return CNV_OK;
}

int cnv_batch_class_info(cnv_batch b, int max_classes, int *out, int *nclasses)
{
    if (!b)
        return fail(CNV_ERR_INVALID, "batch is NULL");
    if (max_classes < 0 || (max_classes > 0 && !out))
        return fail(CNV_ERR_INVALID, "out is NULL");
    const int n = (int)b->classes.size();
    for (int i = 0; i < n && i < max_classes; ++i) {
        const BClass &c = b->classes[(size_t)i];
        out[5 * i] = c.op;
        out[5 * i + 1] = c.edge;
        out[5 * i + 2] = c.wg;
        out[5 * i + 3] = c.zdepth;
        out[5 * i + 4] = (int)c.items.size();
    }
    if (nclasses)
        *nclasses = n;
    return CNV_OK;
}

// For every (operator, edge) class: time the class's launch alone for every
// valid workgroup size and z depth, keep the fastest.  Then, for the items of
// the persistent kernel: per (block edge, ghost width) the z depth, then the
// workgroup size.  results: records of 5 doubles {op, edge, wg, depth,
// microseconds}; persistent-kernel trials carry op = -1 and edge = -1 for
// the workgroup trials.
int cnv_batch_autotune(cnv_batch b, int iters, double *results, int max_results, int *nresults)
{
    if (!b)
        return fail(CNV_ERR_INVALID, "batch is NULL");
    if (iters < 1)
        return fail(CNV_ERR_INVALID, "iters must be >= 1");
    hipStream_t st;
    CNV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    CNV_HIP(hipEventCreate(&e0));
    CNV_HIP(hipEventCreate(&e1));
    int cnt = 0, rc = CNV_OK;
    auto record = [&](double op, double edge, double wg, double depth, double us) {
        if (results && cnt < max_results) {
            double *r = results + 5 * cnt;
            r[0] = op;
            r[1] = edge;
            r[2] = wg;
            r[3] = depth;
            r[4] = us;
        }
        ++cnt;
    };
    auto time_n = [&](auto &&launch, float &us, int n) -> int {
        int r = launch(); // warm
        if (r != CNV_OK)
            return r;
        (void)hipEventRecord(e0, st);
        for (int i = 0; i < n && r == CNV_OK; ++i)
            r = launch();
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess)
            return fail(CNV_ERR_HIP, "autotune: batch kernel failed");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        us = ms * 1000.f / n;
        return r;
    };
    auto time_it = [&](auto &&launch, float &us) -> int { return time_n(launch, us, iters); };
    for (auto &c : b->classes) {
        if (rc != CNV_OK)
            break;
        float best = 1e30f;
        int bwg = c.wg, bz = c.zdepth;
        for (int wg : kWgCand) {
            if (!bm_wg_valid(c.op, c.edge, wg))
                continue;
            for (int z = 2; z <= c.edge && rc == CNV_OK; z *= 2) {
                c.wg = wg;
                c.zdepth = z;
                rc = build_class_items(b, c);
                if (rc != CNV_OK)
                    break;
                float us = 0.f;
                rc = time_it([&]() { return launch_class(b, c, st); }, us);
                if (rc != CNV_OK)
                    break;
                record(c.op, c.edge, wg, z, us);
                if (us < best) {
                    best = us;
                    bwg = wg;
                    bz = z;
                }
            }
        }
        c.wg = bwg;
        c.zdepth = bz;
        if (rc == CNV_OK)
            rc = build_class_items(b, c);
    }
    std::map<std::pair<int, int>, int> keys;
    for (const auto &it : b->src)
        if (!marchable(it))
            keys[{class_key_edge(it.in), it.ghost}] = 1;
    b->wg = 256;
    for (const auto &kv : keys) {
        const auto key = kv.first;
        const int edge = key.first;
        float best = 1e30f;
        int bestz = edge;
        std::map<std::pair<int, int>, int> only{{key, 1}};
        for (int z = 1; z <= edge && rc == CNV_OK; z *= 2) {
            b->zdepth[key] = z;
            rc = build_items(b, &only);
            if (rc != CNV_OK)
                break;
            float us = 0.f;
            rc = time_it([&]() { return launch_generic(b, st); }, us);
            if (rc != CNV_OK)
                break;
            record(-1, edge, key.second, z, us);
            if (us < best) {
                best = us;
                bestz = z;
            }
        }
        b->zdepth[key] = bestz;
    }
    if (rc == CNV_OK)
        rc = build_items(b, nullptr);
    if (!keys.empty()) {
        float best = 1e30f;
        int bestwg = 256;
        const int wgs[3] = {128, 256, 512};
        for (int i = 0; i < 3 && rc == CNV_OK; ++i) {
            b->wg = wgs[i];
            float us = 0.f;
            rc = time_it([&]() { return launch_generic(b, st); }, us);
            if (rc != CNV_OK)
                break;
            record(-1, -1, wgs[i], 0, us);
            if (us < best) {
                best = us;
                bestwg = wgs[i];
            }
        }
        b->wg = bestwg;
    }
    // the class form as a whole: its launches forked and joined as in a run
    // (when the items allow it), and one after another on one stream
    float t_class = 1e30f, t_serial = 1e30f;
    if (rc == CNV_OK) {
        b->one = false;
        b->serial = false;
        rc = time_n([&]() { return launch_batch(b, st); }, t_class, 4 * iters);
        if (rc == CNV_OK)
            record(-3, -1, 0, 0, t_class);
    }
    if (rc == CNV_OK && b->concurrent) {
        b->serial = true;
        rc = time_n([&]() { return launch_batch(b, st); }, t_serial, 4 * iters);
        if (rc == CNV_OK)
            record(-3, -1, 2, 0, t_serial);
    }
    b->serial = t_serial < t_class;
    // the single launch: per (block edge, ghost width) the planes per item,
    // then the workgroup size, each timed on the whole launch
    float t_one = 1e30f;
    if (rc == CNV_OK) {
        b->one = true;
        std::map<std::pair<int, int>, int> okeys;
        for (const auto &it : b->src)
            okeys[{class_key_edge(it.in), it.ghost}] = 1;
        for (const auto &kv : okeys) {
            const auto key = kv.first;
            const int edge = key.first;
            float best = 1e30f;
            int bestz = one_zdepth(b, b->src[0]);
            for (int z = edge >= 16 ? 2 : 1; z <= edge && rc == CNV_OK; z *= 2) {
                b->zmap[key] = z;
                rc = build_one(b);
                if (rc != CNV_OK)
                    break;
                float us = 0.f;
                rc = time_it([&]() { return launch_one(b, st); }, us);
                if (rc != CNV_OK)
                    break;
                record(-2, edge, key.second, z, us);
                if (us < best) {
                    best = us;
                    bestz = z;
                }
            }
            b->zmap[key] = bestz;
        }
        int bestwg = b->one_wg;
        for (int wg : {128, 256}) {
            if (rc != CNV_OK)
                break;
            b->one_wg = wg;
            rc = build_one(b);
            if (rc != CNV_OK)
                break;
            float us = 0.f;
            rc = time_it([&]() { return launch_one(b, st); }, us);
            if (rc != CNV_OK)
                break;
            record(-2, -1, wg, 0, us);
            if (us < t_one) {
                t_one = us;
                bestwg = wg;
            }
        }
        b->one_wg = bestwg;
        if (rc == CNV_OK)
            rc = build_one(b);
        // the whole form timed as long as the class form
        if (rc == CNV_OK)
            rc = time_n([&]() { return launch_one(b, st); }, t_one, 4 * iters);
        if (rc == CNV_OK)
            record(-3, -1, 1, 0, t_one);
    }
    // the fastest form runs from now on
    b->one = t_one <= std::min(t_class, t_serial);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (nresults)
        *nresults = std::min(cnt, max_results);
    return rc;
}

} // extern "C"
