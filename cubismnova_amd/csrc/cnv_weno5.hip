// cnv_weno5.hip -- fifth-order WENO face states over the ghost-extended
// blocks of a grid (Jiang & Shu, J. Comput. Phys. 126:202-228, 1996).
//
// CubismNova advertises a WENO toolbox (README.rst:28-31) but ships none; the
// lab it would read is the FieldLab of Stencil(-3,4,false) with the blocks'
// BC lists, which this kernel assembles on the fly exactly as cnv_lab_load
// does.  For every cell i, direction d and component c it writes
//   u^-_{i+1/2} from cells i-2..i+2 with linear weights (1,6,3)/10, and
//   u^+_{i-1/2} from the mirrored stencils with weights (3,6,1)/10,
// sharing the three smoothness indicators between the two states.
// eps = 1e-6 * m2, m2 = mean of the five squared values + 1e-30, applied in
// the scale-free form  alpha_k = d_k / (1e-6 + beta_k / m2)^2, which keeps
// FP32 free of underflow for any data magnitude; the weights are evaluated
// multiplied by prod_j (1e-6 + beta_j / m2)^2 (three divisions per state
// pair instead of twelve).
//
// Work layout as in the 7-point kernel: a workgroup owns an x-y plane of one
// block of one component and marches through a z-chunk; the plane plus a
// two-cell x/y halo lives in LDS, the z neighbours in a register queue of
// planes loaded ahead.  FP64 grids are evaluated without contraction in the
// operation order of the CPU restatement (bit-identical); FP32 grids let the
// compiler fuse multiply-adds and use the hardware reciprocal.
#include "cnv_march.h"

#include <map>
#include <type_traits>
#include <utility>

namespace cnv
{

constexpr int kWMaxComp = 8;

struct W5Args {
    const char *in;
    char *out;
    Geom G;       // input geometry (cstride of the input grid)
    long long ocstride;
    int ncomp;
    int nzc, zchunk, tiles_per_comp;
    int has_bc[kWMaxComp];
    const DevBCHead *heads[kWMaxComp];
    const DevBC *entries[kWMaxComp];
};

template <typename T, int NX, int NY, int WG>
__global__ __launch_bounds__(WG) void k_weno5(const W5Args a)
{
    using S = W5Shape<T, NX, NY, WG>;
    __shared__ alignas(16) T lds[2 * S::PLANE];

    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int comp = tile / a.tiles_per_comp;
    const int tt = tile - comp * a.tiles_per_comp;
    const long long blk = tt / a.nzc;
    const int zc = tt - (int)blk * a.nzc;
    const int NZ = a.G.n[2];
    const int z0 = zc * a.zchunk;
    const int z1 = min(NZ, z0 + a.zchunk);

    const T *cbase = reinterpret_cast<const T *>(a.in) + (long long)comp * a.G.cstride;
    const T *ub = cbase + blk * a.G.bstride;
    T *ob = reinterpret_cast<T *>(a.out) + (long long)(6 * comp) * a.ocstride + blk * a.G.bstride;

    FaceL<T> F[6];
    const int sb[3] = {-3, -3, -3}, se[3] = {4, 4, 4};
    make_faces<T>(a.G, a.heads[comp], a.entries[comp], a.has_bc[comp], blk, cbase, sb, se, F);
    w5_march<T, NX, NY, WG>(ub, ob, a.ocstride, F, NZ, z0, z1, lds);
}

// Generic fallback for any block shape: one thread per cell, values through
// the resolved faces (ghost width up to 2 read, faces resolved for (-3,4)).
template <typename T>
__global__ __launch_bounds__(256) void k_weno5_generic(const W5Args a)
{
    const long long nbt = (long long)a.G.nb[0] * a.G.nb[1] * a.G.nb[2];
    const long long slot = blockIdx.y;
    const int comp = (int)(slot / nbt);
    const long long blk = slot - comp * nbt;
    const T *cbase = reinterpret_cast<const T *>(a.in) + (long long)comp * a.G.cstride;
    const T *ub = cbase + blk * a.G.bstride;
    T *ob = reinterpret_cast<T *>(a.out) + (long long)(6 * comp) * a.ocstride + blk * a.G.bstride;
    FaceL<T> F[6];
    const int sb[3] = {-3, -3, -3}, se[3] = {4, 4, 4};
    make_faces<T>(a.G, a.heads[comp], a.entries[comp], a.has_bc[comp], blk, cbase, sb, se, F);
    const int *n = a.G.n;
    const long long cells = (long long)n[0] * n[1] * n[2];
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= cells)
        return;
    const int x = (int)(i % n[0]);
    const int y = (int)((i / n[0]) % n[1]);
    const int z = (int)(i / ((long long)n[0] * n[1]));
    const long long nx = n[0], nxy = (long long)n[0] * n[1];
    auto at = [&](int xx, int yy, int zz) -> T {
        if (xx < 0)
            return face_layer(F[0], -xx, 1, nx * yy + nxy * zz);
        if (xx >= n[0])
            return face_layer(F[1], xx - n[0] + 1, 1, nx * yy + nxy * zz);
        if (yy < 0)
            return face_layer(F[2], -yy, nx, xx + nxy * zz);
        if (yy >= n[1])
            return face_layer(F[3], yy - n[1] + 1, nx, xx + nxy * zz);
        if (zz < 0)
            return face_layer(F[4], -zz, nxy, xx + nx * yy);
        if (zz >= n[2])
            return face_layer(F[5], zz - n[2] + 1, nxy, xx + nx * yy);
        return ub[xx + nx * yy + nxy * zz];
    };
    T um, up;
    weno5<T>(at(x - 2, y, z), at(x - 1, y, z), ub[i], at(x + 1, y, z), at(x + 2, y, z), um, up);
    ob[i] = um;
    ob[a.ocstride + i] = up;
    weno5<T>(at(x, y - 2, z), at(x, y - 1, z), ub[i], at(x, y + 1, z), at(x, y + 2, z), um, up);
    ob[2 * a.ocstride + i] = um;
    ob[3 * a.ocstride + i] = up;
    weno5<T>(at(x, y, z - 2), at(x, y, z - 1), ub[i], at(x, y, z + 1), at(x, y, z + 2), um, up);
    ob[4 * a.ocstride + i] = um;
    ob[5 * a.ocstride + i] = up;
}

// ---------------------------------------------------------------- host side

struct W5Tune {
    int wg, zchunk;
};
static std::mutex g_w5_mtx;
static std::map<std::pair<int, int>, W5Tune> g_w5_tune; // (dtype, edge) -> choice

// built-in choices (measured at 512^3 for edge 32)
static W5Tune w5_default(int dtype, int edge)
{
    if (edge == 16)
        return dtype == CNV_F64 ? W5Tune{128, 16} : W5Tune{64, 16};
    if (edge == 32)
        return dtype == CNV_F64 ? W5Tune{256, 32} : W5Tune{256, 32};
    return W5Tune{256, 16};
}

static W5Tune w5_tune(int dtype, int edge)
{
    std::lock_guard<std::mutex> lk(g_w5_mtx);
    auto it = g_w5_tune.find({dtype, edge});
    return it == g_w5_tune.end() ? w5_default(dtype, edge) : it->second;
}

template <typename T, int NX, int WG>
constexpr bool w5_valid()
{
    return W5Shape<T, NX, NX, WG>::valid;
}

template <typename T, int NX, int WG>
static bool w5_launch(const W5Args &a, int ntiles, hipStream_t st)
{
    if constexpr (w5_valid<T, NX, WG>()) {
        hipLaunchKernelGGL((k_weno5<T, NX, NX, WG>), dim3(ntiles), dim3(WG), 0, st, a);
        return true;
    }
    return false;
}

template <typename T, int NX>
static bool w5_dispatch_wg(const W5Args &a, int wg, int ntiles, hipStream_t st)
{
    switch (wg) {
    case 64:
        return w5_launch<T, NX, 64>(a, ntiles, st);
    case 128:
        return w5_launch<T, NX, 128>(a, ntiles, st);
    case 256:
        return w5_launch<T, NX, 256>(a, ntiles, st);
    case 512:
        return w5_launch<T, NX, 512>(a, ntiles, st);
    default:
        return false;
    }
}

template <typename T>
static int w5_run(W5Args &a, hipStream_t st, const W5Tune *over)
{
    const int *n = a.G.n;
    const long long nbt = (long long)a.G.nb[0] * a.G.nb[1] * a.G.nb[2];
    if (n[0] == n[1] && (n[0] == 16 || n[0] == 32 || n[0] == 64)) {
        const W5Tune tu = over ? *over : w5_tune(sizeof(T) == 8 ? CNV_F64 : CNV_F32, n[0]);
        a.zchunk = tu.zchunk < n[2] ? tu.zchunk : n[2];
        a.nzc = (n[2] + a.zchunk - 1) / a.zchunk;
        a.tiles_per_comp = (int)(nbt * a.nzc);
        const long long ntiles = (long long)a.tiles_per_comp * a.ncomp;
        bool ok = false;
        if (ntiles < (1LL << 31)) {
            if (n[0] == 16)
                ok = w5_dispatch_wg<T, 16>(a, tu.wg, (int)ntiles, st);
            else if (n[0] == 32)
                ok = w5_dispatch_wg<T, 32>(a, tu.wg, (int)ntiles, st);
            else
                ok = w5_dispatch_wg<T, 64>(a, tu.wg, (int)ntiles, st);
        }
        if (ok) {
            CNV_HIP(hipGetLastError());
            return CNV_OK;
        }
        if (over)
            return fail(CNV_ERR_UNSUPPORTED, "tile shape not available for this block edge and dtype");
    }
    const long long cells = (long long)n[0] * n[1] * n[2];
    const long long gx = (cells + 255) / 256, gy = nbt * a.ncomp;
    if (gx > 0x7fffffffLL || gy > 65535)
        return fail(CNV_ERR_UNSUPPORTED, "grid too large for the generic WENO5 kernel");
    hipLaunchKernelGGL(k_weno5_generic<T>, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, st, a);
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

static int check_faces_w5(cnv_grid_s *g, int comp)
{
    const int sb[3] = {-3, -3, -3}, se[3] = {4, 4, 4};
    std::vector<DevBC> tmp;
    for (const auto &l : g->bcs[comp]) {
        if (l.empty())
            continue;
        tmp.clear();
        for (const auto &b : l)
            tmp.push_back({b.kind, b.dir, b.side, 0, b.value});
        int kind[6];
        double val[6];
        resolve_faces(tmp.data(), (int)tmp.size(), sb, se, kind, val);
        for (int f = 0; f < 6; ++f)
            if (kind[f] == FACE_UNSET)
                return fail(CNV_ERR_UNSUPPORTED, "BC list leaves ghost cells of a face unwritten for Stencil(-3,4)");
    }
    return CNV_OK;
}

int launch_weno5(cnv_grid_s *in, cnv_grid_s *out, hipStream_t st, const int *tile)
{
    const int nc = in->desc.ncomp;
    if (nc > kWMaxComp)
        return fail(CNV_ERR_UNSUPPORTED, "more than 8 components in one WENO5 call");
    for (int d = 0; d < 3; ++d)
        if (in->desc.block_cells[d] < 2)
            return fail(CNV_ERR_UNSUPPORTED, "WENO5 reads two ghost layers: block edges must be >= 2");
    W5Args a;
    a.in = static_cast<const char *>(in->data);
    a.out = static_cast<char *>(out->data);
    a.G = geom_of(in);
    a.ocstride = (long long)(out->component_bytes / out->elem);
    a.ncomp = nc;
    {
        std::lock_guard<std::mutex> lk(in->mtx);
        for (int c = 0; c < nc; ++c) {
            int rc = ensure_bc_table(in, c);
            if (rc != CNV_OK)
                return rc;
            rc = check_faces_w5(in, c);
            if (rc != CNV_OK)
                return rc;
            a.has_bc[c] = (in->tables[c].any && in->tables[c].heads && in->tables[c].entries) ? 1 : 0;
            a.heads[c] = in->tables[c].heads;
            a.entries[c] = in->tables[c].entries;
        }
    }
    for (int c = nc; c < kWMaxComp; ++c) {
        a.has_bc[c] = 0;
        a.heads[c] = nullptr;
        a.entries[c] = nullptr;
    }
    a.nzc = a.zchunk = a.tiles_per_comp = 0;
    W5Tune over;
    if (tile) {
        over.wg = tile[0];
        over.zchunk = tile[1];
    }
    if (in->desc.dtype == CNV_F64)
        return w5_run<double>(a, st, tile ? &over : nullptr);
    return w5_run<float>(a, st, tile ? &over : nullptr);
}

template <typename T>
static bool w5_valid_rt(int edge, int wg)
{
    auto pick = [&](auto e) -> bool {
        constexpr int E = decltype(e)::value;
        return wg == 64    ? w5_valid<T, E, 64>()
               : wg == 128 ? w5_valid<T, E, 128>()
               : wg == 256 ? w5_valid<T, E, 256>()
               : wg == 512 ? w5_valid<T, E, 512>()
                           : false;
    };
    if (edge == 16)
        return pick(std::integral_constant<int, 16>());
    if (edge == 32)
        return pick(std::integral_constant<int, 32>());
    if (edge == 64)
        return pick(std::integral_constant<int, 64>());
    return false;
}

int weno5_set_tune(int dtype, int edge, int wg, int zchunk)
{
    if (edge != 16 && edge != 32 && edge != 64)
        return fail(CNV_ERR_INVALID, "WENO5 tiles exist for block edges 16, 32 and 64");
    if (zchunk < 1)
        return fail(CNV_ERR_INVALID, "zchunk must be >= 1");
    const bool ok = dtype == CNV_F64 ? w5_valid_rt<double>(edge, wg) : w5_valid_rt<float>(edge, wg);
    if (!ok)
        return fail(CNV_ERR_UNSUPPORTED, "WENO5 tile shape not available for this block edge and dtype");
    std::lock_guard<std::mutex> lk(g_w5_mtx);
    g_w5_tune[{dtype, edge}] = {wg, zchunk};
    return CNV_OK;
}

int weno5_get_tune(int dtype, int edge, int *wg, int *zchunk)
{
    const W5Tune t = w5_tune(dtype, edge);
    *wg = t.wg;
    *zchunk = t.zchunk;
    return CNV_OK;
}

// Times every (workgroup size, z-chunk) candidate on in -> out and keeps the
// fastest for the input's (dtype, block edge).
int autotune_weno5(cnv_grid_s *in, cnv_grid_s *out, int iters, int *best_wg, int *best_z, double *results,
                   int max_results, int *nresults)
{
    const int *n = in->desc.block_cells;
    if (!(n[0] == n[1] && (n[0] == 16 || n[0] == 32 || n[0] == 64)))
        return fail(CNV_ERR_UNSUPPORTED, "autotuning needs x/y block edges of 16, 32 or 64");
    if (iters < 1)
        return fail(CNV_ERR_INVALID, "iters must be >= 1");
    hipStream_t st;
    CNV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipEvent_t e0, e1;
    CNV_HIP(hipEventCreate(&e0));
    CNV_HIP(hipEventCreate(&e1));
    const int wgs[4] = {64, 128, 256, 512};
    const int zcs[5] = {4, 8, 16, 32, 64};
    float best = 1e30f;
    int bw = 0, bz = 0, cnt = 0, rc = CNV_OK;
    for (int wi = 0; wi < 4 && rc == CNV_OK; ++wi)
        for (int zi = 0; zi < 5; ++zi) {
            if (zcs[zi] > n[2])
                continue;
            const int tile[2] = {wgs[wi], zcs[zi]};
            int r = launch_weno5(in, out, st, tile);
            if (r == CNV_ERR_UNSUPPORTED)
                continue;
            if (r != CNV_OK) {
                rc = r;
                break;
            }
            (void)hipEventRecord(e0, st);
            for (int i = 0; i < iters; ++i)
                (void)launch_weno5(in, out, st, tile);
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(CNV_ERR_HIP, "autotune: kernel failed");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const float us = ms * 1000.f / iters;
            if (results && cnt < max_results) {
                results[3 * cnt] = tile[0];
                results[3 * cnt + 1] = tile[1];
                results[3 * cnt + 2] = us;
            }
            ++cnt;
            if (us < best) {
                best = us;
                bw = tile[0];
                bz = tile[1];
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (rc != CNV_OK)
        return rc;
    if (bw == 0)
        return fail(CNV_ERR_UNSUPPORTED, "no tile candidate ran");
    weno5_set_tune(in->desc.dtype, n[0], bw, bz);
    if (best_wg)
        *best_wg = bw;
    if (best_z)
        *best_z = bz;
    if (nresults)
        *nresults = cnt < max_results ? cnt : max_results;
    return CNV_OK;
}

} // namespace cnv
