// cnv_jacobi.hip -- weighted Jacobi on the variable-coefficient 7-point
// operator of a z-stretched grid, and its mixed-precision iterative
// refinement (FP32 sweeps, FP64 defect and correction).
//
// CubismNova has no solver and only names a stretched mesh class
// (include/Cubism/Common.h:86); the operator here is the flux form
//   (A u) = ihx2 ((u[+x] - 2u) + u[-x]) + ihy2 ((u[+y] - 2u) + u[-y])
//         + czp_k (u[+z] - u) - czm_k (u - u[-z]),
// summed left to right, with ghosts from the lab of Stencil(-1,2) and the
// blocks' BC lists (Dirichlet: constant ghost, BC/Dirichlet.h:75-113), and
//   sweep:  u' = u + (f - A u) * wd_k,  wd_k = omega / diag_k,
//           diag_k = -(((2 ihx2 + 2 ihy2) + czm_k) + czp_k)
//   defect: r  = f - A u  (+ FP64 sum of r^2 before any rounding of r).
// The z-march, LDS plane and register queue are those of the 7-point kernel.
#include "cnv_device.h"

#include <cmath>
#include <map>
#include <mutex>

namespace cnv
{

struct JArgs {
    const char *u;  // grid base of u (component cu)
    const char *f;  // grid base of f (component cf)
    char *out;      // grid base of the output (component co)
    Geom G;         // geometry of u
    long long cs_u, cs_f, cs_o; // component strides (elements) of the three grids
    long long bs_f, bs_o;       // block strides of f and out
    int cu, cf, co;
    int nzc, zchunk;
    int has_bc;
    const DevBCHead *heads;
    const DevBC *entries;
    const double *czm, *czp; // per global z cell
    double ihx2, ihy2, omega;
    double *partials;
};

// PF: f of plane z + 1 is loaded while plane z is evaluated (one plane of
// latency cover for the streamed operand, at 4 * R more registers)
template <typename T, typename TF, typename TO, int NX, int WG, int MODE, bool PF>
__global__ __launch_bounds__(WG) void k_jacobi(const JArgs a)
{
    constexpr int NY = NX;
    constexpr int VEC = 16 / (int)sizeof(T);
    constexpr int TX = NX / VEC;
    constexpr int ROWS = WG / TX;
    constexpr int R = NY / ROWS;
    constexpr int P = NX + 2 * VEC;
    constexpr int PLANE = (NY + 2) * P;
    static_assert(NX % VEC == 0 && WG % TX == 0 && NY % ROWS == 0, "tile shape");
    static_assert(WG >= 2 * NY && WG >= 2 * TX, "halo roles need enough threads");
    using V = VecT<T, VEC>;

    __shared__ alignas(16) T lds[2][PLANE];
    __shared__ double red[WG / 64];

    const int tid = threadIdx.x;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const long long blk = tile / a.nzc;
    const int zc = tile - (int)blk * a.nzc;
    const int NZ = a.G.n[2];
    const int z0 = zc * a.zchunk;
    const int z1 = min(NZ, z0 + a.zchunk);
    const int bz = (int)(blk / ((long long)a.G.nb[0] * a.G.nb[1]));

    const T *cbase = reinterpret_cast<const T *>(a.u) + (long long)a.cu * a.cs_u;
    const T *ub = cbase + blk * a.G.bstride;
    const TF *fb = reinterpret_cast<const TF *>(a.f) + (long long)a.cf * a.cs_f + blk * a.bs_f;
    TO *ob = reinterpret_cast<TO *>(a.out) + (long long)a.co * a.cs_o + blk * a.bs_o;

    FaceL<T> F[6];
    const int sb[3] = {-1, -1, -1}, se[3] = {2, 2, 2};
    make_faces<T>(a.G, a.heads, a.entries, a.has_bc, blk, cbase, sb, se, F);

    const int tx = tid % TX, ty = tid / TX;
    const int x0 = tx * VEC;
    constexpr long long SXY = (long long)NX * NY;
    const bool hx_on = tid < 2 * NY;
    const int hx_side = tid / NY, hx_y = tid % NY;
    const bool hy_on = tid < 2 * TX;
    const int hy_side = tid / TX, hy_x = (tid % TX) * VEC;
    const FaceL<T> FX = hx_side ? F[1] : F[0];
    const FaceL<T> FY = hy_side ? F[3] : F[2];
    const FaceL<T> FZ0 = F[4], FZ1 = F[5];

    using VF = VecT<TF, VEC>;
    struct Plane {
        V c[R];
        T hx;
        V hy;
    };
    // halo: plane p is computed by this workgroup (x/y ghosts needed)
    auto load_plane = [&](int p, bool halo, Plane &Q) {
        if (p >= 0 && p < NZ) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                Q.c[r] = *reinterpret_cast<const V *>(ub + x0 + NX * (ty + r * ROWS) + SXY * p);
            if (halo) {
                if (hx_on)
                    Q.hx = face_layer(FX, 1, 1, (long long)NX * hx_y + SXY * p);
                if (hy_on) {
                    if (FY.mode == 2) {
#pragma unroll
                        for (int v = 0; v < VEC; ++v)
                            Q.hy.v[v] = FY.s;
                    } else {
                        Q.hy = *reinterpret_cast<const V *>(FY.p + (long long)(FY.li0 + FY.lstep) * NX + hy_x + SXY * p);
                        if (FY.mode == 1)
#pragma unroll
                            for (int v = 0; v < VEC; ++v)
                                Q.hy.v[v] = FY.s * Q.hy.v[v];
                    }
                }
            }
        } else {
            const FaceL<T> f = p < 0 ? FZ0 : FZ1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (f.mode == 2) {
#pragma unroll
                    for (int v = 0; v < VEC; ++v)
                        Q.c[r].v[v] = f.s;
                } else {
                    Q.c[r] = *reinterpret_cast<const V *>(f.p + (long long)(f.li0 + f.lstep) * SXY + x0 +
                                                          NX * (ty + r * ROWS));
                    if (f.mode == 1)
#pragma unroll
                        for (int v = 0; v < VEC; ++v)
                            Q.c[r].v[v] = f.s * Q.c[r].v[v];
                }
            }
        }
    };

    const T ihx2 = (T)a.ihx2, ihy2 = (T)a.ihy2, w = (T)a.omega;
    double acc = 0.0;
    Plane pm, pc, pn, pnn;
    load_plane(z0 - 1, false, pm);
    load_plane(z0, true, pc);
    load_plane(z0 + 1, z0 + 1 < z1, pn);
    VF fr[R], fn[R];
    if (PF)
#pragma unroll
        for (int r = 0; r < R; ++r)
            fn[r] = load_nt<TF, VEC>(fb + x0 + NX * (ty + r * ROWS) + SXY * z0);
    for (int z = z0; z < z1; ++z) {
        prio_ramp(z - z0, z1 - z0);
        if (PF) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                fr[r] = fn[r];
        }
        if (z + 2 <= z1)
            load_plane(z + 2, z + 2 < z1, pnn);
        if (PF) {
            if (z + 1 < z1)
#pragma unroll
                for (int r = 0; r < R; ++r)
                    fn[r] = load_nt<TF, VEC>(fb + x0 + NX * (ty + r * ROWS) + SXY * (z + 1));
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r)
                fr[r] = load_nt<TF, VEC>(fb + x0 + NX * (ty + r * ROWS) + SXY * z);
        }
        T *L = lds[(z - z0) & 1];
#pragma unroll
        for (int r = 0; r < R; ++r)
            *reinterpret_cast<V *>(L + (ty + r * ROWS + 1) * P + VEC + x0) = pc.c[r];
        if (hx_on)
            L[(hx_y + 1) * P + (hx_side ? VEC + NX : VEC - 1)] = pc.hx;
        if (hy_on)
            *reinterpret_cast<V *>(L + (hy_side ? NY + 1 : 0) * P + VEC + hy_x) = pc.hy;
        __syncthreads();
        const int kz = bz * NZ + z;
        const T cm = (T)a.czm[kz], cp = (T)a.czp[kz];
        const T diag = -(2 * ihx2 + 2 * ihy2 + cm + cp);
        const T wd = w / diag;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int y = ty + r * ROWS;
            const T *row = L + (y + 1) * P + VEC;
            const V ym = *reinterpret_cast<const V *>(row - P + x0);
            const V yp = *reinterpret_cast<const V *>(row + P + x0);
            const T xl = row[x0 - 1];
            const T xr = row[x0 + VEC];
            VecT<TO, VEC> res;
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
                const T c = pc.c[r].v[v];
                const T xm = v == 0 ? xl : pc.c[r].v[v - 1];
                const T xp = v == VEC - 1 ? xr : pc.c[r].v[v + 1];
                const T ax = ihx2 * (fms2(c, xp) + xm);
                const T ay = ihy2 * (fms2(c, yp.v[v]) + ym.v[v]);
                const T az = cp * (pn.c[r].v[v] - c) - cm * (c - pm.c[r].v[v]);
                const T au = ax + ay + az;
                const T rr = (T)fr[r].v[v] - au;
                if (MODE == 0) {
                    res.v[v] = (TO)(c + rr * wd);
                } else {
                    res.v[v] = (TO)rr;
                    acc = __builtin_fma((double)rr, (double)rr, acc);
                }
            }
            store_nt<TO, VEC>(ob + x0 + NX * y + SXY * z, res);
        }
        pm = pc;
        pc = pn;
        pn = pnn;
    }
    if (MODE == 1) {
        for (int o = 32; o > 0; o >>= 1)
            acc += __shfl_xor(acc, o, 64);
        if ((tid & 63) == 0)
            red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int wv = 0; wv < WG / 64; ++wv)
                s += red[wv];
            a.partials[tile] = s;
        }
    }
}

// Sum of n partials in a fixed order (one workgroup of 1024 threads).
__global__ __launch_bounds__(1024) void k_jfinalize(const double *__restrict__ partials, int n, double *__restrict__ out)
{
    __shared__ double red[16];
    const int tid = threadIdx.x;
    double s = 0.0;
    for (int base = 0; base < n; base += 8 * 1024) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int i = base + j * 1024 + tid;
            v[j] = i < n ? partials[i] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
            s += v[j];
    }
    for (int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o, 64);
    if ((tid & 63) == 0)
        red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w)
            t += red[w];
        *out = t;
    }
}

// u64 += e (any dtype of e), and e-fill with zero
template <typename TE>
__global__ __launch_bounds__(256) void k_correct(double *u, const TE *e, long long bs_u, long long bs_e, long long cells,
                                                   long long nbt)
{
    const long long n = cells * nbt;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const long long b = i / cells, c = i - b * cells;
        u[b * bs_u + c] += (double)e[b * bs_e + c];
    }
}

// ---------------------------------------------------------------- host side

struct JTune {
    int wg, zchunk, pf;
};

template <typename T, int NX, int WG>
constexpr bool j_valid()
{
    constexpr int VEC = 16 / (int)sizeof(T);
    constexpr int TX = NX / VEC;
    // more than eight rows per thread spill registers to scratch memory
    return NX % VEC == 0 && WG % TX == 0 && (WG / TX) <= NX && NX % (WG / TX) == 0 && WG >= 2 * NX &&
           WG >= 2 * TX && NX / (WG / TX) <= 8;
}

// the prefetch variant exists for sweeps (MODE 0) only
template <typename T, typename TF, typename TO, int NX, int WG, int MODE>
static bool j_launch(const JArgs &a, int ntiles, bool pf, hipStream_t st)
{
    if constexpr (j_valid<T, NX, WG>()) {
        if constexpr (MODE == 0) {
            if (pf) {
                hipLaunchKernelGGL((k_jacobi<T, TF, TO, NX, WG, MODE, true>), dim3(ntiles), dim3(WG), 0, st, a);
                return true;
            }
        }
        hipLaunchKernelGGL((k_jacobi<T, TF, TO, NX, WG, MODE, false>), dim3(ntiles), dim3(WG), 0, st, a);
        return true;
    }
    return false;
}

template <typename T, typename TF, typename TO, int MODE>
static bool j_dispatch(const JArgs &a, int edge, int wg, bool pf, int ntiles, hipStream_t st)
{
#define J_WG(NX)                                                                                        \
    switch (wg) {                                                                                       \
    case 64:                                                                                            \
        return j_launch<T, TF, TO, NX, 64, MODE>(a, ntiles, pf, st);                                    \
    case 128:                                                                                           \
        return j_launch<T, TF, TO, NX, 128, MODE>(a, ntiles, pf, st);                                   \
    case 256:                                                                                           \
        return j_launch<T, TF, TO, NX, 256, MODE>(a, ntiles, pf, st);                                   \
    case 512:                                                                                           \
        return j_launch<T, TF, TO, NX, 512, MODE>(a, ntiles, pf, st);                                   \
    default:                                                                                            \
        return false;                                                                                   \
    }
    switch (edge) {
    case 16:
        J_WG(16)
    case 32:
        J_WG(32)
    case 64:
        J_WG(64)
    default:
        return false;
    }
#undef J_WG
}

static JTune j_default(int dtype, int edge)
{
    if (edge >= 64)
        return {512, 16, dtype == CNV_F64 ? 1 : 0};
    if (edge >= 32)
        return {256, 16, 0};
    return {dtype == CNV_F64 ? 128 : 64, 16, 0};
}

static std::mutex g_j_mtx;
static std::map<std::pair<int, int>, JTune> g_j_tune; // (dtype, x/y block edge) -> choice

static JTune j_tune(int dtype, int edge)
{
    std::lock_guard<std::mutex> lk(g_j_mtx);
    auto it = g_j_tune.find({dtype, edge});
    return it == g_j_tune.end() ? j_default(dtype, edge) : it->second;
}

// does a Jacobi tile of `wg` threads exist for this dtype and x/y edge
static bool j_tile_ok(int dtype, int edge, int wg)
{
    auto chk = [&](auto tag) {
        using T = decltype(tag);
        switch (edge) {
        case 16:
            return wg == 64 ? j_valid<T, 16, 64>() : wg == 128 ? j_valid<T, 16, 128>()
                   : wg == 256 ? j_valid<T, 16, 256>() : wg == 512 ? j_valid<T, 16, 512>() : false;
        case 32:
            return wg == 64 ? j_valid<T, 32, 64>() : wg == 128 ? j_valid<T, 32, 128>()
                   : wg == 256 ? j_valid<T, 32, 256>() : wg == 512 ? j_valid<T, 32, 512>() : false;
        case 64:
            return wg == 64 ? j_valid<T, 64, 64>() : wg == 128 ? j_valid<T, 64, 128>()
                   : wg == 256 ? j_valid<T, 64, 256>() : wg == 512 ? j_valid<T, 64, 512>() : false;
        default:
            return false;
        }
    };
    return dtype == CNV_F64 ? chk(double()) : chk(float());
}

int jacobi_set_tune(int dtype, int edge, int wg, int zchunk, int prefetch)
{
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    if (zchunk < 1)
        return fail(CNV_ERR_INVALID, "zchunk must be >= 1");
    if (!j_tile_ok(dtype, edge, wg))
        return fail(CNV_ERR_UNSUPPORTED, "no Jacobi tile of this workgroup size for this dtype and block edge");
    std::lock_guard<std::mutex> lk(g_j_mtx);
    g_j_tune[{dtype, edge}] = JTune{wg, zchunk, prefetch ? 1 : 0};
    return CNV_OK;
}

int jacobi_get_tune(int dtype, int edge, int *wg, int *zchunk, int *prefetch)
{
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    const JTune t = j_tune(dtype, edge);
    if (wg)
        *wg = t.wg;
    if (zchunk)
        *zchunk = t.zchunk;
    if (prefetch)
        *prefetch = t.pf;
    return CNV_OK;
}

// metric of the grid's z direction: uniform (1/h^2) unless faces were set
int ensure_zmetric(cnv_grid_s *g)
{
    if (g->czm_dev)
        return CNV_OK;
    const int nz = g->desc.nblocks[2] * g->desc.block_cells[2];
    std::vector<double> zf(nz + 1);
    for (int j = 0; j <= nz; ++j)
        zf[j] = g->desc.begin[2] + j * g->h[2];
    return set_zfaces(g, zf.data(), nz + 1);
}

int set_zfaces(cnv_grid_s *g, const double *zf, int n)
{
    const int nz = g->desc.nblocks[2] * g->desc.block_cells[2];
    if (n != nz + 1)
        return fail(CNV_ERR_SIZE, "need nz + 1 face positions");
    for (int j = 0; j < nz; ++j)
        if (!(zf[j + 1] > zf[j]))
            return fail(CNV_ERR_INVALID, "face positions must increase");
    std::vector<double> dz(nz), zc(nz + 2), czm(nz), czp(nz);
    for (int k = 0; k < nz; ++k) {
        dz[k] = zf[k + 1] - zf[k];
        zc[k + 1] = 0.5 * (zf[k + 1] + zf[k]);
    }
    zc[0] = 2 * zf[0] - zc[1];
    zc[nz + 1] = 2 * zf[nz] - zc[nz];
    for (int k = 0; k < nz; ++k) {
        czm[k] = 1.0 / (dz[k] * (zc[k + 1] - zc[k]));
        czp[k] = 1.0 / (dz[k] * (zc[k + 2] - zc[k + 1]));
    }
    if (!g->czm_dev) {
        CNV_HIP(hipMalloc((void **)&g->czm_dev, sizeof(double) * nz));
        CNV_HIP(hipMalloc((void **)&g->czp_dev, sizeof(double) * nz));
    }
    CNV_HIP(hipMemcpy(g->czm_dev, czm.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    CNV_HIP(hipMemcpy(g->czp_dev, czp.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    return CNV_OK;
}

// mode 0: sweep out = u + omega (f - Au)/diag (u, f, out of one dtype)
// mode 1: defect out = f - A u (u, f FP64 or both FP32; out any dtype) + sumsq
int launch_jacobi(cnv_grid_s *u, int cu, cnv_grid_s *f, int cf, cnv_grid_s *out, int co, int mode, double omega,
                  double *sumsq, hipStream_t st, const int *tile)
{
    const int *n = u->desc.block_cells;
    if (!(n[0] == n[1] && (n[0] == 16 || n[0] == 32 || n[0] == 64)))
        return fail(CNV_ERR_UNSUPPORTED, "Jacobi kernels need x/y block edges of 16, 32 or 64");
    if (f->desc.dtype != u->desc.dtype)
        return fail(CNV_ERR_SIZE, "u and f must share a dtype");
    if (mode == 0 && out->desc.dtype != u->desc.dtype)
        return fail(CNV_ERR_SIZE, "sweep output must share the dtype of u");
    JArgs a;
    {
        std::lock_guard<std::mutex> lk(u->mtx);
        int rc = ensure_zmetric(u);
        if (rc != CNV_OK)
            return rc;
        rc = ensure_bc_table(u, cu);
        if (rc != CNV_OK)
            return rc;
        const int sb[3] = {-1, -1, -1}, se[3] = {2, 2, 2};
        for (const auto &l : u->bcs[cu]) {
            if (l.empty())
                continue;
            std::vector<DevBC> tmp;
            for (const auto &b : l)
                tmp.push_back({b.kind, b.dir, b.side, 0, b.value});
            int kind[6];
            double val[6];
            resolve_faces(tmp.data(), (int)tmp.size(), sb, se, kind, val);
            for (int k = 0; k < 6; ++k)
                if (kind[k] == FACE_UNSET)
                    return fail(CNV_ERR_UNSUPPORTED, "BC list leaves ghost cells unwritten");
        }
        const BCTable &t = u->tables[cu];
        a.has_bc = (t.any && t.heads && t.entries) ? 1 : 0;
        a.heads = t.heads;
        a.entries = t.entries;
    }
    a.u = static_cast<const char *>(u->data);
    a.f = static_cast<const char *>(f->data);
    a.out = static_cast<char *>(out->data);
    a.G = geom_of(u);
    a.cs_u = (long long)(u->component_bytes / u->elem);
    a.cs_f = (long long)(f->component_bytes / f->elem);
    a.cs_o = (long long)(out->component_bytes / out->elem);
    a.bs_f = (long long)(f->block_bytes / f->elem);
    a.bs_o = (long long)(out->block_bytes / out->elem);
    a.cu = cu;
    a.cf = cf;
    a.co = co;
    a.czm = u->czm_dev;
    a.czp = u->czp_dev;
    a.ihx2 = 1. / (u->h[0] * u->h[0]);
    a.ihy2 = 1. / (u->h[1] * u->h[1]);
    a.omega = omega;
    a.partials = u->partials;
    const JTune tu = tile ? JTune{tile[0], tile[1], tile[2]} : j_tune(u->desc.dtype, n[0]);
    a.zchunk = tu.zchunk < n[2] ? tu.zchunk : n[2];
    a.nzc = (n[2] + a.zchunk - 1) / a.zchunk;
    const long long ntiles = u->nblocks_total * a.nzc;
    if (ntiles > (long long)u->partials_cap || ntiles >= (1LL << 31))
        return fail(CNV_ERR_UNSUPPORTED, "grid too large for the Jacobi kernels");
    bool ok;
    const bool dbl = u->desc.dtype == CNV_F64;
    const bool pf = tu.pf != 0;
    if (mode == 0)
        ok = dbl ? j_dispatch<double, double, double, 0>(a, n[0], tu.wg, pf, (int)ntiles, st)
                 : j_dispatch<float, float, float, 0>(a, n[0], tu.wg, pf, (int)ntiles, st);
    else if (dbl)
        ok = out->desc.dtype == CNV_F64 ? j_dispatch<double, double, double, 1>(a, n[0], tu.wg, pf, (int)ntiles, st)
                                        : j_dispatch<double, double, float, 1>(a, n[0], tu.wg, pf, (int)ntiles, st);
    else
        ok = out->desc.dtype == CNV_F64 ? j_dispatch<float, float, double, 1>(a, n[0], tu.wg, pf, (int)ntiles, st)
                                        : j_dispatch<float, float, float, 1>(a, n[0], tu.wg, pf, (int)ntiles, st);
    if (!ok)
        return fail(CNV_ERR_UNSUPPORTED, "no Jacobi tile for this block edge");
    CNV_HIP(hipGetLastError());
    if (mode == 1 && sumsq) {
        hipLaunchKernelGGL(k_jfinalize, dim3(1), dim3(1024), 0, st, (const double *)u->partials, (int)ntiles, sumsq);
        CNV_HIP(hipGetLastError());
    }
    return CNV_OK;
}

// Times every (workgroup size, z-planes per tile, f prefetch) candidate of
// the sweep u -> out (iters launches each), keeps the fastest for (dtype of
// u, x/y block edge).
int autotune_jacobi(cnv_grid_s *u, cnv_grid_s *f, cnv_grid_s *out, int iters, int *best_wg, int *best_z,
                    int *best_pf, double *results, int max_results, int *nresults)
{
    const int *n = u->desc.block_cells;
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
    const int zcs[4] = {8, 16, 32, 64};
    float best = 1e30f;
    int bw = 0, bz = 0, bp = 0, cnt = 0, rc = CNV_OK;
    for (int wi = 0; wi < 4 && rc == CNV_OK; ++wi)
        for (int zi = 0; zi < 4 && rc == CNV_OK; ++zi)
            for (int pf = 0; pf < 2; ++pf) {
                if (zcs[zi] > n[2] || !j_tile_ok(u->desc.dtype, n[0], wgs[wi]))
                    continue;
                const int tile[3] = {wgs[wi], zcs[zi], pf};
                int r = launch_jacobi(u, 0, f, 0, out, 0, 0, 2.0 / 3.0, nullptr, st, tile);
                if (r == CNV_ERR_UNSUPPORTED)
                    continue;
                if (r != CNV_OK) {
                    rc = r;
                    break;
                }
                (void)hipEventRecord(e0, st);
                for (int i = 0; i < iters; ++i)
                    (void)launch_jacobi(u, 0, f, 0, out, 0, 0, 2.0 / 3.0, nullptr, st, tile);
                (void)hipEventRecord(e1, st);
                if (hipEventSynchronize(e1) != hipSuccess) {
                    rc = fail(CNV_ERR_HIP, "autotune: kernel failed");
                    break;
                }
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                const float us = ms * 1000.f / iters;
                if (results && cnt < max_results) {
                    results[4 * cnt] = tile[0];
                    results[4 * cnt + 1] = tile[1];
                    results[4 * cnt + 2] = tile[2];
                    results[4 * cnt + 3] = us;
                }
                ++cnt;
                if (us < best) {
                    best = us;
                    bw = tile[0];
                    bz = tile[1];
                    bp = tile[2];
                }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    if (rc != CNV_OK)
        return rc;
    if (bw == 0)
        return fail(CNV_ERR_UNSUPPORTED, "no tile candidate ran");
    {
        std::lock_guard<std::mutex> lk(g_j_mtx);
        g_j_tune[{u->desc.dtype, n[0]}] = JTune{bw, bz, bp};
    }
    if (best_wg)
        *best_wg = bw;
    if (best_z)
        *best_z = bz;
    if (best_pf)
        *best_pf = bp;
    if (nresults)
        *nresults = cnt < max_results ? cnt : max_results;
    return CNV_OK;
}

int launch_correct(cnv_grid_s *u, int cu, cnv_grid_s *e, int ce, hipStream_t st)
{
    const long long cells = (long long)u->desc.block_cells[0] * u->desc.block_cells[1] * u->desc.block_cells[2];
    const long long nbt = u->nblocks_total;
    const long long nwg = std::min<long long>((cells * nbt + 255) / 256, 8192);
    double *ub = reinterpret_cast<double *>((char *)u->data + (size_t)cu * u->component_bytes);
    const long long bs_u = (long long)(u->block_bytes / 8);
    if (e->desc.dtype == CNV_F64)
        hipLaunchKernelGGL(k_correct<double>, dim3((unsigned)nwg), dim3(256), 0, st, ub,
                           reinterpret_cast<const double *>((char *)e->data + (size_t)ce * e->component_bytes), bs_u,
                           (long long)(e->block_bytes / 8), cells, nbt);
    else
        hipLaunchKernelGGL(k_correct<float>, dim3((unsigned)nwg), dim3(256), 0, st, ub,
                           reinterpret_cast<const float *>((char *)e->data + (size_t)ce * e->component_bytes), bs_u,
                           (long long)(e->block_bytes / 4), cells, nbt);
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

} // namespace cnv

namespace cnv
{
// Everything a Jacobi launch on (g, comp) would set up on first use (z
// metric, BC table), done ahead of a graph capture.
int jacobi_prepare(cnv_grid_s *g, int comp)
{
    std::lock_guard<std::mutex> lk(g->mtx);
    int rc = ensure_zmetric(g);
    if (rc != CNV_OK)
        return rc;
    return ensure_bc_table(g, comp);
}
} // namespace cnv
