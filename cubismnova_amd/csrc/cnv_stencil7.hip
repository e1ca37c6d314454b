// cnv_stencil7.hip -- 7-point Laplacian / forward-Euler diffusion over the
// ghost-extended blocks of a grid, with the lab assembled on the fly.
//
// Layout of the work: a workgroup owns one x-y plane of one block and marches
// through a chunk of z-planes.  Each plane (plus its one-cell x/y halo, read
// straight from the face-neighbour blocks or filled from the block's BCs) is
// staged in LDS; the z neighbours come from a register queue of planes that is
// loaded two planes ahead, so HBM reads of the next planes overlap the
// arithmetic of the current one.  No lab buffer is ever written to HBM.
//
// Arithmetic follows test/sandbox/Benchmarks/Kernel/Laplacian/Laplacian.cpp
// :125-129 operation for operation: ddx = ihx2*((u[+x] - 2u) + u[-x]), same
// for y and z, lap = (ddx + ddy) + ddz; the Euler update is lap*nu_dt then
// u + that (Block/Field.h:457-512, FieldOperator.h:43-114).  The library is
// compiled with -ffp-contract=off so no step is fused into an FMA and FP64
// results are bit-identical to the CPU evaluation of the same order.
#include "cnv_march.h"

#include <cstdlib>
#include <map>
#include <tuple>

namespace cnv
{

constexpr int kMaxComp = 8;

struct S7Args {
    const char *u;   // input grid base (component slabs)
    char *out;       // output grid base
    char *lap;       // optional Laplacian grid base
    Geom G;
    int comp_in, comp_out;
    int ncomp;
    int nzc;         // z-chunks per block
    int zchunk;      // planes per chunk
    int tiles_per_comp;
    int has_bc[kMaxComp];
    const DevBCHead *heads[kMaxComp];
    const DevBC *entries[kMaxComp];
    double ih2[3];
    double nudt;
    double *partials;
    double *sumsq;
    // deferred norm of the previous step of a multi-step run: workgroup 0
    // sums prev_n[k] partials (written by the previous launch) into
    // *prev_out[k], k = 0, 1 (unused sets have prev_p[k] == nullptr)
    const double *prev_p[2];
    int prev_n[2];
    double *prev_out[2];
};

template <typename T, int NX, int NY, int WG, int MODE, bool SUMSQ>
__global__ __launch_bounds__(WG) void k_stencil7(const S7Args a)
{
    using S = S7Shape<T, NX, NY, WG>;
    __shared__ alignas(16) T lds[2 * S::PLANE];
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
    const int NZ = a.G.n[2];
    const int z0 = zc * a.zchunk;
    const int z1 = min(NZ, z0 + a.zchunk);

    const T *cbase = reinterpret_cast<const T *>(a.u) + (long long)(a.comp_in + comp) * a.G.cstride;
    const T *ub = cbase + blk * a.G.bstride;
    T *ob = reinterpret_cast<T *>(a.out) + (long long)(a.comp_out + comp) * a.G.cstride + blk * a.G.bstride;
    T *lb = MODE == 2 ? reinterpret_cast<T *>(a.lap) + (long long)(a.comp_out + comp) * a.G.cstride +
                            blk * a.G.bstride
                      : nullptr;

    FaceSrc<T> F[6];
    block_faces<T>(a.G, a.heads[comp], a.entries[comp], a.has_bc[comp], blk, cbase, F);
    double acc = s7_march<T, NX, NY, WG, MODE, SUMSQ>(ub, ob, lb, F, NZ, z0, z1, (T)a.ih2[0], (T)a.ih2[1],
                                                      (T)a.ih2[2], (T)a.nudt, lds);

    if (SUMSQ) {
        // per-tile partial in a fixed wave tree and wave order; the tiles are
        // summed in tile order by k_finalize_sum after this kernel
        for (int o = 32; o > 0; o >>= 1)
            acc += __shfl_xor(acc, o, 64);
        if ((tid & 63) == 0)
            red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < WG / 64; ++w)
                s += red[w];
            a.partials[tile] = s;
        }
    }
}

// Sum of n partials in a fixed order (one workgroup of 1024 threads): every
// thread issues its loads for a round of 8 before adding them in index order,
// then a fixed wave tree and the waves in order.  Deterministic run to run.
__global__ __launch_bounds__(1024) void k_finalize_sum(const double *__restrict__ partials, int n,
                                                       double *__restrict__ out)
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

// Generic fallback (any block shape): one thread per cell, lab values read
// straight from the block or through the resolved faces.
template <typename T, int MODE, bool SUMSQ>
__global__ __launch_bounds__(256) void k_stencil7_generic(const S7Args a)
{
    __shared__ double red[4];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && blockIdx.y == 0)
        for (int k = 0; k < 2; ++k)
            if (a.prev_p[k])
                wg_sum_partials<256>(a.prev_p[k], a.prev_n[k], a.prev_out[k], red);
    const long long nb_tot = (long long)a.G.nb[0] * a.G.nb[1] * a.G.nb[2];
    const long long slot = blockIdx.y; // comp * nblocks + blk
    const int comp = (int)(slot / nb_tot);
    const long long blk = slot - comp * nb_tot;
    const T *cbase = reinterpret_cast<const T *>(a.u) + (long long)(a.comp_in + comp) * a.G.cstride;
    const T *ub = cbase + blk * a.G.bstride;
    T *ob = reinterpret_cast<T *>(a.out) + (long long)(a.comp_out + comp) * a.G.cstride + blk * a.G.bstride;
    FaceSrc<T> F[6];
    block_faces<T>(a.G, a.heads[comp], a.entries[comp], a.has_bc[comp], blk, cbase, F);
    const int *n = a.G.n;
    const long long cells = (long long)n[0] * n[1] * n[2];
    const long long i = (long long)blockIdx.x * 256 + tid;
    double acc = 0.0;
    if (i < cells) {
        const int x = (int)(i % n[0]);
        const int y = (int)((i / n[0]) % n[1]);
        const int z = (int)(i / ((long long)n[0] * n[1]));
        const long long nx = n[0], nxy = (long long)n[0] * n[1];
        auto at = [&](int xx, int yy, int zz) -> T {
            if (xx < 0)
                return face_val(F[0], nx * yy + nxy * zz);
            if (xx >= n[0])
                return face_val(F[1], nx * yy + nxy * zz);
            if (yy < 0)
                return face_val(F[2], xx + nxy * zz);
            if (yy >= n[1])
                return face_val(F[3], xx + nxy * zz);
            if (zz < 0)
                return face_val(F[4], xx + nx * yy);
            if (zz >= n[2])
                return face_val(F[5], xx + nx * yy);
            return ub[xx + nx * yy + nxy * zz];
        };
        const T ihx2 = (T)a.ih2[0], ihy2 = (T)a.ih2[1], ihz2 = (T)a.ih2[2];
        const T c = ub[i];
        const T ddx = ihx2 * (fms2(c, at(x + 1, y, z)) + at(x - 1, y, z));
        const T ddy = ihy2 * (fms2(c, at(x, y + 1, z)) + at(x, y - 1, z));
        const T ddz = ihz2 * (fms2(c, at(x, y, z + 1)) + at(x, y, z - 1));
        const T lap = ddx + ddy + ddz;
        if (MODE == 0) {
            ob[i] = lap;
        } else {
            const T l = lap * (T)a.nudt;
            const T un = c + l;
            ob[i] = un;
            if (MODE == 2)
                reinterpret_cast<T *>(a.lap)[(long long)(a.comp_out + comp) * a.G.cstride + blk * a.G.bstride + i] = lap;
            acc = (double)un * (double)un;
        }
    }
    if (SUMSQ) {
        for (int o = 32; o > 0; o >>= 1)
            acc += __shfl_xor(acc, o, 64);
        if ((tid & 63) == 0)
            red[tid >> 6] = acc;
        __syncthreads();
        if (tid == 0)
            a.partials[(long long)blockIdx.y * gridDim.x + blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
    }
}

// ---------------------------------------------------------------- host side

// Tile choice per (dtype, block edge): workgroup size and z-planes per tile.
struct S7Tune {
    int wg;
    int zchunk;
};

static std::mutex g_tune_mtx;
static std::map<std::tuple<int, int>, S7Tune> g_tune; // (dtype, edge) -> choice

S7Tune s7_default_tune(int dtype, int edge)
{
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    auto it = g_tune.find({dtype, edge});
    if (it != g_tune.end())
        return it->second;
    if (edge >= 64)
        return {512, 16};
    if (edge >= 32)
        return {256, 8};
    if (dtype == CNV_F64)
        return {128, 8};
    return {64, 8};
}

void s7_set_tune(int dtype, int edge, int wg, int zchunk)
{
    std::lock_guard<std::mutex> lk(g_tune_mtx);
    g_tune[{dtype, edge}] = {wg, zchunk};
}

template <typename T, int NX, int WG>
constexpr bool s7_valid()
{
    return S7Shape<T, NX, NX, WG>::valid;
}

template <typename T, int NX, int WG, int MODE, bool SUMSQ>
static bool launch_spec(const S7Args &a, int ntiles, hipStream_t st)
{
    if constexpr (s7_valid<T, NX, WG>()) {
        hipLaunchKernelGGL((k_stencil7<T, NX, NX, WG, MODE, SUMSQ>), dim3(ntiles), dim3(WG), 0, st, a);
        return true;
    }
    return false;
}

template <typename T, int NX, int MODE, bool SUMSQ>
static bool dispatch_wg(const S7Args &a, int wg, int ntiles, hipStream_t st)
{
    switch (wg) {
    case 64:
        return launch_spec<T, NX, 64, MODE, SUMSQ>(a, ntiles, st);
    case 128:
        return launch_spec<T, NX, 128, MODE, SUMSQ>(a, ntiles, st);
    case 256:
        return launch_spec<T, NX, 256, MODE, SUMSQ>(a, ntiles, st);
    case 512:
        return launch_spec<T, NX, 512, MODE, SUMSQ>(a, ntiles, st);
    default:
        return false;
    }
}

template <typename T, int MODE, bool SUMSQ>
static bool dispatch_spec(const S7Args &a, int edge, int wg, int ntiles, hipStream_t st)
{
    switch (edge) {
    case 16:
        return dispatch_wg<T, 16, MODE, SUMSQ>(a, wg, ntiles, st);
    case 32:
        return dispatch_wg<T, 32, MODE, SUMSQ>(a, wg, ntiles, st);
    case 64:
        return dispatch_wg<T, 64, MODE, SUMSQ>(a, wg, ntiles, st);
    default:
        return false;
    }
}

// cap: partials the buffer at a.partials holds (0: sized by the caller for
// the largest launch)
template <typename T, int MODE, bool SUMSQ>
static int run_t(S7Args &a, int ncomp, hipStream_t st, int force_generic, const S7Tune *over, bool finalize,
                 int *nparts, long long cap = 0)
{
    const int *n = a.G.n;
    const long long nbt = (long long)a.G.nb[0] * a.G.nb[1] * a.G.nb[2];
    const bool cubic_xy = n[0] == n[1];
    if (!force_generic && cubic_xy && (n[0] == 16 || n[0] == 32 || n[0] == 64)) {
        S7Tune tu = over ? *over : s7_default_tune(sizeof(T) == 8 ? CNV_F64 : CNV_F32, n[0]);
        a.zchunk = tu.zchunk < n[2] ? tu.zchunk : n[2];
        a.nzc = (n[2] + a.zchunk - 1) / a.zchunk;
        a.tiles_per_comp = (int)(nbt * a.nzc);
        const long long ntiles = (long long)a.tiles_per_comp * ncomp;
        if (SUMSQ && cap > 0 && ntiles > cap)
            return fail(CNV_ERR_SIZE, "grid needs more norm partials than the scratch array holds");
        if (ntiles < (1LL << 31)) {
            if (dispatch_spec<T, MODE, SUMSQ>(a, n[0], tu.wg, (int)ntiles, st)) {
                CNV_HIP(hipGetLastError());
                if (nparts)
                    *nparts = (int)ntiles;
                if (SUMSQ && finalize) {
                    hipLaunchKernelGGL(k_finalize_sum, dim3(1), dim3(1024), 0, st, a.partials, (int)ntiles, a.sumsq);
                    CNV_HIP(hipGetLastError());
                }
                return CNV_OK;
            }
            if (over)
                return fail(CNV_ERR_UNSUPPORTED, "tile shape not available for this block edge and dtype");
        }
    }
    const long long cells = (long long)n[0] * n[1] * n[2];
    const long long gx = (cells + 255) / 256;
    const long long gy = nbt * ncomp;
    if (gy > 65535 || gx > 65535)
        return fail(CNV_ERR_UNSUPPORTED, "grid too large for the generic stencil kernel");
    if (SUMSQ && cap > 0 && gx * gy > cap)
        return fail(CNV_ERR_SIZE, "grid needs more norm partials than the scratch array holds");
    hipLaunchKernelGGL((k_stencil7_generic<T, MODE, SUMSQ>), dim3((unsigned)gx, (unsigned)gy), dim3(256),
                       0, st, a);
    CNV_HIP(hipGetLastError());
    if (nparts)
        *nparts = (int)(gx * gy);
    if (SUMSQ && finalize) {
        hipLaunchKernelGGL(k_finalize_sum, dim3(1), dim3(1024), 0, st, a.partials, (int)(gx * gy), a.sumsq);
        CNV_HIP(hipGetLastError());
    }
    return CNV_OK;
}

static int g_force_generic = -1;

static bool force_generic()
{
    if (g_force_generic < 0) {
        const char *e = getenv("CNV_STENCIL7_GENERIC");
        g_force_generic = (e && e[0] == '1') ? 1 : 0;
    }
    return g_force_generic == 1;
}

// Host check that every face the operator reads is written by the lab rules.
static int check_faces(cnv_grid_s *g, int comp)
{
    const int sb[3] = {-1, -1, -1}, se[3] = {2, 2, 2};
    DevBC tmp[64];
    for (const auto &l : g->bcs[comp]) {
        if (l.empty())
            continue;
        if (l.size() > 64)
            return fail(CNV_ERR_UNSUPPORTED, "BC list too long");
        for (size_t k = 0; k < l.size(); ++k)
            tmp[k] = {l[k].kind, l[k].dir, l[k].side, 0, l[k].value};
        int kind[6];
        double val[6];
        resolve_faces(tmp, (int)l.size(), sb, se, kind, val);
        for (int f = 0; f < 6; ++f)
            if (kind[f] == FACE_UNSET)
                return fail(CNV_ERR_UNSUPPORTED,
                            "BC list leaves ghost cells of a face unwritten for Stencil(-1,2)");
    }
    return CNV_OK;
}

// State of one launch inside a multi-step run (cnv_diffusion_run).
struct S7Run {
    double *partials;          // this launch's partial buffer (first set)
    double *partials2;         // second set (two-step launches)
    const double *prev[2];     // the previous launch's partial sets, summed by workgroup 0
    int prev_n[2];
    double *prev_out[2];
    int *nparts;               // out: partials written per set by this launch
    size_t cap;                // capacity of each buffer half
};

// Partials of a launch that owns no grid scratch: `cap` doubles at `p`.
struct S7Scratch {
    double *p;
    long long cap;
};

static int launch_stencil7_t(cnv_grid_s *u, cnv_grid_s *out, cnv_grid_s *lap, int comp_in, int comp_out,
                             int ncomp, int mode, double nu_dt, double *sumsq, hipStream_t st,
                             const S7Tune *over, const S7Run *run = nullptr, const S7Scratch *scr = nullptr)
{
    if (ncomp > kMaxComp)
        return fail(CNV_ERR_UNSUPPORTED, "more than 8 components in one stencil call");
    for (int d = 0; d < 3; ++d)
        if (u->desc.block_cells[d] < 1)
            return fail(CNV_ERR_INVALID, "empty block");
    S7Args a;
    a.u = static_cast<const char *>(u->data);
    a.out = static_cast<char *>(out->data);
    a.lap = lap ? static_cast<char *>(lap->data) : nullptr;
    a.G = geom_of(u);
    a.comp_in = comp_in;
    a.comp_out = comp_out;
    a.ncomp = ncomp;
    for (int c = 0; c < ncomp; ++c) {
        a.has_bc[c] = 0;
        a.heads[c] = nullptr;
        a.entries[c] = nullptr;
    }
    if (!u->tables.empty()) { // a grid without tables is periodic on every face
        std::lock_guard<std::mutex> lk(u->mtx);
        for (int c = 0; c < ncomp; ++c) {
            int rc = ensure_bc_table(u, comp_in + c);
            if (rc != CNV_OK)
                return rc;
            rc = check_faces(u, comp_in + c);
            if (rc != CNV_OK)
                return rc;
            const BCTable &t = u->tables[comp_in + c];
            a.has_bc[c] = (t.any && t.heads && t.entries) ? 1 : 0;
            a.heads[c] = t.heads;
            a.entries[c] = t.entries;
        }
    }
    for (int c = ncomp; c < kMaxComp; ++c) {
        a.has_bc[c] = 0;
        a.heads[c] = nullptr;
        a.entries[c] = nullptr;
    }
    for (int d = 0; d < 3; ++d)
        a.ih2[d] = 1. / (u->h[d] * u->h[d]);
    a.nudt = nu_dt;
    a.sumsq = sumsq;
    a.partials = nullptr;
    for (int k = 0; k < 2; ++k) {
        a.prev_p[k] = run ? run->prev[k] : nullptr;
        a.prev_n[k] = run ? run->prev_n[k] : 0;
        a.prev_out[k] = run ? run->prev_out[k] : nullptr;
    }
    a.zchunk = a.nzc = a.tiles_per_comp = 0;
    if (sumsq) {
        const long long nbt = u->nblocks_total;
        const long long cells = (long long)u->desc.block_cells[0] * u->desc.block_cells[1] * u->desc.block_cells[2];
        // upper bound on workgroups: generic kernel uses ceil(cells/256) per block
        const long long need = nbt * ncomp * ((cells + 255) / 256 + u->desc.block_cells[2]);
        if (run) { // fixed buffer halves: never reallocated inside a run
            if (need > (long long)run->cap)
                return fail(CNV_ERR_SIZE, "partial buffer too small for this run");
            a.partials = run->partials;
        } else if (scr) { // fixed scratch, checked against the launch's own tile count in run_t
            a.partials = scr->p;
        } else {
            int rc = ensure_partials(u, (size_t)need);
            if (rc != CNV_OK)
                return rc;
            a.partials = u->partials;
        }
    }
    const bool gen = over ? false : force_generic();
    const bool fin = run == nullptr;
    int *np = run ? run->nparts : nullptr;
    const bool dbl = u->desc.dtype == CNV_F64;
    if (scr) { // the one-step Euler update with its norm (the allocation-free entry point)
        if (mode != 1 || !sumsq)
            return fail(CNV_ERR_INVALID, "scratch partials serve the Euler step with a norm only");
        return dbl ? run_t<double, 1, true>(a, ncomp, st, gen, over, true, nullptr, scr->cap)
                   : run_t<float, 1, true>(a, ncomp, st, gen, over, true, nullptr, scr->cap);
    }
    if (mode == 0)
        return dbl ? run_t<double, 0, false>(a, ncomp, st, gen, over, fin, np)
                   : run_t<float, 0, false>(a, ncomp, st, gen, over, fin, np);
    if (mode == 1) {
        if (sumsq)
            return dbl ? run_t<double, 1, true>(a, ncomp, st, gen, over, fin, np)
                       : run_t<float, 1, true>(a, ncomp, st, gen, over, fin, np);
        return dbl ? run_t<double, 1, false>(a, ncomp, st, gen, over, fin, np)
                   : run_t<float, 1, false>(a, ncomp, st, gen, over, fin, np);
    }
    if (sumsq)
        return dbl ? run_t<double, 2, true>(a, ncomp, st, gen, over, fin, np)
                   : run_t<float, 2, true>(a, ncomp, st, gen, over, fin, np);
    return dbl ? run_t<double, 2, false>(a, ncomp, st, gen, over, fin, np)
               : run_t<float, 2, false>(a, ncomp, st, gen, over, fin, np);
}

int launch_stencil7(cnv_grid_s *u, cnv_grid_s *out, cnv_grid_s *lap, int comp_in, int comp_out,
                    int ncomp, int mode, double nu_dt, double *sumsq, hipStream_t st)
{
    return launch_stencil7_t(u, out, lap, comp_in, comp_out, ncomp, mode, nu_dt, sumsq, st, nullptr);
}

// ------------------------------------------------ allocation-free Euler step
//
// The per-tile FP64 partials of the norm of cnv_diffusion_step_noalloc live
// in `scratch`, a static device array (one copy per GPU, held until the
// process ends).  Host code never passes the array's name to a kernel: on
// the host it is a stand-in, not a device address.  The first call on a
// device looks up that device's address of the array with
// hipGetSymbolAddress, and every launch takes it from g_scratch_dev.  Each
// launch writes every partial it reads before the finalize kernel reads it,
// so nothing needs clearing between calls; calls that overlap on one GPU
// share the array and race.
constexpr long long kNoAllocPartials = 1LL << 17; // 1 MiB of doubles
static __device__ __align__(256) double scratch[kNoAllocPartials];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mtx;
static double *g_scratch_dev[kMaxDevices] = {};

static int scratch_partials(double **out)
{
    int dev = -1;
    CNV_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(CNV_ERR_UNSUPPORTED, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[dev];
    return CNV_OK;
}

int launch_diffusion_noalloc(const cnv_grid_desc *desc, const void *u, void *u_new, double nu_dt, double *sumsq,
                             hipStream_t st)
{
    // grid objects on the stack over the caller's memory: layout fields
    // only, no BC lists or tables (every face periodic), no partials
    cnv_grid_s gu, go;
    int rc = grid_layout(desc, &gu);
    if (rc != CNV_OK)
        return rc;
    if (desc->entity != CNV_CELL)
        return fail(CNV_ERR_UNSUPPORTED, "the stencil operators take cell grids only");
    (void)grid_layout(desc, &go);
    gu.data = const_cast<void *>(u);
    go.data = u_new;
    gu.owns_data = go.owns_data = false;
    if (grids_overlap(&gu, &go))
        return fail(CNV_ERR_INVALID, "u and u_new must not overlap");
    if (!sumsq)
        return launch_stencil7_t(&gu, &go, nullptr, 0, 0, desc->ncomp, 1, nu_dt, nullptr, st, nullptr);
    S7Scratch scr{nullptr, kNoAllocPartials};
    rc = scratch_partials(&scr.p);
    if (rc != CNV_OK)
        return rc;
    return launch_stencil7_t(&gu, &go, nullptr, 0, 0, desc->ncomp, 1, nu_dt, sumsq, st, nullptr, nullptr, &scr);
}

// nsteps Euler steps alternating u -> tmp -> u ...  On periodic grids of
// 16^3 or 32^3 blocks an even number of two-step launches covers the steps
// up to a multiple of four, single steps the rest (so the field ends in u
// for even nsteps, in tmp for odd).  The norms of a launch are reduced by
// workgroup 0 of the next launch from the other half of u's partial buffer,
// those of the last launch by k_finalize_sum.
int launch_diffusion_run(cnv_grid_s *u, cnv_grid_s *tmp, int nsteps, double nu_dt, double *norms, hipStream_t st)
{
    double *P[2] = {u->partials, u->partials + u->partials_cap};
    const size_t cap = u->partials_cap;
    int nfused = 0, xwg = 0, xz = 0;
    if (nsteps >= 4 && !force_generic() && x2_applicable(u) && x2_applicable(tmp)) {
        x2_get_tune(u->desc.dtype, u->desc.block_cells[0], &xwg, &xz);
        if (2 * x2_tiles(u, xz) <= (long long)cap)
            nfused = 2 * (nsteps / 4);
    }
    const double *prev[2] = {nullptr, nullptr};
    int prev_n[2] = {0, 0};
    double *prev_out[2] = {nullptr, nullptr};
    int k = 0, L = 0; // steps and launches done
    auto grid_at = [&](int l) { return l % 2 == 0 ? u : tmp; };
    for (int j = 0; j < nfused; ++j, ++L) {
        double *p1 = norms ? P[L % 2] : nullptr;
        double *p2 = norms ? p1 + x2_tiles(u, xz) : nullptr;
        int nt = 0;
        const int rc = launch_diffusion_x2(grid_at(L), grid_at(L + 1), nu_dt, p1, p2, prev, prev_n, prev_out, xwg, xz,
                                           &nt, st);
        if (rc != CNV_OK)
            return rc;
        if (norms) {
            prev[0] = p1;
            prev[1] = p2;
            prev_n[0] = prev_n[1] = nt;
            prev_out[0] = norms + k;
            prev_out[1] = norms + k + 1;
        }
        k += 2;
    }
    for (; k < nsteps; ++k, ++L) {
        S7Run run;
        run.partials = P[L % 2];
        run.partials2 = nullptr;
        for (int s = 0; s < 2; ++s) {
            run.prev[s] = prev[s];
            run.prev_n[s] = prev_n[s];
            run.prev_out[s] = prev_out[s];
        }
        int nt = 0;
        run.nparts = &nt;
        run.cap = cap;
        const int rc = launch_stencil7_t(grid_at(L), grid_at(L + 1), nullptr, 0, 0, u->desc.ncomp, 1, nu_dt,
                                         norms ? norms + k : nullptr, st, nullptr, &run);
        if (rc != CNV_OK)
            return rc;
        if (norms) {
            prev[0] = P[L % 2];
            prev[1] = nullptr;
            prev_n[0] = nt;
            prev_n[1] = 0;
            prev_out[0] = norms + k;
            prev_out[1] = nullptr;
        }
    }
    for (int s = 0; s < 2; ++s)
        if (prev[s]) {
            hipLaunchKernelGGL(k_finalize_sum, dim3(1), dim3(1024), 0, st, prev[s], prev_n[s], prev_out[s]);
            CNV_HIP(hipGetLastError());
        }
    return CNV_OK;
}

// ---------------------------------------------------------------- autotuner

// Times every (workgroup size, z-planes per tile) candidate of the fused
// diffusion step (with the FP64 norm) on u -> out and keeps the fastest for
// (dtype, block edge).  Candidates that the block shape cannot take are
// skipped by the dispatcher's validity check.
int autotune_stencil7(cnv_grid_s *u, cnv_grid_s *out, int iters, int *best_wg, int *best_z, double *results,
                      int max_results, int *nresults)
{
    const int *n = u->desc.block_cells;
    if (!(n[0] == n[1] && (n[0] == 16 || n[0] == 32 || n[0] == 64)))
        return fail(CNV_ERR_UNSUPPORTED, "autotuning needs x/y block edges of 16, 32 or 64");
    if (iters < 1)
        return fail(CNV_ERR_INVALID, "iters must be >= 1");
    hipStream_t st;
    CNV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    double *norm = nullptr;
    hipEvent_t e0, e1;
    CNV_HIP(hipMalloc((void **)&norm, sizeof(double)));
    CNV_HIP(hipEventCreate(&e0));
    CNV_HIP(hipEventCreate(&e1));
    const int wgs[4] = {64, 128, 256, 512};
    const int zcs[6] = {2, 4, 8, 16, 32, 64};
    float best = 1e30f;
    int bw = 0, bz = 0, cnt = 0;
    int rc = CNV_OK;
    for (int wi = 0; wi < 4 && rc == CNV_OK; ++wi)
        for (int zi = 0; zi < 6 && rc == CNV_OK; ++zi) {
            if (zcs[zi] > n[2])
                continue;
            S7Tune tu{wgs[wi], zcs[zi]};
            int r = launch_stencil7_t(u, out, nullptr, 0, 0, u->desc.ncomp, 1, 1e-9, norm, st, &tu);
            if (r == CNV_ERR_UNSUPPORTED)
                continue;
            if (r != CNV_OK) {
                rc = r;
                break;
            }
            (void)hipEventRecord(e0, st);
            for (int i = 0; i < iters; ++i)
                (void)launch_stencil7_t(u, out, nullptr, 0, 0, u->desc.ncomp, 1, 1e-9, norm, st, &tu);
            (void)hipEventRecord(e1, st);
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(CNV_ERR_HIP, "autotune: kernel failed");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const float us = ms * 1000.f / iters;
            if (results && cnt < max_results) {
                results[3 * cnt] = tu.wg;
                results[3 * cnt + 1] = tu.zchunk;
                results[3 * cnt + 2] = us;
            }
            ++cnt;
            if (us < best) {
                best = us;
                bw = tu.wg;
                bz = tu.zchunk;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(norm);
    (void)hipStreamDestroy(st);
    if (rc != CNV_OK)
        return rc;
    if (bw == 0)
        return fail(CNV_ERR_UNSUPPORTED, "no tile candidate ran");
    s7_set_tune(u->desc.dtype, n[0], bw, bz);
    if (best_wg)
        *best_wg = bw;
    if (best_z)
        *best_z = bz;
    if (nresults)
        *nresults = cnt < max_results ? cnt : max_results;
    return CNV_OK;
}

int get_tune_stencil7(int dtype, int edge, int *wg, int *zchunk)
{
    S7Tune t = s7_default_tune(dtype, edge);
    *wg = t.wg;
    *zchunk = t.zchunk;
    return CNV_OK;
}

// ---------------------------------------------------------------- sums

template <typename T>
__global__ __launch_bounds__(256) void k_sums(const T *p, long long bstride, long long cells, long long nbt,
                                                double *partials)
{
    __shared__ double red[2][4];
    const int tid = threadIdx.x;
    double s = 0.0, q = 0.0;
    for (long long b = blockIdx.x; b < nbt; b += gridDim.x) {
        const T *pb = p + b * bstride;
        for (long long i = tid; i < cells; i += 256) {
            const double v = (double)pb[i];
            s += v;
            q += v * v;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, 64);
        q += __shfl_xor(q, o, 64);
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = s;
        red[1][tid >> 6] = q;
    }
    __syncthreads();
    if (tid == 0) {
        partials[blockIdx.x] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        partials[gridDim.x + blockIdx.x] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
    }
}

int launch_sums(cnv_grid_s *g, int comp, double *out, hipStream_t st)
{
    const long long nbt = g->nblocks_total;
    const long long cells = (long long)g->desc.block_cells[0] * g->desc.block_cells[1] * g->desc.block_cells[2];
    const int nwg = (int)(nbt < 2048 ? nbt : 2048);
    {
        std::lock_guard<std::mutex> lk(g->mtx);
        int rc = ensure_partials(g, 2 * (size_t)nwg);
        if (rc != CNV_OK)
            return rc;
    }
    const Geom G = geom_of(g);
    if (g->desc.dtype == CNV_F64)
        hipLaunchKernelGGL(k_sums<double>, dim3(nwg), dim3(256), 0, st,
                           reinterpret_cast<const double *>((const char *)g->data + comp * g->component_bytes),
                           G.bstride, cells, nbt, g->partials);
    else
        hipLaunchKernelGGL(k_sums<float>, dim3(nwg), dim3(256), 0, st,
                           reinterpret_cast<const float *>((const char *)g->data + comp * g->component_bytes),
                           G.bstride, cells, nbt, g->partials);
    CNV_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finalize_sum, dim3(1), dim3(1024), 0, st, g->partials, nwg, out);
    hipLaunchKernelGGL(k_finalize_sum, dim3(1), dim3(1024), 0, st, g->partials + nwg, nwg, out + 1);
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

} // namespace cnv

namespace cnv
{

// ---------------------------------------------------------------- compare

// Per-workgroup partials of sum (a-b)^2, sum b^2 and max |a-b| over one
// component of two grids of equal shape (any dtypes), evaluated in FP64.
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_compare(const TA *pa, const TB *pb, long long bstride_a, long long bstride_b,
                                                   long long cells, long long nbt, double *partials)
{
    __shared__ double red[3][4];
    const int tid = threadIdx.x;
    double d2 = 0.0, b2 = 0.0, mx = 0.0;
    for (long long b = blockIdx.x; b < nbt; b += gridDim.x) {
        const TA *qa = pa + b * bstride_a;
        const TB *qb = pb + b * bstride_b;
        for (long long i = tid; i < cells; i += 256) {
            const double x = (double)qa[i], y = (double)qb[i];
            const double d = x - y;
            d2 += d * d;
            b2 += y * y;
            mx = fmax(mx, fabs(d));
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        d2 += __shfl_xor(d2, o, 64);
        b2 += __shfl_xor(b2, o, 64);
        mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = d2;
        red[1][tid >> 6] = b2;
        red[2][tid >> 6] = mx;
    }
    __syncthreads();
    if (tid == 0) {
        const int n = gridDim.x;
        partials[blockIdx.x] = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
        partials[n + blockIdx.x] = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
        partials[2 * n + blockIdx.x] = fmax(fmax(red[2][0], red[2][1]), fmax(red[2][2], red[2][3]));
    }
}

__global__ __launch_bounds__(1024) void k_finalize_max(const double *__restrict__ partials, int n,
                                                       double *__restrict__ out)
{
    __shared__ double red[16];
    const int tid = threadIdx.x;
    double s = 0.0;
    for (int i = tid; i < n; i += 1024)
        s = fmax(s, partials[i]);
    for (int o = 32; o > 0; o >>= 1)
        s = fmax(s, __shfl_xor(s, o, 64));
    if ((tid & 63) == 0)
        red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w)
            t = fmax(t, red[w]);
        *out = t;
    }
}

template <typename TA, typename TB>
static void compare_t(const cnv_grid_s *a, int ca, const cnv_grid_s *b, int cb, double *partials, int nwg,
                      double *out, hipStream_t st)
{
    const long long nbt = a->nblocks_total;
    const long long cells = (long long)a->desc.block_cells[0] * a->desc.block_cells[1] * a->desc.block_cells[2];
    const TA *pa = reinterpret_cast<const TA *>((const char *)a->data + (size_t)ca * a->component_bytes);
    const TB *pb = reinterpret_cast<const TB *>((const char *)b->data + (size_t)cb * b->component_bytes);
    hipLaunchKernelGGL((k_compare<TA, TB>), dim3(nwg), dim3(256), 0, st, pa, pb,
                       (long long)(a->block_bytes / a->elem), (long long)(b->block_bytes / b->elem), cells, nbt,
                       partials);
    hipLaunchKernelGGL(k_finalize_sum, dim3(1), dim3(1024), 0, st, partials, nwg, out);
    hipLaunchKernelGGL(k_finalize_sum, dim3(1), dim3(1024), 0, st, partials + nwg, nwg, out + 1);
    hipLaunchKernelGGL(k_finalize_max, dim3(1), dim3(1024), 0, st, partials + 2 * nwg, nwg, out + 2);
}

int launch_compare(const cnv_grid_s *a, int ca, const cnv_grid_s *b, int cb, double *out, hipStream_t st)
{
    const long long nbt = a->nblocks_total;
    const int nwg = (int)(nbt < 1024 ? nbt : 1024);
    // the reduction scratch of `a` holds at least 4096 doubles
    double *partials = a->partials;
    const bool da = a->desc.dtype == CNV_F64, db = b->desc.dtype == CNV_F64;
    if (da && db)
        compare_t<double, double>(a, ca, b, cb, partials, nwg, out, st);
    else if (da)
        compare_t<double, float>(a, ca, b, cb, partials, nwg, out, st);
    else if (db)
        compare_t<float, double>(a, ca, b, cb, partials, nwg, out, st);
    else
        compare_t<float, float>(a, ca, b, cb, partials, nwg, out, st);
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

} // namespace cnv
