// cnv_internal.h -- shared host/device definitions of libcubismnova_amd.
#pragma once

#include "../../include/cubismnova_amd.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace cnv
{

// ------------------------------------------------------------ errors

void set_error(const std::string &msg);
int fail(int status, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define CNV_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::cnv::hip_fail(e_, #call);                                               \
    } while (0)

// ------------------------------------------------------------ device-side tables

// One BC entry as the kernels read it.
struct DevBC {
    int kind;
    int dir;
    int side;
    int pad;
    double value;
};

// Per-block header into the flattened BC entry array.
struct DevBCHead {
    int offset;
    int count;
};

// Face of a block as the non-tensorial operator kernels see it after the
// block's BC list is resolved for a given stencil.
enum FaceKind { FACE_NBR = 0, FACE_ABSORB = 1, FACE_DIRICHLET = 2, FACE_SYMMETRY = 3, FACE_UNSET = 4 };

// Geometry of a grid as passed to kernels.
struct Geom {
    int nb[3];        // blocks per dimension
    int n[3];         // cells per block per dimension
    long long bstride; // elements between blocks
    long long cstride; // elements between component slabs
};

// ------------------------------------------------------------ host grid object

struct BCTable {
    uint64_t version = ~0ull; // BC version this table was built from
    DevBCHead *heads = nullptr;
    DevBC *entries = nullptr;
    int nentries_cap = 0;
    bool any = false; // false: no block of this component carries a BC
};

} // namespace cnv

struct cnv_grid_s {
    cnv_grid_desc desc;
    size_t elem = 0;            // sizeof(T)
    size_t block_bytes = 0;
    size_t component_bytes = 0;
    long long nblocks_total = 0;
    int nslabs = 0;             // ncomp, 3 * ncomp for face grids
    void *data = nullptr;       // device memory, nslabs component slabs
    bool owns_data = true;      // false for cnv_grid_wrap
    double h[3] = {0, 0, 0};    // mesh spacing (extent / cells)

    // BC lists per component per block (host truth)
    std::vector<std::vector<std::vector<cnv_bc>>> bcs;
    uint64_t bc_version = 0;
    std::vector<cnv::BCTable> tables; // device copies per component

    // scratch for deterministic reductions: per-tile partial sums, two halves
    // of partials_cap doubles each
    double *partials = nullptr;

    // z metric of a stretched mesh (per global z cell), set on first use
    double *czm_dev = nullptr;
    double *czp_dev = nullptr;
    size_t partials_cap = 0;

    std::mutex mtx;
};

namespace cnv
{

// Dimensions in which the blocks at the upper end hold one more element:
// bit d for node grids in every d, for face slab s in d = s / ncomp
// (BlockFieldAssembler.h:123-131).
inline int plus_mask(const cnv_grid_s *g, int slab)
{
    if (g->desc.entity == CNV_NODE)
        return 7;
    if (g->desc.entity == CNV_FACE)
        return 1 << (slab / g->desc.ncomp);
    return 0;
}

// True when the device memory of two grids overlaps: handles wrapped over
// the same (or overlapping) memory count as one grid.
inline bool grids_overlap(const cnv_grid_s *a, const cnv_grid_s *b)
{
    const char *pa = static_cast<const char *>(a->data), *pb = static_cast<const char *>(b->data);
    const size_t na = a->component_bytes * (size_t)a->nslabs, nb = b->component_bytes * (size_t)b->nslabs;
    return pa < pb + nb && pb < pa + na;
}

// True when component (slab) ca of a and cb of b share device memory.
inline bool slabs_overlap(const cnv_grid_s *a, int ca, const cnv_grid_s *b, int cb)
{
    const char *pa = static_cast<const char *>(a->data) + (size_t)ca * a->component_bytes;
    const char *pb = static_cast<const char *>(b->data) + (size_t)cb * b->component_bytes;
    return pa < pb + b->component_bytes && pb < pa + a->component_bytes;
}

inline cnv::Geom geom_of(const cnv_grid_s *g)
{
    Geom G;
    for (int d = 0; d < 3; ++d) {
        G.nb[d] = g->desc.nblocks[d];
        G.n[d] = g->desc.block_cells[d];
    }
    G.bstride = (long long)(g->block_bytes / g->elem);
    G.cstride = (long long)(g->component_bytes / g->elem);
    return G;
}

// Check a descriptor and fill the layout fields of g (no allocation).
int grid_layout(const cnv_grid_desc *desc, cnv_grid_s *g);
// Upload (if stale) the BC table of component comp; returns status.
int ensure_bc_table(cnv_grid_s *g, int comp);
// Make sure the reduction scratch holds at least n partials.
int ensure_partials(cnv_grid_s *g, size_t n);

int check_stencil(const cnv_stencil *s);

// Resolve one block's ordered BC list into six faces for a non-tensorial
// stencil; shared by host and device.  Rule (FieldLab.h:206-223 with the
// BC slabs of BC/*.h): the face (dir, side) starts as a neighbour copy unless
// the last BC of `dir` is non-periodic with that side (then unwritten); every
// non-Base BC on that face whose stencil has width in `dir`
// (BC/Base.h:73-78) overwrites the whole slab, so the last one decides.
// Loops run over compile-time face indices so that device code keeps the
// arrays in registers.
__host__ __device__ inline void resolve_faces(const DevBC *list, int n, const int sb[3],
                                              const int se[3], int kind[6], double value[6])
{
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int dir = f >> 1, side = f & 1;
        bool periodic = true;
        int skip = 1;
        int k_f = FACE_NBR;
        double v_f = 0.0;
        const bool valid = !(sb[dir] == 0 || se[dir] == 1);
        for (int k = 0; k < n; ++k) {
            const DevBC b = list[k];
            if (b.dir != dir)
                continue;
            periodic = (b.kind == CNV_BC_BASE);
            skip = (b.side == 0) ? -1 : 1;
            if (valid && b.kind != CNV_BC_BASE && b.side == side) {
                k_f = b.kind; // ABSORBING=1, DIRICHLET=2, SYMMETRY=3 match FaceKind
                v_f = b.value;
            }
        }
        if (k_f == FACE_NBR && !periodic && skip == (side == 0 ? -1 : 1))
            k_f = FACE_UNSET;
        kind[f] = k_f;
        value[f] = v_f;
    }
}

// ------------------------------------------------------------ kernel launchers
// (defined in the .hip files; all return a cnv_status)

int launch_lab(const cnv_grid_s *g, int comp, const cnv_stencil &s, bool apply_bc, bool use_list,
               const DevBC *list, int nlist, const int *block, void *out, hipStream_t st);

// z metric of u (uniform unless set), uploaded on first use; caller holds
// the grid's mutex (cnv_jacobi.hip)
int ensure_zmetric(cnv_grid_s *g);
// two Jacobi sweeps per launch on grids whose axes are periodic or carry one
// Dirichlet constant per domain face (cnv_jacobi2.hip)
bool jacobi2_applies(cnv_grid_s *u, int cu, cnv_grid_s *f, cnv_grid_s *out);
int launch_jacobi2(cnv_grid_s *u, int cu, cnv_grid_s *f, int cf, cnv_grid_s *out, int co, double omega,
                   hipStream_t st, const int *tile);
int jacobi2_set_tune(int dtype, int edge, int wg, int zchunk, int split);
int jacobi2_get_tune(int dtype, int edge, int *wg, int *zchunk, int *split);

int launch_stencil7(cnv_grid_s *u, cnv_grid_s *out, cnv_grid_s *lap, int comp_in, int comp_out,
                    int ncomp, int mode, double nu_dt, double *sumsq, hipStream_t st);

// cnv_diffusion_step on caller memory of a periodic cell grid, norm
// partials in the library's static device array (cnv_stencil7.hip)
int launch_diffusion_noalloc(const cnv_grid_desc *desc, const void *u, void *u_new, double nu_dt, double *sumsq,
                             hipStream_t st);

int launch_sums(cnv_grid_s *g, int comp, double *out, hipStream_t st);
int launch_diffusion_run(cnv_grid_s *u, cnv_grid_s *tmp, int nsteps, double nu_dt, double *norms, hipStream_t st);
// two Euler steps per launch on periodic grids (cnv_stencil7x2.hip)
bool x2_applicable(cnv_grid_s *u);
long long x2_tiles(const cnv_grid_s *u, int zchunk);
int launch_diffusion_x2(cnv_grid_s *u, cnv_grid_s *out, double nu_dt, double *part1, double *part2,
                        const double *const prev_p[2], const int prev_n[2], double *const prev_out[2], int wg,
                        int zchunk, int *nparts, hipStream_t st);
int x2_set_tune(int dtype, int edge, int wg, int zchunk);
int x2_get_tune(int dtype, int edge, int *wg, int *zchunk);
int autotune_x2(cnv_grid_s *u, cnv_grid_s *out, int iters, int *best_wg, int *best_z, double *results,
                int max_results, int *nresults);

int autotune_stencil7(cnv_grid_s *u, cnv_grid_s *out, int iters, int *best_wg, int *best_z, double *results,
                      int max_results, int *nresults);
void s7_set_tune(int dtype, int edge, int wg, int zchunk);
int get_tune_stencil7(int dtype, int edge, int *wg, int *zchunk);

int launch_compare(const cnv_grid_s *a, int ca, const cnv_grid_s *b, int cb, double *out, hipStream_t st);

int set_zfaces(cnv_grid_s *g, const double *zf, int n);
// tile (may be null: the table's choice) = {workgroup size, z-planes per tile, f prefetch}
int launch_jacobi(cnv_grid_s *u, int cu, cnv_grid_s *f, int cf, cnv_grid_s *out, int co, int mode, double omega,
                  double *sumsq, hipStream_t st, const int *tile = nullptr);
int autotune_jacobi(cnv_grid_s *u, cnv_grid_s *f, cnv_grid_s *out, int iters, int *best_wg, int *best_z,
                    int *best_pf, double *results, int max_results, int *nresults);
int jacobi_set_tune(int dtype, int edge, int wg, int zchunk, int prefetch);
int jacobi_get_tune(int dtype, int edge, int *wg, int *zchunk, int *prefetch);
int launch_correct(cnv_grid_s *u, int cu, cnv_grid_s *e, int ce, hipStream_t st);
int jacobi_prepare(cnv_grid_s *g, int comp);

int launch_weno5(cnv_grid_s *in, cnv_grid_s *out, hipStream_t st, const int *tile);
int autotune_weno5(cnv_grid_s *in, cnv_grid_s *out, int iters, int *best_wg, int *best_z, double *results,
                   int max_results, int *nresults);
int weno5_set_tune(int dtype, int edge, int wg, int zchunk);
int weno5_get_tune(int dtype, int edge, int *wg, int *zchunk);

int launch_fill(void *dst, int dtype, double value, size_t count, hipStream_t st);

} // namespace cnv
