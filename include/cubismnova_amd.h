/* cubismnova_amd.h -- C ABI of the MI355X (gfx950) block-structured stencil
 * library.
 *
 * The library keeps CubismNova's block storage on the GPU and exposes the
 * path "FieldLab ghost assembly for every block of a Grid::Cartesian + the
 * stencil operators applied over it" through plain C entry points: opaque
 * grid handles, plain pointers and sizes, int status codes (0 = success, no
 * exceptions cross this boundary) and calls ordered on a caller-supplied HIP
 * stream (NULL selects the default stream).
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the CubismNova source tree).  Memory layout of a grid is the reference's
 * (include/Cubism/Grid/Cartesian.h:490-529 with
 * include/Cubism/Grid/BlockFieldAssembler.h:188-199): component c of block i
 * starts at byte c*component_bytes + i*block_bytes, block_bytes =
 * roundup32(cells*sizeof(T)), component_bytes = block_bytes*nblocks, blocks
 * numbered x-fastest, cells inside a block x-fastest.
 */
#ifndef CUBISMNOVA_AMD_H
#define CUBISMNOVA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ types */

typedef struct cnv_grid_s *cnv_grid; /* opaque grid handle */
typedef void *cnv_stream;            /* a hipStream_t; NULL = default stream */

enum cnv_status {
    CNV_OK = 0,
    CNV_ERR_INVALID = 1,       /* bad argument or descriptor */
    CNV_ERR_ALLOC = 2,         /* device or host allocation failed */
    CNV_ERR_HIP = 3,           /* a HIP runtime call failed */
    CNV_ERR_UNSUPPORTED = 4,   /* valid request this library does not run */
    CNV_ERR_NOT_ALLOCATED = 5, /* lab used before allocation */
    CNV_ERR_SIZE = 6           /* size or shape mismatch */
};

enum cnv_dtype { CNV_F32 = 0, CNV_F64 = 1 };

/* BC kinds: include/Cubism/BC/{Base,Absorbing,Dirichlet,Symmetry}.h */
enum cnv_bc_kind {
    CNV_BC_BASE = 0,      /* periodic, no-op apply (BC/Base.h:45-62) */
    CNV_BC_ABSORBING = 1, /* zero-gradient copy (BC/Absorbing.h:62-104) */
    CNV_BC_DIRICHLET = 2, /* ghost = value (BC/Dirichlet.h:75-113) */
    CNV_BC_SYMMETRY = 3   /* ghost = value * mirror (BC/Symmetry.h:69-114) */
};

/* One boundary condition: BC::*<Lab>(dir, side[, value|sign]).
 * `value` is the Dirichlet constant or the Symmetry sign. */
typedef struct {
    int kind;
    int dir;
    int side;
    int reserved;
    double value;
} cnv_bc;

/* Core::Stencil<3>(begin, end, tensorial) (include/Cubism/Core/Stencil.h:44-67):
 * begin <= 0 inclusive, end >= 1 exclusive, per dimension. */
typedef struct {
    int begin[3];
    int end[3];
    int tensorial;
} cnv_stencil;

/* Data entity of a grid: Cubism::EntityType (include/Cubism/Common.h). */
enum cnv_entity { CNV_CELL = 0, CNV_NODE = 1, CNV_FACE = 2 };

/* Grid::Cartesian(nblocks, block_cells, begin, end) with scalar components
 * (include/Cubism/Grid/Cartesian.h:132-143). ncomp independent scalar
 * components stored as consecutive component slabs.  Node and face grids
 * (Cartesian.h:490-529, BlockFieldAssembler.h:111-185) hold (N+1)^3 elements
 * per block; a block at the upper end of dimension d stores N[d]+1 elements
 * along d (every d for nodes, d = the face direction for faces), the others
 * N[d], each block x-fastest with its own extent as pitch.  A face grid holds
 * 3*ncomp slabs, slab d*ncomp + c the faces normal to d of component c; every
 * call taking a component index takes a slab index there.  The operators
 * (Laplacian, diffusion, WENO5, Jacobi, sums, batches) take cell grids only;
 * labs, BCs, uploads and the field operations take any entity. */
typedef struct {
    int nblocks[3];
    int block_cells[3];
    int ncomp;
    int dtype; /* enum cnv_dtype */
    double begin[3];
    double end[3];
    int entity; /* enum cnv_entity; 0 (cell) when zero-initialised */
    int reserved;
} cnv_grid_desc;

/* --------------------------------------------------------------- errors */

/* Static message for a status code. */
const char *cnv_strerror(int status);
/* Message of the last failing call on this thread ("" if none). */
const char *cnv_last_error(void);

/* ------------------------------------------------------- device helpers */
/* Thin wrappers so that callers without their own HIP binding (ctypes,
 * JNI, ...) can stage buffers; they call the HIP runtime directly. */
int cnv_device_count(int *count);
int cnv_malloc(void **ptr, size_t bytes);
int cnv_free(void *ptr);
int cnv_memcpy_h2d(void *dst, const void *src, size_t bytes, cnv_stream s);
int cnv_memcpy_d2h(void *dst, const void *src, size_t bytes, cnv_stream s);
int cnv_memcpy_d2d(void *dst, const void *src, size_t bytes, cnv_stream s);
int cnv_fill(void *dst, int dtype, double value, size_t count, cnv_stream s);
/* No counterpart in the reference: device-to-device copy with non-temporal
 * 16-byte loads and stores (dst, src and bytes multiples of 16), the
 * streaming read + write ceiling the stencil kernels are measured against. */
int cnv_memcpy_stream(void *dst, const void *src, size_t bytes, cnv_stream s);
/* No counterpart in the reference: `bytes` of src copied to `fanout` slabs
 * of dst, dst_stride bytes apart (every value read once, written fanout
 * times; non-temporal 16-byte accesses; all sizes multiples of 16): the
 * memory pattern of the WENO5 operator (fanout 6) without its arithmetic,
 * the ceiling that kernel is measured against. */
int cnv_memcpy_fanout(void *dst, size_t dst_stride, const void *src, size_t bytes, int fanout, cnv_stream s);
int cnv_stream_create(cnv_stream *s);
int cnv_stream_destroy(cnv_stream s);
int cnv_stream_synchronize(cnv_stream s);

/* ------------------------------------------------------------------ grid */

/* Replaces Grid::Cartesian::Cartesian + initTopology_/alloc_
 * (Grid/Cartesian.h:132-143, 444-529): one device allocation holding every
 * component slab, reference byte layout. Memory is not initialised. */
int cnv_grid_create(const cnv_grid_desc *desc, cnv_grid *out);
/* The same grid over memory the caller owns: the `Alloc` seam of
 * Grid::Cartesian (Cartesian.h:481,527,537; the AlignedBlockAllocator
 * contract of Alloc/AlignedBlockAllocator.h:34-56).  A reference Cartesian
 * built with a device allocator lays its data out exactly as described
 * here, so its allocation can be handed over as is and the operators run on
 * it in place.  device_data: device (or managed) memory, 32-byte aligned,
 * at least `bytes` >= the grid's bytes; cnv_grid_destroy leaves it alone. */
int cnv_grid_wrap(const cnv_grid_desc *desc, void *device_data, size_t bytes, cnv_grid *out);
int cnv_grid_destroy(cnv_grid g);
/* Layout query: block_bytes and component_bytes as in Cartesian.h:490-529. */
int cnv_grid_info(cnv_grid g, cnv_grid_desc *desc, size_t *block_bytes, size_t *component_bytes);
/* Device address of component `comp` (block 0). */
int cnv_grid_data(cnv_grid g, int comp, void **device_ptr);
/* Copy one component slab (component_bytes bytes, reference layout) to or
 * from host memory. Replaces host writes through Block::Field accessors. */
int cnv_grid_upload(cnv_grid g, int comp, const void *host, size_t bytes, cnv_stream s);
int cnv_grid_download(cnv_grid g, int comp, void *host, size_t bytes, cnv_stream s);

/* ------------------------------------------------------------------- BCs */

/* Replaces Block::Field::getBC() (Block/Field.h:437-442): sets the ordered BC
 * list of the scalar field (comp, block). Order matters exactly as in
 * FieldLab::loadData (Block/FieldLab.h:206-223): the last BC per direction
 * decides periodicity and the skipped side; BCs are applied in list order. */
int cnv_grid_set_block_bcs(cnv_grid g, int comp, const int block[3], const cnv_bc *bcs, int n);
/* Appends `bc` to the lists of every block that touches the grid face
 * (bc->dir, bc->side), for one component or all (comp = -1).  This is the
 * loop a CubismNova application writes over its boundary blocks. */
int cnv_grid_add_face_bc(cnv_grid g, int comp, const cnv_bc *bc);
int cnv_grid_clear_bcs(cnv_grid g, int comp);

/* ------------------------------------------------------------------- lab */

/* Extent of the active lab range of a block for stencil s:
 * N + end - 1 - begin per dimension (Block/FieldLab.h:198-202).  On node
 * and face grids N is the largest block's extent over all slabs. */
int cnv_lab_extent(cnv_grid g, const cnv_stencil *s, int extent[3]);
/* Lab geometry of slab `comp` for stencil s: pitch = the lab extent of the
 * slab's largest block (the per-block layout of cnv_lab_load_all), active =
 * the lab extent of block `block` (the layout cnv_lab_load writes; in
 * cnv_lab_load_all the block's lab occupies the low corner [0, active) of
 * its pitch-sized slot).  block may be NULL (active = pitch).  Cell grids:
 * both N + end - 1 - begin. */
int cnv_lab_geometry(cnv_grid g, int comp, const int block[3], const cnv_stencil *s, int pitch[3],
                     int active[3]);

/* Replaces Grid::Cartesian::loadLab + FieldLab::loadData
 * (Grid/Cartesian.h:411-427, Block/FieldLab.h:180-248, FieldLabLoader.h
 * :129-287): assembles the lab of block `block` of component `comp` and
 * writes it to device memory `out` in compact x-fastest layout over the
 * active lab range [begin, N + end - 1), N the block's own extent (node and
 * face blocks below the upper end are one smaller; their ghosts below the
 * block count back from the neighbour's own end, FieldLabLoader.h:217-219). Cells the reference leaves unwritten
 * (edges/corners of non-tensorial labs, skipped sides without BC) are left
 * untouched in `out`. n_extern < 0 uses the block's own BC list, otherwise
 * the `extern_bcs` list (FieldLab.h:184,206,241-248). */
int cnv_lab_load(cnv_grid g, int comp, const int block[3], const cnv_stencil *s, int apply_bc,
                 const cnv_bc *extern_bcs, int n_extern, void *out, cnv_stream st);
/* The same for every block of the grid, blocks consecutive in block order
 * (the reference's per-block loop over a grid, Laplacian.cpp:110-117), one
 * slot of pitch cells each (cnv_lab_geometry). */
int cnv_lab_load_all(cnv_grid g, int comp, const cnv_stencil *s, void *out, cnv_stream st);

/* ------------------------------------------------------------- operators */

/* 7-point Laplacian of `in` (component comp_in) into `out` (comp_out),
 * res = ihx2*((u[+x] - 2u) + u[-x]) + ihy2*(...) + ihz2*(...), ih2 = 1/(h*h)
 * from the grid mesh, ghosts from the lab with Stencil(-1,2,false) and the
 * blocks' BC lists. Replaces the loop of
 * test/sandbox/Benchmarks/Kernel/Laplacian/Laplacian.cpp:110-131. */
int cnv_laplacian(cnv_grid in, int comp_in, cnv_grid out, int comp_out, cnv_stream s);

/* One forward-Euler diffusion step fused with the Laplacian:
 * lap *= nu_dt; u_new = u + lap (two roundings, Field.h:457-512 with
 * FieldOperator.h:43-114), for every component of `u` into `u_new`.
 * If lap_out is non-NULL the Laplacian is also stored there.
 * If sumsq is non-NULL (device address of one double) the FP64 sum of
 * u_new^2 over all cells and components is written to it (deterministic,
 * fixed-order reduction). Calls on the same `u` must not overlap. */
int cnv_diffusion_step(cnv_grid u, cnv_grid u_new, double nu_dt, cnv_grid lap_out, double *sumsq,
                       cnv_stream s);

/* No counterpart in the reference: cnv_diffusion_step for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  u and u_new are caller-owned
 * device memory (32-byte aligned, distinct, not overlapping) laid out as the
 * cell grid `desc` describes, every face periodic (no BC lists); every
 * component takes one Euler step, bit-identical to cnv_diffusion_step on a
 * grid of that layout, and sumsq (device double, may be NULL) receives the
 * same FP64 norm.  No grid handle is made and nothing is allocated: the
 * norm's per-tile partials live in a static device array of 2^17 doubles
 * (1 MiB, one per GPU, held until the process ends); CNV_ERR_SIZE when a
 * launch needs more partials than that (one per tile: blocks x components x
 * z-chunks of the tile shape in use, e.g. 16384 for 512^3 cells in 32^3
 * blocks with the built-in 8-plane tiles).  Calls that overlap on one GPU
 * share that array and race. */
int cnv_diffusion_step_noalloc(const cnv_grid_desc *desc, const void *u, void *u_new, double nu_dt, double *sumsq,
                               cnv_stream s);

/* nsteps forward-Euler steps alternating u -> tmp -> u -> ... (the same
 * arithmetic, bit for bit, as nsteps calls of cnv_diffusion_step; the field
 * ends in u when nsteps is even, in tmp when odd).  tmp must carry the same
 * BC lists as u.  If norms is non-NULL (device array of nsteps doubles)
 * norms[k] is the FP64 sum of squares after step k; the reductions of a
 * launch run inside the next launch (one finalize launch at the end).  On
 * periodic grids (no BC list) of 16^3 or 32^3 blocks, steps go two per launch
 * (one HBM read and write for two steps) up to a multiple of four.  The loop
 * of the reference's time stepping over Field.h:457-512 / FieldOperator.h. */
int cnv_diffusion_run(cnv_grid u, cnv_grid tmp, int nsteps, double nu_dt, double *norms, cnv_stream s);

/* Fifth-order WENO face states (Jiang & Shu, J. Comput. Phys. 126:202-228,
 * 1996).  CubismNova advertises a WENO toolbox (README.rst:28-31) but ships
 * none; the lab read is the FieldLab of Stencil(-3,4,false) with the blocks'
 * BC lists (Block/FieldLab.h:180-224).  For every cell i, direction d and
 * component c of `in`: out component 6c+2d   = u^-_{i+1/2} (linear weights
 * (1,6,3)/10), out component 6c+2d+1 = u^+_{i-1/2} (weights (3,6,1)/10), from
 * cells i-2..i+2.  eps = 1e-6 * m2 with m2 the mean squared value of the five
 * cells (+1e-30), applied as alpha_k = d_k / (1e-6 + beta_k/m2)^2 up to
 * factors common to the three weights.  Both precisions evaluate the same
 * difference form with explicit fused multiply-adds: FP64 with IEEE
 * quotients (bit-identical to the CPU restatement), FP32 with hardware
 * reciprocals.  `out` has the same shape and dtype and at least
 * 6*ncomp(in) components. */
int cnv_weno5(cnv_grid in, cnv_grid out, cnv_stream s);

/* FP64 sum and sum of squares of one component into device doubles out[0],
 * out[1] (deterministic). */
int cnv_grid_sums(cnv_grid g, int comp, double *out, cnv_stream s);

/* No counterpart in the reference: error of one grid component against
 * another of the same shape (any dtypes), in FP64 on the device:
 * out[0] = sum (a-b)^2, out[1] = sum b^2, out[2] = max |a-b| (device doubles).
 * rel-L2 = sqrt(out[0]/out[1]).  Deterministic. */
int cnv_grid_compare(cnv_grid a, int comp_a, cnv_grid b, int comp_b, double *out, cnv_stream s);

/* ---------------------------------------------------- stretched-z Jacobi */
/* No counterpart in the reference: CubismNova names MeshClass::Stretched
 * (include/Cubism/Common.h:86) but ships neither a stretched mesh nor a
 * solver.  Operator, per global z cell k with faces zf[k], zf[k+1]:
 *   A u = ihx2 ((u[+x] - 2u) + u[-x]) + ihy2 ((u[+y] - 2u) + u[-y])
 *       + czp_k (u[+z] - u) - czm_k (u - u[-z]),
 *   czp_k = 1/(dz_k (zc_{k+1} - zc_k)), czm_k = 1/(dz_k (zc_k - zc_{k-1})),
 * ghost centres mirrored about the boundary faces; ghosts from the lab of
 * Stencil(-1,2) with the BC lists of u (Dirichlet = constant ghost,
 * BC/Dirichlet.h:75-113).  Without cnv_grid_set_zfaces the z faces are
 * uniform. */
int cnv_grid_set_zfaces(cnv_grid g, const double *zfaces, int n /* nz_total + 1 */);
/* weighted Jacobi sweep: out = u + (f - A u) * wd_k, wd_k = omega / diag_k,
 * diag_k = -(((2 ihx2 + 2 ihy2) + czm_k) + czp_k); u, f, out of one dtype */
int cnv_jacobi_sweep(cnv_grid u, int cu, cnv_grid f, int cf, cnv_grid out, int co, double omega, cnv_stream s);
/* defect r = f - A u in u's dtype, stored in r's dtype; if sumsq is
 * non-NULL the FP64 sum of r^2 (before r is rounded) is written there */
int cnv_jacobi_defect(cnv_grid u, int cu, cnv_grid f, int cf, cnv_grid r, int cr, double *sumsq, cnv_stream s);
/* No counterpart in the reference (see cnv_jacobi_sweep).  Two sweeps in
 * one launch: out = S(S(u)), S the sweep above, bit-identical to two calls
 * of cnv_jacobi_sweep, reading u and f once.  Needs u's BC lists (component
 * cu) to give each domain face one rule on all its blocks: periodic on both
 * sides of an axis, or per side Dirichlet, Symmetry (nonzero sign) or
 * Absorbing (BC/Dirichlet.h, BC/Symmetry.h, BC/Absorbing.h); x/y block
 * edges of 16, 32 or 64; u, f, out of one dtype; CNV_ERR_UNSUPPORTED
 * otherwise.  out must not alias u or f. */
int cnv_jacobi_sweep2(cnv_grid u, int cu, cnv_grid f, int cf, cnv_grid out, int co, double omega, cnv_stream s);
/* Tile of the two-sweep launch per (dtype, x/y block edge): workgroup size,
 * z planes per tile and ysplit, the tiles per block plane (1: whole planes;
 * 2 or 4: parts in y, for less LDS per workgroup). */
int cnv_tune_set_jacobi2(int dtype, int edge, int wg, int zchunk, int ysplit);
int cnv_tune_get_jacobi2(int dtype, int edge, int *wg, int *zchunk, int *ysplit);
/* u (FP64) += e (FP32 or FP64), elementwise */
int cnv_add_correction(cnv_grid u, int cu, cnv_grid e, int ce, cnv_stream s);
/* No counterpart in the reference.  Times every (workgroup size, z-planes
 * per tile, f prefetch) candidate of the sweep u -> out (component 0 of
 * each, `iters` launches each; `out` is overwritten), keeps the fastest in
 * the process-wide table the sweeps and defects use for (dtype, x/y block
 * edge) and returns it.  Prefetch (sweeps only) loads f one plane ahead.
 * results (may be NULL) receives up to max_results quadruples {wg, zchunk,
 * prefetch, microseconds per sweep}. */
int cnv_autotune_jacobi(cnv_grid u, cnv_grid f, cnv_grid out, int iters, int *wg, int *zchunk, int *prefetch,
                        double *results, int max_results, int *nresults);
int cnv_tune_set_jacobi(int dtype, int edge, int wg, int zchunk, int prefetch);
int cnv_tune_get_jacobi(int dtype, int edge, int *wg, int *zchunk, int *prefetch);

typedef struct {
    int refine;           /* 1: FP32 sweeps on A e = r, FP64 defect and correction */
    int sweeps_per_cycle; /* even; 50 in the refined Poisson workload */
    int max_cycles;
    double omega;         /* 2/3 */
    double tol;           /* stop when ||f - A u||_2 / ||f||_2 <= tol */
} cnv_solve_opts;

typedef struct {
    int cycles;
    long long sweeps;
    double rel_residual;
} cnv_solve_info;

/* Solve A u = f for component 0 of u (in place) by weighted Jacobi.
 * refine = 1: u and f FP64, work FP32 with 3 components (correction pair and
 * FP32 defect); each cycle zeroes e, runs sweeps_per_cycle FP32 sweeps on
 * A e = r with u's BCs made homogeneous, adds e to u in FP64 and recomputes
 * the FP64 defect.  refine = 0: plain sweeps in u's dtype, work of u's dtype
 * with 2 components.  One cycle is recorded once into a HIP graph and
 * replayed; the relative residual after cycle c goes to history[c]
 * (history[0] = initial).  Calls on the same grids must not overlap. */
int cnv_jacobi_solve(cnv_grid u, cnv_grid f, cnv_grid work, const cnv_solve_opts *opts, cnv_solve_info *info,
                     double *history, int history_cap, cnv_stream s);

/* ------------------------------------------------------------ field ops */
/* Element-wise arithmetic on device memory (a grid's component slab from
 * cnv_grid_data, or cnv_malloc memory), one IEEE operation per element:
 * op2: dst[i] = a[i] OP b[i];  op1: dst[i] = a[i] OP scalar;
 * rcp: dst[i] = scalar * (1 / a[i]) (two roundings).  The loops of
 * include/Cubism/Block/FieldOperator.h:11-30, 44-141, stream-ordered. */
enum cnv_field_op {
    CNV_FIELD_ADD = 0,
    CNV_FIELD_SUB = 1,
    CNV_FIELD_MUL = 2,
    CNV_FIELD_DIV = 3,
    CNV_FIELD_RCP = 4, /* cnv_grid_field_op only */
    CNV_FIELD_NEG = 5  /* cnv_grid_field_op only */
};
int cnv_field_op2(int op, int dtype, const void *a, const void *b, void *dst, size_t n, cnv_stream s);
int cnv_field_op1(int op, int dtype, const void *a, double scalar, void *dst, size_t n, cnv_stream s);
int cnv_field_rcp(int dtype, const void *a, double scalar, void *dst, size_t n, cnv_stream s);

/* Field and FieldContainer arithmetic over whole grids, one launch: every
 * block over its own elements only (range_.size(): N^3 for cells, the
 * block's extent for node and face slabs), as Field::operator+= and the
 * other Field operators (Block/Field.h:444-556) applied to each block of a
 * grid, and FieldContainer's operators (Field.h:1039-1154) to each of its
 * components:
 *   CNV_FIELD_ADD..DIV: dst = a OP b (b a grid) or a OP scalar (b NULL);
 *   CNV_FIELD_RCP: dst = scalar * (1 / a), two roundings (Field::reciprocal);
 *   CNV_FIELD_NEG: dst = -a (Field::operator-()).
 * cd, ca, cb are slab indices, or all -1 for every slab.  The grids have one
 * shape, entity, dtype and slab count; dst may be a or b.  Elements past a
 * block's own extent (padding) are not touched. */
int cnv_grid_field_op(int op, cnv_grid dst, int cd, cnv_grid a, int ca, cnv_grid b, int cb, double scalar,
                      cnv_stream s);

/* The reference's own C-ABI hooks, same names and signatures as the extern
 * "C" declarations of FieldOperator.h:149-227 (CUBISM_OPTIMIZED_FIELD_OP):
 * linking this library provides them.  Pointers must be device-accessible;
 * each call runs on the null stream and returns when done, like the host
 * loops it replaces.  Failures are readable with cnv_last_error. */
void fieldAdd2_double(const double *src0, const double *src1, double *dst, size_t n);
void fieldSub2_double(const double *src0, const double *src1, double *dst, size_t n);
void fieldMul2_double(const double *src0, const double *src1, double *dst, size_t n);
void fieldDiv2_double(const double *src0, const double *src1, double *dst, size_t n);
void fieldAdd2_float(const float *src0, const float *src1, float *dst, size_t n);
void fieldSub2_float(const float *src0, const float *src1, float *dst, size_t n);
void fieldMul2_float(const float *src0, const float *src1, float *dst, size_t n);
void fieldDiv2_float(const float *src0, const float *src1, float *dst, size_t n);
void fieldAdd1_double(const double *src0, double src1, double *dst, size_t n);
void fieldSub1_double(const double *src0, double src1, double *dst, size_t n);
void fieldMul1_double(const double *src0, double src1, double *dst, size_t n);
void fieldDiv1_double(const double *src0, double src1, double *dst, size_t n);
void fieldAdd1_float(const float *src0, float src1, float *dst, size_t n);
void fieldSub1_float(const float *src0, float src1, float *dst, size_t n);
void fieldMul1_float(const float *src0, float src1, float *dst, size_t n);
void fieldDiv1_float(const float *src0, float src1, float *dst, size_t n);
void fieldRcp_double(const double *src0, double src1, double *dst, size_t n);
void fieldRcp_float(const float *src0, float src1, float *dst, size_t n);

/* ------------------------------------------------------------ on-disk format */
/* Replaces IO::CartesianWriteHDF / CartesianReadHDF (IO/CartesianHDF.h:46-215,
 * src/IO/HDFDriver.cpp:26-148, IO/XDMFDriver.h:29-160): <fname>.h5 holds the
 * dataset "data" of shape (Z, Y, X, ncomp) over the global index span of the
 * grid's entity (cells, nodes, or the faces normal to face_dir), components
 * interleaved per point (IO/FieldAOS.h), elements of file_dtype (CNV_F32 or
 * CNV_F64, converted from or to the grid's dtype); <fname>.xmf is the XDMF
 * wrapper (cell and node grids, when create_xdmf).  begin / end (global
 * indices, NULL for the whole span) clip the written or read box, as the
 * reference's sub-mesh argument does.  The reordering between block layout
 * and file order runs on the GPU; the HDF5 C library is loaded at the first
 * call (CNV_ERR_UNSUPPORTED without it; CNV_HDF5_LIBRARY names its path).
 * Both calls return when the file, or the grid, is complete. */
int cnv_grid_write_hdf(cnv_grid g, const char *fname, const char *aname, int file_dtype, double time, int face_dir,
                       int create_xdmf, const int begin[3], const int end[3], cnv_stream s);
int cnv_grid_read_hdf(cnv_grid g, const char *fname, int file_dtype, int face_dir, const int begin[3],
                      const int end[3], cnv_stream s);

/* ------------------------------------------------------------ batches */
/* No counterpart in the reference (its applications loop over one grid's
 * blocks on the host, Laplacian.cpp:110-131).  Many independent FP32 grids
 * of uneven block counts, block sizes, component counts, ghost widths and
 * BCs, each with its own operator, processed together over a device work
 * list of (grid, component, block, z-range) items.  A run is one persistent
 * launch over a list of work units sorted by cost: an item of a cubic 32^3
 * block (the single-grid tile body of its operator, whole workgroup), up to
 * WG/64 items of cubic 16^3 blocks (one per 64-thread slice, same tile
 * bodies), or an item of any other block shape (staged in LDS by a generic
 * body).  The other form, one launch per (operator, edge) class of 16^3/32^3
 * blocks plus a persistent launch for the rest (forked onto streams of the
 * batch when no item's output grid is another item's output or any item's
 * input), stays available: cnv_batch_autotune times both and keeps the
 * faster.  Ghosts follow the lab rules of cnv_lab_load for the grid's
 * Stencil(-ghost, ghost+1) and BC lists.  A run is ordered on the caller's
 * stream and capturable in a HIP graph (cnv_graph_begin/end) after one plain
 * run. */
enum cnv_op {
    CNV_OP_DIFFUSION = 0, /* out = in + nu_dt * laplacian(in), as cnv_diffusion_step */
    CNV_OP_WENO5 = 1      /* out = WENO5 states of in, as cnv_weno5 */
};

typedef struct {
    cnv_grid in;
    cnv_grid out;
    int op;       /* enum cnv_op */
    int ghost;    /* g of the grid's Stencil(-g, g+1); >= 2 for CNV_OP_WENO5 */
    double nu_dt; /* Euler step (CNV_OP_DIFFUSION) */
} cnv_batch_item;

typedef struct cnv_batch_s *cnv_batch;
int cnv_batch_create(const cnv_batch_item *items, int n, cnv_batch *out);
/* Autotuner.  The single launch: per (block edge, ghost width) the planes per
 * item, then the workgroup size (128 or 256), each timed on the whole
 * launch.  The class form: per (operator, edge) class the workgroup size and
 * z depth of its launch, for the other items per (block edge, ghost width)
 * the z depth, then the workgroup size of their launch.  Then the forms as
 * a whole (the class form with and without overlapping launches); the
 * fastest runs from then on.  results: up to max_results
 * records of 5 doubles {op, edge, wg, depth, microseconds}: class trials
 * carry the class's op; trials of the class form's persistent launch op = -1
 * (edge = -1 for its workgroup trials, wg then the workgroup size and depth 0;
 * for its depth trials wg holds the ghost width); single-launch trials
 * op = -2 (depth trials: edge, wg = ghost width, depth; workgroup trials:
 * edge = -1, wg, depth 0); the forms as a whole op = -3 with wg = 1 for the
 * single launch, 0 for the class form with its launches overlapping on the
 * batch's streams and 2 for the class form with its launches one after
 * another on the caller's stream. */
int cnv_batch_autotune(cnv_batch b, int iters, double *results, int max_results, int *nresults);
int cnv_batch_run(cnv_batch b, cnv_stream s);
/* nitems: all work items; wg, lds_bytes: of the persistent launch. */
int cnv_batch_info(cnv_batch b, int *nitems, int *wg, size_t *lds_bytes);
/* Per class 5 ints {op, edge, wg, z depth, items}, up to max_classes (the
 * class form's launches). */
int cnv_batch_class_info(cnv_batch b, int max_classes, int *out, int *nclasses);
/* Which form runs (1: the single persistent launch; 0: the class form, its
 * launches overlapping on the batch's streams when no item's output is
 * another's input or output; 2: the class form, its launches one after
 * another on the caller's stream) and the number of work units of the
 * single launch. */
int cnv_batch_form(cnv_batch b, int *form, int *nunits);
/* Pins the form (0, 1 or 2 as above) until the next cnv_batch_autotune. */
int cnv_batch_set_form(cnv_batch b, int form);
int cnv_batch_destroy(cnv_batch b);

/* ------------------------------------------------------------ autotuner */
/* No counterpart in the reference.  Times every valid (workgroup size,
 * z-planes per tile) pair of the fused diffusion kernel on u -> out (`iters`
 * launches each; `out` is overwritten), keeps the fastest in the
 * process-wide table used by cnv_laplacian / cnv_diffusion_step for
 * (dtype, block edge) and returns it.  results (may be NULL) receives up to
 * max_results triples {wg, zchunk, microseconds per launch}. */
int cnv_autotune_stencil7(cnv_grid u, cnv_grid out, int iters, int *wg, int *zchunk, double *results,
                          int max_results, int *nresults);
int cnv_tune_set_stencil7(int dtype, int edge, int wg, int zchunk);
int cnv_tune_get_stencil7(int dtype, int edge, int *wg, int *zchunk);
/* The same for the two-step kernel that cnv_diffusion_run uses on periodic
 * grids (no BC list on any block) of 16^3 or 32^3 blocks: `tmp` is
 * overwritten; CNV_ERR_UNSUPPORTED for other grids. */
int cnv_autotune_diffusion_run(cnv_grid u, cnv_grid tmp, int iters, int *wg, int *zchunk, double *results,
                               int max_results, int *nresults);
int cnv_tune_set_diffusion_run(int dtype, int edge, int wg, int zchunk);
int cnv_tune_get_diffusion_run(int dtype, int edge, int *wg, int *zchunk);
/* The same for the WENO5 kernel (candidates: wg 64/128/256/512, z-chunks 4..64),
 * keyed by (dtype, block edge); edges 16, 32 and 64 have tiles. */
int cnv_autotune_weno5(cnv_grid in, cnv_grid out, int iters, int *wg, int *zchunk, double *results,
                       int max_results, int *nresults);
int cnv_tune_set_weno5(int dtype, int edge, int wg, int zchunk);
int cnv_tune_get_weno5(int dtype, int edge, int *wg, int *zchunk);

/* ------------------------------------------------------------ HIP graphs */
/* No counterpart in the reference (its loop is host code per block): record
 * the calls issued on stream `s` between begin and end into an executable
 * graph, replayed with cnv_graph_launch.  Every entry point above is
 * capturable after it has run once on the same grids (first calls upload
 * BC tables synchronously); reductions use plain stores and a fixed-order
 * finalize kernel, so no memset or counter node is ever recorded.  The
 * capture runs in HIP's relaxed mode: in the thread-local mode a capture
 * begun after grids had been destroyed earlier in the process failed at its
 * first kernel launch. */
/* Does now what the first operator call on g would do synchronously (device
 * BC tables of every component, the z metric): call it before recording
 * calls on g into a graph.  A change of g's BC lists rebuilds its tables
 * (after waiting for the device), which may move them: a graph recorded
 * before the change must be recorded again. */
int cnv_grid_prepare(cnv_grid g);

typedef struct cnv_graph_s *cnv_graph;
int cnv_graph_begin(cnv_stream s);
int cnv_graph_end(cnv_stream s, cnv_graph *out);
int cnv_graph_launch(cnv_graph g, cnv_stream s);
int cnv_graph_destroy(cnv_graph g);

#ifdef __cplusplus
}
#endif

#endif /* CUBISMNOVA_AMD_H */
