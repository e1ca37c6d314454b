// cnv_api.hip -- host side of the C ABI: errors, device helpers, grid
// objects, BC lists and argument checking in front of the kernels.
#include "cnv_internal.h"

#include <cstdio>
#include <cstring>
#include <new>

namespace cnv
{

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int status, const std::string &msg)
{
    set_error(msg);
    return status;
}

int hip_fail(hipError_t e, const char *what)
{
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
    return fail(CNV_ERR_HIP, buf);
}

int check_stencil(const cnv_stencil *s)
{
    if (!s)
        return fail(CNV_ERR_INVALID, "stencil is NULL");
    // Core::Stencil::check_ (Core/Stencil.h:99-107)
    for (int d = 0; d < 3; ++d) {
        if (s->begin[d] > 0)
            return fail(CNV_ERR_INVALID, "Stencil: begin_ must be <= 0");
        if (s->end[d] < 1)
            return fail(CNV_ERR_INVALID, "Stencil: end_ must be > 0");
    }
    return CNV_OK;
}

int ensure_bc_table(cnv_grid_s *g, int comp)
{
    BCTable &t = g->tables[comp];
    if (t.version == g->bc_version)
        return CNV_OK;
    bool none = true;
    for (const auto &l : g->bcs[comp])
        if (!l.empty()) {
            none = false;
            break;
        }
    if (none) { // kernels read no table when no block carries a BC
        t.any = false;
        t.version = g->bc_version;
        return CNV_OK;
    }
    const long long nb = g->nblocks_total;
    std::vector<DevBCHead> heads((size_t)nb);
    std::vector<DevBC> entries;
    bool any = false;
    for (long long b = 0; b < nb; ++b) {
        const auto &lst = g->bcs[comp][(size_t)b];
        heads[(size_t)b].offset = (int)entries.size();
        heads[(size_t)b].count = (int)lst.size();
        for (const auto &bc : lst) {
            entries.push_back({bc.kind, bc.dir, bc.side, 0, bc.value});
            any = true;
        }
    }
    if (t.heads || t.entries) {
        // the table is rewritten (BC lists changed): kernels still queued on
        // any stream may read it, so they finish first
        CNV_HIP(hipDeviceSynchronize());
    }
    if (!t.heads)
        CNV_HIP(hipMalloc((void **)&t.heads, sizeof(DevBCHead) * (size_t)nb));
    if ((int)entries.size() > t.nentries_cap) {
        if (t.entries)
            CNV_HIP(hipFree(t.entries));
        t.entries = nullptr;
        t.nentries_cap = (int)entries.size();
        CNV_HIP(hipMalloc((void **)&t.entries, sizeof(DevBC) * (size_t)t.nentries_cap));
    }
    CNV_HIP(hipMemcpy(t.heads, heads.data(), sizeof(DevBCHead) * (size_t)nb, hipMemcpyHostToDevice));
    if (!entries.empty())
        CNV_HIP(hipMemcpy(t.entries, entries.data(), sizeof(DevBC) * entries.size(),
                          hipMemcpyHostToDevice));
    t.any = any;
    t.version = g->bc_version;
    return CNV_OK;
}

int ensure_partials(cnv_grid_s *g, size_t n)
{
    if (n > g->partials_cap) {
        if (g->partials)
            CNV_HIP(hipFree(g->partials));
        g->partials = nullptr;
        CNV_HIP(hipMalloc((void **)&g->partials, 2 * sizeof(double) * n));
        g->partials_cap = n;
    }
    return CNV_OK;
}

static int check_grid(const cnv_grid_s *g)
{
    if (!g)
        return fail(CNV_ERR_INVALID, "grid handle is NULL");
    return CNV_OK;
}

static int check_comp(const cnv_grid_s *g, int comp)
{
    if (comp < 0 || comp >= g->nslabs)
        return fail(CNV_ERR_INVALID, "component index out of range");
    return CNV_OK;
}

// the stencil operators are defined on cell-centred data
static int check_cell(const cnv_grid_s *g)
{
    if (g && g->desc.entity != CNV_CELL)
        return fail(CNV_ERR_UNSUPPORTED, "operator takes cell-centred grids only");
    return CNV_OK;
}

// Checks a descriptor and fills the layout fields of g (desc, element size,
// block and component bytes, block count, slab count, mesh spacing); no
// allocation.
int grid_layout(const cnv_grid_desc *desc, cnv_grid_s *g)
{
    if (!desc)
        return fail(CNV_ERR_INVALID, "descriptor is NULL");
    for (int d = 0; d < 3; ++d) {
        if (desc->nblocks[d] < 1 || desc->block_cells[d] < 1)
            return fail(CNV_ERR_INVALID, "nblocks and block_cells must be >= 1");
        if (!(desc->end[d] > desc->begin[d]))
            return fail(CNV_ERR_INVALID, "mesh end must exceed begin");
    }
    if (desc->ncomp < 1)
        return fail(CNV_ERR_INVALID, "ncomp must be >= 1");
    if (desc->dtype != CNV_F32 && desc->dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    if (desc->entity != CNV_CELL && desc->entity != CNV_NODE && desc->entity != CNV_FACE)
        return fail(CNV_ERR_INVALID, "entity must be CNV_CELL, CNV_NODE or CNV_FACE");
    const long long cells = (long long)desc->block_cells[0] * desc->block_cells[1] * desc->block_cells[2];
    // Cartesian::alloc_ (Grid/Cartesian.h:490-503): (N+1)^3 elements per
    // block for node and face data, 3 face slabs per component
    const long long elements = desc->entity == CNV_CELL
                                   ? cells
                                   : (long long)(desc->block_cells[0] + 1) * (desc->block_cells[1] + 1) *
                                         (desc->block_cells[2] + 1);
    const long long nbt = (long long)desc->nblocks[0] * desc->nblocks[1] * desc->nblocks[2];
    if (cells > (1LL << 30) || nbt > (1LL << 30))
        return fail(CNV_ERR_INVALID, "grid too large");
    g->desc = *desc;
    g->elem = desc->dtype == CNV_F64 ? 8 : 4;
    // Cartesian::alloc_ (Grid/Cartesian.h:490-529): 32-byte aligned blocks
    g->block_bytes = ((size_t)elements * g->elem + 31) / 32 * 32;
    g->nblocks_total = nbt;
    g->nslabs = desc->entity == CNV_FACE ? 3 * desc->ncomp : desc->ncomp;
    g->component_bytes = g->block_bytes * (size_t)nbt;
    for (int d = 0; d < 3; ++d)
        g->h[d] = (desc->end[d] - desc->begin[d]) / (double)(desc->nblocks[d] * desc->block_cells[d]);
    return CNV_OK;
}

} // namespace cnv

using namespace cnv;

#define CNV_CHECK(expr)                                                                      \
    do {                                                                                     \
        int rc_ = (expr);                                                                    \
        if (rc_ != CNV_OK)                                                                   \
            return rc_;                                                                      \
    } while (0)

#define CNV_GUARD_BEGIN try {
#define CNV_GUARD_END                                                                        \
    }                                                                                        \
    catch (const std::bad_alloc &)                                                           \
    {                                                                                        \
        return fail(CNV_ERR_ALLOC, "host allocation failed");                                \
    }                                                                                        \
    catch (...)                                                                              \
    {                                                                                        \
        return fail(CNV_ERR_INVALID, "unexpected exception");                                \
    }

extern "C" {

const char *cnv_strerror(int status)
{
    switch (status) {
    case CNV_OK:
        return "success";
    case CNV_ERR_INVALID:
        return "invalid argument";
    case CNV_ERR_ALLOC:
        return "allocation failed";
    case CNV_ERR_HIP:
        return "HIP runtime error";
    case CNV_ERR_UNSUPPORTED:
        return "unsupported configuration";
    case CNV_ERR_NOT_ALLOCATED:
        return "FieldLab: can not load lab data when not allocated first";
    case CNV_ERR_SIZE:
        return "size mismatch";
    default:
        return "unknown status";
    }
}

const char *cnv_last_error(void) { return g_last_error.c_str(); }

int cnv_device_count(int *count)
{
    if (!count)
        return fail(CNV_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return CNV_OK;
}

int cnv_malloc(void **ptr, size_t bytes)
{
    if (!ptr)
        return fail(CNV_ERR_INVALID, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0)
        return CNV_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess)
        return fail(CNV_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    return CNV_OK;
}

int cnv_free(void *ptr)
{
    if (ptr)
        CNV_HIP(hipFree(ptr));
    return CNV_OK;
}

int cnv_memcpy_h2d(void *dst, const void *src, size_t bytes, cnv_stream s)
{
    CNV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s));
    CNV_HIP(hipStreamSynchronize((hipStream_t)s));
    return CNV_OK;
}

int cnv_memcpy_d2h(void *dst, const void *src, size_t bytes, cnv_stream s)
{
    CNV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s));
    CNV_HIP(hipStreamSynchronize((hipStream_t)s));
    return CNV_OK;
}

int cnv_memcpy_d2d(void *dst, const void *src, size_t bytes, cnv_stream s)
{
    CNV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s));
    return CNV_OK;
}

int cnv_fill(void *dst, int dtype, double value, size_t count, cnv_stream s)
{
    if (!dst && count)
        return fail(CNV_ERR_INVALID, "dst is NULL");
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    return launch_fill(dst, dtype, value, count, (hipStream_t)s);
}

int cnv_stream_create(cnv_stream *s)
{
    if (!s)
        return fail(CNV_ERR_INVALID, "stream pointer is NULL");
    hipStream_t st;
    CNV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *s = (cnv_stream)st;
    return CNV_OK;
}

int cnv_stream_destroy(cnv_stream s)
{
    if (s)
        CNV_HIP(hipStreamDestroy((hipStream_t)s));
    return CNV_OK;
}

int cnv_stream_synchronize(cnv_stream s)
{
    CNV_HIP(hipStreamSynchronize((hipStream_t)s));
    return CNV_OK;
}

// ---------------------------------------------------------------- grids

// Grid over `data` (device memory of at least the grid's bytes) or, with
// data == nullptr, over a new allocation the grid owns.
static int grid_make(const cnv_grid_desc *desc, void *data, size_t data_bytes, cnv_grid *out)
{
    if (!desc || !out)
        return fail(CNV_ERR_INVALID, "descriptor or output handle is NULL");
    *out = nullptr;
    cnv_grid_s lay;
    CNV_CHECK(grid_layout(desc, &lay));
    const long long cells = (long long)desc->block_cells[0] * desc->block_cells[1] * desc->block_cells[2];
    const long long nbt = lay.nblocks_total;
    const int nslabs = lay.nslabs;
    const size_t all = lay.component_bytes * (size_t)nslabs;
    if (data) {
        if ((reinterpret_cast<uintptr_t>(data) & 31) != 0)
            return fail(CNV_ERR_INVALID, "wrapped grid memory must be 32-byte aligned");
        if (data_bytes < all)
            return fail(CNV_ERR_SIZE, "wrapped memory smaller than the grid");
        hipPointerAttribute_t attr;
        if (hipPointerGetAttributes(&attr, data) != hipSuccess ||
            (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged)) {
            (void)hipGetLastError();
            return fail(CNV_ERR_INVALID, "wrapped grid memory is not device memory");
        }
    }

    cnv_grid_s *g = new cnv_grid_s();
    CNV_CHECK(grid_layout(desc, g));
    g->bcs.assign((size_t)nslabs, std::vector<std::vector<cnv_bc>>((size_t)nbt));
    g->tables.resize((size_t)nslabs);
    if (data) {
        g->data = data;
        g->owns_data = false;
    } else {
        const hipError_t e = hipMalloc(&g->data, all);
        if (e != hipSuccess) {
            delete g;
            return fail(CNV_ERR_ALLOC, std::string("grid allocation: ") + hipGetErrorString(e));
        }
    }
    // reduction scratch sized for the largest launch of any operator, so that
    // no call allocates later (calls may be recorded into a HIP graph)
    const size_t nparts = (size_t)nbt * desc->ncomp * ((size_t)(cells + 255) / 256 + desc->block_cells[2]) + 4096;
    const hipError_t e = hipMalloc((void **)&g->partials, 2 * sizeof(double) * nparts); // two halves
    if (e != hipSuccess) {
        if (g->owns_data)
            (void)hipFree(g->data);
        delete g;
        return fail(CNV_ERR_ALLOC, std::string("scratch allocation: ") + hipGetErrorString(e));
    }
    g->partials_cap = nparts;
    *out = g;
    return CNV_OK;
}

int cnv_grid_create(const cnv_grid_desc *desc, cnv_grid *out)
{
    CNV_GUARD_BEGIN
    return grid_make(desc, nullptr, 0, out);
    CNV_GUARD_END
}

int cnv_grid_wrap(const cnv_grid_desc *desc, void *device_data, size_t bytes, cnv_grid *out)
{
    CNV_GUARD_BEGIN
    if (!device_data)
        return fail(CNV_ERR_INVALID, "wrapped grid memory is NULL");
    return grid_make(desc, device_data, bytes, out);
    CNV_GUARD_END
}

int cnv_grid_destroy(cnv_grid g)
{
    if (!g)
        return CNV_OK;
    if (g->owns_data)
        (void)hipFree(g->data);
    for (auto &t : g->tables) {
        if (t.heads)
            (void)hipFree(t.heads);
        if (t.entries)
            (void)hipFree(t.entries);
    }
    if (g->partials)
        (void)hipFree(g->partials);
    if (g->czm_dev)
        (void)hipFree(g->czm_dev);
    if (g->czp_dev)
        (void)hipFree(g->czp_dev);
    delete g;
    return CNV_OK;
}

int cnv_grid_info(cnv_grid g, cnv_grid_desc *desc, size_t *block_bytes, size_t *component_bytes)
{
    CNV_CHECK(check_grid(g));
    if (desc)
        *desc = g->desc;
    if (block_bytes)
        *block_bytes = g->block_bytes;
    if (component_bytes)
        *component_bytes = g->component_bytes;
    return CNV_OK;
}

int cnv_grid_data(cnv_grid g, int comp, void **device_ptr)
{
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_comp(g, comp));
    if (!device_ptr)
        return fail(CNV_ERR_INVALID, "device_ptr is NULL");
    *device_ptr = (char *)g->data + (size_t)comp * g->component_bytes;
    return CNV_OK;
}

int cnv_grid_upload(cnv_grid g, int comp, const void *host, size_t bytes, cnv_stream s)
{
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_comp(g, comp));
    if (bytes != g->component_bytes)
        return fail(CNV_ERR_SIZE, "upload size must equal component_bytes");
    if (!host)
        return fail(CNV_ERR_INVALID, "host pointer is NULL");
    CNV_HIP(hipMemcpyAsync((char *)g->data + (size_t)comp * g->component_bytes, host, bytes,
                           hipMemcpyHostToDevice, (hipStream_t)s));
    CNV_HIP(hipStreamSynchronize((hipStream_t)s));
    return CNV_OK;
}

int cnv_grid_download(cnv_grid g, int comp, void *host, size_t bytes, cnv_stream s)
{
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_comp(g, comp));
    if (bytes != g->component_bytes)
        return fail(CNV_ERR_SIZE, "download size must equal component_bytes");
    if (!host)
        return fail(CNV_ERR_INVALID, "host pointer is NULL");
    CNV_HIP(hipMemcpyAsync(host, (char *)g->data + (size_t)comp * g->component_bytes, bytes,
                           hipMemcpyDeviceToHost, (hipStream_t)s));
    CNV_HIP(hipStreamSynchronize((hipStream_t)s));
    return CNV_OK;
}

// ---------------------------------------------------------------- BCs

static int check_bc(const cnv_bc &b)
{
    if (b.kind < CNV_BC_BASE || b.kind > CNV_BC_SYMMETRY)
        return fail(CNV_ERR_INVALID, "unknown BC kind");
    if (b.dir < 0 || b.dir > 2)
        return fail(CNV_ERR_INVALID, "BC dir must be 0, 1 or 2");
    if (b.side != 0 && b.side != 1)
        return fail(CNV_ERR_INVALID, "BC side must be 0 or 1");
    return CNV_OK;
}

int cnv_grid_set_block_bcs(cnv_grid g, int comp, const int block[3], const cnv_bc *bcs, int n)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_comp(g, comp));
    if (!block || n < 0 || (n > 0 && !bcs))
        return fail(CNV_ERR_INVALID, "bad block index or BC list");
    for (int d = 0; d < 3; ++d)
        if (block[d] < 0 || block[d] >= g->desc.nblocks[d])
            return fail(CNV_ERR_INVALID, "block index out of range");
    for (int k = 0; k < n; ++k)
        CNV_CHECK(check_bc(bcs[k]));
    std::lock_guard<std::mutex> lk(g->mtx);
    const long long b = block[0] + (long long)g->desc.nblocks[0] * (block[1] + (long long)g->desc.nblocks[1] * block[2]);
    g->bcs[comp][(size_t)b].assign(bcs, bcs + n);
    ++g->bc_version;
    return CNV_OK;
    CNV_GUARD_END
}

int cnv_grid_add_face_bc(cnv_grid g, int comp, const cnv_bc *bc)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(g));
    if (!bc)
        return fail(CNV_ERR_INVALID, "bc is NULL");
    CNV_CHECK(check_bc(*bc));
    if (comp != -1)
        CNV_CHECK(check_comp(g, comp));
    std::lock_guard<std::mutex> lk(g->mtx);
    const int *nb = g->desc.nblocks;
    for (int c = 0; c < g->nslabs; ++c) {
        if (comp != -1 && c != comp)
            continue;
        for (int bz = 0; bz < nb[2]; ++bz)
            for (int by = 0; by < nb[1]; ++by)
                for (int bx = 0; bx < nb[0]; ++bx) {
                    const int bi[3] = {bx, by, bz};
                    const int want = bc->side == 0 ? 0 : nb[bc->dir] - 1;
                    if (bi[bc->dir] != want)
                        continue;
                    const long long b = bx + (long long)nb[0] * (by + (long long)nb[1] * bz);
                    g->bcs[c][(size_t)b].push_back(*bc);
                }
    }
    ++g->bc_version;
    return CNV_OK;
    CNV_GUARD_END
}

int cnv_grid_clear_bcs(cnv_grid g, int comp)
{
    CNV_CHECK(check_grid(g));
    if (comp != -1)
        CNV_CHECK(check_comp(g, comp));
    std::lock_guard<std::mutex> lk(g->mtx);
    for (int c = 0; c < g->nslabs; ++c)
        if (comp == -1 || c == comp)
            for (auto &l : g->bcs[c])
                l.clear();
    ++g->bc_version;
    return CNV_OK;
}

// ---------------------------------------------------------------- lab

int cnv_lab_extent(cnv_grid g, const cnv_stencil *s, int extent[3])
{
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_stencil(s));
    if (!extent)
        return fail(CNV_ERR_INVALID, "extent is NULL");
    const int plus = g->desc.entity == CNV_CELL ? 0 : 1;
    for (int d = 0; d < 3; ++d)
        extent[d] = g->desc.block_cells[d] + plus + s->end[d] - 1 - s->begin[d];
    return CNV_OK;
}

int cnv_lab_geometry(cnv_grid g, int comp, const int block[3], const cnv_stencil *s, int pitch[3], int active[3])
{
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_comp(g, comp));
    CNV_CHECK(check_stencil(s));
    if (block)
        for (int d = 0; d < 3; ++d)
            if (block[d] < 0 || block[d] >= g->desc.nblocks[d])
                return fail(CNV_ERR_INVALID, "block index out of range");
    const int pm = plus_mask(g, comp);
    for (int d = 0; d < 3; ++d) {
        const int p = (pm >> d) & 1;
        const int wide = p && (!block || block[d] == g->desc.nblocks[d] - 1);
        const int w = s->end[d] - 1 - s->begin[d];
        if (pitch)
            pitch[d] = g->desc.block_cells[d] + p + w;
        if (active)
            active[d] = g->desc.block_cells[d] + wide + w;
    }
    return CNV_OK;
}

static int check_lab_widths(const cnv_grid_s *g, const cnv_stencil *s)
{
    for (int d = 0; d < 3; ++d)
        if (-s->begin[d] > g->desc.block_cells[d] || s->end[d] - 1 > g->desc.block_cells[d])
            return fail(CNV_ERR_UNSUPPORTED, "ghost width larger than the block extent");
    return CNV_OK;
}

int cnv_lab_load(cnv_grid g, int comp, const int block[3], const cnv_stencil *s, int apply_bc,
                 const cnv_bc *extern_bcs, int n_extern, void *out, cnv_stream st)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_comp(g, comp));
    CNV_CHECK(check_stencil(s));
    CNV_CHECK(check_lab_widths(g, s));
    if (!block || !out)
        return fail(CNV_ERR_INVALID, "block index or output is NULL");
    for (int d = 0; d < 3; ++d)
        if (block[d] < 0 || block[d] >= g->desc.nblocks[d])
            return fail(CNV_ERR_INVALID, "block index out of range");
    const std::vector<cnv_bc> *lst = nullptr;
    std::vector<cnv_bc> ext;
    if (n_extern >= 0) {
        if (n_extern > 0 && !extern_bcs)
            return fail(CNV_ERR_INVALID, "extern BC list is NULL");
        for (int k = 0; k < n_extern; ++k)
            CNV_CHECK(check_bc(extern_bcs[k]));
        ext.assign(extern_bcs, extern_bcs + n_extern);
        lst = &ext;
    }
    DevBC list[16];
    int nl = 0;
    {
        // the block's own list is read under the grid's lock (other threads
        // may set BC lists meanwhile)
        std::unique_lock<std::mutex> lk(g->mtx, std::defer_lock);
        if (!lst) {
            lk.lock();
            const long long b =
                block[0] + (long long)g->desc.nblocks[0] * (block[1] + (long long)g->desc.nblocks[1] * block[2]);
            lst = &g->bcs[comp][(size_t)b];
        }
        if (lst->size() > 16)
            return fail(CNV_ERR_UNSUPPORTED, "more than 16 BCs on one block");
        for (size_t k = 0; k < lst->size(); ++k)
            list[k] = {(*lst)[k].kind, (*lst)[k].dir, (*lst)[k].side, 0, (*lst)[k].value};
        nl = (int)lst->size();
    }
    return launch_lab(g, comp, *s, apply_bc != 0, true, list, nl, block, out, (hipStream_t)st);
    CNV_GUARD_END
}

int cnv_lab_load_all(cnv_grid g, int comp, const cnv_stencil *s, void *out, cnv_stream st)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_comp(g, comp));
    CNV_CHECK(check_stencil(s));
    CNV_CHECK(check_lab_widths(g, s));
    if (!out)
        return fail(CNV_ERR_INVALID, "output is NULL");
    {
        std::lock_guard<std::mutex> lk(g->mtx);
        CNV_CHECK(ensure_bc_table(g, comp));
    }
    return launch_lab(g, comp, *s, true, false, nullptr, 0, nullptr, out, (hipStream_t)st);
    CNV_GUARD_END
}

// ---------------------------------------------------------------- operators

static int same_shape(const cnv_grid_s *a, const cnv_grid_s *b)
{
    for (int d = 0; d < 3; ++d)
        if (a->desc.nblocks[d] != b->desc.nblocks[d] || a->desc.block_cells[d] != b->desc.block_cells[d])
            return fail(CNV_ERR_SIZE, "grids differ in block count or block size");
    if (a->desc.dtype != b->desc.dtype)
        return fail(CNV_ERR_SIZE, "grids differ in dtype");
    return CNV_OK;
}

int cnv_laplacian(cnv_grid in, int comp_in, cnv_grid out, int comp_out, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(in));
    CNV_CHECK(check_cell(in));
    CNV_CHECK(check_grid(out));
    CNV_CHECK(check_cell(out));
    CNV_CHECK(check_comp(in, comp_in));
    CNV_CHECK(check_comp(out, comp_out));
    CNV_CHECK(same_shape(in, out));
    if (grids_overlap(in, out))
        return fail(CNV_ERR_INVALID, "input and output grids must not share memory");
    return launch_stencil7(in, out, nullptr, comp_in, comp_out, 1, 0, 0.0, nullptr, (hipStream_t)s);
    CNV_GUARD_END
}

int cnv_diffusion_step(cnv_grid u, cnv_grid u_new, double nu_dt, cnv_grid lap_out, double *sumsq,
                       cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(u));
    CNV_CHECK(check_cell(u));
    CNV_CHECK(check_grid(u_new));
    CNV_CHECK(check_cell(u_new));
    CNV_CHECK(same_shape(u, u_new));
    if (grids_overlap(u, u_new))
        return fail(CNV_ERR_INVALID, "u and u_new must not share memory");
    if (u_new->desc.ncomp < u->desc.ncomp)
        return fail(CNV_ERR_SIZE, "u_new has fewer components than u");
    if (lap_out) {
        CNV_CHECK(same_shape(u, lap_out));
        if (lap_out->desc.ncomp < u->desc.ncomp || grids_overlap(lap_out, u) || grids_overlap(lap_out, u_new))
            return fail(CNV_ERR_INVALID, "lap_out must be a separate grid with enough components");
    }
    return launch_stencil7(u, u_new, lap_out, 0, 0, u->desc.ncomp, lap_out ? 2 : 1, nu_dt, sumsq,
                           (hipStream_t)s);
    CNV_GUARD_END
}

int cnv_diffusion_step_noalloc(const cnv_grid_desc *desc, const void *u, void *u_new, double nu_dt, double *sumsq,
                               cnv_stream s)
{
    CNV_GUARD_BEGIN
    if (!u || !u_new)
        return fail(CNV_ERR_INVALID, "u or u_new is NULL");
    if (u == u_new)
        return fail(CNV_ERR_INVALID, "u and u_new must be different memory");
    if (((reinterpret_cast<uintptr_t>(u) | reinterpret_cast<uintptr_t>(u_new)) & 31) != 0)
        return fail(CNV_ERR_INVALID, "grid memory must be 32-byte aligned");
    return launch_diffusion_noalloc(desc, u, u_new, nu_dt, sumsq, (hipStream_t)s);
    CNV_GUARD_END
}

int cnv_diffusion_run(cnv_grid u, cnv_grid tmp, int nsteps, double nu_dt, double *norms, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(u));
    CNV_CHECK(check_cell(u));
    CNV_CHECK(check_grid(tmp));
    CNV_CHECK(check_cell(tmp));
    CNV_CHECK(same_shape(u, tmp));
    if (grids_overlap(u, tmp))
        return fail(CNV_ERR_INVALID, "u and tmp must not share memory");
    if (tmp->desc.ncomp < u->desc.ncomp)
        return fail(CNV_ERR_SIZE, "tmp has fewer components than u");
    if (nsteps < 0)
        return fail(CNV_ERR_INVALID, "nsteps must be >= 0");
    // tmp is read as the source of every other step: same BCs as u
    if (nsteps > 1) {
        std::lock_guard<std::mutex> lk(tmp->mtx);
        bool same = true;
        for (int c = 0; c < u->desc.ncomp && same; ++c)
            for (long long b = 0; b < u->nblocks_total && same; ++b) {
                const auto &x = u->bcs[c][(size_t)b], &y = tmp->bcs[c][(size_t)b];
                same = x.size() == y.size();
                for (size_t i = 0; same && i < x.size(); ++i)
                    same = x[i].kind == y[i].kind && x[i].dir == y[i].dir && x[i].side == y[i].side &&
                           x[i].value == y[i].value;
            }
        if (!same)
            return fail(CNV_ERR_INVALID, "u and tmp must carry the same BC lists");
    }
    return launch_diffusion_run(u, tmp, nsteps, nu_dt, norms, (hipStream_t)s);
    CNV_GUARD_END
}

int cnv_grid_sums(cnv_grid g, int comp, double *out, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_cell(g));
    CNV_CHECK(check_comp(g, comp));
    if (!out)
        return fail(CNV_ERR_INVALID, "out is NULL");
    return launch_sums(g, comp, out, (hipStream_t)s);
    CNV_GUARD_END
}

int cnv_grid_compare(cnv_grid a, int comp_a, cnv_grid b, int comp_b, double *out, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(a));
    CNV_CHECK(check_cell(a));
    CNV_CHECK(check_grid(b));
    CNV_CHECK(check_cell(b));
    CNV_CHECK(check_comp(a, comp_a));
    CNV_CHECK(check_comp(b, comp_b));
    if (!out)
        return fail(CNV_ERR_INVALID, "out is NULL");
    for (int d = 0; d < 3; ++d)
        if (a->desc.nblocks[d] != b->desc.nblocks[d] || a->desc.block_cells[d] != b->desc.block_cells[d])
            return fail(CNV_ERR_SIZE, "grids differ in block count or block size");
    return launch_compare(a, comp_a, b, comp_b, out, (hipStream_t)s);
    CNV_GUARD_END
}

int cnv_weno5(cnv_grid in, cnv_grid out, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(in));
    CNV_CHECK(check_cell(in));
    CNV_CHECK(check_grid(out));
    CNV_CHECK(check_cell(out));
    CNV_CHECK(same_shape(in, out));
    if (grids_overlap(in, out))
        return fail(CNV_ERR_INVALID, "input and output grids must not share memory");
    if (out->desc.ncomp < 6 * in->desc.ncomp)
        return fail(CNV_ERR_SIZE, "out needs 6 components per input component");
    return launch_weno5(in, out, (hipStream_t)s, nullptr);
    CNV_GUARD_END
}

int cnv_autotune_weno5(cnv_grid in, cnv_grid out, int iters, int *wg, int *zchunk, double *results,
                       int max_results, int *nresults)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(in));
    CNV_CHECK(check_cell(in));
    CNV_CHECK(check_grid(out));
    CNV_CHECK(check_cell(out));
    CNV_CHECK(same_shape(in, out));
    if (grids_overlap(in, out) || out->desc.ncomp < 6 * in->desc.ncomp)
        return fail(CNV_ERR_INVALID, "out must be a separate grid with 6 components per input component");
    return autotune_weno5(in, out, iters, wg, zchunk, results, max_results, nresults);
    CNV_GUARD_END
}

int cnv_tune_set_weno5(int dtype, int edge, int wg, int zchunk)
{
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    if (wg != 64 && wg != 128 && wg != 256 && wg != 512)
        return fail(CNV_ERR_INVALID, "workgroup size must be 64, 128, 256 or 512");
    return weno5_set_tune(dtype, edge, wg, zchunk);
}

int cnv_tune_get_weno5(int dtype, int edge, int *wg, int *zchunk)
{
    if (!wg || !zchunk)
        return fail(CNV_ERR_INVALID, "output pointer is NULL");
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    return weno5_get_tune(dtype, edge, wg, zchunk);
}

// ---------------------------------------------------------------- Jacobi

int cnv_grid_set_zfaces(cnv_grid g, const double *zfaces, int n)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(g));
    CNV_CHECK(check_cell(g));
    if (!zfaces)
        return fail(CNV_ERR_INVALID, "zfaces is NULL");
    std::lock_guard<std::mutex> lk(g->mtx);
    return set_zfaces(g, zfaces, n);
    CNV_GUARD_END
}

static int check_jacobi_grids(const cnv_grid_s *u, int cu, const cnv_grid_s *f, int cf, const cnv_grid_s *o, int co)
{
    CNV_CHECK(check_grid(u));
    CNV_CHECK(check_cell(u));
    CNV_CHECK(check_grid(f));
    CNV_CHECK(check_cell(f));
    CNV_CHECK(check_grid(o));
    CNV_CHECK(check_cell(o));
    CNV_CHECK(check_comp(u, cu));
    CNV_CHECK(check_comp(f, cf));
    CNV_CHECK(check_comp(o, co));
    for (int d = 0; d < 3; ++d)
        if (u->desc.nblocks[d] != f->desc.nblocks[d] || u->desc.block_cells[d] != f->desc.block_cells[d] ||
            u->desc.nblocks[d] != o->desc.nblocks[d] || u->desc.block_cells[d] != o->desc.block_cells[d])
            return fail(CNV_ERR_SIZE, "grids differ in block count or block size");
    if (slabs_overlap(u, cu, o, co) || slabs_overlap(f, cf, o, co))
        return fail(CNV_ERR_INVALID, "output must not alias an input component");
    return CNV_OK;
}

int cnv_jacobi_sweep(cnv_grid u, int cu, cnv_grid f, int cf, cnv_grid out, int co, double omega, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_jacobi_grids(u, cu, f, cf, out, co));
    return launch_jacobi(u, cu, f, cf, out, co, 0, omega, nullptr, (hipStream_t)s);
    CNV_GUARD_END
}

int cnv_jacobi_defect(cnv_grid u, int cu, cnv_grid f, int cf, cnv_grid r, int cr, double *sumsq, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_jacobi_grids(u, cu, f, cf, r, cr));
    return launch_jacobi(u, cu, f, cf, r, cr, 1, 0.0, sumsq, (hipStream_t)s);
    CNV_GUARD_END
}

int cnv_autotune_jacobi(cnv_grid u, cnv_grid f, cnv_grid out, int iters, int *wg, int *zchunk, int *prefetch,
                        double *results, int max_results, int *nresults)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_jacobi_grids(u, 0, f, 0, out, 0));
    if (f->desc.dtype != u->desc.dtype || out->desc.dtype != u->desc.dtype)
        return fail(CNV_ERR_SIZE, "u, f and out must share a dtype");
    if (slabs_overlap(u, 0, f, 0))
        return fail(CNV_ERR_INVALID, "autotune needs u and f in separate memory");
    return autotune_jacobi(u, f, out, iters, wg, zchunk, prefetch, results, max_results, nresults);
    CNV_GUARD_END
}

int cnv_tune_set_jacobi(int dtype, int edge, int wg, int zchunk, int prefetch)
{
    return jacobi_set_tune(dtype, edge, wg, zchunk, prefetch);
}

int cnv_tune_get_jacobi(int dtype, int edge, int *wg, int *zchunk, int *prefetch)
{
    if (!wg || !zchunk || !prefetch)
        return fail(CNV_ERR_INVALID, "output pointer is NULL");
    return jacobi_get_tune(dtype, edge, wg, zchunk, prefetch);
}

int cnv_jacobi_sweep2(cnv_grid u, int cu, cnv_grid f, int cf, cnv_grid out, int co, double omega, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_jacobi_grids(u, cu, f, cf, out, co));
    return launch_jacobi2(u, cu, f, cf, out, co, omega, (hipStream_t)s, nullptr);
    CNV_GUARD_END
}

int cnv_tune_set_jacobi2(int dtype, int edge, int wg, int zchunk, int ysplit)
{
    return jacobi2_set_tune(dtype, edge, wg, zchunk, ysplit);
}

int cnv_tune_get_jacobi2(int dtype, int edge, int *wg, int *zchunk, int *ysplit)
{
    if (!wg || !zchunk || !ysplit)
        return fail(CNV_ERR_INVALID, "output pointer is NULL");
    return jacobi2_get_tune(dtype, edge, wg, zchunk, ysplit);
}

int cnv_add_correction(cnv_grid u, int cu, cnv_grid e, int ce, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(u));
    CNV_CHECK(check_cell(u));
    CNV_CHECK(check_grid(e));
    CNV_CHECK(check_cell(e));
    CNV_CHECK(check_comp(u, cu));
    CNV_CHECK(check_comp(e, ce));
    if (u->desc.dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "the corrected grid must be FP64");
    for (int d = 0; d < 3; ++d)
        if (u->desc.nblocks[d] != e->desc.nblocks[d] || u->desc.block_cells[d] != e->desc.block_cells[d])
            return fail(CNV_ERR_SIZE, "grids differ in block count or block size");
    return launch_correct(u, cu, e, ce, (hipStream_t)s);
    CNV_GUARD_END
}

// BC lists of `src` component 0 copied to `dst` component c; homogeneous
// copies zero the Dirichlet values (the correction equation A e = r).
static void copy_bcs(cnv_grid_s *src, cnv_grid_s *dst, int c, bool homogeneous)
{
    std::lock_guard<std::mutex> lk(dst->mtx);
    for (long long b = 0; b < dst->nblocks_total; ++b) {
        auto l = src->bcs[0][(size_t)b];
        if (homogeneous)
            for (auto &x : l)
                if (x.kind == CNV_BC_DIRICHLET)
                    x.value = 0.0;
        dst->bcs[c][(size_t)b] = l;
    }
    ++dst->bc_version;
}

int cnv_jacobi_solve(cnv_grid u, cnv_grid f, cnv_grid work, const cnv_solve_opts *o, cnv_solve_info *info,
                     double *history, int history_cap, cnv_stream s)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(u));
    CNV_CHECK(check_cell(u));
    CNV_CHECK(check_grid(f));
    CNV_CHECK(check_cell(f));
    CNV_CHECK(check_grid(work));
    CNV_CHECK(check_cell(work));
    if (!o || !info)
        return fail(CNV_ERR_INVALID, "options or info is NULL");
    if (o->sweeps_per_cycle < 2 || o->sweeps_per_cycle % 2 != 0)
        return fail(CNV_ERR_INVALID, "sweeps_per_cycle must be even and >= 2");
    if (o->max_cycles < 0)
        return fail(CNV_ERR_INVALID, "max_cycles must be >= 0");
    if (u->desc.dtype != f->desc.dtype)
        return fail(CNV_ERR_SIZE, "u and f must share a dtype");
    const bool refine = o->refine != 0;
    if (refine && (u->desc.dtype != CNV_F64 || work->desc.dtype != CNV_F32 || work->desc.ncomp < 3))
        return fail(CNV_ERR_INVALID, "refinement needs FP64 u and f and an FP32 work grid with 3 components");
    if (!refine && (work->desc.dtype != u->desc.dtype || work->desc.ncomp < 2))
        return fail(CNV_ERR_INVALID, "plain Jacobi needs a work grid of u's dtype with 2 components");
    CNV_CHECK(check_jacobi_grids(u, 0, f, 0, work, 0));
    // the correction e lives in work components 0/1 with homogeneous BCs;
    // plain sweeps ping-pong between u and work component 0 with u's BCs
    copy_bcs(u, work, 0, refine);
    if (refine)
        copy_bcs(u, work, 1, true);
    // set up metric and BC tables now: nothing may allocate inside the capture
    CNV_CHECK(jacobi_prepare(u, 0));
    {
        std::lock_guard<std::mutex> lk(work->mtx);
        const int nz = u->desc.nblocks[2] * u->desc.block_cells[2];
        if (!work->czm_dev) {
            CNV_HIP(hipMalloc((void **)&work->czm_dev, sizeof(double) * nz));
            CNV_HIP(hipMalloc((void **)&work->czp_dev, sizeof(double) * nz));
        }
        CNV_HIP(hipMemcpy(work->czm_dev, u->czm_dev, sizeof(double) * nz, hipMemcpyDeviceToDevice));
        CNV_HIP(hipMemcpy(work->czp_dev, u->czp_dev, sizeof(double) * nz, hipMemcpyDeviceToDevice));
    }
    CNV_CHECK(jacobi_prepare(work, 0));
    CNV_CHECK(jacobi_prepare(work, 1));
    hipStream_t st = (hipStream_t)s;
    hipStream_t cs;
    CNV_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    double *dev = nullptr; // [0] = |f|^2 pieces, [2] = |r|^2
    int rc = CNV_OK;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double hst[3] = {0, 0, 0};
    const int S = o->sweeps_per_cycle;
    long long sweeps = 0;
    int cycles = 0;
    double rel = 0.0;
    auto defect = [&](hipStream_t q) -> int {
        // refine: r (FP32, work comp 2) = f - A u in FP64; plain: work comp 1
        return launch_jacobi(u, 0, f, 0, work, refine ? 2 : 1, 1, 0.0, dev + 2, q);
    };
    do {
        if (hipMalloc((void **)&dev, 3 * sizeof(double)) != hipSuccess) {
            rc = fail(CNV_ERR_ALLOC, "solver scratch");
            break;
        }
        if (st)
            CNV_HIP(hipStreamSynchronize(st));
        rc = launch_sums(f, 0, dev, cs);
        if (rc != CNV_OK)
            break;
        rc = defect(cs);
        if (rc != CNV_OK)
            break;
        CNV_HIP(hipMemcpyAsync(hst, dev, 3 * sizeof(double), hipMemcpyDeviceToHost, cs));
        CNV_HIP(hipStreamSynchronize(cs));
        const double fn = std::sqrt(hst[1]);
        if (!(fn > 0.0)) {
            rel = 0.0;
            break;
        }
        rel = std::sqrt(hst[2]) / fn;
        if (history && history_cap > 0)
            history[0] = rel;
        // one cycle, recorded once and replayed (relaxed capture: see
        // cnv_graph_begin)
        CNV_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
        // sweep pairs in one launch each where the faces allow it (the same
        // values as single sweeps, u and f read once per pair)
        int crc = CNV_OK;
        if (refine) {
            const bool two = jacobi2_applies(work, 0, work, work) && jacobi2_applies(work, 1, work, work);
            crc = launch_fill((char *)work->data, CNV_F32, 0.0, work->component_bytes / 4, cs);
            const int L = two ? S / 2 : S;
            for (int k = 0; k < L && crc == CNV_OK; ++k)
                crc = two ? launch_jacobi2(work, k % 2, work, 2, work, 1 - k % 2, o->omega, cs, nullptr)
                          : launch_jacobi(work, k % 2, work, 2, work, 1 - k % 2, 0, o->omega, nullptr, cs);
            if (crc == CNV_OK)
                crc = launch_correct(u, 0, work, L % 2, cs);
        } else {
            // an even number of pair launches (u -> work -> u ...), then the
            // remaining two sweeps, if any, one at a time: the cycle ends in u
            const bool two = jacobi2_applies(u, 0, f, work) && jacobi2_applies(work, 0, f, u);
            const int pairs = two ? (S / 2) & ~1 : 0;
            for (int k = 0; k < pairs + (S - 2 * pairs) && crc == CNV_OK; ++k) {
                cnv_grid_s *a = k % 2 == 0 ? u : work, *b = k % 2 == 0 ? work : u;
                crc = k < pairs ? launch_jacobi2(a, 0, f, 0, b, 0, o->omega, cs, nullptr)
                                : launch_jacobi(a, 0, f, 0, b, 0, 0, o->omega, nullptr, cs);
            }
        }
        if (crc == CNV_OK)
            crc = defect(cs);
        hipError_t ce = hipStreamEndCapture(cs, &graph);
        if (crc != CNV_OK) {
            rc = crc;
            break;
        }
        if (ce != hipSuccess) {
            rc = hip_fail(ce, "hipStreamEndCapture");
            break;
        }
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            rc = fail(CNV_ERR_HIP, "hipGraphInstantiate failed");
            break;
        }
        while (cycles < o->max_cycles && rel > o->tol) {
            CNV_HIP(hipGraphLaunch(exec, cs));
            CNV_HIP(hipMemcpyAsync(hst + 2, dev + 2, sizeof(double), hipMemcpyDeviceToHost, cs));
            CNV_HIP(hipStreamSynchronize(cs));
            ++cycles;
            sweeps += S;
            rel = std::sqrt(hst[2]) / fn;
            if (history && cycles < history_cap)
                history[cycles] = rel;
            if (!std::isfinite(rel)) {
                rc = fail(CNV_ERR_INVALID, "solver diverged (non-finite residual)");
                break;
            }
        }
    } while (false);
    if (exec)
        (void)hipGraphExecDestroy(exec);
    if (graph)
        (void)hipGraphDestroy(graph);
    if (dev)
        (void)hipFree(dev);
    (void)hipStreamDestroy(cs);
    info->cycles = cycles;
    info->sweeps = sweeps;
    info->rel_residual = rel;
    return rc;
    CNV_GUARD_END
}

int cnv_grid_prepare(cnv_grid g)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(g));
    for (int c = 0; c < g->nslabs; ++c) {
        if (g->desc.entity == CNV_CELL) {
            CNV_CHECK(jacobi_prepare(g, c));
        } else {
            std::lock_guard<std::mutex> lk(g->mtx);
            CNV_CHECK(ensure_bc_table(g, c));
        }
    }
    return CNV_OK;
    CNV_GUARD_END
}

// ---------------------------------------------------------------- tuning

int cnv_autotune_stencil7(cnv_grid u, cnv_grid out, int iters, int *wg, int *zchunk, double *results,
                          int max_results, int *nresults)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(u));
    CNV_CHECK(check_cell(u));
    CNV_CHECK(check_grid(out));
    CNV_CHECK(check_cell(out));
    CNV_CHECK(same_shape(u, out));
    if (grids_overlap(u, out))
        return fail(CNV_ERR_INVALID, "autotune needs two grids that do not share memory");
    if (out->desc.ncomp < u->desc.ncomp)
        return fail(CNV_ERR_SIZE, "out has fewer components than u");
    return autotune_stencil7(u, out, iters, wg, zchunk, results, max_results, nresults);
    CNV_GUARD_END
}

int cnv_autotune_diffusion_run(cnv_grid u, cnv_grid tmp, int iters, int *wg, int *zchunk, double *results,
                               int max_results, int *nresults)
{
    CNV_GUARD_BEGIN
    CNV_CHECK(check_grid(u));
    CNV_CHECK(check_cell(u));
    CNV_CHECK(check_grid(tmp));
    CNV_CHECK(check_cell(tmp));
    CNV_CHECK(same_shape(u, tmp));
    if (grids_overlap(u, tmp))
        return fail(CNV_ERR_INVALID, "autotune needs two grids that do not share memory");
    if (tmp->desc.ncomp < u->desc.ncomp)
        return fail(CNV_ERR_SIZE, "tmp has fewer components than u");
    if (iters < 1)
        return fail(CNV_ERR_INVALID, "iters must be >= 1");
    if (!x2_applicable(u))
        return fail(CNV_ERR_UNSUPPORTED, "the two-step kernel needs periodic grids of 16^3 or 32^3 blocks");
    return autotune_x2(u, tmp, iters, wg, zchunk, results, max_results, nresults);
    CNV_GUARD_END
}

int cnv_tune_set_diffusion_run(int dtype, int edge, int wg, int zchunk)
{
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    return x2_set_tune(dtype, edge, wg, zchunk);
}

int cnv_tune_get_diffusion_run(int dtype, int edge, int *wg, int *zchunk)
{
    if (!wg || !zchunk)
        return fail(CNV_ERR_INVALID, "output pointer is NULL");
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    if (edge != 16 && edge != 32)
        return fail(CNV_ERR_INVALID, "two-step kernel: block edge must be 16 or 32");
    return x2_get_tune(dtype, edge, wg, zchunk);
}

int cnv_tune_set_stencil7(int dtype, int edge, int wg, int zchunk)
{
    if (dtype != CNV_F32 && dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
    if (wg != 64 && wg != 128 && wg != 256 && wg != 512)
        return fail(CNV_ERR_INVALID, "workgroup size must be 64, 128, 256 or 512");
    if (zchunk < 1 || edge < 1)
        return fail(CNV_ERR_INVALID, "edge and zchunk must be >= 1");
    s7_set_tune(dtype, edge, wg, zchunk);
    return CNV_OK;
}

int cnv_tune_get_stencil7(int dtype, int edge, int *wg, int *zchunk)
{
    if (!wg || !zchunk)
        return fail(CNV_ERR_INVALID, "output pointer is NULL");
    return get_tune_stencil7(dtype, edge, wg, zchunk);
}

// ---------------------------------------------------------------- graphs

struct cnv_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int cnv_graph_begin(cnv_stream s)
{
    if (!s)
        return fail(CNV_ERR_INVALID, "graph capture needs a non-default stream");
    // Relaxed: with the thread-local mode, a capture begun after grids were
    // destroyed earlier in the process failed at its first kernel launch
    // ("operation failed due to a previous error during capture") while the
    // same calls captured in a fresh process did not; the entry points make
    // no allocating or synchronising call while capturing once their grids
    // are prepared, so the relaxed mode records the same work.
    CNV_HIP(hipStreamBeginCapture((hipStream_t)s, hipStreamCaptureModeRelaxed));
    return CNV_OK;
}

int cnv_graph_end(cnv_stream s, cnv_graph *out)
{
    CNV_GUARD_BEGIN
    if (!s || !out)
        return fail(CNV_ERR_INVALID, "stream or output is NULL");
    *out = nullptr;
    hipGraph_t gr = nullptr;
    CNV_HIP(hipStreamEndCapture((hipStream_t)s, &gr));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(gr);
        return hip_fail(e, "hipGraphInstantiate");
    }
    cnv_graph_s *g = new cnv_graph_s();
    g->graph = gr;
    g->exec = ex;
    *out = g;
    return CNV_OK;
    CNV_GUARD_END
}

int cnv_graph_launch(cnv_graph g, cnv_stream s)
{
    if (!g)
        return fail(CNV_ERR_INVALID, "graph is NULL");
    CNV_HIP(hipGraphLaunch(g->exec, (hipStream_t)s));
    return CNV_OK;
}

int cnv_graph_destroy(cnv_graph g)
{
    if (!g)
        return CNV_OK;
    if (g->exec)
        (void)hipGraphExecDestroy(g->exec);
    if (g->graph)
        (void)hipGraphDestroy(g->graph);
    delete g;
    return CNV_OK;
}

} // extern "C"
