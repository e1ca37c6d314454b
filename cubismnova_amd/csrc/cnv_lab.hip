// cnv_lab.hip -- FieldLab assembly on the GPU, one thread per lab cell.
//
// Instead of the reference's sequence "inner copy, 26 neighbour copies, then
// each BC in list order" (Block/FieldLab.h:180-224), every lab cell is
// evaluated independently: walking the block's BC list backwards, the last BC
// whose slab holds the cell decides it; Dirichlet ends the walk, Absorbing
// and Symmetry move the cell to their source (mirror) and continue with the
// BCs before them.  A cell no BC touches is the block's own value, a
// periodic neighbour's value, or unwritten (edges/corners of non-tensorial
// labs; sides skipped by a non-periodic BC).  That is the same result as the
// sequential passes because every BC pass reads only cells it does not write
// (BC/Absorbing.h:94-102, BC/Symmetry.h:107-113).
//
// Node and face slabs: a block at the upper end of a "plus" dimension holds
// one more element there, so each block's extent (and its lab range) is its
// own, and ghosts below a block count back from the neighbour's own end
// (FieldLabLoader.h:217-219, 236-241).  The all-blocks launch lays every lab
// out in the largest block's geometry, the block's lab in the low corner.
#include "cnv_internal.h"

namespace cnv
{

constexpr int kMaxBC = 16;

template <typename T>
struct LabArgs {
    const T *base; // component slab
    Geom G;
    int sb[3], se[3];
    int tens;
    int apply_bc;
    int plus;       // bit d: upper-end blocks one element larger along d
    int L[3];       // lab slot extent (the largest block's lab)
    long long lcells; // L0*L1*L2
    T *out;
    // single-block mode: block index and the list by value
    int single;
    int block[3];
    int nlist;
    DevBC list[kMaxBC];
    // all-blocks mode: per-block tables (null when no block carries a BC)
    int has_tab;
    const DevBCHead *heads;
    const DevBC *entries;
};

// extent of block B along d (node / face slabs: one more at the upper end)
template <typename T>
__device__ inline int blk_extent(const LabArgs<T> &a, const int *B, int d)
{
    return a.G.n[d] + (((a.plus >> d) & 1) && B[d] == a.G.nb[d] - 1 ? 1 : 0);
}

template <typename T>
__device__ bool lab_value(const LabArgs<T> &a, const int B[3], const DevBC *list, int n,
                          long long q0, long long q1, long long q2, T &out)
{
    const int N[3] = {blk_extent(a, B, 0), blk_extent(a, B, 1), blk_extent(a, B, 2)};
    long long q[3] = {q0, q1, q2};
    bool periodic[3] = {true, true, true};
    int skip[3] = {1, 1, 1};
    for (int k = 0; k < n; ++k) {
        periodic[list[k].dir] = (list[k].kind == CNV_BC_BASE);
        skip[list[k].dir] = (list[k].side == 0) ? -1 : 1;
    }
    T signs[kMaxBC];
    int depth = 0;
    int k = a.apply_bc ? n : 0;
    bool have_const = false;
    T cval = T(0);
    while (k > 0) {
        int j = k - 1;
        for (; j >= 0; --j) {
            const DevBC &b = list[j];
            if (b.kind == CNV_BC_BASE)
                continue;
            const int d = b.dir;
            if (a.sb[d] == 0 || a.se[d] == 1)
                continue; // BC::Base::isValidStencil_ (BC/Base.h:73-78)
            bool in = (b.side == 0) ? (q[d] >= a.sb[d] && q[d] < 0)
                                    : (q[d] >= N[d] && q[d] < N[d] + a.se[d] - 1);
            for (int e = 0; e < 3 && in; ++e) {
                if (e == d)
                    continue;
                if (a.tens)
                    in = q[e] >= a.sb[e] && q[e] < N[e] + a.se[e] - 1;
                else
                    in = q[e] >= 0 && q[e] < N[e];
            }
            if (in)
                break;
        }
        if (j < 0)
            break;
        const DevBC &b = list[j];
        const int d = b.dir;
        if (b.kind == CNV_BC_DIRICHLET) {
            have_const = true;
            cval = (T)b.value;
            break;
        } else if (b.kind == CNV_BC_ABSORBING) {
            q[d] = (b.side == 0) ? 0 : N[d] - 1;
        } else { // symmetry
            signs[depth++] = (T)b.value;
            q[d] = (b.side == 0) ? (-1 - q[d]) : (2LL * N[d] - 1 - q[d]);
        }
        k = j;
    }
    T v;
    if (have_const) {
        v = cval;
    } else {
        int bi[3];
        int l1 = 0;
        for (int d = 0; d < 3; ++d) {
            bi[d] = q[d] < 0 ? -1 : (q[d] >= N[d] ? 1 : 0);
            l1 += bi[d] != 0;
        }
        int nb[3];
        for (int d = 0; d < 3; ++d) {
            if (!periodic[d] && bi[d] == skip[d])
                return false; // neighbour skipped, never written
            nb[d] = (B[d] + bi[d] + a.G.nb[d]) % a.G.nb[d];
        }
        if (!a.tens && l1 > 1)
            return false; // edge/corner of a non-tensorial lab
        const long long bf = nb[0] + (long long)a.G.nb[0] * (nb[1] + (long long)a.G.nb[1] * nb[2]);
        // source in the neighbour's own extent NE; below the block it counts
        // back from the neighbour's end (shift NE - N)
        long long s[3];
        int NE[3];
        for (int d = 0; d < 3; ++d) {
            NE[d] = blk_extent(a, nb, d);
            s[d] = q[d] - (long long)bi[d] * N[d] + (bi[d] < 0 ? NE[d] - N[d] : 0);
        }
        v = a.base[bf * a.G.bstride + s[0] + (long long)NE[0] * (s[1] + (long long)NE[1] * s[2])];
    }
    for (int i = depth - 1; i >= 0; --i)
        v = signs[i] * v;
    out = v;
    return true;
}

template <typename T>
__global__ __launch_bounds__(256) void k_lab(const LabArgs<T> a)
{
    __shared__ DevBC s_list[kMaxBC];
    __shared__ int s_n;
    int B[3];
    long long slot;
    if (a.single) {
        B[0] = a.block[0];
        B[1] = a.block[1];
        B[2] = a.block[2];
        slot = 0;
        if (threadIdx.x < a.nlist)
            s_list[threadIdx.x] = a.list[threadIdx.x];
        if (threadIdx.x == 0)
            s_n = a.nlist;
    } else {
        slot = blockIdx.y;
        B[0] = (int)(slot % a.G.nb[0]);
        B[1] = (int)((slot / a.G.nb[0]) % a.G.nb[1]);
        B[2] = (int)(slot / ((long long)a.G.nb[0] * a.G.nb[1]));
        int n = 0;
        if (a.has_tab) {
            const DevBCHead h = a.heads[slot];
            n = h.count < kMaxBC ? h.count : kMaxBC;
            if ((int)threadIdx.x < n)
                s_list[threadIdx.x] = a.entries[h.offset + threadIdx.x];
        }
        if (threadIdx.x == 0)
            s_n = n;
    }
    __syncthreads();
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.lcells)
        return;
    const long long x = c % a.L[0];
    const long long y = (c / a.L[0]) % a.L[1];
    const long long z = c / ((long long)a.L[0] * a.L[1]);
    if (a.plus) { // smaller blocks fill the low corner of the slot
        if (x >= blk_extent(a, B, 0) + a.se[0] - 1 - a.sb[0] || y >= blk_extent(a, B, 1) + a.se[1] - 1 - a.sb[1] ||
            z >= blk_extent(a, B, 2) + a.se[2] - 1 - a.sb[2])
            return;
    }
    T v;
    if (lab_value(a, B, s_list, s_n, x + a.sb[0], y + a.sb[1], z + a.sb[2], v))
        a.out[slot * a.lcells + c] = v;
}

template <typename T>
static int launch_lab_t(const cnv_grid_s *g, int comp, const cnv_stencil &s, bool apply_bc,
                        bool single, const DevBC *list, int nlist, const int *block, void *out,
                        hipStream_t st)
{
    LabArgs<T> a;
    a.G = geom_of(g);
    a.base = reinterpret_cast<const T *>((const char *)g->data + (size_t)comp * g->component_bytes);
    a.plus = plus_mask(g, comp);
    for (int d = 0; d < 3; ++d) {
        a.sb[d] = s.begin[d];
        a.se[d] = s.end[d];
        // one block: its own lab extent; all blocks: the largest block's
        const int wide = ((a.plus >> d) & 1) && (!single || block[d] == a.G.nb[d] - 1);
        a.L[d] = a.G.n[d] + wide + s.end[d] - 1 - s.begin[d];
    }
    a.tens = s.tensorial ? 1 : 0;
    a.apply_bc = apply_bc ? 1 : 0;
    a.lcells = (long long)a.L[0] * a.L[1] * a.L[2];
    a.out = static_cast<T *>(out);
    a.single = single ? 1 : 0;
    a.nlist = 0;
    a.has_tab = 0;
    a.heads = nullptr;
    a.entries = nullptr;
    long long nslots = 1;
    if (single) {
        for (int d = 0; d < 3; ++d)
            a.block[d] = block[d];
        a.nlist = nlist;
        for (int k = 0; k < nlist; ++k)
            a.list[k] = list[k];
    } else {
        const BCTable &t = g->tables[comp];
        a.has_tab = (t.any && t.heads && t.entries) ? 1 : 0;
        a.heads = t.heads;
        a.entries = t.entries;
        nslots = g->nblocks_total;
        for (int d = 0; d < 3; ++d)
            a.block[d] = 0;
        // per-block lists are read into LDS; longer lists are not supported
        for (const auto &l : g->bcs[comp])
            if (l.size() > (size_t)kMaxBC)
                return fail(CNV_ERR_UNSUPPORTED, "more than 16 BCs on one block");
    }
    if (nslots > 65535)
        return fail(CNV_ERR_UNSUPPORTED, "too many blocks for one lab launch");
    const long long gx = (a.lcells + 255) / 256;
    if (gx > 0x7fffffffLL)
        return fail(CNV_ERR_UNSUPPORTED, "lab too large");
    hipLaunchKernelGGL(k_lab<T>, dim3((unsigned)gx, (unsigned)nslots), dim3(256), 0, st, a);
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

int launch_lab(const cnv_grid_s *g, int comp, const cnv_stencil &s, bool apply_bc, bool single,
               const DevBC *list, int nlist, const int *block, void *out, hipStream_t st)
{
    if (g->desc.dtype == CNV_F64)
        return launch_lab_t<double>(g, comp, s, apply_bc, single, list, nlist, block, out, st);
    return launch_lab_t<float>(g, comp, s, apply_bc, single, list, nlist, block, out, st);
}

// ---------------------------------------------------------------- fill

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *dst, T v, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

int launch_fill(void *dst, int dtype, double value, size_t count, hipStream_t st)
{
    if (count == 0)
        return CNV_OK;
    size_t blocks = (count + 255) / 256;
    if (blocks > 65536)
        blocks = 65536;
    if (dtype == CNV_F64)
        hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)blocks), dim3(256), 0, st, (double *)dst,
                           value, count);
    else
        hipLaunchKernelGGL(k_fill<float>, dim3((unsigned)blocks), dim3(256), 0, st, (float *)dst,
                           (float)value, count);
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

} // namespace cnv
