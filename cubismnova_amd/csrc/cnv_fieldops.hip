// cnv_fieldops.hip -- element-wise block field arithmetic on device memory.
//
// CubismNova's only C-ABI seam is the optional set of extern "C" hooks of
// include/Cubism/Block/FieldOperator.h:149-227 (CUBISM_OPTIMIZED_FIELD_OP):
//   field{Add,Sub,Mul,Div}2_{double,float}(const T*, const T*, T*, size_t)
//   field{Add,Sub,Mul,Div}1_{double,float}(const T*, T, T*, size_t)
//   fieldRcp_{double,float}(const T*, T, T*, size_t)
// with the loops of FieldOperator.h:11-30: dst[i] = a[i] OP b[i],
// dst[i] = a[i] OP s, and dst[i] = s * (1 / a[i]) (two roundings).  This file
// provides those symbols for pointers the GPU can address (the data of a
// cnv_grid, cnv_malloc memory), blocking like the reference's loops, and
// stream-ordered variants with a status (cnv_field_op2 / op1 / rcp).  The
// Field and FieldContainer operators of whole grids (Field.h:444-556,
// 1039-1154: each block over its own range_.size() elements, each
// component) run as one launch over (slab, block) rows (cnv_grid_field_op).  Every
// element is one IEEE operation (the library is built without contraction
// and with correctly rounded division), so results equal the host loops bit
// for bit.
#include "cnv_device.h"

#include <algorithm>

namespace cnv
{

enum { FOP_ADD = 0, FOP_SUB = 1, FOP_MUL = 2, FOP_DIV = 3, FOP_RCP = 4, FOP_NEG = 5 };

template <typename T, int OP>
__device__ inline T fop(T a, T b)
{
    if constexpr (OP == FOP_ADD)
        return a + b;
    else if constexpr (OP == FOP_SUB)
        return a - b;
    else if constexpr (OP == FOP_MUL)
        return a * b;
    else if constexpr (OP == FOP_DIV)
        return a / b;
    else if constexpr (OP == FOP_RCP)
        return b * ((T)1 / a); // s * (1 / a)
    else
        return -a;
}

// dst[i] = a[i] OP (b ? b[i] : s); four elements per thread per pass
template <typename T, int OP>
__global__ __launch_bounds__(256) void k_fieldop(const T *__restrict__ a, const T *__restrict__ b, T s,
                                                 T *__restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t base = (size_t)blockIdx.x * 256 * 4; base < n; base += stride * 4) {
        T va[4], vb[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = base + (size_t)j * 256 + threadIdx.x;
            va[j] = i < n ? a[i] : (T)1;
            vb[j] = b ? (i < n ? b[i] : (T)1) : s;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = base + (size_t)j * 256 + threadIdx.x;
            if (i < n)
                dst[i] = fop<T, OP>(va[j], vb[j]);
        }
    }
}

template <typename T>
static int fieldop_t(int op, const T *a, const T *b, T s, T *dst, size_t n, hipStream_t st)
{
    if (n == 0)
        return CNV_OK;
    int dev = 0, ncu = 0;
    CNV_HIP(hipGetDevice(&dev));
    CNV_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t per = (size_t)256 * 4;
    size_t nwg = (n + per - 1) / per;
    const size_t cap = (size_t)ncu * 16;
    if (nwg > cap)
        nwg = cap;
    const dim3 g((unsigned)nwg), blk(256);
    switch (op) {
    case FOP_ADD:
        hipLaunchKernelGGL((k_fieldop<T, FOP_ADD>), g, blk, 0, st, a, b, s, dst, n);
        break;
    case FOP_SUB:
        hipLaunchKernelGGL((k_fieldop<T, FOP_SUB>), g, blk, 0, st, a, b, s, dst, n);
        break;
    case FOP_MUL:
        hipLaunchKernelGGL((k_fieldop<T, FOP_MUL>), g, blk, 0, st, a, b, s, dst, n);
        break;
    case FOP_DIV:
        hipLaunchKernelGGL((k_fieldop<T, FOP_DIV>), g, blk, 0, st, a, b, s, dst, n);
        break;
    case FOP_RCP:
        hipLaunchKernelGGL((k_fieldop<T, FOP_RCP>), g, blk, 0, st, a, nullptr, s, dst, n);
        break;
    default:
        return fail(CNV_ERR_INVALID, "unknown field operation");
    }
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

static int fieldop(int op, int dtype, const void *a, const void *b, double s, void *dst, size_t n, hipStream_t st)
{
    if (n && (!a || !dst))
        return fail(CNV_ERR_INVALID, "field pointer is NULL");
    if (dtype == CNV_F64)
        return fieldop_t<double>(op, (const double *)a, (const double *)b, s, (double *)dst, n, st);
    if (dtype == CNV_F32)
        return fieldop_t<float>(op, (const float *)a, (const float *)b, (float)s, (float *)dst, n, st);
    return fail(CNV_ERR_INVALID, "dtype must be CNV_F32 or CNV_F64");
}

// the reference's hooks return nothing and complete before returning
static void fieldop_blocking(int op, int dtype, const void *a, const void *b, double s, void *dst, size_t n)
{
    if (fieldop(op, dtype, a, b, s, dst, n, nullptr) == CNV_OK) {
        const hipError_t e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess)
            hip_fail(e, "field operation");
    }
}

// ------------------------------------------------------------ whole grids

template <typename T>
struct GridOpArgs {
    T *dst;
    const T *a, *b; // b null: the scalar operand
    T s;
    long long cstride, bstride; // elements between slabs, between blocks
    int cd, ca, cb;             // first slab of each operand
    int nb[3], n[3];
    int entity, ncomp;
    long long nbt, rows; // rows = slabs * blocks
};

// one workgroup row per (slab, block); 4 elements per thread per pass over
// the block's own extent, 16- or 32-byte vectors where the block has them
template <typename T, int OP>
__global__ __launch_bounds__(256) void k_grid_fieldop(const GridOpArgs<T> g)
{
    using V = VecT<T, 4>;
    for (long long r = blockIdx.y; r < g.rows; r += gridDim.y) {
        const int sl = (int)(r / g.nbt);
        const long long blk = r - (long long)sl * g.nbt;
        const int B[3] = {(int)(blk % g.nb[0]), (int)((blk / g.nb[0]) % g.nb[1]),
                          (int)(blk / ((long long)g.nb[0] * g.nb[1]))};
        // plus dimensions of this slab (BlockFieldAssembler.h:123-131)
        const int pm = g.entity == CNV_NODE ? 7 : (g.entity == CNV_FACE ? 1 << ((g.cd + sl) / g.ncomp) : 0);
        long long ne = 1;
        for (int d = 0; d < 3; ++d)
            ne *= g.n[d] + (((pm >> d) & 1) && B[d] == g.nb[d] - 1 ? 1 : 0);
        T *dst = g.dst + (g.cd + sl) * g.cstride + blk * g.bstride;
        const T *a = g.a + (g.ca + sl) * g.cstride + blk * g.bstride;
        const T *b = g.b ? g.b + (g.cb + sl) * g.cstride + blk * g.bstride : nullptr;
        for (long long o = ((long long)blockIdx.x * 256 + threadIdx.x) * 4; o < ne;
             o += (long long)gridDim.x * 256 * 4) {
            if (o + 4 <= ne) {
                const V va = load_vec<T, 4>(a + o);
                V vb;
                if (b)
                    vb = load_vec<T, 4>(b + o);
                V r4;
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    r4.v[j] = fop<T, OP>(va.v[j], b ? vb.v[j] : g.s);
                store_vec<T, 4>(dst + o, r4);
            } else {
                for (long long i = o; i < ne; ++i)
                    dst[i] = fop<T, OP>(a[i], b ? b[i] : g.s);
            }
        }
    }
}

template <typename T>
static int grid_fieldop_t(int op, cnv_grid_s *dst, int cd, const cnv_grid_s *a, int ca, const cnv_grid_s *b, int cb,
                          double s, int nsl, hipStream_t st)
{
    GridOpArgs<T> g;
    g.dst = reinterpret_cast<T *>(dst->data);
    g.a = reinterpret_cast<const T *>(a->data);
    g.b = b ? reinterpret_cast<const T *>(b->data) : nullptr;
    g.s = (T)s;
    g.cstride = (long long)(a->component_bytes / a->elem);
    g.bstride = (long long)(a->block_bytes / a->elem);
    g.cd = cd;
    g.ca = ca;
    g.cb = cb;
    for (int d = 0; d < 3; ++d) {
        g.nb[d] = a->desc.nblocks[d];
        g.n[d] = a->desc.block_cells[d];
    }
    g.entity = a->desc.entity;
    g.ncomp = a->desc.ncomp;
    g.nbt = a->nblocks_total;
    g.rows = g.nbt * nsl;
    long long maxe = 1;
    for (int d = 0; d < 3; ++d)
        maxe *= g.n[d] + (g.entity == CNV_CELL ? 0 : 1);
    const long long gx = std::min<long long>((maxe + 1023) / 1024, 64);
    const long long gy = std::min<long long>(g.rows, 65535);
    const dim3 grid((unsigned)gx, (unsigned)gy), blk(256);
    switch (op) {
    case FOP_ADD:
        hipLaunchKernelGGL((k_grid_fieldop<T, FOP_ADD>), grid, blk, 0, st, g);
        break;
    case FOP_SUB:
        hipLaunchKernelGGL((k_grid_fieldop<T, FOP_SUB>), grid, blk, 0, st, g);
        break;
    case FOP_MUL:
        hipLaunchKernelGGL((k_grid_fieldop<T, FOP_MUL>), grid, blk, 0, st, g);
        break;
    case FOP_DIV:
        hipLaunchKernelGGL((k_grid_fieldop<T, FOP_DIV>), grid, blk, 0, st, g);
        break;
    case FOP_RCP:
        hipLaunchKernelGGL((k_grid_fieldop<T, FOP_RCP>), grid, blk, 0, st, g);
        break;
    default:
        hipLaunchKernelGGL((k_grid_fieldop<T, FOP_NEG>), grid, blk, 0, st, g);
        break;
    }
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

static int same_layout(const cnv_grid_s *x, const cnv_grid_s *y)
{
    for (int d = 0; d < 3; ++d)
        if (x->desc.nblocks[d] != y->desc.nblocks[d] || x->desc.block_cells[d] != y->desc.block_cells[d])
            return fail(CNV_ERR_SIZE, "grids differ in block count or block size");
    if (x->desc.dtype != y->desc.dtype || x->desc.entity != y->desc.entity || x->nslabs != y->nslabs)
        return fail(CNV_ERR_SIZE, "grids differ in dtype, entity or slab count");
    return CNV_OK;
}

// Streaming copy of n16 16-byte words: non-temporal loads and stores, four
// words in flight per thread.  The practical read + write ceiling of HBM
// that the operator kernels are compared with.
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_nt(const u4v *__restrict__ src, u4v *__restrict__ dst, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * 1024;
    for (size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x; base < n16; base += stride) {
        u4v v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = base + (size_t)j * 256;
            if (i < n16)
                v[j] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = base + (size_t)j * 256;
            if (i < n16)
                __builtin_nontemporal_store(v[j], dst + i);
        }
    }
}

// One read, `fanout` writes: every 16-byte word of src goes to the same
// offset of `fanout` destination slabs, dst_stride16 words apart, with
// non-temporal loads and stores (the memory pattern of the WENO5 operator,
// one input component to six state components, without its arithmetic).
__global__ __launch_bounds__(256) void k_fanout_nt(const u4v *__restrict__ src, u4v *__restrict__ dst, size_t n16,
                                                   int fanout, size_t dst_stride16)
{
    const size_t stride = (size_t)gridDim.x * 1024;
    for (size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x; base < n16; base += stride) {
        u4v v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = base + (size_t)j * 256;
            if (i < n16)
                v[j] = __builtin_nontemporal_load(src + i);
        }
        for (int k = 0; k < fanout; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const size_t i = base + (size_t)j * 256;
                if (i < n16)
                    __builtin_nontemporal_store(v[j], dst + (size_t)k * dst_stride16 + i);
            }
    }
}

} // namespace cnv

using namespace cnv;

extern "C" {

int cnv_memcpy_fanout(void *dst, size_t dst_stride, const void *src, size_t bytes, int fanout, cnv_stream s)
{
    if (bytes == 0 || fanout == 0)
        return CNV_OK;
    if (!dst || !src || fanout < 0)
        return fail(CNV_ERR_INVALID, "fanout copy: NULL pointer or negative fanout");
    if (((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src) | bytes | dst_stride) & 15) != 0)
        return fail(CNV_ERR_INVALID, "fanout copy: pointers, size and stride must be multiples of 16 bytes");
    if (fanout > 1 && dst_stride < bytes)
        return fail(CNV_ERR_INVALID, "fanout copy: destination slabs overlap");
    const size_t n16 = bytes / 16;
    const size_t want = (n16 + 1023) / 1024;
    const unsigned nwg = (unsigned)(want < 8192 ? want : 8192);
    hipLaunchKernelGGL(k_fanout_nt, dim3(nwg), dim3(256), 0, (hipStream_t)s, static_cast<const u4v *>(src),
                       static_cast<u4v *>(dst), n16, fanout, dst_stride / 16);
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

int cnv_memcpy_stream(void *dst, const void *src, size_t bytes, cnv_stream s)
{
    if (bytes == 0)
        return CNV_OK;
    if (!dst || !src)
        return fail(CNV_ERR_INVALID, "copy: NULL pointer");
    if (((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src) | bytes) & 15) != 0)
        return fail(CNV_ERR_INVALID, "copy: pointers and size must be multiples of 16 bytes");
    const size_t n16 = bytes / 16;
    const size_t want = (n16 + 1023) / 1024;
    const unsigned nwg = (unsigned)(want < 8192 ? want : 8192);
    hipLaunchKernelGGL(k_copy_nt, dim3(nwg), dim3(256), 0, (hipStream_t)s, static_cast<const u4v *>(src),
                       static_cast<u4v *>(dst), n16);
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

int cnv_field_op2(int op, int dtype, const void *a, const void *b, void *dst, size_t n, cnv_stream s)
{
    if (op < FOP_ADD || op > FOP_DIV)
        return fail(CNV_ERR_INVALID, "op must be CNV_FIELD_ADD, _SUB, _MUL or _DIV");
    if (n && !b)
        return fail(CNV_ERR_INVALID, "field pointer is NULL");
    return fieldop(op, dtype, a, b, 0.0, dst, n, (hipStream_t)s);
}

int cnv_field_op1(int op, int dtype, const void *a, double scalar, void *dst, size_t n, cnv_stream s)
{
    if (op < FOP_ADD || op > FOP_DIV)
        return fail(CNV_ERR_INVALID, "op must be CNV_FIELD_ADD, _SUB, _MUL or _DIV");
    return fieldop(op, dtype, a, nullptr, scalar, dst, n, (hipStream_t)s);
}

int cnv_field_rcp(int dtype, const void *a, double scalar, void *dst, size_t n, cnv_stream s)
{
    return fieldop(FOP_RCP, dtype, a, nullptr, scalar, dst, n, (hipStream_t)s);
}

int cnv_grid_field_op(int op, cnv_grid dst, int cd, cnv_grid a, int ca, cnv_grid b, int cb, double scalar,
                      cnv_stream s)
{
    if (!dst || !a)
        return fail(CNV_ERR_INVALID, "grid handle is NULL");
    if (op < FOP_ADD || op > FOP_NEG)
        return fail(CNV_ERR_INVALID, "unknown field operation");
    if (b && op >= FOP_RCP)
        return fail(CNV_ERR_INVALID, "reciprocal and negation take no second grid");
    int rc = same_layout(dst, a);
    if (rc == CNV_OK && b)
        rc = same_layout(a, b);
    if (rc != CNV_OK)
        return rc;
    int nsl = 1;
    if (cd == -1 && ca == -1 && (!b || cb == -1)) {
        cd = ca = cb = 0;
        nsl = a->nslabs;
    } else {
        if (cd < 0 || cd >= dst->nslabs || ca < 0 || ca >= a->nslabs || (b && (cb < 0 || cb >= b->nslabs)))
            return fail(CNV_ERR_INVALID, "component index out of range");
        if (a->desc.entity == CNV_FACE &&
            (cd / a->desc.ncomp != ca / a->desc.ncomp || (b && cb / a->desc.ncomp != ca / a->desc.ncomp)))
            return fail(CNV_ERR_SIZE, "face slabs of different directions differ in shape");
        if (!b)
            cb = 0;
    }
    if (a->desc.dtype == CNV_F64)
        return grid_fieldop_t<double>(op, dst, cd, a, ca, b, cb, scalar, nsl, (hipStream_t)s);
    return grid_fieldop_t<float>(op, dst, cd, a, ca, b, cb, scalar, nsl, (hipStream_t)s);
}

#define CNV_FIELD2(NAME, OPC, TYPE, DT)                                                                 \
    void NAME(const TYPE *src0, const TYPE *src1, TYPE *dst, const size_t n)                            \
    {                                                                                                   \
        fieldop_blocking(OPC, DT, src0, src1, 0.0, dst, n);                                             \
    }
#define CNV_FIELD1(NAME, OPC, TYPE, DT)                                                                 \
    void NAME(const TYPE *src0, const TYPE src1, TYPE *dst, const size_t n)                             \
    {                                                                                                   \
        fieldop_blocking(OPC, DT, src0, nullptr, (double)src1, dst, n);                                 \
    }

CNV_FIELD2(fieldAdd2_double, FOP_ADD, double, CNV_F64)
CNV_FIELD2(fieldSub2_double, FOP_SUB, double, CNV_F64)
CNV_FIELD2(fieldMul2_double, FOP_MUL, double, CNV_F64)
CNV_FIELD2(fieldDiv2_double, FOP_DIV, double, CNV_F64)
CNV_FIELD2(fieldAdd2_float, FOP_ADD, float, CNV_F32)
CNV_FIELD2(fieldSub2_float, FOP_SUB, float, CNV_F32)
CNV_FIELD2(fieldMul2_float, FOP_MUL, float, CNV_F32)
CNV_FIELD2(fieldDiv2_float, FOP_DIV, float, CNV_F32)
CNV_FIELD1(fieldAdd1_double, FOP_ADD, double, CNV_F64)
CNV_FIELD1(fieldSub1_double, FOP_SUB, double, CNV_F64)
CNV_FIELD1(fieldMul1_double, FOP_MUL, double, CNV_F64)
CNV_FIELD1(fieldDiv1_double, FOP_DIV, double, CNV_F64)
CNV_FIELD1(fieldAdd1_float, FOP_ADD, float, CNV_F32)
CNV_FIELD1(fieldSub1_float, FOP_SUB, float, CNV_F32)
CNV_FIELD1(fieldMul1_float, FOP_MUL, float, CNV_F32)
CNV_FIELD1(fieldDiv1_float, FOP_DIV, float, CNV_F32)
CNV_FIELD1(fieldRcp_double, FOP_RCP, double, CNV_F64)
CNV_FIELD1(fieldRcp_float, FOP_RCP, float, CNV_F32)

#undef CNV_FIELD2
#undef CNV_FIELD1

} // extern "C"
