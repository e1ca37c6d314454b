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
// stream-ordered variants with a status (cnv_field_op2 / op1 / rcp).  Every
// element is one IEEE operation (the library is built without contraction
// and with correctly rounded division), so results equal the host loops bit
// for bit.
#include "cnv_internal.h"

namespace cnv
{

enum { FOP_ADD = 0, FOP_SUB = 1, FOP_MUL = 2, FOP_DIV = 3, FOP_RCP = 4 };

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
    else
        return b * ((T)1 / a); // RCP: s * (1 / a)
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

} // namespace cnv

using namespace cnv;

extern "C" {

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
