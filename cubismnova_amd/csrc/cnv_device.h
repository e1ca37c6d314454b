// cnv_device.h -- device helpers shared by the z-marching stencil kernels.
#pragma once

#include "cnv_internal.h"

namespace cnv
{

template <typename T, int N>
struct alignas(16) VecT {
    T v[N];
};

// p - 2c with one rounding: 2c is exact (a power-of-two scaling, finite for
// |c| below half the largest value), so the fused form rounds exactly like
// the subtraction p - (2c) of the 7-point operator
template <typename T>
__device__ inline T fms2(T c, T p)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_fmaf(-2.f, c, p);
    else
        return __builtin_fma(-2.0, c, p);
}

// Wave priority from 3 at the start of a workgroup's z loop down to 0 at
// its end (done = planes finished, total = planes of the tile): on a compute
// unit the waves that lag behind are issued first, so the workgroups that
// share it finish together instead of the oldest first.
__device__ inline void prio_ramp(int done, int total)
{
    const int lvl = (total - 1 - done) * 4 / total;
    if (lvl >= 3)
        __builtin_amdgcn_s_setprio(3);
    else if (lvl == 2)
        __builtin_amdgcn_s_setprio(2);
    else if (lvl == 1)
        __builtin_amdgcn_s_setprio(1);
    else
        __builtin_amdgcn_s_setprio(0);
}

// Plain vector load / store of exactly N elements (N * sizeof(T) bytes,
// naturally aligned), whatever the padding of VecT.
template <typename T, int N>
__device__ inline VecT<T, N> load_vec(const T *p)
{
    typedef T vt __attribute__((ext_vector_type(N)));
    const vt x = *reinterpret_cast<const vt *>(p);
    VecT<T, N> r;
#pragma unroll
    for (int i = 0; i < N; ++i)
        r.v[i] = x[i];
    return r;
}

template <typename T, int N>
__device__ inline void store_vec(T *p, const VecT<T, N> &v)
{
    typedef T vt __attribute__((ext_vector_type(N)));
    vt x;
#pragma unroll
    for (int i = 0; i < N; ++i)
        x[i] = v.v[i];
    *reinterpret_cast<vt *>(p) = x;
}

// Raw buffer access: a buffer resource over [base, base + bytes) in scalar
// registers, addressed by a 32-bit per-lane byte offset plus a 32-bit uniform
// one, so no lane computes a 64-bit address.  Accesses past `bytes` read 0
// and write nothing.  The resource's fourth word is the gfx9 data format
// for raw (untyped) access.
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

template <typename T, int N>
__device__ inline VecT<T, N> buf_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    static_assert(sizeof(T) * N == 16 || sizeof(T) * N == 8, "8- or 16-byte vectors");
    typedef T vt __attribute__((ext_vector_type(N)));
    vt x;
    if constexpr (sizeof(T) * N == 16)
        x = __builtin_bit_cast(vt, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
    else
        x = __builtin_bit_cast(vt, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
    VecT<T, N> o;
#pragma unroll
    for (int i = 0; i < N; ++i)
        o.v[i] = x[i];
    return o;
}

template <typename T>
__device__ inline T buf_load1(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

template <typename T, int N>
__device__ inline void buf_store(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, const VecT<T, N> &v)
{
    static_assert(sizeof(T) * N == 16 || sizeof(T) * N == 8, "8- or 16-byte vectors");
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    typedef T vt __attribute__((ext_vector_type(N)));
    vt x;
#pragma unroll
    for (int i = 0; i < N; ++i)
        x[i] = v.v[i];
    if constexpr (sizeof(T) * N == 16)
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, x), r, (int)voff, (int)soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, x), r, (int)voff, (int)soff, 0);
}

// Streaming (non-temporal) vector load of exactly N elements, for data read
// once per launch.
template <typename T, int N>
__device__ inline VecT<T, N> load_nt(const T *p)
{
    typedef T vt __attribute__((ext_vector_type(N)));
    const vt x = __builtin_nontemporal_load(reinterpret_cast<const vt *>(p));
    VecT<T, N> r;
#pragma unroll
    for (int i = 0; i < N; ++i)
        r.v[i] = x[i];
    return r;
}

// Streaming (non-temporal) 16-byte store of a VecT.
template <typename T, int N>
__device__ inline void store_nt(T *p, const VecT<T, N> &v)
{
    typedef T vt __attribute__((ext_vector_type(N)));
    vt x;
#pragma unroll
    for (int i = 0; i < N; ++i)
        x[i] = v.v[i];
    __builtin_nontemporal_store(x, reinterpret_cast<vt *>(p));
}

// Per-face source description for one block (uniform in a workgroup).
template <typename T>
struct FaceSrc {
    const T *p;  // base such that ghost(i) = p[i] (copy) or sign*p[i]
    int mode;    // 0 copy, 1 scaled, 2 constant
    T s;         // sign or constant
};

template <typename T>
__device__ inline T face_val(const FaceSrc<T> &f, long long i)
{
    if (f.mode == 2)
        return f.s;
    const T v = f.p[i];
    return f.mode == 1 ? f.s * v : v;
}

// Resolve the six one-layer faces of block `blk` (ghost layer 1 of the lab,
// the only one the 7-point operators read; the same for any ghost width).
template <typename T>
__device__ inline void block_faces(const Geom &G, const DevBCHead *heads, const DevBC *entries, int has_bc,
                                   long long blk, const T *cbase, FaceSrc<T> F[6])
{
    const int *nb = G.nb, *n = G.n;
    const int bidx[3] = {(int)(blk % nb[0]), (int)((blk / nb[0]) % nb[1]),
                         (int)(blk / ((long long)nb[0] * nb[1]))};
    int kind[6];
    double val[6];
    const int sb[3] = {-1, -1, -1}, se[3] = {2, 2, 2};
    if (has_bc) {
        const DevBCHead h = heads[blk];
        resolve_faces(entries + h.offset, h.count, sb, se, kind, val);
    } else {
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            kind[f] = FACE_NBR;
            val[f] = 0.0;
        }
    }
    const long long nx = n[0], nxy = (long long)n[0] * n[1];
    const T *self = cbase + blk * G.bstride;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int dir = f >> 1, side = f & 1;
        int b[3] = {bidx[0], bidx[1], bidx[2]};
        b[dir] = (b[dir] + (side ? 1 : -1) + nb[dir]) % nb[dir];
        const long long nbr = b[0] + (long long)nb[0] * (b[1] + (long long)nb[1] * b[2]);
        const T *src = kind[f] == FACE_NBR ? cbase + nbr * G.bstride : self;
        // index of the source line/plane along dir
        const int li = kind[f] == FACE_NBR ? (side ? 0 : n[dir] - 1) : (side ? n[dir] - 1 : 0);
        const long long step = dir == 0 ? 1 : (dir == 1 ? nx : nxy);
        F[f].p = src + li * step;
        F[f].mode = kind[f] == FACE_DIRICHLET ? 2 : (kind[f] == FACE_SYMMETRY ? 1 : 0);
        F[f].s = (T)val[f];
    }
}

// Fixed-order sum of n partials by one workgroup: strided per-thread sums
// issued eight loads at a time, a fixed wave tree, the waves in order.
template <int WG>
__device__ inline void wg_sum_partials(const double *__restrict__ p, int n, double *out, double *red)
{
    const int tid = threadIdx.x;
    double s = 0.0;
    for (int base = 0; base < n; base += 8 * WG) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int i = base + j * WG + tid;
            v[j] = i < n ? p[i] : 0.0;
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
        for (int w = 0; w < WG / 64; ++w)
            t += red[w];
        *out = t;
    }
    __syncthreads();
}

// Workgroup i is dispatched to XCD i % 8; give each XCD a contiguous range
// of tiles so that the tiles sharing halo lines meet in the same L2.
__device__ inline int xcd_remap(int i, int n)
{
    const int x = i & 7, local = i >> 3;
    const int q = n >> 3, r = n & 7;
    return x * q + (x < r ? x : r) + local;
}

// Source of the ghost layers of one block face.  Ghost layer l (1 = next to
// the face) at in-face offset i reads p[(li0 + lstep*l) * dstride + i], then
// copies it (mode 0), scales it by s (mode 1, Symmetry), or is the constant s
// (mode 2, Dirichlet).  Derivation from the lab rules:
//   neighbour, side 0: layer l is the neighbour's index N-l  (li0 = N,  lstep = -1)
//   neighbour, side 1: layer l is the neighbour's index l-1  (li0 = -1, lstep = +1)
//   absorbing, side 0/1: index 0 / N-1 for every layer       (lstep = 0)
//   symmetry,  side 0: mirror -1-(-l) = l-1                  (li0 = -1, lstep = +1)
//   symmetry,  side 1: mirror 2N-1-(N-1+l) = N-l             (li0 = N,  lstep = -1)
template <typename T>
struct FaceL {
    const T *p;
    int li0;
    int lstep;
    int mode;
    T s;
};

template <typename T>
__device__ inline T face_layer(const FaceL<T> &f, int l, long long dstride, long long i)
{
    if (f.mode == 2)
        return f.s;
    const T v = f.p[(long long)(f.li0 + f.lstep * l) * dstride + i];
    return f.mode == 1 ? f.s * v : v;
}

// Six faces of block `blk` for a non-tensorial stencil (sb, se) from the
// block's BC list (heads/entries may be null when no block carries a BC).
template <typename T>
__device__ inline void make_faces(const Geom &G, const DevBCHead *heads, const DevBC *entries, int has_bc,
                                  long long blk, const T *cbase, const int sb[3], const int se[3], FaceL<T> F[6])
{
    const int *nb = G.nb, *n = G.n;
    const int bidx[3] = {(int)(blk % nb[0]), (int)((blk / nb[0]) % nb[1]),
                         (int)(blk / ((long long)nb[0] * nb[1]))};
    int kind[6];
    double val[6];
    if (has_bc) {
        const DevBCHead h = heads[blk];
        resolve_faces(entries + h.offset, h.count, sb, se, kind, val);
    } else {
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            kind[f] = FACE_NBR;
            val[f] = 0.0;
        }
    }
    const T *self = cbase + blk * G.bstride;
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int dir = f >> 1, side = f & 1;
        int b[3] = {bidx[0], bidx[1], bidx[2]};
        b[dir] = (b[dir] + (side ? 1 : -1) + nb[dir]) % nb[dir];
        const long long nbr = b[0] + (long long)nb[0] * (b[1] + (long long)nb[1] * b[2]);
        const int N = n[dir];
        FaceL<T> F_;
        F_.s = (T)val[f];
        if (kind[f] == FACE_NBR) {
            F_.p = cbase + nbr * G.bstride;
            F_.li0 = side ? -1 : N;
            F_.lstep = side ? 1 : -1;
            F_.mode = 0;
        } else if (kind[f] == FACE_ABSORB) {
            F_.p = self;
            F_.li0 = side ? N - 1 : 0;
            F_.lstep = 0;
            F_.mode = 0;
        } else if (kind[f] == FACE_SYMMETRY) {
            F_.p = self;
            F_.li0 = side ? N : -1;
            F_.lstep = side ? -1 : 1;
            F_.mode = 1;
        } else { // Dirichlet (unwritten faces are rejected on the host)
            F_.p = self;
            F_.li0 = 0;
            F_.lstep = 0;
            F_.mode = 2;
        }
        F[f] = F_;
    }
}

} // namespace cnv
