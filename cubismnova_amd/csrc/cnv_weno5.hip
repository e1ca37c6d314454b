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
#include "cnv_device.h"

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

// IEEE evaluation in the operation order of the CPU restatement.  The
// weights alpha_k = d_k / (eps' + beta_k / m2)^2 are scaled by the common
// factor prod_j (eps' + beta_j / m2)^2, so only 1/m2 and the two
// normalisations divide; the 1/6 of the candidate polynomials goes into the
// normalisation.
template <typename R>
__device__ inline void weno5_exact(R a, R b, R c, R d, R e, R &um, R &up)
{
    const R m2 = (a * a + b * b + c * c + d * d + e * e) * (R)0.2 + (R)1e-30;
    const R im2 = (R)1 / m2;
    const R t0 = a - 2 * b + c, s0 = a - 4 * b + 3 * c;
    const R t1 = b - 2 * c + d, s1 = b - d;
    const R t2 = c - 2 * d + e, s2 = 3 * c - 4 * d + e;
    const R k13 = (R)13 / (R)12;
    const R b0 = (k13 * t0 * t0 + (R)0.25 * s0 * s0) * im2 + (R)1e-6;
    const R b1 = (k13 * t1 * t1 + (R)0.25 * s1 * s1) * im2 + (R)1e-6;
    const R b2 = (k13 * t2 * t2 + (R)0.25 * s2 * s2) * im2 + (R)1e-6;
    const R g0 = b0 * b0, g1 = b1 * b1, g2 = b2 * b2;
    const R p12 = g1 * g2, p02 = g0 * g2, p01 = g0 * g1;
    {
        const R w0 = (R)0.1 * p12, w1 = (R)0.6 * p02, w2 = (R)0.3 * p01;
        const R q0 = 2 * a - 7 * b + 11 * c;
        const R q1 = -b + 5 * c + 2 * d;
        const R q2 = 2 * c + 5 * d - e;
        um = (w0 * q0 + w1 * q1 + w2 * q2) / ((R)6 * (w0 + w1 + w2));
    }
    {
        const R w0 = (R)0.3 * p12, w1 = (R)0.6 * p02, w2 = (R)0.1 * p01;
        const R q0 = -a + 5 * b + 2 * c;
        const R q1 = 2 * b + 5 * c - d;
        const R q2 = 11 * c - 7 * d + 2 * e;
        up = (w0 * q0 + w1 * q1 + w2 * q2) / ((R)6 * (w0 + w1 + w2));
    }
}

__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// FP32 evaluation: same formulas, contracted, divisions by the hardware
// reciprocal.  Every scaled weight stays within [1e-25, 1e10] for any data
// (beta_k / m2 <= 170 since max u^2 <= 5 m2), so no product over- or
// underflows in FP32.
__device__ inline void weno5_fast(float a, float b, float c, float d, float e, float &um, float &up)
{
#pragma clang fp contract(fast)
    const float m2 = (a * a + b * b + c * c + d * d + e * e) * 0.2f + 1e-30f;
    const float im2 = rcpf(m2);
    const float t0 = a - 2.f * b + c, s0 = a - 4.f * b + 3.f * c;
    const float t1 = b - 2.f * c + d, s1 = b - d;
    const float t2 = c - 2.f * d + e, s2 = 3.f * c - 4.f * d + e;
    const float k13 = 13.f / 12.f;
    const float b0 = (k13 * t0 * t0 + 0.25f * s0 * s0) * im2 + 1e-6f;
    const float b1 = (k13 * t1 * t1 + 0.25f * s1 * s1) * im2 + 1e-6f;
    const float b2 = (k13 * t2 * t2 + 0.25f * s2 * s2) * im2 + 1e-6f;
    const float g0 = b0 * b0, g1 = b1 * b1, g2 = b2 * b2;
    const float p12 = g1 * g2, p02 = g0 * g2, p01 = g0 * g1;
    {
        const float w0 = 0.1f * p12, w1 = 0.6f * p02, w2 = 0.3f * p01;
        const float q0 = 2.f * a - 7.f * b + 11.f * c;
        const float q1 = -b + 5.f * c + 2.f * d;
        const float q2 = 2.f * c + 5.f * d - e;
        um = (w0 * q0 + w1 * q1 + w2 * q2) * rcpf(6.f * (w0 + w1 + w2));
    }
    {
        const float w0 = 0.3f * p12, w1 = 0.6f * p02, w2 = 0.1f * p01;
        const float q0 = -a + 5.f * b + 2.f * c;
        const float q1 = 2.f * b + 5.f * c - d;
        const float q2 = 11.f * c - 7.f * d + 2.f * e;
        up = (w0 * q0 + w1 * q1 + w2 * q2) * rcpf(6.f * (w0 + w1 + w2));
    }
}

template <typename T>
__device__ inline void weno5(T a, T b, T c, T d, T e, T &um, T &up)
{
    if constexpr (sizeof(T) == 4)
        weno5_fast(a, b, c, d, e, um, up);
    else
        weno5_exact<T>(a, b, c, d, e, um, up);
}

// Tile geometry of the z-marching WENO5 kernels (two ghost layers read).
template <typename T, int NX, int NY, int WG>
struct W5Shape {
    static constexpr int G = 2;          // ghost layers read (i-2..i+2)
    static constexpr int VEC = 16 / (int)sizeof(T);
    static constexpr int TX = NX / VEC;
    static constexpr int ROWS = WG / TX;
    static constexpr int R = NY / ROWS;
    static constexpr int P = NX + 2 * VEC; // x halo of 2 fits in the VEC padding
    static constexpr int PLANE = (NY + 2 * G) * P;
    static constexpr int NHX = 2 * G * NY; // x-ghost scalars per plane
    static constexpr int NHY = 2 * G * TX; // y-ghost vectors per plane
    static constexpr int HX = (NHX + WG - 1) / WG;
    static constexpr int HY = (NHY + WG - 1) / WG;
    static constexpr bool valid = NX % VEC == 0 && WG % TX == 0 && ROWS <= NY && NY % ROWS == 0 && VEC >= G;
};

// One tile of WENO5: planes [z0, z1) of the block at `ub` (NZ planes), faces
// F resolved for the lab's stencil, the six state components at ob + k *
// ocstride.  `lds` holds two planes (2 * PLANE elements).
template <typename T, int NX, int NY, int WG>
__device__ inline void w5_march(const T *__restrict__ ub, T *__restrict__ ob, long long ocstride,
                                const FaceL<T> F[6], int NZ, int z0, int z1, T *lds)
{
    using S = W5Shape<T, NX, NY, WG>;
    constexpr int G = S::G, VEC = S::VEC, TX = S::TX, ROWS = S::ROWS, R = S::R, P = S::P, PLANE = S::PLANE;
    constexpr int NHX = S::NHX, NHY = S::NHY, HX = S::HX, HY = S::HY;
    static_assert(S::valid, "tile shape");
    using V = VecT<T, VEC>;
    const int tid = threadIdx.x;
    const FaceL<T> FX0 = F[0], FX1 = F[1], FY0 = F[2], FY1 = F[3], FZ0 = F[4], FZ1 = F[5];

    const int tx = tid % TX, ty = tid / TX;
    const int x0 = tx * VEC;
    constexpr long long SXY = (long long)NX * NY;

    struct Plane {
        V c[R];
        T hx[HX];
        V hy[HY];
    };

    // x-ghost slot s: side = s / (G*NY), layer = 1 + (s / NY) % G, row = s % NY
    // y-ghost slot s: side = s / (G*TX), layer = 1 + (s / TX) % G, x = (s % TX) * VEC
    auto load_plane = [&](int p, bool halo, Plane &Q) {
        if (p >= 0 && p < NZ) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                Q.c[r] = *reinterpret_cast<const V *>(ub + x0 + NX * (ty + r * ROWS) + SXY * p);
            if (halo) {
#pragma unroll
                for (int h = 0; h < HX; ++h) {
                    const int s = tid + h * WG;
                    if (s < NHX) {
                        const int side = s / (G * NY), layer = 1 + (s / NY) % G, y = s % NY;
                        const FaceL<T> f = side ? FX1 : FX0;
                        Q.hx[h] = face_layer(f, layer, 1, (long long)NX * y + SXY * p);
                    }
                }
#pragma unroll
                for (int h = 0; h < HY; ++h) {
                    const int s = tid + h * WG;
                    if (s < NHY) {
                        const int side = s / (G * TX), layer = 1 + (s / TX) % G, x = (s % TX) * VEC;
                        const FaceL<T> f = side ? FY1 : FY0;
                        if (f.mode == 2) {
#pragma unroll
                            for (int v = 0; v < VEC; ++v)
                                Q.hy[h].v[v] = f.s;
                        } else {
                            Q.hy[h] = *reinterpret_cast<const V *>(f.p + (long long)(f.li0 + f.lstep * layer) * NX +
                                                                    x + SXY * p);
                            if (f.mode == 1)
#pragma unroll
                                for (int v = 0; v < VEC; ++v)
                                    Q.hy[h].v[v] = f.s * Q.hy[h].v[v];
                        }
                    }
                }
            }
        } else {
            const FaceL<T> f = p < 0 ? FZ0 : FZ1;
            const int layer = p < 0 ? -p : p - NZ + 1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (f.mode == 2) {
#pragma unroll
                    for (int v = 0; v < VEC; ++v)
                        Q.c[r].v[v] = f.s;
                } else {
                    Q.c[r] = *reinterpret_cast<const V *>(f.p + (long long)(f.li0 + f.lstep * layer) * SXY + x0 +
                                                          NX * (ty + r * ROWS));
                    if (f.mode == 1)
#pragma unroll
                        for (int v = 0; v < VEC; ++v)
                            Q.c[r].v[v] = f.s * Q.c[r].v[v];
                }
            }
        }
    };

    // register queue: q[0..4] = planes z-2..z+2, q5 = z+3 in flight
    Plane q0, q1, q2, q3, q4, q5;
    load_plane(z0 - 2, false, q0);
    load_plane(z0 - 1, false, q1);
    load_plane(z0, true, q2);
    load_plane(z0 + 1, z0 + 1 < z1, q3);
    load_plane(z0 + 2, z0 + 2 < z1, q4);

    for (int z = z0; z < z1; ++z) {
        if (z + 3 <= z1 + 1)
            load_plane(z + 3, z + 3 < z1, q5);
        T *L = lds + ((z - z0) & 1) * PLANE;
#pragma unroll
        for (int r = 0; r < R; ++r)
            *reinterpret_cast<V *>(L + (ty + r * ROWS + G) * P + VEC + x0) = q2.c[r];
#pragma unroll
        for (int h = 0; h < HX; ++h) {
            const int s = tid + h * WG;
            if (s < NHX) {
                const int side = s / (G * NY), layer = 1 + (s / NY) % G, y = s % NY;
                L[(y + G) * P + (side ? VEC + NX - 1 + layer : VEC - layer)] = q2.hx[h];
            }
        }
#pragma unroll
        for (int h = 0; h < HY; ++h) {
            const int s = tid + h * WG;
            if (s < NHY) {
                const int side = s / (G * TX), layer = 1 + (s / TX) % G, x = (s % TX) * VEC;
                const int row = side ? NY - 1 + layer + G : G - layer;
                *reinterpret_cast<V *>(L + row * P + VEC + x) = q2.hy[h];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int y = ty + r * ROWS;
            const T *row = L + (y + G) * P + VEC;
            const V ym2 = *reinterpret_cast<const V *>(row - 2 * P + x0);
            const V ym1 = *reinterpret_cast<const V *>(row - P + x0);
            const V yp1 = *reinterpret_cast<const V *>(row + P + x0);
            const V yp2 = *reinterpret_cast<const V *>(row + 2 * P + x0);
            T xs[VEC + 4];
            xs[0] = row[x0 - 2];
            xs[1] = row[x0 - 1];
#pragma unroll
            for (int v = 0; v < VEC; ++v)
                xs[2 + v] = q2.c[r].v[v];
            xs[VEC + 2] = row[x0 + VEC];
            xs[VEC + 3] = row[x0 + VEC + 1];
            V o[6];
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
                weno5<T>(xs[v], xs[v + 1], xs[v + 2], xs[v + 3], xs[v + 4], o[0].v[v], o[1].v[v]);
                weno5<T>(ym2.v[v], ym1.v[v], q2.c[r].v[v], yp1.v[v], yp2.v[v], o[2].v[v], o[3].v[v]);
                weno5<T>(q0.c[r].v[v], q1.c[r].v[v], q2.c[r].v[v], q3.c[r].v[v], q4.c[r].v[v], o[4].v[v],
                         o[5].v[v]);
            }
            const long long off = x0 + NX * y + SXY * z;
#pragma unroll
            for (int k = 0; k < 6; ++k)
                store_nt<T, VEC>(ob + k * ocstride + off, o[k]);
        }
        q0 = q1;
        q1 = q2;
        q2 = q3;
        q3 = q4;
        q4 = q5;
    }
}

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
static W5Tune g_w5_tune[2] = {{256, 8}, {256, 8}}; // [dtype]

template <typename T, int NX, int WG>
constexpr bool w5_valid()
{
    constexpr int VEC = 16 / (int)sizeof(T);
    constexpr int TX = NX / VEC;
    return NX % VEC == 0 && WG % TX == 0 && (WG / TX) <= NX && NX % (WG / TX) == 0;
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
        W5Tune tu;
        {
            std::lock_guard<std::mutex> lk(g_w5_mtx);
            tu = over ? *over : g_w5_tune[sizeof(T) == 8 ? 1 : 0];
        }
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

int weno5_set_tune(int dtype, int wg, int zchunk)
{
    std::lock_guard<std::mutex> lk(g_w5_mtx);
    g_w5_tune[dtype == CNV_F64 ? 1 : 0] = {wg, zchunk};
    return CNV_OK;
}

int weno5_get_tune(int dtype, int *wg, int *zchunk)
{
    std::lock_guard<std::mutex> lk(g_w5_mtx);
    *wg = g_w5_tune[dtype == CNV_F64 ? 1 : 0].wg;
    *zchunk = g_w5_tune[dtype == CNV_F64 ? 1 : 0].zchunk;
    return CNV_OK;
}

// Times every (workgroup size, z-chunk) candidate on in -> out and keeps the
// fastest for the input dtype.
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
    const int wgs[3] = {64, 128, 256};
    const int zcs[5] = {4, 8, 16, 32, 64};
    float best = 1e30f;
    int bw = 0, bz = 0, cnt = 0, rc = CNV_OK;
    for (int wi = 0; wi < 3 && rc == CNV_OK; ++wi)
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
    weno5_set_tune(in->desc.dtype, bw, bz);
    if (best_wg)
        *best_wg = bw;
    if (best_z)
        *best_z = bz;
    if (nresults)
        *nresults = cnt < max_results ? cnt : max_results;
    return CNV_OK;
}

} // namespace cnv
