// cnv_march.h -- the z-marching tile bodies shared by the single-grid
// kernels and the batched launch: the fused 7-point Laplacian / Euler step
// (s7_march) and the WENO5 face states (w5_march).
//
// A workgroup owns one x-y plane of one block of one component and marches
// through a chunk of z-planes.  The plane plus its x/y halo (read from the
// face-neighbour blocks or filled by the block's BC rules) is staged in LDS,
// double-buffered; the z neighbours come from registers holding the next
// planes, loaded ahead so that HBM reads overlap the arithmetic.
#pragma once

#include "cnv_device.h"

namespace cnv
{

#ifndef S7_PF
#define S7_PF 2 // planes loaded ahead of the plane being computed
#endif

// Tile geometry of the z-marching 7-point kernels for an NX x NY plane and
// WG threads: VEC-wide rows, R rows per thread, LDS planes of pitch P.
template <typename T, int NX, int NY, int WG>
struct S7Shape {
    static constexpr int VEC = 16 / (int)sizeof(T);
    static constexpr int TX = NX / VEC;  // threads per row
    static constexpr int ROWS = WG / TX; // rows per pass
    static constexpr int R = NY / ROWS;  // rows per thread
    static constexpr int P = NX + 2 * VEC; // LDS row pitch (interior starts at VEC)
    static constexpr int PLANE = (NY + 2) * P;
    static constexpr bool valid = NX % VEC == 0 && WG % TX == 0 && ROWS <= NY && NY % ROWS == 0 &&
                                  WG >= 2 * NY && WG >= 2 * TX;
};

// One tile of the fused 7-point kernel: planes [z0, z1) of the block at `ub`
// (NZ planes), one-layer faces F, output at `ob` (and the Laplacian at `lb`
// for MODE 2).  `lds` holds two planes (2 * PLANE elements).  Returns this
// thread's FP64 sum of u_new^2 (SUMSQ).  SUB = 0: the tile is the whole
// workgroup of WG threads; SUB = WG: the tile is one WG-thread slice of a
// larger workgroup (threads threadIdx.x / WG * WG ...), every slice running
// a tile of the same z-range so that the workgroup's barriers match.
template <typename T, int NX, int NY, int WG, int MODE, bool SUMSQ, int SUB = 0>
__device__ inline double s7_march(const T *__restrict__ ub, T *__restrict__ ob, T *__restrict__ lb,
                                  const FaceSrc<T> F[6], int NZ, int z0, int z1, T ihx2, T ihy2, T ihz2, T nudt,
                                  T *lds)
{
    using S = S7Shape<T, NX, NY, WG>;
    constexpr int VEC = S::VEC, TX = S::TX, ROWS = S::ROWS, R = S::R, P = S::P, PLANE = S::PLANE;
    static_assert(S::valid, "tile shape");
    static_assert(SUB == 0 || SUB == WG, "a slice is one tile");
    using V = VecT<T, VEC>;
    const int tid = SUB ? (int)(threadIdx.x % SUB) : (int)threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int x0 = tx * VEC;
    constexpr long long SXY = (long long)NX * NY;

    // halo roles: x-ghost scalar (tid < 2NY), y-ghost vector (tid < 2TX);
    // faces are picked by value so no array is indexed at run time
    const bool hx_on = tid < 2 * NY;
    const int hx_side = tid / NY, hx_y = tid % NY;
    const bool hy_on = tid < 2 * TX;
    const int hy_side = tid / TX, hy_x = (tid % TX) * VEC;
    const FaceSrc<T> FX = hx_side ? F[1] : F[0];
    const FaceSrc<T> FY = hy_side ? F[3] : F[2];
    const FaceSrc<T> FZ0 = F[4], FZ1 = F[5];

    struct Plane {
        V c[R];
        T hx;
        V hy;
    };

    auto load_plane = [&](int p, bool halo, Plane &P_) {
        if (p >= 0 && p < NZ) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                P_.c[r] = *reinterpret_cast<const V *>(ub + x0 + NX * (ty + r * ROWS) + SXY * p);
            if (halo) {
                if (hx_on)
                    P_.hx = face_val(FX, (long long)NX * hx_y + SXY * p);
                if (hy_on) {
                    const FaceSrc<T> f = FY;
                    if (f.mode == 2) {
#pragma unroll
                        for (int v = 0; v < VEC; ++v)
                            P_.hy.v[v] = f.s;
                    } else {
                        P_.hy = *reinterpret_cast<const V *>(f.p + hy_x + SXY * p);
                        if (f.mode == 1)
#pragma unroll
                            for (int v = 0; v < VEC; ++v)
                                P_.hy.v[v] = f.s * P_.hy.v[v];
                    }
                }
            }
        } else {
            const FaceSrc<T> f = p < 0 ? FZ0 : FZ1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (f.mode == 2) {
#pragma unroll
                    for (int v = 0; v < VEC; ++v)
                        P_.c[r].v[v] = f.s;
                } else {
                    P_.c[r] = *reinterpret_cast<const V *>(f.p + x0 + NX * (ty + r * ROWS));
                    if (f.mode == 1)
#pragma unroll
                        for (int v = 0; v < VEC; ++v)
                            P_.c[r].v[v] = f.s * P_.c[r].v[v];
                }
            }
        }
    };

    double acc = 0.0;

    // Ring of NQ planes in registers: plane p lives in slot (p - z0 + 1) % NQ.
    // The z loop is unrolled by NQ so every slot index is a constant (no
    // register moves); plane z + S7_PF is loaded while plane z is computed.
    constexpr int NQ = S7_PF + 2;
    Plane Q[NQ];
#pragma unroll
    for (int k = 0; k <= S7_PF; ++k)
        if (z0 - 1 + k <= z1)
            load_plane(z0 - 1 + k, k >= 1 && z0 - 1 + k < z1, Q[k]);

    for (int zb = z0; zb < z1; zb += NQ) {
        prio_ramp(zb - z0, z1 - z0);
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
            const int z = zb + u;
            if (z >= z1)
                break;
            const Plane &pm = Q[u % NQ];
            const Plane &pc = Q[(u + 1) % NQ];
            const Plane &pn = Q[(u + 2) % NQ];
            if (z + S7_PF <= z1)
                load_plane(z + S7_PF, z + S7_PF < z1, Q[(u + NQ - 1) % NQ]);
            T *L = lds + ((z - z0) & 1) * PLANE;
#pragma unroll
            for (int r = 0; r < R; ++r)
                *reinterpret_cast<V *>(L + (ty + r * ROWS + 1) * P + VEC + x0) = pc.c[r];
            if (hx_on)
                L[(hx_y + 1) * P + (hx_side ? VEC + NX : VEC - 1)] = pc.hx;
            if (hy_on)
                *reinterpret_cast<V *>(L + (hy_side ? NY + 1 : 0) * P + VEC + hy_x) = pc.hy;
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int y = ty + r * ROWS;
                const T *row = L + (y + 1) * P + VEC;
                const V ym = *reinterpret_cast<const V *>(row - P + x0);
                const V yp = *reinterpret_cast<const V *>(row + P + x0);
                const T xl = row[x0 - 1];
                const T xr = row[x0 + VEC];
                V res, lapv;
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    const T c = pc.c[r].v[v];
                    const T xm = v == 0 ? xl : pc.c[r].v[v - 1];
                    const T xp = v == VEC - 1 ? xr : pc.c[r].v[v + 1];
                    // p - 2c as one fma: 2c is exact, so it rounds like the subtraction
                    const T ddx = ihx2 * (fms2(c, xp) + xm);
                    const T ddy = ihy2 * (fms2(c, yp.v[v]) + ym.v[v]);
                    const T ddz = ihz2 * (fms2(c, pn.c[r].v[v]) + pm.c[r].v[v]);
                    const T lap = ddx + ddy + ddz;
                    lapv.v[v] = lap;
                    if (MODE == 0) {
                        res.v[v] = lap;
                    } else {
                        const T l = lap * nudt;
                        const T un = c + l;
                        res.v[v] = un;
                        if (SUMSQ) // one rounding of un^2 + acc
                            acc = __builtin_fma((double)un, (double)un, acc);
                    }
                }
                const long long off = x0 + NX * y + SXY * z;
                *reinterpret_cast<V *>(ob + off) = res;
                if (MODE == 2)
                    *reinterpret_cast<V *>(lb + off) = lapv;
            }
        }
    }

    return acc;
}

__device__ inline float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ inline float w5_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double w5_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
// The indicators are bk = (4 beta_k + 0.8e-6 S) r with r = 1/S up to a
// factor common to the three of them.  FP32: r = the hardware reciprocal of S.
// FP64: with S = m 2^e, m in [0.5, 1), r = 2^-e exactly and the epsilon term
// 0.8e-6 S r = 0.8e-6 m, which replaces a division by two exponent
// operations and a product.
__device__ inline void w5_scale(float S, float &r, float &eps)
{
    r = rcpf(S);
    eps = 0.8e-6f;
}
__device__ inline void w5_scale(double S, double &r, double &eps)
{
    r = __builtin_ldexp(1.0, -__builtin_amdgcn_frexp_exp(S));
    eps = 0.8e-6 * __builtin_amdgcn_frexp_mant(S);
}
// The two states c + nm / dm and c + np / dp: a multiply-add with the
// hardware reciprocal each in FP32; in FP64 one IEEE quotient q = 1/(dm dp)
// shared by both, c + (nm dp) q and c + (np dm) q
__device__ inline void w5_fin(float nm, float dm, float np, float dp, float c, float &um, float &up)
{
    um = __builtin_fmaf(nm, rcpf(dm), c);
    up = __builtin_fmaf(np, rcpf(dp), c);
}
__device__ inline void w5_fin(double nm, double dm, double np, double dp, double c, double &um, double &up)
{
    const double q = 1.0 / (dm * dp);
    um = c + (nm * dp) * q;
    up = c + (np * dm) * q;
}

// WENO5 (Jiang & Shu weights) in difference form.  With the stencil
// (a, b, c, d, e) = u_{i-2..i+2}: first differences dm2 = b - a,
// dm1 = c - b, d0 = d - c, dp1 = e - d; second differences tm = dm1 - dm2,
// t0 = d0 - dm1, tp = dp1 - d0 and Tk = (13/3) tk^2; S = sum of
// (u_j^2 + 1e-30) over the stencil, so that 0.8 beta_k / m2 = 4 beta_k / S
// with 4 beta_k = Tk + sk^2, s0 = a - 4b + 3c = 3 dm1 - dm2, s1 = d - b
// (sign free) = dm1 + d0, s2 = 3c - 4d + e = dp1 - 3 d0.  The nonlinear
// weights depend on the ratios of the bk = 4 beta_k / S + 0.8e-6 only, so
// this common factor 0.8 of the indicators and the epsilon changes nothing
// (nor does FP64's factor m in [0.5, 1), see w5_scale).
// The candidates enter as offsets from c; with p_k the product of the other
// two squared indicators and the linear weights (1, 6, 3) / 10:
//   u- = c + ((p12/6) C0 + p02 C1 + (p01/2) C2) / (p12 + 6 p02 + 3 p01)
//        C0 = 2 s0 - dm1, C1 = dm1 + 2 d0, C2 = 4 d0 - dp1
//   u+ = c + ((p12/2) E0 + p02 E1 + (p01/6) E2) / (3 p12 + 6 p02 + p01)
//        E0 = dm2 - 4 dm1, E1 = -2 dm1 - d0, E2 = 2 s2 + d0
// (6 (q_k - c) of Jiang & Shu's candidates q_k); FP64 divides once for both
// states, by the product of the two denominators.  Every operation is
// explicit (one fma or one rounding), so any caller that forms the same
// differences gets the same bits: cells of one row share the differences,
// and the FP64 evaluation is the CPU restatement's operation for operation.
// The indicators stay within [0.4e-6, 136] and their products within
// [1e-27, 1e10], so nothing over- or underflows at any data magnitude.
template <typename R>
__device__ inline R w5_sqr(R u) { return w5_fma(u, u, (R)1e-30); }
template <typename R>
__device__ inline R w5_T(R t) { return (((R)13 / (R)3) * t) * t; }
// S of a stencil from its five squares, in one fixed association
template <typename R>
__device__ inline R w5_S(R sa, R sb, R sc, R sd, R se) { return ((sa + sb) + (sc + sd)) + se; }

template <typename R>
__device__ inline void weno5_core(R c, R dm2, R dm1, R d0, R dp1, R Tm, R T0, R Tp, R S, R &um, R &up)
{
    R r, eps;
    w5_scale(S, r, eps);
    const R s0 = w5_fma((R)3, dm1, -dm2);
    const R s1 = dm1 + d0;
    const R s2 = w5_fma((R)-3, d0, dp1);
    const R b0 = w5_fma(w5_fma(s0, s0, Tm), r, eps);
    const R b1 = w5_fma(w5_fma(s1, s1, T0), r, eps);
    const R b2 = w5_fma(w5_fma(s2, s2, Tp), r, eps);
    const R g0 = b0 * b0, g1 = b1 * b1, g2 = b2 * b2;
    const R p12 = g1 * g2, p02 = g0 * g2, p01 = g0 * g1;
    const R C0 = w5_fma((R)2, s0, -dm1);
    const R C1 = w5_fma((R)2, d0, dm1);
    const R C2 = w5_fma((R)4, d0, -dp1);
    const R nm = w5_fma(p12 * ((R)1 / (R)6), C0, w5_fma(p02, C1, ((R)0.5 * p01) * C2));
    const R dm = w5_fma((R)6, p02, w5_fma((R)3, p01, p12));
    const R E0 = w5_fma((R)-4, dm1, dm2);
    const R E1 = w5_fma((R)-2, dm1, -d0);
    const R E2 = w5_fma((R)2, s2, d0);
    const R np = w5_fma((R)0.5 * p12, E0, w5_fma(p02, E1, (p01 * ((R)1 / (R)6)) * E2));
    const R dp = w5_fma((R)6, p02, w5_fma((R)3, p12, p01));
    w5_fin(nm, dm, np, dp, c, um, up);
}

// one cell from its stencil and the stencil's squares (sc shared between directions)
template <typename R>
__device__ inline void weno5_sq(R a, R b, R c, R d, R e, R sa, R sb, R sc, R sd, R se, R &um, R &up)
{
    const R dm2 = b - a, dm1 = c - b, d0 = d - c, dp1 = e - d;
    const R tm = dm1 - dm2, t0 = d0 - dm1, tp = dp1 - d0;
    weno5_core<R>(c, dm2, dm1, d0, dp1, w5_T(tm), w5_T(t0), w5_T(tp), w5_S(sa, sb, sc, sd, se), um, up);
}

template <typename T>
__device__ inline void weno5(T a, T b, T c, T d, T e, T &um, T &up)
{
    weno5_sq<T>(a, b, c, d, e, w5_sqr(a), w5_sqr(b), w5_sqr(c), w5_sqr(d), w5_sqr(e), um, up);
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
    // more than eight rows per thread spill registers to scratch memory
    static constexpr bool valid =
        NX % VEC == 0 && WG % TX == 0 && ROWS <= NY && NY % ROWS == 0 && VEC >= G && R <= 8;
};

// One tile of WENO5: planes [z0, z1) of the block at `ub` (NZ planes), faces
// F resolved for the lab's stencil, the six state components at ob + k *
// ocstride.  `lds` holds two planes (2 * PLANE elements).  SUB as in
// s7_march.
template <typename T, int NX, int NY, int WG, int SUB = 0>
__device__ inline void w5_march(const T *__restrict__ ub, T *__restrict__ ob, long long ocstride,
                                const FaceL<T> F[6], int NZ, int z0, int z1, T *lds)
{
    using S = W5Shape<T, NX, NY, WG>;
    constexpr int G = S::G, VEC = S::VEC, TX = S::TX, ROWS = S::ROWS, R = S::R, P = S::P, PLANE = S::PLANE;
    constexpr int NHX = S::NHX, NHY = S::NHY, HX = S::HX, HY = S::HY;
    static_assert(S::valid, "tile shape");
    static_assert(SUB == 0 || SUB == WG, "a slice is one tile");
    using V = VecT<T, VEC>;
    const int tid = SUB ? (int)(threadIdx.x % SUB) : (int)threadIdx.x;
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
        prio_ramp(z - z0, z1 - z0);
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
            // x: differences, second differences and squares shared along the row
            T sq[VEC + 4], dx[VEC + 3], tx[VEC + 3], Tx[VEC + 3], pp[VEC + 3];
#pragma unroll
            for (int j = 0; j < VEC + 4; ++j)
                sq[j] = w5_sqr(xs[j]);
#pragma unroll
            for (int j = 0; j < VEC + 3; ++j) {
                dx[j] = xs[j + 1] - xs[j];
                pp[j] = sq[j] + sq[j + 1];
            }
#pragma unroll
            for (int j = 1; j < VEC + 3; ++j) {
                tx[j] = dx[j] - dx[j - 1];
                Tx[j] = w5_T(tx[j]);
            }
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
                const int i = v + 2;
                weno5_core<T>(xs[i], dx[i - 2], dx[i - 1], dx[i], dx[i + 1], Tx[i - 1], Tx[i], Tx[i + 1],
                              (pp[i - 2] + pp[i]) + sq[i + 2], o[0].v[v], o[1].v[v]);
                const T c = q2.c[r].v[v], sc = sq[i];
                weno5_sq<T>(ym2.v[v], ym1.v[v], c, yp1.v[v], yp2.v[v], w5_sqr(ym2.v[v]), w5_sqr(ym1.v[v]), sc,
                            w5_sqr(yp1.v[v]), w5_sqr(yp2.v[v]), o[2].v[v], o[3].v[v]);
                weno5_sq<T>(q0.c[r].v[v], q1.c[r].v[v], c, q3.c[r].v[v], q4.c[r].v[v], w5_sqr(q0.c[r].v[v]),
                            w5_sqr(q1.c[r].v[v]), sc, w5_sqr(q3.c[r].v[v]), w5_sqr(q4.c[r].v[v]), o[4].v[v],
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

} // namespace cnv
