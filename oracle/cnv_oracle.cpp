// cnv_oracle.cpp -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
//
// CPU restatement of CubismNova's block-structured stencil path, written from
// the reference's algorithm so that the HIP kernels can be checked against it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only to check results or as the timed CPU baseline.
// The product library (libcubismnova_amd.so) never links or calls it.
//
// Functions and the reference code they restate (paths relative to the
// CubismNova source tree):
//   OLab::allocate      include/Cubism/Block/FieldLab.h:104-165
//   OLab::loadData      include/Cubism/Block/FieldLab.h:180-224
//   OLab::loadInner     include/Cubism/Block/FieldLabLoader.h:129-173
//   OLab::loadGhosts    include/Cubism/Block/FieldLabLoader.h:175-287
//   wrap_block          include/Cubism/Block/FieldLabLoader.h:405-411
//   applyAbsorbing      include/Cubism/BC/Absorbing.h:62-104
//   applyDirichlet      include/Cubism/BC/Dirichlet.h:75-113
//   applySymmetry       include/Cubism/BC/Symmetry.h:69-114
//   bcValidStencil      include/Cubism/BC/Base.h:73-78
//   laplacianBlock      test/sandbox/Benchmarks/Kernel/Laplacian/Laplacian.cpp:75-80,123-130
//   eulerBlock          include/Cubism/Block/Field.h:457-512 with
//                       include/Cubism/Block/FieldOperator.h:11-23,43-114
//   weno5 / jacobi      no reference implementation exists (README.rst:28-31
//                       advertises WENO only); restated from Jiang & Shu,
//                       J. Comput. Phys. 126:202-228 (1996), and from the
//                       textbook weighted-Jacobi iteration.  WENO5 has two
//                       statements: the textbook one (weno5_cell_textbook)
//                       and the difference form the kernels evaluate
//                       (weno5_cell), checked against each other.
//
// Pinning (what checks these restatements independently of their own text):
// the reference's own unit-test properties (FieldLabTest.cpp, BC/*Test.cpp,
// CartesianTest.cpp), an independent global-index formulation of periodic
// ghosts, closed-form Fourier eigenvalues of the 7-point operator, WENO5
// polynomial exactness / mirror symmetry / convergence order and agreement of
// its two statements to rounding, and manufactured
// solutions for the Jacobi operator.  All of it lives in tests/.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// BC kinds; numbering shared with include/cubismnova_amd.h
enum { ORC_BC_BASE = 0, ORC_BC_ABSORBING = 1, ORC_BC_DIRICHLET = 2, ORC_BC_SYMMETRY = 3 };

typedef struct {
    int kind;
    int dir;
    int side;
    int reserved;
    double value; // Dirichlet constant or Symmetry sign
} orc_bc;

typedef struct {
    void *data;            // first element of block 0 of this component
    int dtype;             // 0 = float, 1 = double
    int nblocks[3];
    int cells[3];
    int plus;               // bit d set: blocks at the upper end of dim d hold one
                            // more element (node / face fields, BlockFieldAssembler.h:121-131)
    long long block_stride; // elements between consecutive blocks
    const int *bc_offsets;  // nblocks_total + 1 prefix offsets, or NULL
    const orc_bc *bcs;      // flattened per-block BC lists
} orc_field;

typedef struct {
    int begin[3];
    int end[3];
    int tensorial;
} orc_stencil;
}

namespace
{

struct Idx3 {
    long long v[3];
};

inline long long prod3(const int *e) { return (long long)e[0] * e[1] * e[2]; }

inline int wrap_block(int p, int n) { return (p + n) % n; }

inline long long block_flat(const int b[3], const int nb[3])
{
    return b[0] + (long long)nb[0] * (b[1] + (long long)nb[1] * b[2]);
}

// element extent of block b in dim d: cells, plus one at the upper end of
// dims flagged in f.plus (node fields, or face fields in their direction)
inline int block_extent(const orc_field &f, const int b[3], int d)
{
    return f.cells[d] + (((f.plus >> d) & 1) && b[d] == f.nblocks[d] - 1 ? 1 : 0);
}

template <typename T>
struct BCT {
    int kind, dir, side;
    T value;
};

// Ghost-extended lab with the reference's allocation geometry.
template <typename T>
struct OLab {
    int lab_begin[3] = {0, 0, 0}; // offset of local index 0 in the allocation
    int ext[3] = {0, 0, 0};       // allocation extent
    int sb[3], se[3];
    bool tens = false;
    int N[3] = {0, 0, 0}; // active (loaded) range extent
    std::vector<T> buf;

    T &at(long long x, long long y, long long z)
    {
        return buf[(x + lab_begin[0]) +
                   (long long)ext[0] *
                       ((y + lab_begin[1]) + (long long)ext[1] * (z + lab_begin[2]))];
    }

    void allocate(const orc_stencil &s, const int max_req_extent[3], T poison)
    {
        const int n_per_align = 32 / (int)sizeof(T);
        int max_extent[3];
        for (int d = 0; d < 3; ++d) {
            sb[d] = s.begin[d];
            se[d] = s.end[d];
            max_extent[d] = max_req_extent[d] + 2;
            lab_begin[d] = -sb[d];
        }
        tens = s.tensorial != 0;
        lab_begin[0] = ((lab_begin[0] + n_per_align - 1) / n_per_align) * n_per_align;
        for (int d = 0; d < 3; ++d)
            max_extent[d] += se[d] - 1;
        max_extent[0] = ((max_extent[0] + n_per_align - 1) / n_per_align) * n_per_align;
        for (int d = 0; d < 3; ++d)
            ext[d] = lab_begin[d] + max_extent[d];
        buf.assign((size_t)prod3(ext), poison);
    }

    void loadInner(const T *src)
    {
        for (int iz = 0; iz < N[2]; ++iz)
            for (int iy = 0; iy < N[1]; ++iy)
                std::memcpy(&at(0, iy, iz), src + (size_t)N[0] * (iy + (size_t)N[1] * iz),
                            sizeof(T) * N[0]);
    }

    // 26-neighbour ghost copy; neighbours found through periodic block wrap.
    void loadGhosts(const orc_field &f, const int bidx[3], const bool periodic[3],
                    const int skip[3])
    {
        const T *base = static_cast<const T *>(f.data);
        int halo[3];
        for (int d = 0; d < 3; ++d)
            halo[d] = N[d] + se[d] - 1;
        for (int i = 0; i < 27; ++i) {
            if (i == 13)
                continue;
            const int bi[3] = {i % 3 - 1, (i / 3) % 3 - 1, i / 9 - 1};
            bool skip_it = false;
            for (int d = 0; d < 3; ++d)
                if (!periodic[d] && bi[d] == skip[d])
                    skip_it = true;
            if (skip_it)
                continue;
            if (!tens && (std::abs(bi[0]) + std::abs(bi[1]) + std::abs(bi[2]) > 1))
                continue;
            int nb[3], ne[3], shift[3];
            for (int d = 0; d < 3; ++d) {
                nb[d] = wrap_block(bidx[d] + bi[d], f.nblocks[d]);
                ne[d] = block_extent(f, nb, d);
                shift[d] = bi[d] < 0 ? ne[d] - N[d] : 0; // FieldLabLoader.h:217-219
            }
            const T *src = base + block_flat(nb, f.nblocks) * f.block_stride;
            int b[3], e[3];
            for (int d = 0; d < 3; ++d) {
                if (bi[d] < 0) {
                    b[d] = sb[d];
                    e[d] = 0;
                } else if (bi[d] == 0) {
                    b[d] = 0;
                    e[d] = N[d];
                } else {
                    b[d] = N[d];
                    e[d] = halo[d];
                }
            }
            const int nx = e[0] - b[0];
            if (nx <= 0)
                continue;
            // source index p - bi N + shift in the neighbour's own extent
            // (FieldLabLoader.h:236-241)
            for (int iz = b[2]; iz < e[2]; ++iz)
                for (int iy = b[1]; iy < e[1]; ++iy) {
                    const long long sx = b[0] - (long long)bi[0] * N[0] + shift[0];
                    const long long sy = iy - (long long)bi[1] * N[1] + shift[1];
                    const long long sz = iz - (long long)bi[2] * N[2] + shift[2];
                    std::memcpy(&at(b[0], iy, iz), src + sx + (long long)ne[0] * (sy + (long long)ne[1] * sz),
                                sizeof(T) * nx);
                }
        }
    }

    bool bcValidStencil(int dir) const { return !(sb[dir] == 0 || se[dir] == 1); }

    void applyAbsorbing(int dir, int side)
    {
        if (!bcValidStencil(dir))
            return;
        int slice[3], begin[3] = {0, 0, 0};
        for (int d = 0; d < 3; ++d)
            slice[d] = tens ? (N[d] + se[d] - 1 - sb[d]) : N[d];
        if (tens)
            for (int d = 0; d < 3; ++d)
                begin[d] = sb[d];
        int src[3] = {0, 0, 0};
        src[dir] = (side == 0) ? 1 : -1;
        const int sdir = (side == 0) ? -1 : 1;
        int end;
        if (side == 0) {
            begin[dir] = -1;
            end = -sb[dir];
        } else {
            begin[dir] = N[dir];
            end = se[dir] - 1;
        }
        slice[dir] = 1;
        for (int pz = 0; pz < slice[2]; ++pz)
            for (int py = 0; py < slice[1]; ++py)
                for (int px = 0; px < slice[0]; ++px) {
                    long long q[3] = {px + begin[0], py + begin[1], pz + begin[2]};
                    const T val = at(q[0] + src[0], q[1] + src[1], q[2] + src[2]);
                    at(q[0], q[1], q[2]) = val;
                    for (int i = 1; i < end; ++i) {
                        q[dir] += sdir;
                        at(q[0], q[1], q[2]) = val;
                    }
                }
    }

    // region shared by Dirichlet and Symmetry
    void slabRegion(int dir, int side, int extent[3], int start[3]) const
    {
        for (int d = 0; d < 3; ++d) {
            extent[d] = tens ? (N[d] + se[d] - 1 - sb[d]) : N[d];
            start[d] = 0;
        }
        if (side == 0) {
            extent[dir] = -sb[dir];
            if (tens)
                for (int d = 0; d < 3; ++d)
                    start[d] = sb[d];
            else
                start[dir] = sb[dir];
        } else {
            extent[dir] = se[dir] - 1;
            if (tens)
                for (int d = 0; d < 3; ++d)
                    start[d] = sb[d];
            start[dir] = N[dir];
        }
    }

    void applyDirichlet(int dir, int side, T value)
    {
        if (!bcValidStencil(dir))
            return;
        int extent[3], start[3];
        slabRegion(dir, side, extent, start);
        for (int pz = 0; pz < extent[2]; ++pz)
            for (int py = 0; py < extent[1]; ++py)
                for (int px = 0; px < extent[0]; ++px)
                    at(px + start[0], py + start[1], pz + start[2]) = value;
    }

    void applySymmetry(int dir, int side, T sign)
    {
        if (!bcValidStencil(dir))
            return;
        int extent[3], start[3];
        slabRegion(dir, side, extent, start);
        const long long roffset = (side == 0) ? -1 : 2LL * N[dir] - 1;
        for (int pz = 0; pz < extent[2]; ++pz)
            for (int py = 0; py < extent[1]; ++py)
                for (int px = 0; px < extent[0]; ++px) {
                    const long long q[3] = {px + start[0], py + start[1], pz + start[2]};
                    long long r[3] = {q[0], q[1], q[2]};
                    r[dir] = roffset - r[dir];
                    at(q[0], q[1], q[2]) = sign * at(r[0], r[1], r[2]);
                }
    }

    void loadData(const orc_field &f, const int bidx[3], const std::vector<BCT<T>> &bcs,
                  bool apply_bc)
    {
        for (int d = 0; d < 3; ++d)
            N[d] = block_extent(f, bidx, d);
        const T *base = static_cast<const T *>(f.data);
        loadInner(base + block_flat(bidx, f.nblocks) * f.block_stride);
        bool periodic[3] = {true, true, true};
        int skip[3] = {1, 1, 1};
        for (const auto &bc : bcs) {
            periodic[bc.dir] = (bc.kind == ORC_BC_BASE);
            skip[bc.dir] = (bc.side == 0) ? -1 : 1;
        }
        loadGhosts(f, bidx, periodic, skip);
        if (apply_bc) {
            for (const auto &bc : bcs) {
                switch (bc.kind) {
                case ORC_BC_ABSORBING:
                    applyAbsorbing(bc.dir, bc.side);
                    break;
                case ORC_BC_DIRICHLET:
                    applyDirichlet(bc.dir, bc.side, bc.value);
                    break;
                case ORC_BC_SYMMETRY:
                    applySymmetry(bc.dir, bc.side, bc.value);
                    break;
                default:
                    break; // Base: no-op
                }
            }
        }
    }

    // copy the active lab range [sb, N + se - 1) into a compact x-fastest array
    void extract(T *out)
    {
        const int L[3] = {N[0] + se[0] - 1 - sb[0], N[1] + se[1] - 1 - sb[1],
                          N[2] + se[2] - 1 - sb[2]};
        for (int z = 0; z < L[2]; ++z)
            for (int y = 0; y < L[1]; ++y)
                for (int x = 0; x < L[0]; ++x)
                    out[x + (size_t)L[0] * (y + (size_t)L[1] * z)] = at(x + sb[0], y + sb[1], z + sb[2]);
    }
};

template <typename T>
std::vector<BCT<T>> blockBCs(const orc_field &f, long long bflat)
{
    std::vector<BCT<T>> v;
    if (!f.bc_offsets)
        return v;
    for (int k = f.bc_offsets[bflat]; k < f.bc_offsets[bflat + 1]; ++k) {
        const orc_bc &b = f.bcs[k];
        v.push_back({b.kind, b.dir, b.side, (T)b.value});
    }
    return v;
}

template <typename T>
std::vector<BCT<T>> listBCs(const orc_bc *bcs, int n)
{
    std::vector<BCT<T>> v;
    for (int k = 0; k < n; ++k)
        v.push_back({bcs[k].kind, bcs[k].dir, bcs[k].side, (T)bcs[k].value});
    return v;
}

inline void block_index(long long i, const int nb[3], int b[3])
{
    b[0] = (int)(i % nb[0]);
    b[1] = (int)((i / nb[0]) % nb[1]);
    b[2] = (int)(i / ((long long)nb[0] * nb[1]));
}

template <typename T>
void lab_load(const orc_field &f, const int bidx[3], const orc_stencil &s, int apply_bc,
              const orc_bc *ext, int next, double poison, void *out)
{
    OLab<T> lab;
    lab.allocate(s, f.cells, (T)poison);
    const long long bflat = block_flat(bidx, f.nblocks);
    const auto bcs = (next >= 0) ? listBCs<T>(ext, next) : blockBCs<T>(f, bflat);
    lab.loadData(f, bidx, bcs, apply_bc != 0);
    lab.extract(static_cast<T *>(out));
}

// res = ihx2*((l[+x] - 2 l) + l[-x]) + ihy2*(...) + ihz2*(...), summed left to right
template <typename T>
void laplacianBlock(OLab<T> &lab, T ihx2, T ihy2, T ihz2, T *res)
{
    const int *N = lab.N;
    for (int iz = 0; iz < N[2]; ++iz)
        for (int iy = 0; iy < N[1]; ++iy)
            for (int ix = 0; ix < N[0]; ++ix) {
                const T c = lab.at(ix, iy, iz);
                const T ddx = ihx2 * (lab.at(ix + 1, iy, iz) - 2 * c + lab.at(ix - 1, iy, iz));
                const T ddy = ihy2 * (lab.at(ix, iy + 1, iz) - 2 * c + lab.at(ix, iy - 1, iz));
                const T ddz = ihz2 * (lab.at(ix, iy, iz + 1) - 2 * c + lab.at(ix, iy, iz - 1));
                res[ix + (size_t)N[0] * (iy + (size_t)N[1] * iz)] = ddx + ddy + ddz;
            }
}

// Field ops: lap *= s (scalar), then u += lap (field), each a separate rounding
template <typename T>
void eulerBlock(const T *u, T *lap, T s, T *unew, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        lap[i] = lap[i] * s;
    for (size_t i = 0; i < n; ++i)
        unew[i] = u[i] + lap[i];
}

template <typename T>
double diffusion(const orc_field &in, void *out_v, long long out_stride, void *lap_v,
                 const double h[3], double nu_dt, int mode, int nthreads)
{
    // mode 0: Laplacian only into out; mode 1: Euler update into out
    const T ihx2 = (T)(1. / (h[0] * h[0]));
    const T ihy2 = (T)(1. / (h[1] * h[1]));
    const T ihz2 = (T)(1. / (h[2] * h[2]));
    const T s = (T)nu_dt;
    const long long nbt = prod3(in.nblocks);
    const size_t n = (size_t)prod3(in.cells);
    T *out = static_cast<T *>(out_v);
    T *lapo = static_cast<T *>(lap_v);
    const T *u = static_cast<const T *>(in.data);
    std::vector<long double> partial((size_t)nbt, 0.0L);
    const orc_stencil st = {{-1, -1, -1}, {2, 2, 2}, 0};
#pragma omp parallel num_threads(nthreads)
    {
        OLab<T> lab; // thread-local lab
        lab.allocate(st, in.cells, (T)0);
        std::vector<T> lap(n);
#pragma omp for schedule(dynamic, 1)
        for (long long b = 0; b < nbt; ++b) {
            int bidx[3];
            block_index(b, in.nblocks, bidx);
            lab.loadData(in, bidx, blockBCs<T>(in, b), true);
            laplacianBlock(lab, ihx2, ihy2, ihz2, lap.data());
            T *ob = out + b * out_stride;
            if (mode == 0) {
                std::memcpy(ob, lap.data(), n * sizeof(T));
            } else {
                if (lapo)
                    std::memcpy(lapo + b * out_stride, lap.data(), n * sizeof(T));
                eulerBlock(u + b * in.block_stride, lap.data(), s, ob, n);
                long double acc = 0.0L;
                for (size_t i = 0; i < n; ++i)
                    acc += (long double)((double)ob[i] * (double)ob[i]);
                partial[(size_t)b] = acc;
            }
        }
    }
    long double tot = 0.0L;
    for (auto p : partial)
        tot += p;
    return (double)tot;
}

// ---------------------------------------------------------------- WENO5
// Per-cell left/right states with Jiang-Shu weights; eps = 1e-6 * m2 where m2
// is the mean of the squared values of the five stencil cells (plus 1e-30),
// evaluated in the scale-free form beta_k / m2.  The textbook statement: it
// pins the difference form below (agreement to rounding).
template <typename R>
inline void weno5_cell_textbook(R a, R b, R c, R d, R e, R &um, R &up)
{
    // Jiang & Shu weights alpha_k = d_k / (eps' + beta_k / m2)^2, written
    // with the common factor prod_j (eps' + beta_j / m2)^2 multiplied out so
    // that only 1/m2 and the two normalisations divide; the 1/6 of the
    // candidate polynomials goes into the normalisation.
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
    // u^- at i+1/2: linear weights (1,6,3)/10 on stencils 0,1,2
    {
        const R w0 = (R)0.1 * p12, w1 = (R)0.6 * p02, w2 = (R)0.3 * p01;
        const R q0 = 2 * a - 7 * b + 11 * c;
        const R q1 = -b + 5 * c + 2 * d;
        const R q2 = 2 * c + 5 * d - e;
        um = (w0 * q0 + w1 * q1 + w2 * q2) / ((R)6 * (w0 + w1 + w2));
    }
    // u^+ at i-1/2: mirrored stencils, weights (3,6,1)/10
    {
        const R w0 = (R)0.3 * p12, w1 = (R)0.6 * p02, w2 = (R)0.1 * p01;
        const R q0 = -a + 5 * b + 2 * c;
        const R q1 = 2 * b + 5 * c - d;
        const R q2 = 11 * c - 7 * d + 2 * e;
        up = (w0 * q0 + w1 * q1 + w2 * q2) / ((R)6 * (w0 + w1 + w2));
    }
}

// The same states in difference form, the evaluation the GPU kernels follow
// operation for operation in FP64.  With dm2 = b - a, dm1 = c - b,
// d0 = d - c, dp1 = e - d, tk the second differences and S the sum of
// (u^2 + 1e-30) over the stencil: 4 beta_k = (13/3) tk^2 + sk^2 with
// s0 = 3 dm1 - dm2, s1 = dm1 + d0, s2 = dp1 - 3 d0, and the weights use
// bk = 4 beta_k / S + 0.8e-6 = 0.8 (beta_k / m2 + 1e-6), evaluated as
// (4 beta_k + 0.8e-6 S) 2^-e with S = m 2^e, m in [0.5, 1): common factors
// of the three indicators, which the normalised weights do not see.  The
// candidates enter as offsets from c:
//   u- = c + ((p12/6) C0 + p02 C1 + (p01/2) C2) / (p12 + 6 p02 + 3 p01)
//   u+ = c + ((p12/2) E0 + p02 E1 + (p01/6) E2) / (3 p12 + 6 p02 + p01)
// with C = (2 s0 - dm1, dm1 + 2 d0, 4 d0 - dp1) and
// E = (dm2 - 4 dm1, -2 dm1 - d0, 2 s2 + d0), i.e. 6 (q_k - c) of the
// textbook candidates, and p_k the product of the other two bj^2; both
// states divide once, by the product of the two denominators.
template <typename R>
inline R w5_T(R t)
{
    return (((R)13 / (R)3) * t) * t;
}

template <typename R>
inline void weno5_cell(R a, R b, R c, R d, R e, R &um, R &up)
{
    const R sa = std::fma(a, a, (R)1e-30), sb = std::fma(b, b, (R)1e-30), sc = std::fma(c, c, (R)1e-30);
    const R sd = std::fma(d, d, (R)1e-30), se = std::fma(e, e, (R)1e-30);
    const R S = ((sa + sb) + (sc + sd)) + se;
    const R dm2 = b - a, dm1 = c - b, d0 = d - c, dp1 = e - d;
    const R Tm = w5_T(dm1 - dm2), T0 = w5_T(d0 - dm1), Tp = w5_T(dp1 - d0);
    // S = m 2^e, m in [0.5, 1): the indicators times the common factor m
    int ex = 0;
    const R m = std::frexp(S, &ex);
    const R r = std::ldexp((R)1, -ex);
    const R eps = (R)0.8e-6 * m;
    const R s0 = std::fma((R)3, dm1, -dm2);
    const R s1 = dm1 + d0;
    const R s2 = std::fma((R)-3, d0, dp1);
    const R b0 = std::fma(std::fma(s0, s0, Tm), r, eps);
    const R b1 = std::fma(std::fma(s1, s1, T0), r, eps);
    const R b2 = std::fma(std::fma(s2, s2, Tp), r, eps);
    const R g0 = b0 * b0, g1 = b1 * b1, g2 = b2 * b2;
    const R p12 = g1 * g2, p02 = g0 * g2, p01 = g0 * g1;
    const R C0 = std::fma((R)2, s0, -dm1);
    const R C1 = std::fma((R)2, d0, dm1);
    const R C2 = std::fma((R)4, d0, -dp1);
    const R nm = std::fma(p12 * ((R)1 / (R)6), C0, std::fma(p02, C1, ((R)0.5 * p01) * C2));
    const R dm = std::fma((R)6, p02, std::fma((R)3, p01, p12));
    const R E0 = std::fma((R)-4, dm1, dm2);
    const R E1 = std::fma((R)-2, dm1, -d0);
    const R E2 = std::fma((R)2, s2, d0);
    const R np = std::fma((R)0.5 * p12, E0, std::fma(p02, E1, (p01 * ((R)1 / (R)6)) * E2));
    const R dp = std::fma((R)6, p02, std::fma((R)3, p12, p01));
    // one quotient for both states
    const R q = (R)1 / (dm * dp);
    um = c + (nm * dp) * q;
    up = c + (np * dm) * q;
}

// in: field of one component (input type T); out: 6 slabs (xm,xp,ym,yp,zm,zp)
// in grid layout, computed in precision R and stored as R.
template <typename T, typename R>
void weno5_field(const orc_field &f, R *const *out, long long out_stride, int nthreads,
                 const long long *sel, int nsel)
{
    // sel == NULL: every block, output at its block slot; otherwise only the
    // listed blocks, the j-th one written at slot j
    const long long nbt = sel ? nsel : prod3(f.nblocks);
    const orc_stencil st = {{-3, -3, -3}, {4, 4, 4}, 0};
#pragma omp parallel num_threads(nthreads)
    {
        OLab<T> lab;
        lab.allocate(st, f.cells, (T)0);
#pragma omp for schedule(dynamic, 1)
        for (long long j = 0; j < nbt; ++j) {
            const long long b = sel ? sel[j] : j;
            int bidx[3];
            block_index(b, f.nblocks, bidx);
            lab.loadData(f, bidx, blockBCs<T>(f, b), true);
            const int *N = lab.N;
            for (int iz = 0; iz < N[2]; ++iz)
                for (int iy = 0; iy < N[1]; ++iy)
                    for (int ix = 0; ix < N[0]; ++ix) {
                        const size_t o = j * out_stride + ix + (size_t)N[0] * (iy + (size_t)N[1] * iz);
                        R um, up;
                        weno5_cell<R>(lab.at(ix - 2, iy, iz), lab.at(ix - 1, iy, iz), lab.at(ix, iy, iz),
                                      lab.at(ix + 1, iy, iz), lab.at(ix + 2, iy, iz), um, up);
                        out[0][o] = um;
                        out[1][o] = up;
                        weno5_cell<R>(lab.at(ix, iy - 2, iz), lab.at(ix, iy - 1, iz), lab.at(ix, iy, iz),
                                      lab.at(ix, iy + 1, iz), lab.at(ix, iy + 2, iz), um, up);
                        out[2][o] = um;
                        out[3][o] = up;
                        weno5_cell<R>(lab.at(ix, iy, iz - 2), lab.at(ix, iy, iz - 1), lab.at(ix, iy, iz),
                                      lab.at(ix, iy, iz + 1), lab.at(ix, iy, iz + 2), um, up);
                        out[4][o] = um;
                        out[5][o] = up;
                    }
        }
    }
}

// ---------------------------------------------------------------- Jacobi
// Variable-coefficient 7-point operator A u = ihx2 (u+x - 2u + u-x)
// + ihy2 (...) + czm[k] u-z - (czm[k] + czp[k]) u + czp[k] u+z, where k is the
// global z cell index; ghosts come from the lab (Dirichlet constant ghosts).
template <typename T>
inline T applyA(OLab<T> &lab, int ix, int iy, int iz, T ihx2, T ihy2, T czm, T czp)
{
    const T c = lab.at(ix, iy, iz);
    const T ax = ihx2 * (lab.at(ix + 1, iy, iz) - 2 * c + lab.at(ix - 1, iy, iz));
    const T ay = ihy2 * (lab.at(ix, iy + 1, iz) - 2 * c + lab.at(ix, iy - 1, iz));
    const T az = czp * (lab.at(ix, iy, iz + 1) - c) - czm * (c - lab.at(ix, iy, iz - 1));
    return ax + ay + az;
}

template <typename T>
double jacobi(const orc_field &u, const void *f_v, long long fstride, void *out_v, long long ostride,
              const double *czm, const double *czp, const double h[3], double omega, int mode,
              int nthreads)
{
    // mode 0: weighted Jacobi sweep out = u + (f - A u) * (omega / diag(A))
    // mode 1: defect out = f - A u; returns sum of squared defect (long double)
    const T ihx2 = (T)(1. / (h[0] * h[0]));
    const T ihy2 = (T)(1. / (h[1] * h[1]));
    const T w = (T)omega;
    const long long nbt = prod3(u.nblocks);
    const T *f = static_cast<const T *>(f_v);
    T *out = static_cast<T *>(out_v);
    const orc_stencil st = {{-1, -1, -1}, {2, 2, 2}, 0};
    std::vector<long double> partial((size_t)nbt, 0.0L);
#pragma omp parallel num_threads(nthreads)
    {
        OLab<T> lab;
        lab.allocate(st, u.cells, (T)0);
#pragma omp for schedule(dynamic, 1)
        for (long long b = 0; b < nbt; ++b) {
            int bidx[3];
            block_index(b, u.nblocks, bidx);
            lab.loadData(u, bidx, blockBCs<T>(u, b), true);
            const int *N = lab.N;
            long double acc = 0.0L;
            for (int iz = 0; iz < N[2]; ++iz) {
                const int kz = bidx[2] * N[2] + iz;
                const T cm = (T)czm[kz], cp = (T)czp[kz];
                const T diag = -(2 * ihx2 + 2 * ihy2 + cm + cp);
                const T wd = w / diag; // one division per z plane
                for (int iy = 0; iy < N[1]; ++iy)
                    for (int ix = 0; ix < N[0]; ++ix) {
                        const size_t li = ix + (size_t)N[0] * (iy + (size_t)N[1] * iz);
                        const T au = applyA(lab, ix, iy, iz, ihx2, ihy2, cm, cp);
                        const T r = f[b * fstride + li] - au;
                        if (mode == 0) {
                            out[b * ostride + li] = lab.at(ix, iy, iz) + r * wd;
                        } else {
                            out[b * ostride + li] = r;
                            acc += (long double)((double)r * (double)r);
                        }
                    }
            }
            partial[(size_t)b] = acc;
        }
    }
    long double tot = 0.0L;
    for (auto p : partial)
        tot += p;
    return (double)tot;
}

} // namespace

extern "C" {

int orc_version(void) { return 1; }

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Lab of one block in compact layout over the active lab range
// [sbegin, N + send - 1).  Cells the reference never writes keep `poison`.
// n_extern < 0 uses the block's own BC list; otherwise `extern_bcs`.
void orc_lab_load(const orc_field *f, const int *bidx, const orc_stencil *s, int apply_bc,
                  const orc_bc *extern_bcs, int n_extern, double poison, void *out)
{
    if (f->dtype == 1)
        lab_load<double>(*f, bidx, *s, apply_bc, extern_bcs, n_extern, poison, out);
    else
        lab_load<float>(*f, bidx, *s, apply_bc, extern_bcs, n_extern, poison, out);
}

// Laplacian into `out` (mode 0) or Euler step u + nu_dt*lap (mode 1) with
// the optional Laplacian copy in `lap_out`.  Returns sum of squares of the
// Euler result (mode 1), accumulated in long double.
double orc_diffusion(const orc_field *in, void *out, long long out_stride, void *lap_out,
                     const double *h, double nu_dt, int mode, int nthreads)
{
    if (in->dtype == 1)
        return diffusion<double>(*in, out, out_stride, lap_out, h, nu_dt, mode, nthreads);
    return diffusion<float>(*in, out, out_stride, lap_out, h, nu_dt, mode, nthreads);
}

// WENO5 states of one component in double arithmetic and outputs (input
// dtype from the field); prec must be 1 (the FP32 kernels are checked
// against this FP64 evaluation).  Returns 0, or -1 for another prec.
int orc_weno5(const orc_field *f, void *const *out6, long long out_stride, int prec, int nthreads,
              const long long *sel, int nsel)
{
    if (prec != 1)
        return -1;
    double *const *o = reinterpret_cast<double *const *>(out6);
    if (f->dtype == 1)
        weno5_field<double, double>(*f, o, out_stride, nthreads, sel, nsel);
    else
        weno5_field<float, double>(*f, o, out_stride, nthreads, sel, nsel);
    return 0;
}

// Scalar WENO5 of one 5-point stencil, double precision.
void orc_weno5_point(const double *v, double *um, double *up)
{
    weno5_cell<double>(v[0], v[1], v[2], v[3], v[4], *um, *up);
}

// The textbook statement of the same stencil (pins the difference form).
void orc_weno5_point_textbook(const double *v, double *um, double *up)
{
    weno5_cell_textbook<double>(v[0], v[1], v[2], v[3], v[4], *um, *up);
}

// Jacobi sweep (mode 0) or defect (mode 1, returns sum of squares).
double orc_jacobi(const orc_field *u, const void *f, long long fstride, void *out, long long ostride,
                  const double *czm, const double *czp, const double *h, double omega, int mode,
                  int nthreads)
{
    if (u->dtype == 1)
        return jacobi<double>(*u, f, fstride, out, ostride, czm, czp, h, omega, mode, nthreads);
    return jacobi<float>(*u, f, fstride, out, ostride, czm, czp, h, omega, mode, nthreads);
}

} // extern "C"
