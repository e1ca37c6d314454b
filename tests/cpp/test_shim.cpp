// test_shim.cpp -- the reference's own unit tests for the stencil path,
// written against the C++ drop-in (Cubism/AMD.h) and run on the GPU.
// Mirrors test/UnitTests/Block/FieldLabTest.cpp, test/UnitTests/BC/*Test.cpp
// and test/UnitTests/Grid/CartesianTest.cpp:172-329, plus the Laplacian loop
// of test/sandbox/Benchmarks/Kernel/Laplacian/Laplacian.cpp:110-131 against
// the device operator.
#include "Cubism/AMD.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <thread>

using namespace Cubism;

static int g_fail = 0, g_checks = 0;
#define EXPECT(cond)                                                                                         \
    do {                                                                                                     \
        ++g_checks;                                                                                          \
        if (!(cond)) {                                                                                       \
            ++g_fail;                                                                                        \
            if (g_fail < 20)                                                                                 \
                std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                                \
        }                                                                                                    \
    } while (0)

using Grid64 = Grid::Cartesian<double>;
using Field64 = Grid64::BaseType;
using Lab64 = Grid64::FieldLabType;
using Stencil = Core::Stencil<3>;

// FieldLabTest.cpp:16-115 (3-D cells): iota field, self-periodic one-block grid
static void fieldlab_ghosts(bool tensorial)
{
    Grid64 grid({1, 1, 1}, {16, 16, 16});
    Field64 &f = grid[0];
    double k = 0;
    for (Index z = 0; z < 16; ++z)
        for (Index y = 0; y < 16; ++y)
            for (Index x = 0; x < 16; ++x)
                f(x, y, z) = k++;
    Lab64 lab;
    const Stencil s(-2, 3, tensorial);
    lab.allocate(s, {16, 16, 16});
    grid.loadLab(f, lab);
    for (Index z = 0; z < 16; ++z)
        for (Index y = 0; y < 16; ++y)
            for (Index x = 0; x < 16; ++x)
                EXPECT(lab(x, y, z) == f(x, y, z));
    const Index n = 16;
    for (Index z = -2; z < n + 2; ++z)
        for (Index y = -2; y < n + 2; ++y)
            for (Index x = -2; x < n + 2; ++x) {
                const MIndex q{x, y, z};
                int out = 0;
                for (int d = 0; d < 3; ++d)
                    out += (q[d] < 0 || q[d] >= n);
                if (out == 0 || (!tensorial && out > 1))
                    continue;
                MIndex p;
                for (int d = 0; d < 3; ++d)
                    p[d] = (q[d] + n) % n;
                EXPECT(lab[q] == lab[p]);
            }
}

// BC/AbsorbingTest.cpp, DirichletTest.cpp, SymmetryTest.cpp, BaseTest.cpp
static void bc_tests()
{
    Grid64 grid({1, 1, 1}, {16, 16, 16});
    Field64 &f = grid[0];
    double k = 0;
    for (Index z = 0; z < 16; ++z)
        for (Index y = 0; y < 16; ++y)
            for (Index x = 0; x < 16; ++x)
                f(x, y, z) = k++;
    auto fields = grid.getIndexFunctor();
    const Index n = 16;
    {
        Field64::BCVector bcv;
        for (size_t d = 0; d < 3; ++d)
            for (size_t side = 0; side < 2; ++side)
                bcv.push_back(new BC::Absorbing<Lab64>(d, side));
        EXPECT(bcv[0]->name() == "Zeroth-Order Absorbing");
        Lab64 lab;
        lab.allocate(Stencil(-3, 4), {n, n, n});
        lab.loadData(MIndex{0, 0, 0}, fields, bcv);
        for (int d = 0; d < 3; ++d)
            for (Index a = 0; a < n; ++a)
                for (Index b = 0; b < n; ++b)
                    for (Index l = 1; l <= 3; ++l) {
                        MIndex q0{a, b, 0}, q1{a, b, 0}, r0{a, b, 0}, r1{a, b, 0};
                        // put the slab index in dimension d and (a, b) in the others
                        const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
                        q0[e1] = q1[e1] = r0[e1] = r1[e1] = a;
                        q0[e2] = q1[e2] = r0[e2] = r1[e2] = b;
                        q0[d] = -l;
                        r0[d] = 0;
                        q1[d] = n - 1 + l;
                        r1[d] = n - 1;
                        EXPECT(lab[q0] == lab[r0]);
                        EXPECT(lab[q1] == lab[r1]);
                    }
        for (auto *bc : bcv)
            delete bc;
    }
    {
        Field64::BCVector bcv;
        for (size_t d = 0; d < 3; ++d)
            for (size_t side = 0; side < 2; ++side)
                bcv.push_back(new BC::Dirichlet<Lab64>(d, side, 10.0 + 2 * d + side));
        EXPECT(bcv[0]->name() == "Dirichlet");
        Lab64 lab;
        lab.allocate(Stencil(-3, 4), {n, n, n});
        lab.loadData(MIndex{0, 0, 0}, fields, bcv);
        for (int d = 0; d < 3; ++d)
            for (Index a = 0; a < n; ++a)
                for (Index l = 1; l <= 3; ++l) {
                    MIndex q0{a, a, a}, q1{a, a, a};
                    q0[d] = -l;
                    q1[d] = n - 1 + l;
                    EXPECT(lab[q0] == 10.0 + 2 * d);
                    EXPECT(lab[q1] == 11.0 + 2 * d);
                }
        for (auto *bc : bcv)
            delete bc;
    }
    {
        Field64::BCVector bcv;
        for (size_t d = 0; d < 3; ++d) {
            bcv.push_back(new BC::Symmetry<Lab64>(d, 0, 1.0));
            bcv.push_back(new BC::Symmetry<Lab64>(d, 1, -1.0));
        }
        EXPECT(bcv[0]->name() == "Symmetry");
        Lab64 lab;
        lab.allocate(Stencil(-3, 4), {n, n, n});
        lab.loadData(MIndex{0, 0, 0}, fields, bcv);
        for (int d = 0; d < 3; ++d)
            for (Index a = 0; a < n; ++a)
                for (Index l = 1; l <= 3; ++l) {
                    MIndex q0{a, a, a}, r0{a, a, a}, q1{a, a, a}, r1{a, a, a};
                    q0[d] = -l;
                    r0[d] = l - 1;
                    q1[d] = n - 1 + l;
                    r1[d] = n - l;
                    EXPECT(lab[q0] == 1.0 * lab[r0]);
                    EXPECT(lab[q1] == -1.0 * lab[r1]);
                }
        for (auto *bc : bcv)
            delete bc;
    }
    {
        Lab64 lab;
        bool threw = false;
        try {
            lab.getActiveField();
        } catch (const std::runtime_error &e) {
            threw = std::string(e.what()) == "FieldLab: no field loaded";
        }
        EXPECT(threw);
        threw = false;
        try {
            lab.loadData(MIndex{0, 0, 0}, fields);
        } catch (const std::runtime_error &e) {
            threw = std::string(e.what()) == "FieldLab: can not load lab data when not allocated first";
        }
        EXPECT(threw);
        threw = false;
        try {
            Stencil bad(1, 2);
        } catch (const std::runtime_error &e) {
            threw = std::string(e.what()) == "Stencil: begin_ must be <= 0";
        }
        EXPECT(threw);
    }
}

// FieldLabTest.cpp:117-152 (Reuse): a lab allocated for a larger stencil
// keeps its buffer for a smaller one and gets a new one for a much larger
// one; the buffer is 32-byte aligned every time
static void fieldlab_reuse()
{
    Lab64 lab;
    const MIndex range{16, 16, 16};
    lab.allocate(Stencil(-3, 4), range);
    const void *p0 = lab.getBlockPtr();
    const size_t b0 = lab.getBlockBytes();
    EXPECT(reinterpret_cast<size_t>(p0) % 32 == 0);
    lab.allocate(Stencil(-1, 2), range);
    const void *p1 = lab.getBlockPtr();
    const size_t b1 = lab.getBlockBytes();
    EXPECT(reinterpret_cast<size_t>(p1) % 32 == 0);
    EXPECT(p0 == p1);
    EXPECT(b0 == b1);
    lab.allocate(Stencil(-128, 129), range);
    const void *p2 = lab.getBlockPtr();
    const size_t b2 = lab.getBlockBytes();
    EXPECT(reinterpret_cast<size_t>(p2) % 32 == 0);
    EXPECT(p2 != p1);
    EXPECT(b2 != b1);
}

// FieldLabTest.cpp:157-221 (PeriodicIndexer), cell grids: block indices
// -3 .. 5 map to the block of index mod 3, for every component's functor
static void periodic_indexer()
{
    for (size_t ncomp : {size_t(1), size_t(3)}) {
        Grid64 grid({3, 3, 3}, {8, 8, 8}, Grid64::PointType(0), Grid64::PointType(1), ncomp);
        for (size_t c = 0; c < ncomp; ++c) {
            auto fields = grid.getIndexFunctor(c);
            for (Index z = -3; z < 6; ++z)
                for (Index y = -3; y < 6; ++y)
                    for (Index x = -3; x < 6; ++x) {
                        const MIndex q{x, y, z};
                        const MIndex s{(x + 3) % 3, (y + 3) % 3, (z + 3) % 3};
                        const Field64 &f = fields(q);
                        EXPECT(f.getState().block_index == s);
                        EXPECT(f.component() == c);
                    }
        }
    }
}

// CartesianTest.cpp:141-170 (BlockAccess) on a 2 x 2 x 1 grid: the periodic
// block functor at sums and differences of block indices
static void block_access()
{
    Grid64 grid({2, 2, 1}, {8, 8, 8});
    auto fields = grid.getIndexFunctor();
    const MIndex f00{0, 0, 0}, f01{0, 1, 0}, f10{1, 0, 0}, f11{1, 1, 0};
    EXPECT(fields(f00 - f10).getState().block_index == f10);
    EXPECT(fields(f10 - f10).getState().block_index == f00);
    EXPECT(fields(f00 + f10).getState().block_index == f10);
    EXPECT(fields(f10 + f10).getState().block_index == f00);
    EXPECT(fields(f10 - f01).getState().block_index == f11);
    EXPECT(fields(f11 - f01).getState().block_index == f10);
    EXPECT(fields(f11 + f01).getState().block_index == f10);
    EXPECT(fields(f10 + f01).getState().block_index == f11);
    EXPECT(fields(f11 + f11).getState().block_index == f00);
    EXPECT(fields(f00 - f11).getState().block_index == f11);
}

// Cartesian.h:132-143: the reference's constructor with the global domain
// (gbegin, gend); the spacing comes from [begin, end], and an integer fifth
// argument is the drop-in's component count
static void cartesian_constructors()
{
    using P = Grid64::PointType;
    Grid64 g({2, 2, 2}, {8, 8, 8}, P(0), P(2), P(-1), P(3));
    EXPECT(g.getMesh().getCellSize(0)[0] == 2.0 / 16);
    EXPECT(g.getGlobalBegin()[0] == -1.0 && g.getGlobalEnd()[2] == 3.0);
    EXPECT(g.getGlobalSize() == MIndex({2, 2, 2}));
    Grid64 c({2, 2, 2}, {8, 8, 8}, P(0), P(1), 3);
    EXPECT(c.getNumComponents() == 3);
    EXPECT(c.getGlobalBegin()[1] == 0.0 && c.getGlobalEnd()[1] == 1.0);
}

// CartesianTest.cpp:172-329 (3-D cells): every lab cell of every block equals
// the analytic field at its coordinate to 8 eps
static void cartesian_lab_analytic()
{
    Grid64 grid({3, 3, 3}, {8, 8, 8});
    const double h = 1.0 / 24;
    auto fexact = [](double x, double y, double z) {
        const double fac = 2 * M_PI;
        return std::sin(fac * x) * std::cos(fac * y) * std::sin(fac * z);
    };
    for (size_t i = 0; i < grid.size(); ++i) {
        Field64 &f = grid[i];
        const MIndex bi = f.getState().block_index;
        for (Index z = 0; z < 8; ++z)
            for (Index y = 0; y < 8; ++y)
                for (Index x = 0; x < 8; ++x)
                    f(x, y, z) = fexact((bi[0] * 8 + x + 0.5) * h, (bi[1] * 8 + y + 0.5) * h, (bi[2] * 8 + z + 0.5) * h);
    }
    Lab64 lab;
    lab.allocate(Stencil(-2, 3, true), {8, 8, 8});
    const double eps = std::numeric_limits<double>::epsilon();
    for (size_t i = 0; i < grid.size(); ++i) {
        Field64 &f = grid[i];
        grid.loadLab(f, lab);
        const MIndex bi = f.getState().block_index;
        for (Index z = -2; z < 10; ++z)
            for (Index y = -2; y < 10; ++y)
                for (Index x = -2; x < 10; ++x) {
                    const double v =
                        fexact((bi[0] * 8 + x + 0.5) * h, (bi[1] * 8 + y + 0.5) * h, (bi[2] * 8 + z + 0.5) * h);
                    EXPECT(std::fabs(lab(x, y, z) - v) <= 8 * eps);
                }
    }
    // CartesianTest.cpp:283-316: two threads, each with a lab of its own,
    // load the blocks of the same grid at the same time
    int bad[2] = {0, 0};
    auto work = [&](int t) {
        Lab64 tlab;
        tlab.allocate(Stencil(-2, 3, true), {8, 8, 8});
        for (size_t i = (size_t)t; i < grid.size(); i += 2) {
            Field64 &f = grid[i];
            grid.loadLab(f, tlab, Vector::X, Dir::X); // the reference's call form
            const MIndex bi = f.getState().block_index;
            for (Index z = -2; z < 10; ++z)
                for (Index y = -2; y < 10; ++y)
                    for (Index x = -2; x < 10; ++x) {
                        const double v = fexact((bi[0] * 8 + x + 0.5) * h, (bi[1] * 8 + y + 0.5) * h,
                                                (bi[2] * 8 + z + 0.5) * h);
                        bad[t] += !(std::fabs(tlab(x, y, z) - v) <= 8 * eps);
                    }
        }
    };
    std::thread t0(work, 0), t1(work, 1);
    t0.join();
    t1.join();
    EXPECT(bad[0] == 0);
    EXPECT(bad[1] == 0);
}

// Laplacian.cpp:110-131 on the host through the lab vs the device operator
static void laplacian_vs_host_loop()
{
    Grid64 u({2, 2, 2}, {16, 16, 16});
    Grid64 lap_u({2, 2, 2}, {16, 16, 16});
    for (size_t i = 0; i < u.size(); ++i) {
        Field64 &f = u[i];
        const MIndex bi = f.getState().block_index;
        for (Index z = 0; z < 16; ++z)
            for (Index y = 0; y < 16; ++y)
                for (Index x = 0; x < 16; ++x)
                    f(x, y, z) = std::sin(0.3 * (bi[0] * 16 + x)) * std::cos(0.2 * (bi[1] * 16 + y) + 0.1 * z) +
                                 1e-3 * ((x * 7 + y * 13 + z * 29) % 17);
    }
    u.upload(0);
    Device::laplacian(u, lap_u);
    lap_u.download(0);
    const std::array<double, 3> h = u.getCellSize();
    const double ihx2 = 1. / (h[0] * h[0]), ihy2 = 1. / (h[1] * h[1]), ihz2 = 1. / (h[2] * h[2]);
    Lab64 lab;
    lab.allocate(Stencil(-1, 2, false), {16, 16, 16});
    for (size_t i = 0; i < u.size(); ++i) {
        Field64 &f = u[i];
        u.loadLab(f, lab);
        Field64 &res = lap_u[f.getState().block_index];
        for (Index z = 0; z < 16; ++z)
            for (Index y = 0; y < 16; ++y)
                for (Index x = 0; x < 16; ++x) {
                    const double ddx = ihx2 * (lab(x + 1, y, z) - 2 * lab(x, y, z) + lab(x - 1, y, z));
                    const double ddy = ihy2 * (lab(x, y + 1, z) - 2 * lab(x, y, z) + lab(x, y - 1, z));
                    const double ddz = ihz2 * (lab(x, y, z + 1) - 2 * lab(x, y, z) + lab(x, y, z - 1));
                    EXPECT(res(x, y, z) == ddx + ddy + ddz);
                }
    }
    // the per-block loop moved no slab: the one explicit upload is all
    EXPECT(u.uploadCount() == 1);
    // a host edit is seen by the next lab load (one upload), and a device
    // result by the next host read (no explicit transfer)
    u[3](5, 6, 7) = 42.0;
    u.loadLab(u[3], lab);
    EXPECT(lab(5, 6, 7) == 42.0);
    EXPECT(u.uploadCount() == 2);
    u.loadLab(u[4], lab);
    EXPECT(u.uploadCount() == 2);
    Grid64 lap2({2, 2, 2}, {16, 16, 16});
    Device::laplacian(u, lap2);
    const Field64 &r3 = lap2[3];
    EXPECT(r3(5, 6, 7) == ihx2 * (u[3](6, 6, 7) - 2 * 42.0 + u[3](4, 6, 7)) +
                              ihy2 * (u[3](5, 7, 7) - 2 * 42.0 + u[3](5, 5, 7)) +
                              ihz2 * (u[3](5, 6, 8) - 2 * 42.0 + u[3](5, 6, 6)));
    // FieldLab::allocate keeps its buffers for a smaller request, takes new
    // ones when forced or when the request grows (FieldLab.h:147-158), and
    // reports the ranges of Block/FieldLab.h:403-456
    lab.allocate(Stencil(-1, 2, false), MIndex{8, 8, 8});
    EXPECT(lab.getMaximumRange().getExtent() == MIndex(8));
    lab.allocate(Stencil(-3, 4, false), MIndex{16, 16, 16}, true);
    u.loadLab(u[0], lab);
    EXPECT(lab.getActiveRange().getExtent() == MIndex(16));
    EXPECT(lab.getActiveLabRange().getBegin() == MIndex(-3) && lab.getActiveLabRange().getEnd() == MIndex(19));
    EXPECT(lab(-3, 0, 0) == u[1](13, 0, 0) && lab(18, 0, 0) == u[1](2, 0, 0));
}

// Device::diffusionRun (two steps per launch on this periodic grid) against
// repeated Device::diffusionStep: identical fields, norms to 1e-14
static void diffusion_run_vs_steps()
{
    using G = Grid::Cartesian<double>;
    G a({2, 2, 2}, {16, 16, 16}), b({2, 2, 2}, {16, 16, 16}), c({2, 2, 2}, {16, 16, 16}), d({2, 2, 2}, {16, 16, 16});
    for (size_t i = 0; i < a.size(); ++i)
        for (Index z = 0; z < 16; ++z)
            for (Index y = 0; y < 16; ++y)
                for (Index x = 0; x < 16; ++x)
                    a[i](x, y, z) = c[i](x, y, z) = std::sin(0.1 * (x + 3 * y + 7 * z) + (double)i);
    a.upload(0);
    c.upload(0);
    const double dt = 1e-6;
    const std::vector<double> n = Device::diffusionRun(a, b, 4, dt);
    std::vector<double> m;
    m.push_back(Device::diffusionStep(c, d, dt));
    m.push_back(Device::diffusionStep(d, c, dt));
    m.push_back(Device::diffusionStep(c, d, dt));
    m.push_back(Device::diffusionStep(d, c, dt));
    a.download(0);
    c.download(0);
    EXPECT(n.size() == 4);
    for (int k = 0; k < 4; ++k)
        EXPECT(std::fabs(n[(size_t)k] - m[(size_t)k]) <= 1e-14 * m[(size_t)k]);
    for (size_t i = 0; i < a.size(); ++i)
        for (Index z = 0; z < 16; ++z)
            for (Index y = 0; y < 16; ++y)
                for (Index x = 0; x < 16; ++x)
                    EXPECT(a[i](x, y, z) == c[i](x, y, z));
}

// The Alloc seam: a grid laid out in memory from DeviceBlockAllocator, wrapped
// without a copy, gives the same Laplacian as the library's own grid
static void device_allocator_wrap()
{
    Grid64 u({2, 2, 2}, {16, 16, 16});
    for (size_t i = 0; i < u.size(); ++i)
        for (Index z = 0; z < 16; ++z)
            for (Index y = 0; y < 16; ++y)
                for (Index x = 0; x < 16; ++x)
                    u[i](x, y, z) = std::cos(0.2 * (x + 5 * y) - 0.3 * z + (double)i);
    u.upload(0);
    DeviceBlockAllocator<double> alloc;
    size_t bytes = 8 * 16 * 16 * 16 * sizeof(double); // 8 blocks of 32 KiB, no padding
    double *p = alloc.allocate(bytes);
    EXPECT(bytes % DeviceBlockAllocator<double>::Alignment == 0);
    void *src = nullptr;
    check(cnv_grid_data(u.handle(), 0, &src), "cnv_grid_data");
    check(cnv_memcpy_d2d(p, src, 8 * 16 * 16 * 16 * sizeof(double), nullptr), "cnv_memcpy_d2d");
    Grid64 a({2, 2, 2}, {16, 16, 16}), b({2, 2, 2}, {16, 16, 16});
    {
        DeviceGrid<double> w(p, bytes, {2, 2, 2}, {16, 16, 16});
        check(cnv_laplacian(w.handle(), 0, a.handle(), 0, nullptr), "cnv_laplacian");
    }
    Device::laplacian(u, b);
    a.download(0);
    b.download(0);
    for (size_t i = 0; i < a.size(); ++i)
        for (Index z = 0; z < 16; ++z)
            for (Index y = 0; y < 16; ++y)
                for (Index x = 0; x < 16; ++x)
                    EXPECT(a[i](x, y, z) == b[i](x, y, z));
    // the wrapped memory outlives its grid
    double first = 0;
    check(cnv_memcpy_d2h(&first, p, sizeof(double), nullptr), "cnv_memcpy_d2h");
    EXPECT(first == u[0](0, 0, 0));
    bool threw = false;
    try {
        DeviceGrid<double> small(p, 1024, {2, 2, 2}, {16, 16, 16});
    } catch (const std::runtime_error &) {
        threw = true;
    }
    EXPECT(threw);
    alloc.deallocate(p);
}

// CartesianHDFTest.cpp:221-252 through the C++ mirror: write, read back,
// every block equal to its index
static void hdf_read_back()
{
    Grid64 src({3, 3, 3}, {8, 8, 8}), dst({3, 3, 3}, {8, 8, 8});
    for (size_t i = 0; i < src.size(); ++i)
        for (Index z = 0; z < 8; ++z)
            for (Index y = 0; y < 8; ++y)
                for (Index x = 0; x < 8; ++x)
                    src[i](x, y, z) = (double)i;
    src.upload(0);
    const std::string fn = "shim_framen";
    IO::CartesianWriteHDF<double>(fn, "ramen", src, 0.0);
    IO::CartesianReadHDF<double>(fn, dst);
    dst.download(0);
    for (size_t i = 0; i < dst.size(); ++i)
        for (Index z = 0; z < 8; ++z)
            for (Index y = 0; y < 8; ++y)
                for (Index x = 0; x < 8; ++x)
                    EXPECT(dst[i](x, y, z) == (double)i);
    std::remove((fn + ".h5").c_str());
    std::remove((fn + ".xmf").c_str());
    bool threw = false;
    try {
        IO::CartesianReadHDF<double>("shim_no_such_file", dst);
    } catch (const std::runtime_error &e) {
        threw = std::string(e.what()).find("does not exist") != std::string::npos;
    }
    EXPECT(threw);
}

// Device::fieldOp against the host Field arithmetic (Field.h:457-556)
static void device_field_ops()
{
    Grid64 a({2, 1, 2}, {8, 8, 8}, {0, 0, 0}, {1, 1, 1}, 2), b({2, 1, 2}, {8, 8, 8}, {0, 0, 0}, {1, 1, 1}, 2),
        d({2, 1, 2}, {8, 8, 8}, {0, 0, 0}, {1, 1, 1}, 2);
    for (size_t c = 0; c < 2; ++c)
        for (size_t i = 0; i < a.size(); ++i)
            for (Index z = 0; z < 8; ++z)
                for (Index y = 0; y < 8; ++y)
                    for (Index x = 0; x < 8; ++x) {
                        a.field(c, i)(x, y, z) = std::sin(0.3 * x + y - 0.7 * z + (double)(i + 3 * c));
                        b.field(c, i)(x, y, z) = 1.5 + std::cos(0.1 * x * y + (double)c);
                    }
    for (size_t c = 0; c < 2; ++c) {
        a.upload(c);
        b.upload(c);
    }
    Device::fieldOp(CNV_FIELD_DIV, d, a, b);
    for (size_t c = 0; c < 2; ++c)
        d.download(c);
    for (size_t c = 0; c < 2; ++c)
        for (size_t i = 0; i < a.size(); ++i)
            for (Index z = 0; z < 8; ++z)
                for (Index y = 0; y < 8; ++y)
                    for (Index x = 0; x < 8; ++x)
                        EXPECT(d.field(c, i)(x, y, z) == a.field(c, i)(x, y, z) / b.field(c, i)(x, y, z));
    Device::fieldOp(CNV_FIELD_RCP, d, b, 2.0, 1);
    d.download(1);
    for (size_t i = 0; i < a.size(); ++i)
        EXPECT(d.field(1, i)(3, 4, 5) == 2.0 * (1.0 / b.field(1, i)(3, 4, 5)));
}

// Device::jacobiSolve: the mixed-precision refined solve reaches 1e-12 and
// agrees with the all-FP64 solve; Dirichlet 0 on every face, tanh-stretched z
static void jacobi_refined_vs_fp64()
{
    using G32 = Grid::Cartesian<float>;
    const Index n = 16, N = 32;
    Grid64 u({2, 2, 2}, {n, n, n}), f({2, 2, 2}, {n, n, n}), u2({2, 2, 2}, {n, n, n});
    G32 w32({2, 2, 2}, {n, n, n}, {0, 0, 0}, {1, 1, 1}, 3);
    Grid64 w64({2, 2, 2}, {n, n, n}, {0, 0, 0}, {1, 1, 1}, 2);
    std::vector<double> zf((size_t)N + 1);
    for (Index k = 0; k <= N; ++k) // faces clustered at both ends
        zf[(size_t)k] = 0.5 * (1.0 + std::tanh(2.0 * (2.0 * k / N - 1.0)) / std::tanh(2.0));
    std::vector<Field64::BCVector *> lists;
    std::vector<BC::Dirichlet<Lab64> *> owned;
    for (Grid64 *g : {&u, &u2})
        for (size_t i = 0; i < g->size(); ++i) {
            Field64 &fl = (*g)[i];
            const MIndex bi = fl.getState().block_index;
            for (size_t d = 0; d < 3; ++d)
                for (size_t side = 0; side < 2; ++side)
                    if (bi[d] == (side ? 1 : 0)) {
                        owned.push_back(new BC::Dirichlet<Lab64>(d, side, 0.0));
                        fl.getBC().push_back(owned.back());
                    }
        }
    double mean = 0;
    for (size_t i = 0; i < f.size(); ++i)
        for (Index z = 0; z < n; ++z)
            for (Index y = 0; y < n; ++y)
                for (Index x = 0; x < n; ++x) {
                    const MIndex bi = f[i].getState().block_index;
                    const double gx = (bi[0] * n + x + 0.5) / N, gy = (bi[1] * n + y + 0.5) / N;
                    const double gz = (bi[2] * n + z + 0.5) / N;
                    f[i](x, y, z) = std::exp(-((gx - 0.4) * (gx - 0.4) + (gy - 0.6) * (gy - 0.6) +
                                               (gz - 0.5) * (gz - 0.5)) / 0.02);
                    mean += f[i](x, y, z);
                    u[i](x, y, z) = u2[i](x, y, z) = 0.0;
                }
    mean /= (double)(N * N * N);
    for (size_t i = 0; i < f.size(); ++i)
        for (Index z = 0; z < n; ++z)
            for (Index y = 0; y < n; ++y)
                for (Index x = 0; x < n; ++x)
                    f[i](x, y, z) -= mean;
    u.upload(0);
    u2.upload(0);
    f.upload(0);
    for (Grid64 *g : {&u, &f, &u2})
        Device::setZFaces(*g, zf);
    cnv_solve_opts o{1, 50, 4000, 2.0 / 3.0, 1e-12};
    const Device::SolveResult r = Device::jacobiSolve(u, f, w32, o);
    o.refine = 0;
    const Device::SolveResult r2 = Device::jacobiSolve(u2, f, w64, o);
    EXPECT(r.rel_residual <= 1e-12 && r2.rel_residual <= 1e-12);
    EXPECT(r.history.size() == (size_t)r.cycles + 1 && std::fabs(r.history[0] - 1.0) <= 1e-12);
    EXPECT(r.sweeps <= (long long)(1.2 * (double)r2.sweeps) + 100);
    u.download(0);
    u2.download(0);
    double scale = 0, diff = 0;
    for (size_t i = 0; i < u.size(); ++i)
        for (Index z = 0; z < n; ++z)
            for (Index y = 0; y < n; ++y)
                for (Index x = 0; x < n; ++x) {
                    scale = std::max(scale, std::fabs(u2[i](x, y, z)));
                    diff = std::max(diff, std::fabs(u[i](x, y, z) - u2[i](x, y, z)));
                }
    EXPECT(scale > 0 && diff <= 1e-9 * scale);
    for (auto *b : owned)
        delete b;
}

// Device::Batch: a diffusion grid and a WENO5 grid in one batch give their
// single-grid entry points' results (diffusion bit for bit, WENO5 states
// within 1e-6), before and after autotuning
static void batch_vs_single_grids()
{
    using G32 = Grid::Cartesian<float>;
    G32 a({2, 2, 1}, {16, 16, 16}), ao({2, 2, 1}, {16, 16, 16}), ar({2, 2, 1}, {16, 16, 16});
    G32 w({1, 2, 2}, {16, 16, 16}), wo({1, 2, 2}, {16, 16, 16}, {0, 0, 0}, {1, 1, 1}, 6),
        wr({1, 2, 2}, {16, 16, 16}, {0, 0, 0}, {1, 1, 1}, 6);
    for (size_t i = 0; i < a.size(); ++i)
        for (Index z = 0; z < 16; ++z)
            for (Index y = 0; y < 16; ++y)
                for (Index x = 0; x < 16; ++x) {
                    a[i](x, y, z) = (float)std::sin(0.2 * x + 0.3 * y - 0.1 * z + (double)i);
                    w[i](x, y, z) = (float)(1.0 + 0.5 * std::tanh(0.7 * (x - 8.0) + 0.1 * y * z + (double)i));
                }
    a.upload(0);
    w.upload(0);
    const double nudt = 1e-4;
    Device::diffusionStep(a, ar, nudt);
    Device::weno5(w, wr);
    ar.download(0);
    std::vector<std::vector<float>> wref;
    for (size_t c = 0; c < 6; ++c) {
        wr.download(c);
        std::vector<float> v;
        for (size_t i = 0; i < wr.size(); ++i)
            v.insert(v.end(), wr.field(c, i).getData(), wr.field(c, i).getData() + 16 * 16 * 16);
        wref.push_back(v);
    }
    Device::Batch b({{&a, &ao, CNV_OP_DIFFUSION, 1, nudt}, {&w, &wo, CNV_OP_WENO5, 2, 0.0}});
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1)
            b.autotune(2);
        b.run();
        ao.download(0);
        for (size_t i = 0; i < a.size(); ++i)
            for (Index z = 0; z < 16; ++z)
                for (Index y = 0; y < 16; ++y)
                    for (Index x = 0; x < 16; ++x)
                        EXPECT(ao[i](x, y, z) == ar[i](x, y, z));
        for (size_t c = 0; c < 6; ++c) {
            wo.download(c);
            size_t k = 0;
            for (size_t i = 0; i < wo.size(); ++i)
                for (size_t e = 0; e < 16 * 16 * 16; ++e, ++k)
                    EXPECT(std::fabs(wo.field(c, i).getData()[e] - wref[c][k]) <= 1e-6f * (1.0f + std::fabs(wref[c][k])));
        }
    }
}

// The grid, initial field and Laplacian of the reference's Laplacian
// benchmark (test/sandbox/Benchmarks/Kernel/Laplacian/Laplacian.cpp, which
// builds against this header unchanged) computed here with Device::laplacian
// and written as laplacian_device.h5 (FP32), the file the benchmark writes
// as laplacian.h5 from its per-block host loop over the labs.
static int laplacian_device(const int nb)
{
    using Mesh = Mesh::StructuredUniform<double, 3>;
    using ScalarGrid = Grid::Cartesian<double, Mesh, EntityType::Cell, 0>;
    ScalarGrid u(Mesh::MultiIndex(nb), Mesh::MultiIndex(32));
    ScalarGrid lap_u(Mesh::MultiIndex(nb), Mesh::MultiIndex(32));
    for (auto bf : u) {
        const Mesh &bm = *bf->getState().mesh;
        for (auto &ci : bm[EntityType::Cell]) {
            const Mesh::PointType x = bm.getCoordsCell(ci);
            (*bf)[ci] = std::sin(2.0 * M_PI * x[0]) * std::cos(6.0 * M_PI * x[1]) * std::sin(4.0 * M_PI * x[2]);
        }
    }
    Device::laplacian(u, lap_u);
    IO::CartesianWriteHDF<float>("laplacian_device", "div(grad(U))", lap_u, 0);
    std::printf("laplacian_device.h5 written (%d^3 blocks of 32^3)\n", nb);
    return 0;
}

int main(int argc, char *argv[])
{
    if (argc == 3 && std::string(argv[1]) == "laplacian-device")
        return laplacian_device(std::atoi(argv[2]));
    fieldlab_ghosts(false);
    fieldlab_ghosts(true);
    bc_tests();
    fieldlab_reuse();
    periodic_indexer();
    block_access();
    cartesian_constructors();
    cartesian_lab_analytic();
    laplacian_vs_host_loop();
    diffusion_run_vs_steps();
    device_allocator_wrap();
    hdf_read_back();
    device_field_ops();
    jacobi_refined_vs_fp64();
    batch_vs_single_grids();
    std::printf("%d checks, %d failed\n", g_checks, g_fail);
    return g_fail == 0 ? 0 : 1;
}
