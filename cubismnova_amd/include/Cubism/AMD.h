// Cubism/AMD.h -- C++ drop-in for CubismNova's stencil-path API on the
// MI355X library.
//
// Mirrors the names, arguments and error behaviour of
//   Core::Stencil            include/Cubism/Core/Stencil.h:20-108
//   BC::Base/Absorbing/...   include/Cubism/BC/{Base,Absorbing,Dirichlet,Symmetry}.h
//   Block::Field (per block) include/Cubism/Block/Field.h:57-594 (getBC, state)
//   Block::FieldLab          include/Cubism/Block/FieldLab.h:30-492
//   Grid::Cartesian          include/Cubism/Grid/Cartesian.h:44-550
// for cell-centred scalar components.  Storage is device-resident
// (cnv_grid); a host mirror in the reference byte layout backs the
// element accessors and is moved with upload()/download().  Every compute
// call goes through the C ABI of include/cubismnova_amd.h; a non-zero status
// becomes std::runtime_error, the reference's error convention.
#ifndef CUBISM_AMD_H
#define CUBISM_AMD_H

#include "../../../include/cubismnova_amd.h"

#include <array>
#include <cstddef>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace Cubism
{

using Index = int;
using MIndex = std::array<Index, 3>;

inline void check(int rc, const char *what)
{
    if (rc != CNV_OK)
        throw std::runtime_error(std::string(what) + ": " + cnv_last_error());
}

inline size_t flat(const MIndex &p, const MIndex &n) { return p[0] + (size_t)n[0] * (p[1] + (size_t)n[1] * p[2]); }

namespace Core
{
template <size_t DIM = 3>
class Stencil
{
public:
    Stencil() : begin_{0, 0, 0}, end_{1, 1, 1}, tensorial_(false) {}
    Stencil(const Index b, const Index e, const bool tensorial = false)
        : begin_{b, b, b}, end_{e, e, e}, tensorial_(tensorial)
    {
        check_();
    }
    Stencil(const MIndex &b, const MIndex &e, const bool tensorial = false) : begin_(b), end_(e), tensorial_(tensorial)
    {
        check_();
    }
    MIndex getBegin() const { return begin_; }
    MIndex getEnd() const { return end_; }
    bool isTensorial() const { return tensorial_; }
    cnv_stencil c() const
    {
        cnv_stencil s;
        for (int d = 0; d < 3; ++d) {
            s.begin[d] = begin_[d];
            s.end[d] = end_[d];
        }
        s.tensorial = tensorial_ ? 1 : 0;
        return s;
    }

private:
    MIndex begin_, end_;
    bool tensorial_;
    void check_() const
    {
        // all-component comparisons (Core/Vector.h:392-420)
        if (begin_[0] > 0 && begin_[1] > 0 && begin_[2] > 0)
            throw std::runtime_error("Stencil: begin_ must be <= 0");
        if (end_[0] < 1 && end_[1] < 1 && end_[2] < 1)
            throw std::runtime_error("Stencil: end_ must be > 0");
    }
};
} // namespace Core

namespace BC
{
struct BoundaryInfo {
    size_t dir;
    size_t side;
    bool is_periodic;
};

template <typename Lab>
class Base
{
public:
    Base(const size_t dir, const size_t side) : kind_(CNV_BC_BASE), value_(0)
    {
        binfo_.dir = dir;
        binfo_.side = side;
        binfo_.is_periodic = true;
    }
    virtual ~Base() {}
    const BoundaryInfo &getBoundaryInfo() const { return binfo_; }
    virtual std::string name() const { return std::string("Base"); }
    cnv_bc c() const { return cnv_bc{kind_, (int)binfo_.dir, (int)binfo_.side, 0, value_}; }

protected:
    BoundaryInfo binfo_;
    int kind_;
    double value_;
};

template <typename Lab>
class Absorbing : public Base<Lab>
{
public:
    Absorbing(const size_t dir, const size_t side) : Base<Lab>(dir, side)
    {
        this->binfo_.is_periodic = false;
        this->kind_ = CNV_BC_ABSORBING;
    }
    std::string name() const override { return std::string("Zeroth-Order Absorbing"); }
};

template <typename Lab>
class Dirichlet : public Base<Lab>
{
public:
    using DataType = typename Lab::DataType;
    Dirichlet(const size_t dir, const size_t side, const DataType value = 0) : Base<Lab>(dir, side)
    {
        this->binfo_.is_periodic = false;
        this->kind_ = CNV_BC_DIRICHLET;
        this->value_ = (double)value;
    }
    std::string name() const override { return std::string("Dirichlet"); }
    DataType getValue() const { return (DataType)this->value_; }
    void setValue(const DataType v) { this->value_ = (double)v; }
};

template <typename Lab>
class Symmetry : public Base<Lab>
{
public:
    using DataType = typename Lab::DataType;
    Symmetry(const size_t dir, const size_t side, const DataType sign = 1) : Base<Lab>(dir, side)
    {
        this->binfo_.is_periodic = false;
        this->kind_ = CNV_BC_SYMMETRY;
        this->value_ = (double)sign;
    }
    std::string name() const override { return std::string("Symmetry"); }
};
} // namespace BC

namespace Grid
{
template <typename T>
class Cartesian;
}

namespace Block
{
template <typename FieldT>
class FieldLab;

// View of one scalar block field of a Cartesian grid: host-mirror element
// access, block state and the block's BC list (non-owning, as in the
// reference, Block/Field.h:73-74,437-442).
template <typename T>
class Field
{
public:
    using DataType = T;
    using FieldType = Field;
    using BCVector = std::vector<BC::Base<FieldLab<Field>> *>;
    struct State {
        MIndex block_index;
    };

    Field(Grid::Cartesian<T> *g, size_t comp, MIndex bi, T *host) : grid_(g), comp_(comp), host_(host)
    {
        state_.block_index = bi;
    }
    T &operator[](const MIndex &p) { return host_[flat(p, extent())]; }
    const T &operator[](const MIndex &p) const { return host_[flat(p, extent())]; }
    T &operator()(Index x, Index y = 0, Index z = 0) { return (*this)[MIndex{x, y, z}]; }
    const State &getState() const { return state_; }
    BCVector &getBC() { return bc_; }
    const BCVector &getBC() const { return bc_; }
    MIndex extent() const;
    T *getData() { return host_; }
    size_t component() const { return comp_; }
    Grid::Cartesian<T> *grid() const { return grid_; }

private:
    Grid::Cartesian<T> *grid_;
    size_t comp_;
    T *host_;
    State state_;
    BCVector bc_;
};

// FieldLab: allocate(stencil, max range) then loadData; the lab is assembled
// on the GPU (cnv_lab_load) and copied to a host buffer over the active lab
// range [begin, N + end - 1) for operator() / operator[].
template <typename FieldT>
class FieldLab
{
public:
    using DataType = typename FieldT::DataType;
    using FieldType = FieldT;
    using StencilType = Core::Stencil<3>;
    using BCVector = typename FieldT::BCVector;

    FieldLab() = default;
    FieldLab(const FieldLab &) = delete;
    FieldLab &operator=(const FieldLab &) = delete;

    void allocate(const StencilType &s, const MIndex &max_extent, const bool force = false)
    {
        (void)force;
        stencil_ = s;
        max_ = max_extent;
        allocated_ = true;
    }
    bool isAllocated() const { return allocated_; }

    // FieldLab::loadData(fid, id2field, apply_bc, extern_bc) (Block/FieldLab.h:180-224)
    template <typename Functor>
    void loadData(const MIndex &fid, Functor &id2field, const bool apply_bc = true,
                  const BCVector *extern_bc = nullptr)
    {
        if (!allocated_)
            throw std::runtime_error("FieldLab: can not load lab data when not allocated first");
        field_ = &id2field(fid);
        auto *g = field_->grid();
        N_ = field_->extent();
        for (int d = 0; d < 3; ++d)
            if (N_[d] > max_[d] + 2)
                throw std::runtime_error("FieldLab: field larger than the allocated maximum range");
        g->upload(field_->component()); // the device copy must hold the host edits
        g->syncBCs(field_->component());
        const cnv_stencil cs = stencil_.c();
        int ext[3];
        check(cnv_lab_extent(g->handle(), &cs, ext), "cnv_lab_extent");
        for (int d = 0; d < 3; ++d)
            L_[d] = ext[d];
        const size_t n = (size_t)ext[0] * ext[1] * ext[2];
        data_.assign(n, DataType(0));
        void *dev = nullptr;
        check(cnv_malloc(&dev, n * sizeof(DataType)), "cnv_malloc");
        std::unique_ptr<void, int (*)(void *)> guard(dev, cnv_free);
        check(cnv_memcpy_h2d(dev, data_.data(), n * sizeof(DataType), nullptr), "cnv_memcpy_h2d");
        std::vector<cnv_bc> ext_bc;
        if (extern_bc)
            for (auto *b : *extern_bc)
                ext_bc.push_back(b->c());
        const int bi[3] = {fid[0], fid[1], fid[2]};
        check(cnv_lab_load(g->handle(), (int)field_->component(), bi, &cs, apply_bc ? 1 : 0,
                           extern_bc ? ext_bc.data() : nullptr, extern_bc ? (int)ext_bc.size() : -1, dev, nullptr),
              "cnv_lab_load");
        check(cnv_memcpy_d2h(data_.data(), dev, n * sizeof(DataType), nullptr), "cnv_memcpy_d2h");
    }
    template <typename Functor>
    void loadData(const MIndex &fid, Functor &id2field, const BCVector &boundaries, const bool apply_bc = true)
    {
        loadData(fid, id2field, apply_bc, &boundaries);
    }

    DataType &operator()(Index x, Index y = 0, Index z = 0) { return at(x, y, z); }
    const DataType &operator()(Index x, Index y = 0, Index z = 0) const { return at(x, y, z); }
    DataType &operator[](const MIndex &p) { return at(p[0], p[1], p[2]); }
    const DataType &operator[](const MIndex &p) const { return at(p[0], p[1], p[2]); }

    const StencilType &getActiveStencil() const { return stencil_; }
    MIndex getActiveRangeExtent() const { return N_; }
    FieldT &getActiveField()
    {
        if (!field_)
            throw std::runtime_error("FieldLab: no field loaded");
        return *field_;
    }
    DataType *getInnerData() { return &at(0, 0, 0); }

private:
    StencilType stencil_;
    MIndex max_{0, 0, 0}, N_{0, 0, 0}, L_{0, 0, 0};
    bool allocated_ = false;
    FieldT *field_ = nullptr;
    std::vector<DataType> data_;

    DataType &at(Index x, Index y, Index z)
    {
        const MIndex b = stencil_.getBegin();
        return data_[(x - b[0]) + (size_t)L_[0] * ((y - b[1]) + (size_t)L_[1] * (z - b[2]))];
    }
    const DataType &at(Index x, Index y, Index z) const { return const_cast<FieldLab *>(this)->at(x, y, z); }
};
} // namespace Block

namespace Grid
{
// Cartesian grid of cell-centred scalar components on the GPU.
template <typename T>
class Cartesian
{
public:
    using DataType = T;
    using BaseType = Block::Field<T>;
    using FieldLabType = Block::FieldLab<BaseType>;

    Cartesian(const MIndex &nblocks, const MIndex &block_cells, const std::array<double, 3> &begin = {0, 0, 0},
              const std::array<double, 3> &end = {1, 1, 1}, const size_t ncomp = 1)
        : nblocks_(nblocks), cells_(block_cells), ncomp_(ncomp)
    {
        cnv_grid_desc d{}; // entity: cell
        for (int k = 0; k < 3; ++k) {
            d.nblocks[k] = nblocks[k];
            d.block_cells[k] = block_cells[k];
            d.begin[k] = begin[k];
            d.end[k] = end[k];
        }
        d.ncomp = (int)ncomp;
        d.dtype = sizeof(T) == 8 ? CNV_F64 : CNV_F32;
        check(cnv_grid_create(&d, &g_), "cnv_grid_create");
        check(cnv_grid_info(g_, nullptr, &block_bytes_, &comp_bytes_), "cnv_grid_info");
        host_.assign(ncomp, std::vector<T>(comp_bytes_ / sizeof(T), T(0)));
        const size_t nbt = (size_t)nblocks[0] * nblocks[1] * nblocks[2];
        for (size_t c = 0; c < ncomp; ++c)
            for (size_t i = 0; i < nbt; ++i) {
                const MIndex bi{(Index)(i % nblocks[0]), (Index)((i / nblocks[0]) % nblocks[1]),
                                (Index)(i / ((size_t)nblocks[0] * nblocks[1]))};
                fields_.emplace_back(new BaseType(this, c, bi, host_[c].data() + i * (block_bytes_ / sizeof(T))));
            }
        for (int k = 0; k < 3; ++k)
            h_[k] = (end[k] - begin[k]) / (nblocks[k] * block_cells[k]);
    }
    Cartesian(const Cartesian &) = delete;
    Cartesian &operator=(const Cartesian &) = delete;
    ~Cartesian() { cnv_grid_destroy(g_); }

    size_t size() const { return (size_t)nblocks_[0] * nblocks_[1] * nblocks_[2]; }
    MIndex getSize() const { return nblocks_; }
    MIndex getBlockCells() const { return cells_; }
    std::array<double, 3> getCellSize() const { return h_; }
    BaseType &operator[](const size_t i) { return *fields_[i]; }
    BaseType &operator[](const MIndex &p) { return *fields_[flat(p, nblocks_)]; }
    BaseType &field(const size_t comp, const size_t i) { return *fields_[comp * size() + i]; }

    // Grid::Cartesian::getIndexFunctor (Cartesian.h:383-390): periodic block map
    std::function<BaseType &(const MIndex &)> getIndexFunctor(const size_t comp = 0)
    {
        return [this, comp](const MIndex &p) -> BaseType & {
            MIndex q;
            for (int d = 0; d < 3; ++d)
                q[d] = (p[d] + nblocks_[d]) % nblocks_[d];
            return field(comp, flat(q, nblocks_));
        };
    }
    // Grid::Cartesian::loadLab (Cartesian.h:411-427)
    void loadLab(const BaseType &f, FieldLabType &lab, const size_t comp = 0)
    {
        auto fn = getIndexFunctor(comp);
        lab.loadData(f.getState().block_index, fn);
    }

    void upload(const size_t comp)
    {
        check(cnv_grid_upload(g_, (int)comp, host_[comp].data(), comp_bytes_, nullptr), "cnv_grid_upload");
    }
    void download(const size_t comp)
    {
        check(cnv_grid_download(g_, (int)comp, host_[comp].data(), comp_bytes_, nullptr), "cnv_grid_download");
    }
    // push every block field's BC list of `comp` to the device grid
    void syncBCs(const size_t comp)
    {
        for (size_t i = 0; i < size(); ++i) {
            BaseType &f = field(comp, i);
            std::vector<cnv_bc> l;
            for (auto *b : f.getBC())
                l.push_back(b->c());
            const MIndex bi = f.getState().block_index;
            const int b3[3] = {bi[0], bi[1], bi[2]};
            check(cnv_grid_set_block_bcs(g_, (int)comp, b3, l.data(), (int)l.size()), "cnv_grid_set_block_bcs");
        }
    }
    cnv_grid handle() const { return g_; }
    MIndex blockExtent() const { return cells_; }
    size_t getNumComponents() const { return ncomp_; }

private:
    MIndex nblocks_, cells_;
    size_t ncomp_;
    cnv_grid g_ = nullptr;
    size_t block_bytes_ = 0, comp_bytes_ = 0;
    std::vector<std::vector<T>> host_;
    std::vector<std::unique_ptr<BaseType>> fields_;
    std::array<double, 3> h_;
};
} // namespace Grid

template <typename T>
MIndex Block::Field<T>::extent() const
{
    return grid_->blockExtent();
}

// ------------------------------------------------------------ device storage
// The Alloc seam of the reference's Grid::Cartesian (Cartesian.h:481,527,537),
// with the contract of Alloc/AlignedBlockAllocator.h:20-57: allocate()
// rounds `bytes` up to the alignment and returns the block memory, here in
// device memory.  A reference Cartesian<T, Mesh, Cell, 0, State,
// DeviceBlockAllocator> then keeps its blocks in HBM (its constructor does
// not touch the data, Cartesian.h:466-469); host code must not dereference
// them, and a DeviceGrid over the same allocation runs the operators in place.
template <typename T>
class DeviceBlockAllocator
{
public:
    using DataType = T;
    static constexpr size_t Alignment = 256;

    DataType *allocate(size_t &bytes) const
    {
        bytes = (bytes + Alignment - 1) / Alignment * Alignment;
        void *p = nullptr;
        check(cnv_malloc(&p, bytes), "cnv_malloc");
        return static_cast<DataType *>(p);
    }
    void deallocate(DataType *block) const
    {
        if (block != nullptr)
            cnv_free(block);
    }
};

// Non-owning device grid over an allocation in Grid::Cartesian's layout
// (component slabs of 32-byte rounded blocks, blocks x-fastest).
template <typename T>
class DeviceGrid
{
public:
    DeviceGrid(T *data, const size_t bytes, const MIndex &nblocks, const MIndex &block_cells, const size_t ncomp = 1,
               const std::array<double, 3> &begin = {0, 0, 0}, const std::array<double, 3> &end = {1, 1, 1},
               const int entity = CNV_CELL)
    {
        cnv_grid_desc d{};
        for (int k = 0; k < 3; ++k) {
            d.nblocks[k] = nblocks[k];
            d.block_cells[k] = block_cells[k];
            d.begin[k] = begin[k];
            d.end[k] = end[k];
        }
        d.ncomp = (int)ncomp;
        d.dtype = sizeof(T) == 8 ? CNV_F64 : CNV_F32;
        d.entity = entity;
        check(cnv_grid_wrap(&d, data, bytes, &g_), "cnv_grid_wrap");
    }
    DeviceGrid(const DeviceGrid &) = delete;
    DeviceGrid &operator=(const DeviceGrid &) = delete;
    ~DeviceGrid() { cnv_grid_destroy(g_); }
    cnv_grid handle() const { return g_; }

private:
    cnv_grid g_ = nullptr;
};

// ------------------------------------------------------------ device operators
// The per-block loops of the reference applications as single device calls.
namespace Device
{
// Laplacian.cpp:110-131: lap_u = 7-point Laplacian of u (Stencil(-1,2))
template <typename T>
void laplacian(Grid::Cartesian<T> &u, Grid::Cartesian<T> &lap_u, const size_t comp = 0)
{
    u.syncBCs(comp);
    check(cnv_laplacian(u.handle(), (int)comp, lap_u.handle(), (int)comp, nullptr), "cnv_laplacian");
}

// u_new = u + nu_dt * lap(u) for every component; returns sum of u_new^2 (FP64)
template <typename T>
double diffusionStep(Grid::Cartesian<T> &u, Grid::Cartesian<T> &u_new, const double nu_dt)
{
    double *dev = nullptr;
    check(cnv_malloc((void **)&dev, sizeof(double)), "cnv_malloc");
    std::unique_ptr<void, int (*)(void *)> guard(dev, cnv_free);
    check(cnv_diffusion_step(u.handle(), u_new.handle(), nu_dt, nullptr, dev, nullptr), "cnv_diffusion_step");
    double s = 0;
    check(cnv_memcpy_d2h(&s, dev, sizeof(double), nullptr), "cnv_memcpy_d2h");
    return s;
}

// nsteps Euler steps u -> tmp -> u ...; returns the sum of squares after each
// step; the field ends in u for even nsteps, in tmp for odd.
template <typename T>
std::vector<double> diffusionRun(Grid::Cartesian<T> &u, Grid::Cartesian<T> &tmp, const int nsteps, const double nu_dt)
{
    std::vector<double> norms((size_t)(nsteps > 0 ? nsteps : 0));
    if (norms.empty())
        return norms;
    for (size_t c = 0; c < u.getNumComponents(); ++c) {
        u.syncBCs(c);
        tmp.syncBCs(c);
    }
    double *dev = nullptr;
    check(cnv_malloc((void **)&dev, sizeof(double) * norms.size()), "cnv_malloc");
    std::unique_ptr<void, int (*)(void *)> guard(dev, cnv_free);
    check(cnv_diffusion_run(u.handle(), tmp.handle(), nsteps, nu_dt, dev, nullptr), "cnv_diffusion_run");
    check(cnv_memcpy_d2h(norms.data(), dev, sizeof(double) * norms.size(), nullptr), "cnv_memcpy_d2h");
    return norms;
}
// WENO5 face states of every component of `in` into `out` (6 per component:
// u- and u+ in x, y, z)
template <typename T>
void weno5(Grid::Cartesian<T> &in, Grid::Cartesian<T> &out)
{
    for (size_t c = 0; c < in.getNumComponents(); ++c)
        in.syncBCs(c);
    check(cnv_weno5(in.handle(), out.handle(), nullptr), "cnv_weno5");
}

// Field / FieldContainer arithmetic on whole grids (Block/Field.h:444-556,
// 1039-1154): dst = a OP b for every block (and every component when
// comp = -1); op one of CNV_FIELD_ADD, _SUB, _MUL, _DIV
template <typename T>
void fieldOp(const int op, Grid::Cartesian<T> &dst, const Grid::Cartesian<T> &a, const Grid::Cartesian<T> &b,
             const int comp = -1)
{
    check(cnv_grid_field_op(op, dst.handle(), comp, a.handle(), comp, b.handle(), comp, 0.0, nullptr),
          "cnv_grid_field_op");
}
// dst = a OP s; CNV_FIELD_RCP: dst = s * (1 / a); CNV_FIELD_NEG: dst = -a
template <typename T>
void fieldOp(const int op, Grid::Cartesian<T> &dst, const Grid::Cartesian<T> &a, const T s, const int comp = -1)
{
    check(cnv_grid_field_op(op, dst.handle(), comp, a.handle(), comp, nullptr, comp, (double)s, nullptr),
          "cnv_grid_field_op");
}

// No counterpart in the reference (it names MeshClass::Stretched,
// Common.h:86, but ships no stretched mesh or solver): the z faces of a
// stretched mesh, nz_total + 1 values, used by the Jacobi operator of g
template <typename T>
void setZFaces(Grid::Cartesian<T> &g, const std::vector<double> &zfaces)
{
    check(cnv_grid_set_zfaces(g.handle(), zfaces.data(), (int)zfaces.size()), "cnv_grid_set_zfaces");
}

struct SolveResult {
    int cycles = 0;
    long long sweeps = 0;
    double rel_residual = 0;
    std::vector<double> history; // relative residual before cycle 1 and after every cycle
};

// Weighted-Jacobi solve of A u = f (component 0 of u, in place) with u's BC
// lists.  opts.refine = 1: u and f FP64, work an FP32 grid of 3 components
// (FP32 sweeps, FP64 defect and correction); 0: work of u's dtype with 2
// components.
template <typename TU, typename TW>
SolveResult jacobiSolve(Grid::Cartesian<TU> &u, Grid::Cartesian<TU> &f, Grid::Cartesian<TW> &work,
                        const cnv_solve_opts &opts)
{
    u.syncBCs(0);
    SolveResult r;
    r.history.assign((size_t)(opts.max_cycles > 0 ? opts.max_cycles : 0) + 1, 0.0);
    cnv_solve_info info{};
    check(cnv_jacobi_solve(u.handle(), f.handle(), work.handle(), &opts, &info, r.history.data(),
                           (int)r.history.size(), nullptr),
          "cnv_jacobi_solve");
    r.cycles = info.cycles;
    r.sweeps = info.sweeps;
    r.rel_residual = info.rel_residual;
    r.history.resize((size_t)info.cycles + 1);
    return r;
}

// No counterpart in the reference: many independent FP32 grids, each with
// its operator (CNV_OP_DIFFUSION: out = in + nu_dt lap(in); CNV_OP_WENO5: 6
// states per component), processed together.  The grids must outlive the
// batch; BC lists are taken from the grids when the batch is built.
class Batch
{
public:
    struct Item {
        Grid::Cartesian<float> *in;
        Grid::Cartesian<float> *out;
        int op;
        int ghost;
        double nu_dt;
    };
    explicit Batch(const std::vector<Item> &items)
    {
        std::vector<cnv_batch_item> v;
        for (const Item &it : items) {
            for (size_t c = 0; c < it.in->getNumComponents(); ++c)
                it.in->syncBCs(c);
            v.push_back(cnv_batch_item{it.in->handle(), it.out->handle(), it.op, it.ghost, it.nu_dt});
        }
        check(cnv_batch_create(v.data(), (int)v.size(), &b_), "cnv_batch_create");
    }
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;
    ~Batch() { cnv_batch_destroy(b_); }
    void run(cnv_stream s = nullptr) { check(cnv_batch_run(b_, s), "cnv_batch_run"); }
    // times the launch choices on the batch's own grids and keeps the fastest
    void autotune(const int iters = 3)
    {
        int n = 0;
        check(cnv_batch_autotune(b_, iters, nullptr, 0, &n), "cnv_batch_autotune");
    }
    cnv_batch handle() const { return b_; }

private:
    cnv_batch b_ = nullptr;
};
} // namespace Device

// ------------------------------------------------------------ on-disk format
// IO::CartesianWriteHDF / CartesianReadHDF (IO/CartesianHDF.h:120-136,
// 203-210): the device grid's data (upload() it first) to <fname>.h5 +
// <fname>.xmf, and back (download() afterwards for the host mirror).
namespace IO
{
template <typename FileDataType, typename T>
void CartesianWriteHDF(const std::string &fname, const std::string &aname, const Grid::Cartesian<T> &grid,
                       const double time, const size_t face_dir = 0, const bool create_xdmf = true)
{
    static_assert(sizeof(FileDataType) == 4 || sizeof(FileDataType) == 8, "FP32 or FP64 files");
    check(cnv_grid_write_hdf(grid.handle(), fname.c_str(), aname.c_str(), sizeof(FileDataType) == 8 ? CNV_F64 : CNV_F32,
                             time, (int)face_dir, create_xdmf ? 1 : 0, nullptr, nullptr, nullptr),
          "CartesianWriteHDF");
}

template <typename FileDataType, typename T>
void CartesianReadHDF(const std::string &fname, Grid::Cartesian<T> &grid, const size_t face_dir = 0)
{
    static_assert(sizeof(FileDataType) == 4 || sizeof(FileDataType) == 8, "FP32 or FP64 files");
    check(cnv_grid_read_hdf(grid.handle(), fname.c_str(), sizeof(FileDataType) == 8 ? CNV_F64 : CNV_F32,
                            (int)face_dir, nullptr, nullptr, nullptr),
          "CartesianReadHDF");
}
} // namespace IO

} // namespace Cubism

#endif // CUBISM_AMD_H
