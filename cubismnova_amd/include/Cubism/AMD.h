// Cubism/AMD.h -- C++ drop-in for CubismNova's stencil-path API on the
// MI355X library.
//
// Mirrors the names, template parameters, arguments and error behaviour of
//   Core::Vector / IndexRange  include/Cubism/Core/Vector.h, Core/Index.h:130-305
//   Core::Stencil              include/Cubism/Core/Stencil.h:20-108
//   Mesh::StructuredUniform    include/Cubism/Mesh/StructuredUniform.h:26-444
//   BC::Base/Absorbing/...     include/Cubism/BC/{Base,Absorbing,Dirichlet,Symmetry}.h
//   Block::Field (per block)   include/Cubism/Block/Field.h:57-594 (getBC, state, ranges)
//   Block::FieldLab            include/Cubism/Block/FieldLab.h:30-492
//   Grid::Cartesian            include/Cubism/Grid/Cartesian.h:44-550
// for cell-centred scalar components, so that application code written
// against those headers (the per-block loops of Laplacian.cpp:24-131)
// compiles against this one; the headers under Cubism/ with the reference's
// file names forward here.
//
// Storage is device-resident (cnv_grid).  A host mirror in the reference
// byte layout backs the element accessors.  Each component keeps two flags:
// host edits (any non-const element access) are uploaded before the next
// device use of the component (lab loads, operators, file writes), and the
// outputs of device operators are downloaded before the next host element
// access.  BC lists are pushed to the device only after getBC() was taken
// non-const, and only the blocks whose list changed.  Every compute call
// goes through the C ABI of include/cubismnova_amd.h; a non-zero status
// becomes std::runtime_error, the reference's error convention.
#ifndef CUBISM_AMD_H
#define CUBISM_AMD_H

#include "../../../include/cubismnova_amd.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <functional>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

namespace Cubism
{

using Index = int;

inline void check(int rc, const char *what)
{
    if (rc != CNV_OK)
        throw std::runtime_error(std::string(what) + ": " + cnv_last_error());
}

// Cubism::EntityType (include/Cubism/Common.h)
enum class EntityType { Cell = 0, Node, Face, Undefined };
// Cubism::Dir, Vector, Tensor (include/Cubism/Common.h, three dimensions):
// direction and tensor-component indices, cast to size_t where they are used
enum class Dir { X = 0, Y, Z };
using Vector = Dir;
enum class Tensor { XX = 0, XY, YX, YY, XZ, YZ, ZX, ZY, ZZ };

namespace Core
{
// Fixed-size vector with element-wise arithmetic and all-component
// comparisons (Core/Vector.h:60-200, 392-420, 698-718).
template <typename T, size_t DIM>
class Vector
{
public:
    using DataType = T;
    using ArrayType = std::array<T, DIM>;
    static constexpr size_t Dim = DIM;

    Vector() : a_{} {}
    template <typename U>
    explicit Vector(const U scalar)
    {
        a_.fill(static_cast<T>(scalar));
    }
    template <typename U>
    Vector(std::initializer_list<U> l) : a_{}
    {
        size_t i = 0;
        for (const U &v : l)
            if (i < DIM)
                a_[i++] = static_cast<T>(v);
    }
    template <typename U, size_t DIMU>
    Vector(const Vector<U, DIMU> &c) : a_{}
    {
        for (size_t i = 0; i < DIM && i < DIMU; ++i)
            a_[i] = static_cast<T>(c[i]);
    }
    template <typename U, size_t DIMU>
    Vector(const std::array<U, DIMU> &c) : a_{}
    {
        for (size_t i = 0; i < DIM && i < DIMU; ++i)
            a_[i] = static_cast<T>(c[i]);
    }
    static Vector getUnitVector(const size_t d)
    {
        Vector v;
        v[d] = 1;
        return v;
    }

    T &operator[](const size_t i) { return a_[i]; }
    const T &operator[](const size_t i) const { return a_[i]; }
    constexpr size_t size() const { return DIM; }
    T *data() { return a_.data(); }
    const T *data() const { return a_.data(); }
    T prod() const
    {
        T p = 1;
        for (size_t i = 0; i < DIM; ++i)
            p *= a_[i];
        return p;
    }
    T sum() const
    {
        T s = 0;
        for (size_t i = 0; i < DIM; ++i)
            s += a_[i];
        return s;
    }

#define CNV_VEC_OP(OP)                                                                                          \
    Vector &operator OP##=(const Vector & o)                                                                    \
    {                                                                                                           \
        for (size_t i = 0; i < DIM; ++i)                                                                        \
            a_[i] OP## = o.a_[i];                                                                               \
        return *this;                                                                                           \
    }                                                                                                           \
    Vector &operator OP##=(const T s)                                                                           \
    {                                                                                                           \
        for (size_t i = 0; i < DIM; ++i)                                                                        \
            a_[i] OP## = s;                                                                                     \
        return *this;                                                                                           \
    }                                                                                                           \
    friend Vector operator OP(Vector l, const Vector &r) { return l OP## = r; }                                \
    friend Vector operator OP(Vector l, const T s) { return l OP## = s; }
    CNV_VEC_OP(+)
    CNV_VEC_OP(-)
    CNV_VEC_OP(*)
    CNV_VEC_OP(/)
#undef CNV_VEC_OP
    friend Vector operator+(const T s, Vector r) { return r += s; }
    friend Vector operator*(const T s, Vector r) { return r *= s; }
    Vector operator-() const
    {
        Vector r(*this);
        for (size_t i = 0; i < DIM; ++i)
            r.a_[i] = -r.a_[i];
        return r;
    }

    // all-component comparisons (Core/Vector.h:392-420)
    bool operator==(const Vector &o) const { return a_ == o.a_; }
    bool operator!=(const Vector &o) const { return !(*this == o); }
    bool operator<(const Vector &o) const { return all_(o, [](T x, T y) { return x < y; }); }
    bool operator<=(const Vector &o) const { return all_(o, [](T x, T y) { return x <= y; }); }
    bool operator>(const Vector &o) const { return all_(o, [](T x, T y) { return x > y; }); }
    bool operator>=(const Vector &o) const { return all_(o, [](T x, T y) { return x >= y; }); }

private:
    ArrayType a_;
    template <typename F>
    bool all_(const Vector &o, F f) const
    {
        for (size_t i = 0; i < DIM; ++i)
            if (!f(a_[i], o.a_[i]))
                return false;
        return true;
    }
};

// Index range [begin, end) whose iteration yields local multi-indices,
// x fastest, from 0 to the extent (Core/Index.h:130-305, 372-467).
template <size_t DIM = 3>
class IndexRange
{
public:
    using MultiIndex = Vector<Index, DIM>;
    static constexpr size_t Dim = DIM;

    IndexRange() : b_(0), e_(0) {}
    explicit IndexRange(const Index e) : b_(0), e_(e) {}
    explicit IndexRange(const MultiIndex &e) : b_(0), e_(e) {}
    IndexRange(const MultiIndex &b, const MultiIndex &e) : b_(b), e_(e) {}

    MultiIndex getBegin() const { return b_; }
    MultiIndex getEnd() const { return e_; }
    MultiIndex getExtent() const { return e_ - b_; }
    size_t size() const
    {
        const MultiIndex x = getExtent();
        size_t n = 1;
        for (size_t d = 0; d < DIM; ++d)
            n *= (size_t)(x[d] > 0 ? x[d] : 0);
        return n;
    }
    size_t sizeDim(const size_t d) const { return (size_t)(e_[d] - b_[d]); }
    // local index p (0 <= p < extent)
    bool isIndex(const MultiIndex &p) const { return MultiIndex(0) <= p && p < getExtent(); }
    bool contains(const MultiIndex &p) const { return b_ <= p && p < e_; }
    size_t getFlatIndex(const MultiIndex &p) const
    {
        const MultiIndex x = getExtent();
        size_t f = 0;
        for (size_t d = DIM; d-- > 0;)
            f = f * (size_t)x[d] + (size_t)p[d];
        return f;
    }
    MultiIndex getMultiIndex(size_t i) const
    {
        const MultiIndex x = getExtent();
        MultiIndex p;
        for (size_t d = 0; d < DIM; ++d) {
            p[d] = (Index)(i % (size_t)x[d]);
            i /= (size_t)x[d];
        }
        return p;
    }

    class iterator
    {
    public:
        iterator(const MultiIndex &ext, size_t i, const MultiIndex &p) : ext_(ext), i_(i), p_(p) {}
        const MultiIndex &operator*() const { return p_; }
        const MultiIndex *operator->() const { return &p_; }
        iterator &operator++()
        {
            ++i_;
            for (size_t d = 0; d < DIM; ++d) {
                if (++p_[d] < ext_[d] || d == DIM - 1)
                    break;
                p_[d] = 0;
            }
            return *this;
        }
        bool operator==(const iterator &o) const { return i_ == o.i_; }
        bool operator!=(const iterator &o) const { return i_ != o.i_; }
        size_t getFlatIndex() const { return i_; }

    private:
        MultiIndex ext_;
        size_t i_;
        MultiIndex p_;
    };
    iterator begin() const { return iterator(getExtent(), 0, MultiIndex(0)); }
    iterator end() const { return iterator(getExtent(), size(), MultiIndex(0)); }

private:
    MultiIndex b_, e_;
};

template <size_t DIM = 3>
using MultiIndex = Vector<Index, DIM>;
} // namespace Core

using MIndex = Core::Vector<Index, 3>;

inline size_t flat(const MIndex &p, const MIndex &n) { return p[0] + (size_t)n[0] * (p[1] + (size_t)n[1] * p[2]); }

namespace Core
{
template <size_t DIM = 3>
class Stencil
{
public:
    using MultiIndex = Vector<Index, DIM>;
    Stencil() : begin_(0), end_(1), tensorial_(false) {}
    Stencil(const Index b, const Index e, const bool tensorial = false) : begin_(b), end_(e), tensorial_(tensorial)
    {
        check_();
    }
    Stencil(const MultiIndex &b, const MultiIndex &e, const bool tensorial = false)
        : begin_(b), end_(e), tensorial_(tensorial)
    {
        check_();
    }
    MultiIndex getBegin() const { return begin_; }
    MultiIndex getEnd() const { return end_; }
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
    MultiIndex begin_, end_;
    bool tensorial_;
    void check_() const
    {
        // all-component comparisons (Core/Vector.h:392-420, Stencil.h:99-107)
        if (begin_ > MultiIndex(0))
            throw std::runtime_error("Stencil: begin_ must be <= 0");
        if (end_ < MultiIndex(1))
            throw std::runtime_error("Stencil: end_ must be > 0");
    }
};
} // namespace Core

namespace Mesh
{
// Uniform Cartesian mesh over [begin, end) with `cells` cells
// (Mesh/StructuredUniform.h:26-444): cell size h = extent / cells (:61) and
// cell centre begin + (p + 1/2) h for local cell index p (:361-365).  The
// block meshes of a grid (BlockFieldAssembler.h:121-141) are meshes of
// their block's sub-range.
template <typename TReal, size_t DIM = 3>
class StructuredUniform
{
public:
    using RealType = TReal;
    using PointType = Core::Vector<TReal, DIM>;
    using MultiIndex = Core::Vector<Index, DIM>;
    using IndexRangeType = Core::IndexRange<DIM>;
    static constexpr size_t Dim = DIM;

    StructuredUniform(const PointType &begin, const PointType &end, const MultiIndex &cells)
        : begin_(begin), end_(end), cells_(cells), h_((end - begin) / PointType(cells))
    {
    }
    explicit StructuredUniform(const MultiIndex &cells) : StructuredUniform(PointType(0), PointType(1), cells) {}

    PointType getBegin() const { return begin_; }
    PointType getEnd() const { return end_; }
    PointType getExtent() const { return end_ - begin_; }
    PointType getCellSize(const size_t) const { return h_; }
    PointType getCellSize(const MultiIndex &) const { return h_; }
    RealType getCellVolume(const size_t = 0) const { return h_.prod(); }
    PointType getCoordsCell(const MultiIndex &p) const
    {
        const PointType c = PointType(p) + (TReal)0.5;
        return begin_ + c * h_;
    }
    PointType getCoordsNode(const MultiIndex &p) const { return begin_ + PointType(p) * h_; }
    IndexRangeType getIndexRange(const EntityType t = EntityType::Cell, const size_t = 0) const
    {
        if (t != EntityType::Cell)
            throw std::runtime_error("StructuredUniform: the drop-in mesh serves cell ranges only");
        return IndexRangeType(cells_);
    }
    // iterable local cell indices (StructuredBase.h:320-332)
    IndexRangeType operator[](const EntityType t) const { return getIndexRange(t); }
    size_t size(const EntityType t = EntityType::Cell) const { return getIndexRange(t).size(); }

private:
    PointType begin_, end_;
    MultiIndex cells_;
    PointType h_;
};
} // namespace Mesh

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

// The reference's default block allocator (Alloc/AlignedBlockAllocator.h:20-57):
// host memory, `bytes` rounded up to the alignment.  The drop-in grid takes
// the host mirror of its blocks from its Alloc parameter.
template <typename T>
class AlignedBlockAllocator
{
public:
    using DataType = T;
    static constexpr size_t Alignment = 32;
    DataType *allocate(size_t &bytes) const
    {
        bytes = (bytes + Alignment - 1) / Alignment * Alignment;
        void *p = std::aligned_alloc(Alignment, bytes);
        if (!p)
            throw std::runtime_error("AlignedBlockAllocator: allocation failed");
        return static_cast<DataType *>(p);
    }
    void deallocate(DataType *block) const { std::free(block); }
};

namespace Grid
{
// The device side of a grid as its block fields and labs see it (one object
// per Cartesian, whatever its template arguments).
class GridLink
{
public:
    virtual ~GridLink() {}
    virtual cnv_grid handle() const = 0;
    // device copy of component c current: host edits uploaded, BC lists pushed
    virtual void pushComp(const size_t c) = 0;
    // host mirror of component c current: device results downloaded
    virtual void pullComp(const size_t c) = 0;

    // per-component coherence state
    struct Sync {
        bool host_dirty = false;   // host mirror edited since the last upload
        bool device_newer = false; // a device operator wrote the component since the last download
        bool bc_dirty = true;      // a BC list may have changed since the last push
    };
    Sync &sync(const size_t c) { return sync_[c]; }

protected:
    std::vector<Sync> sync_;
};
} // namespace Grid

namespace Block
{
template <typename FieldT>
class FieldLab;

// Default user extension of a block's state (Block/Field.h:39)
struct FieldState {
};

// View of one scalar block field of a grid: element access into the host
// mirror, block state (block index, block mesh, user extension), the block's
// local index range and its BC list (non-owning, Block/Field.h:73-74,
// 437-442).
template <typename T, typename MeshT = Mesh::StructuredUniform<double, 3>, typename UserState = FieldState>
class Field
{
public:
    using DataType = T;
    using FieldType = Field;
    using MeshType = MeshT;
    using IndexRangeType = Core::IndexRange<3>;
    using MultiIndex = MIndex;
    using BCVector = std::vector<BC::Base<FieldLab<Field>> *>;
    struct FieldStateType {
        MultiIndex block_index;
        MeshT *mesh;
        UserState user;
    };
    using State = FieldStateType;

    Field(Grid::GridLink *g, size_t comp, const MIndex &bi, const MIndex &extent, T *host, MeshT *mesh)
        : grid_(g), comp_(comp), host_(host), extent_(extent)
    {
        state_.block_index = bi;
        state_.mesh = mesh;
    }
    Field(const Field &) = delete;
    Field &operator=(const Field &) = delete;

    // local multi-index access, x fastest (Block/Data.h:312-372)
    T &operator[](const MIndex &p) { return mutable_()[flat(p, extent_)]; }
    const T &operator[](const MIndex &p) const { return const_()[flat(p, extent_)]; }
    T &operator()(Index x, Index y = 0, Index z = 0) { return (*this)[MIndex{x, y, z}]; }
    const T &operator()(Index x, Index y = 0, Index z = 0) const { return (*this)[MIndex{x, y, z}]; }
    T *getData() { return mutable_(); }
    const T *getData() const { return const_(); }

    IndexRangeType getIndexRange(const size_t = 0) const { return IndexRangeType(extent_); }
    size_t getBlockSize() const { return getIndexRange().size(); }
    FieldStateType &getState() { return state_; }
    const FieldStateType &getState() const { return state_; }
    BCVector &getBC()
    {
        grid_->sync(comp_).bc_dirty = true;
        return bc_;
    }
    const BCVector &getBC() const { return bc_; }
    MIndex extent() const { return extent_; }
    size_t component() const { return comp_; }
    Grid::GridLink *grid() const { return grid_; }

private:
    Grid::GridLink *grid_;
    size_t comp_;
    T *host_;
    MIndex extent_;
    FieldStateType state_;
    BCVector bc_;

    T *mutable_()
    {
        Grid::GridLink::Sync &s = grid_->sync(comp_);
        if (s.device_newer)
            grid_->pullComp(comp_);
        s.host_dirty = true;
        return host_;
    }
    const T *const_() const
    {
        if (grid_->sync(comp_).device_newer)
            grid_->pullComp(comp_);
        return host_;
    }
};

// FieldLab: allocate(stencil, maximum range) once, then loadData per block;
// the lab is assembled on the GPU (cnv_lab_load) into a device buffer the
// lab keeps, and copied to a host buffer over the active lab range
// [begin, N + end - 1) for operator() / operator[].  Cells the reference
// leaves unwritten keep what the buffer held before, as in the reference's
// reused lab (Block/FieldLab.h:147-158, Block/Data.h:452-462).
template <typename FieldT>
class FieldLab
{
public:
    using DataType = typename FieldT::DataType;
    using FieldType = FieldT;
    using StencilType = Core::Stencil<3>;
    using IndexRangeType = Core::IndexRange<3>;
    using MultiIndex = MIndex;
    using BCVector = typename FieldT::BCVector;

    FieldLab() = default;
    FieldLab(const FieldLab &) = delete;
    FieldLab &operator=(const FieldLab &) = delete;
    ~FieldLab()
    {
        if (dev_)
            cnv_free(dev_);
    }

    // FieldLab::allocate (Block/FieldLab.h:104-165): the buffers are kept
    // when they are large enough for the new stencil and range, unless
    // `force` asks for new ones
    void allocate(const StencilType &s, const IndexRangeType &max_range, const bool force = false)
    {
        stencil_ = s;
        max_range_ = max_range;
        const MIndex ext = max_range.getExtent() + s.getEnd() - 1 - s.getBegin();
        const size_t n = (size_t)ext.prod();
        if (force || n > cap_) {
            if (dev_)
                check(cnv_free(dev_), "cnv_free");
            dev_ = nullptr;
            cap_ = 0;
            void *p = nullptr;
            check(cnv_malloc(&p, n * sizeof(DataType)), "cnv_malloc");
            dev_ = p;
            cap_ = n;
            // host buffer as Block::Data allocates it (Block/Data.h:89-521):
            // 32-byte aligned, its bytes rounded up to a multiple of 32
            size_t bytes = n * sizeof(DataType);
            data_.reset(AlignedBlockAllocator<DataType>().allocate(bytes));
            bytes_ = bytes;
            std::fill_n(data_.get(), bytes / sizeof(DataType), DataType(0));
            check(cnv_memcpy_h2d(dev_, data_.get(), n * sizeof(DataType), nullptr), "cnv_memcpy_h2d");
        }
        allocated_ = true;
    }
    void allocate(const StencilType &s, const MIndex &max_extent, const bool force = false)
    {
        allocate(s, IndexRangeType(max_extent), force);
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
        Grid::GridLink *g = field_->grid();
        N_ = field_->extent();
        if (!(N_ <= max_range_.getExtent()))
            throw std::runtime_error("FieldLab: field larger than the allocated maximum range");
        g->pushComp(field_->component());
        const cnv_stencil cs = stencil_.c();
        int ext[3];
        check(cnv_lab_extent(g->handle(), &cs, ext), "cnv_lab_extent");
        for (int d = 0; d < 3; ++d)
            L_[d] = ext[d];
        const size_t n = (size_t)ext[0] * ext[1] * ext[2];
        if (n > cap_)
            throw std::runtime_error("FieldLab: lab larger than its allocation");
        std::vector<cnv_bc> ext_bc;
        if (extern_bc)
            for (auto *b : *extern_bc)
                ext_bc.push_back(b->c());
        const int bi[3] = {fid[0], fid[1], fid[2]};
        check(cnv_lab_load(g->handle(), (int)field_->component(), bi, &cs, apply_bc ? 1 : 0,
                           extern_bc ? ext_bc.data() : nullptr, extern_bc ? (int)ext_bc.size() : -1, dev_, nullptr),
              "cnv_lab_load");
        check(cnv_memcpy_d2h(data_.get(), dev_, n * sizeof(DataType), nullptr), "cnv_memcpy_d2h");
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

    // accessors of Block/FieldLab.h:403-456
    const StencilType &getActiveStencil() const { return stencil_; }
    // index range of the loaded field: [0, N)
    IndexRangeType getActiveRange() const { return IndexRangeType(N_); }
    // lab range of the loaded field: [begin, N + end - 1)
    IndexRangeType getActiveLabRange() const
    {
        return IndexRangeType(stencil_.getBegin(), N_ + stencil_.getEnd() - 1);
    }
    // the maximum field range the lab was allocated for
    IndexRangeType getMaximumRange() const { return max_range_; }
    MIndex getActiveRangeExtent() const { return N_; }
    FieldT &getActiveField()
    {
        if (!field_)
            throw std::runtime_error("FieldLab: no field loaded");
        return *field_;
    }
    DataType *getInnerData() { return &at(0, 0, 0); }
    // Block::Data::getBlockPtr / getBlockBytes (Block/Data.h:220-225): the
    // lab's host buffer and its size; kept by a later allocate that fits
    void *getBlockPtr() { return data_.get(); }
    const void *getBlockPtr() const { return data_.get(); }
    size_t getBlockBytes() const { return bytes_; }

private:
    StencilType stencil_;
    IndexRangeType max_range_;
    MIndex N_{0, 0, 0}, L_{0, 0, 0};
    bool allocated_ = false;
    FieldT *field_ = nullptr;
    struct FreeBlock {
        void operator()(DataType *p) const { AlignedBlockAllocator<DataType>().deallocate(p); }
    };
    std::unique_ptr<DataType[], FreeBlock> data_;
    size_t bytes_ = 0;
    void *dev_ = nullptr;
    size_t cap_ = 0;

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
// Grid::Cartesian<T, Mesh, Entity, RANK, UserState, Alloc> (Cartesian.h:44-550):
// cell-centred scalar components (RANK 0) on the GPU, a uniform mesh, the
// host mirror of the blocks from Alloc.  `ncomp` (beyond the reference's
// arguments) gives several scalar components in one grid, as consecutive
// component slabs; iteration and operator[] visit the blocks of component 0.
template <typename T,
          typename Mesh = Mesh::StructuredUniform<double, 3>,
          EntityType Entity = EntityType::Cell,
          size_t RANK = 0,
          typename UserState = Block::FieldState,
          template <typename> class Alloc = AlignedBlockAllocator>
class Cartesian : public GridLink
{
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                  "the device grid holds float or double data");
    static_assert(Entity == EntityType::Cell, "the drop-in grid holds cell-centred data");
    static_assert(RANK == 0, "the drop-in grid holds scalar fields (several components: the ncomp argument)");
    static_assert(Mesh::Dim == 3, "three-dimensional meshes only");
    static_assert(std::is_same<Mesh, Cubism::Mesh::StructuredUniform<typename Mesh::RealType, 3>>::value,
                  "the drop-in grid takes a uniform mesh");

public:
    using MeshType = Mesh;
    using IndexRangeType = typename Mesh::IndexRangeType;
    using MultiIndex = typename Mesh::MultiIndex;
    using PointType = typename Mesh::PointType;
    using RealType = typename Mesh::RealType;
    using DataType = T;
    using BaseType = Block::Field<T, Mesh, UserState>;
    using FieldState = typename BaseType::FieldStateType;
    using FieldLabType = Block::FieldLab<BaseType>;
    using FieldContainer = std::vector<BaseType *>;
    using iterator = typename FieldContainer::iterator;
    using const_iterator = typename FieldContainer::const_iterator;
    static constexpr size_t Dim = 3;
    static constexpr size_t Rank = RANK;
    static constexpr EntityType EntityType_ = Entity;

    // the reference's main constructor (Cartesian.h:132-143): gbegin / gend
    // are the global domain, which differs from [begin, end] only for a
    // sub-domain of a distributed grid; the mesh spacing comes from
    // [begin, end] in both
    Cartesian(const MultiIndex &nblocks, const MultiIndex &block_cells, const PointType &begin = PointType(0),
              const PointType &end = PointType(1), const PointType &gbegin = PointType(0),
              const PointType &gend = PointType(1))
        : Cartesian(nblocks, block_cells, begin, end, size_t(1))
    {
        gbegin_ = gbegin;
        gend_ = gend;
    }
    // several scalar components in one grid (consecutive component slabs)
    Cartesian(const MultiIndex &nblocks, const MultiIndex &block_cells, const PointType &begin,
              const PointType &end, const size_t ncomp)
        : nblocks_(nblocks), cells_(block_cells), ncomp_(ncomp),
          mesh_(begin, end, MultiIndex(nblocks * block_cells)), gbegin_(begin), gend_(end)
    {
        cnv_grid_desc d{}; // entity: cell
        for (int k = 0; k < 3; ++k) {
            d.nblocks[k] = nblocks[k];
            d.block_cells[k] = block_cells[k];
            d.begin[k] = (double)begin[k];
            d.end[k] = (double)end[k];
        }
        d.ncomp = (int)ncomp;
        d.dtype = sizeof(T) == 8 ? CNV_F64 : CNV_F32;
        check(cnv_grid_create(&d, &g_), "cnv_grid_create");
        check(cnv_grid_info(g_, nullptr, &block_bytes_, &comp_bytes_), "cnv_grid_info");
        sync_.assign(ncomp, Sync());
        const size_t nbt = size();
        // block meshes: the sub-range of block bi (BlockFieldAssembler.h:121-141)
        const PointType bext = (end - begin) / PointType(nblocks);
        for (size_t i = 0; i < nbt; ++i) {
            const MultiIndex bi{(Index)(i % nblocks[0]), (Index)((i / nblocks[0]) % nblocks[1]),
                                (Index)(i / ((size_t)nblocks[0] * nblocks[1]))};
            const PointType bstart = begin + PointType(bi) * bext;
            block_meshes_.emplace_back(new Mesh(bstart, bstart + bext, block_cells));
        }
        for (size_t c = 0; c < ncomp; ++c) {
            size_t bytes = comp_bytes_;
            T *h = Alloc<T>().allocate(bytes);
            std::fill(h, h + comp_bytes_ / sizeof(T), T(0));
            host_.push_back(h);
            pushed_.emplace_back(nbt);
            for (size_t i = 0; i < nbt; ++i) {
                const MultiIndex bi{(Index)(i % nblocks[0]), (Index)((i / nblocks[0]) % nblocks[1]),
                                    (Index)(i / ((size_t)nblocks[0] * nblocks[1]))};
                fields_.emplace_back(
                    new BaseType(this, c, bi, block_cells, h + i * (block_bytes_ / sizeof(T)), block_meshes_[i].get()));
            }
        }
        for (size_t i = 0; i < nbt; ++i)
            view_.push_back(fields_[i].get());
        for (int k = 0; k < 3; ++k)
            h_[k] = (double)mesh_.getCellSize(0)[k];
        // the mirror starts as zeros: the device copy gets them at first use
        for (auto &s : sync_)
            s.host_dirty = true;
    }
    Cartesian(const Cartesian &) = delete;
    Cartesian &operator=(const Cartesian &) = delete;
    ~Cartesian() override
    {
        cnv_grid_destroy(g_);
        for (T *h : host_)
            Alloc<T>().deallocate(h);
    }

    // block field iteration (Cartesian.h:185-228): the elements are pointers
    iterator begin() noexcept { return view_.begin(); }
    iterator end() noexcept { return view_.end(); }
    const_iterator begin() const noexcept { return view_.begin(); }
    const_iterator end() const noexcept { return view_.end(); }

    size_t size() const { return (size_t)nblocks_[0] * nblocks_[1] * nblocks_[2]; }
    MultiIndex getSize() const { return nblocks_; }
    // Cartesian::getGlobalSize (Cartesian.h:397): the blocks of the whole grid
    MultiIndex getGlobalSize() const { return nblocks_; }
    // the global domain given at construction ([begin, end] by default)
    const PointType &getGlobalBegin() const { return gbegin_; }
    const PointType &getGlobalEnd() const { return gend_; }
    MultiIndex getBlockCells() const { return cells_; }
    IndexRangeType getBlockRange() const { return IndexRangeType(nblocks_); }
    const MeshType &getMesh() const { return mesh_; }
    const MeshType &getGlobalMesh() const { return mesh_; }
    std::array<double, 3> getCellSize() const { return h_; }
    FieldContainer &getFields() { return view_; }
    BaseType &operator[](const size_t i) { return *fields_[i]; }
    const BaseType &operator[](const size_t i) const { return *fields_[i]; }
    BaseType &operator[](const MultiIndex &p) { return *fields_[flat(p, nblocks_)]; }
    const BaseType &operator[](const MultiIndex &p) const { return *fields_[flat(p, nblocks_)]; }
    BaseType &field(const size_t comp, const size_t i) { return *fields_[comp * size() + i]; }

    // Grid::Cartesian::getIndexFunctor (Cartesian.h:383-390): periodic block
    // map of component c (any type cast to size_t, e.g. Vector::X); the face
    // direction d has no meaning for cell data and is not used
    template <typename Comp = size_t, typename DirT = size_t>
    std::function<BaseType &(const MultiIndex &)> getIndexFunctor(const Comp c = 0, const DirT d = 0)
    {
        (void)d;
        const size_t comp = static_cast<size_t>(c);
        return [this, comp](const MultiIndex &p) -> BaseType & {
            MultiIndex q;
            for (int d = 0; d < 3; ++d)
                q[d] = (p[d] % nblocks_[d] + nblocks_[d]) % nblocks_[d];
            return field(comp, flat(q, nblocks_));
        };
    }
    // Grid::Cartesian::loadLab (Cartesian.h:411-427), component c and face
    // direction d as in getIndexFunctor
    template <typename Comp = size_t, typename DirT = size_t>
    void loadLab(const BaseType &f, FieldLabType &lab, const Comp c = 0, const DirT d = 0)
    {
        if (!lab.isAllocated())
            throw std::runtime_error("FieldLab: can not load lab data when not allocated first");
        auto fn = getIndexFunctor(c, d);
        lab.loadData(f.getState().block_index, fn);
    }

    // explicit transfers of one component's slab (the flags make them
    // optional: device use uploads host edits, host access downloads results)
    void upload(const size_t comp)
    {
        check(cnv_grid_upload(g_, (int)comp, host_[comp], comp_bytes_, nullptr), "cnv_grid_upload");
        sync_[comp].host_dirty = false;
        sync_[comp].device_newer = false;
        ++uploads_;
    }
    // component slabs uploaded so far (explicitly or before a device use)
    size_t uploadCount() const { return uploads_; }
    void download(const size_t comp)
    {
        check(cnv_grid_download(g_, (int)comp, host_[comp], comp_bytes_, nullptr), "cnv_grid_download");
        sync_[comp].device_newer = false;
        sync_[comp].host_dirty = false;
    }
    // push the BC lists of component comp whose contents changed since the
    // last push (Block::Field::getBC, Block/Field.h:437-442)
    void syncBCs(const size_t comp)
    {
        if (!sync_[comp].bc_dirty)
            return;
        for (size_t i = 0; i < size(); ++i) {
            const BaseType &f = field(comp, i);
            std::vector<cnv_bc> l;
            for (auto *b : f.getBC())
                l.push_back(b->c());
            std::vector<cnv_bc> &p = pushed_[comp][i];
            const bool same = p.size() == l.size() &&
                              std::equal(l.begin(), l.end(), p.begin(), [](const cnv_bc &a, const cnv_bc &b) {
                                  return a.kind == b.kind && a.dir == b.dir && a.side == b.side && a.value == b.value;
                              });
            if (same)
                continue;
            const MultiIndex bi = f.getState().block_index;
            const int b3[3] = {bi[0], bi[1], bi[2]};
            check(cnv_grid_set_block_bcs(g_, (int)comp, b3, l.data(), (int)l.size()), "cnv_grid_set_block_bcs");
            p = l;
        }
        sync_[comp].bc_dirty = false;
    }

    // GridLink
    cnv_grid handle() const override { return g_; }
    // serialised: threads with labs of their own may load them from one grid
    // at once (CartesianTest.cpp:283-316)
    void pushComp(const size_t c) override
    {
        std::lock_guard<std::mutex> lk(push_mtx_);
        if (sync_[c].host_dirty)
            upload(c);
        syncBCs(c);
    }
    void pullComp(const size_t c) override
    {
        std::lock_guard<std::mutex> lk(push_mtx_);
        if (sync_[c].device_newer)
            download(c);
    }
    // every component's device copy current (before a device operator reads the grid)
    void pushAll()
    {
        for (size_t c = 0; c < ncomp_; ++c)
            pushComp(c);
    }
    // after a device operator wrote the grid: host reads download first
    void deviceWrote()
    {
        for (size_t c = 0; c < ncomp_; ++c) {
            sync_[c].device_newer = true;
            sync_[c].host_dirty = false;
        }
    }

    MIndex blockExtent() const { return cells_; }
    size_t getNumComponents() const { return ncomp_; }

private:
    MultiIndex nblocks_, cells_;
    size_t ncomp_;
    MeshType mesh_;
    PointType gbegin_, gend_;
    cnv_grid g_ = nullptr;
    size_t block_bytes_ = 0, comp_bytes_ = 0;
    std::vector<T *> host_;
    std::vector<std::unique_ptr<MeshType>> block_meshes_;
    std::vector<std::unique_ptr<BaseType>> fields_;
    FieldContainer view_;
    std::vector<std::vector<std::vector<cnv_bc>>> pushed_; // BC lists on the device, per component and block
    std::array<double, 3> h_;
    size_t uploads_ = 0;
    std::mutex push_mtx_;
};
} // namespace Grid

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
    void pushAll() {}
    void deviceWrote() {}

private:
    cnv_grid g_ = nullptr;
};

// ------------------------------------------------------------ device operators
// The per-block loops of the reference applications as single device calls.
// Inputs are made current on the device first (host edits uploaded, BC lists
// pushed); outputs are marked so that host element access downloads them.
namespace Device
{
// Laplacian.cpp:110-131: lap_u = 7-point Laplacian of u (Stencil(-1,2))
template <typename GU, typename GL>
void laplacian(GU &u, GL &lap_u, const size_t comp = 0)
{
    u.pushComp(comp);
    check(cnv_laplacian(u.handle(), (int)comp, lap_u.handle(), (int)comp, nullptr), "cnv_laplacian");
    lap_u.deviceWrote();
}

// u_new = u + nu_dt * lap(u) for every component; returns sum of u_new^2 (FP64)
template <typename G>
double diffusionStep(G &u, G &u_new, const double nu_dt)
{
    u.pushAll();
    double *dev = nullptr;
    check(cnv_malloc((void **)&dev, sizeof(double)), "cnv_malloc");
    std::unique_ptr<void, int (*)(void *)> guard(dev, cnv_free);
    check(cnv_diffusion_step(u.handle(), u_new.handle(), nu_dt, nullptr, dev, nullptr), "cnv_diffusion_step");
    u_new.deviceWrote();
    double s = 0;
    check(cnv_memcpy_d2h(&s, dev, sizeof(double), nullptr), "cnv_memcpy_d2h");
    return s;
}

// nsteps Euler steps u -> tmp -> u ...; returns the sum of squares after each
// step; the field ends in u for even nsteps, in tmp for odd.
template <typename G>
std::vector<double> diffusionRun(G &u, G &tmp, const int nsteps, const double nu_dt)
{
    std::vector<double> norms((size_t)(nsteps > 0 ? nsteps : 0));
    if (norms.empty())
        return norms;
    u.pushAll();
    for (size_t c = 0; c < u.getNumComponents(); ++c)
        tmp.syncBCs(c);
    double *dev = nullptr;
    check(cnv_malloc((void **)&dev, sizeof(double) * norms.size()), "cnv_malloc");
    std::unique_ptr<void, int (*)(void *)> guard(dev, cnv_free);
    check(cnv_diffusion_run(u.handle(), tmp.handle(), nsteps, nu_dt, dev, nullptr), "cnv_diffusion_run");
    u.deviceWrote();
    tmp.deviceWrote();
    check(cnv_memcpy_d2h(norms.data(), dev, sizeof(double) * norms.size(), nullptr), "cnv_memcpy_d2h");
    return norms;
}
// WENO5 face states of every component of `in` into `out` (6 per component:
// u- and u+ in x, y, z)
template <typename GI, typename GO>
void weno5(GI &in, GO &out)
{
    in.pushAll();
    check(cnv_weno5(in.handle(), out.handle(), nullptr), "cnv_weno5");
    out.deviceWrote();
}

// Field / FieldContainer arithmetic on whole grids (Block/Field.h:444-556,
// 1039-1154): dst = a OP b for every block (and every component when
// comp = -1); op one of CNV_FIELD_ADD, _SUB, _MUL, _DIV
template <typename G>
void fieldOp(const int op, G &dst, G &a, G &b, const int comp = -1)
{
    a.pushAll();
    b.pushAll();
    dst.pushAll(); // components the call does not write stay current
    check(cnv_grid_field_op(op, dst.handle(), comp, a.handle(), comp, b.handle(), comp, 0.0, nullptr),
          "cnv_grid_field_op");
    dst.deviceWrote();
}
// dst = a OP s; CNV_FIELD_RCP: dst = s * (1 / a); CNV_FIELD_NEG: dst = -a
template <typename G, typename T, typename = typename std::enable_if<std::is_arithmetic<T>::value>::type>
void fieldOp(const int op, G &dst, G &a, const T s, const int comp = -1)
{
    a.pushAll();
    dst.pushAll();
    check(cnv_grid_field_op(op, dst.handle(), comp, a.handle(), comp, nullptr, comp, (double)s, nullptr),
          "cnv_grid_field_op");
    dst.deviceWrote();
}

// No counterpart in the reference (it names MeshClass::Stretched,
// Common.h:86, but ships no stretched mesh or solver): the z faces of a
// stretched mesh, nz_total + 1 values, used by the Jacobi operator of g
template <typename G>
void setZFaces(G &g, const std::vector<double> &zfaces)
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
template <typename GU, typename GW>
SolveResult jacobiSolve(GU &u, GU &f, GW &work, const cnv_solve_opts &opts)
{
    u.pushAll();
    f.pushAll();
    SolveResult r;
    r.history.assign((size_t)(opts.max_cycles > 0 ? opts.max_cycles : 0) + 1, 0.0);
    cnv_solve_info info{};
    check(cnv_jacobi_solve(u.handle(), f.handle(), work.handle(), &opts, &info, r.history.data(),
                           (int)r.history.size(), nullptr),
          "cnv_jacobi_solve");
    u.deviceWrote();
    work.deviceWrote();
    r.cycles = info.cycles;
    r.sweeps = info.sweeps;
    r.rel_residual = info.rel_residual;
    r.history.resize((size_t)info.cycles + 1);
    return r;
}

// No counterpart in the reference: many independent FP32 grids, each with
// its operator (CNV_OP_DIFFUSION: out = in + nu_dt lap(in); CNV_OP_WENO5: 6
// states per component), processed together.  The grids must outlive the
// batch; BC lists are taken from the grids when the batch is built, and the
// inputs' host edits are uploaded at every run.
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
    explicit Batch(const std::vector<Item> &items) : items_(items)
    {
        std::vector<cnv_batch_item> v;
        for (const Item &it : items) {
            it.in->pushAll();
            v.push_back(cnv_batch_item{it.in->handle(), it.out->handle(), it.op, it.ghost, it.nu_dt});
        }
        check(cnv_batch_create(v.data(), (int)v.size(), &b_), "cnv_batch_create");
    }
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;
    ~Batch() { cnv_batch_destroy(b_); }
    void run(cnv_stream s = nullptr)
    {
        for (const Item &it : items_)
            it.in->pushAll();
        check(cnv_batch_run(b_, s), "cnv_batch_run");
        for (const Item &it : items_)
            it.out->deviceWrote();
    }
    // times the launch choices on the batch's own grids and keeps the fastest
    void autotune(const int iters = 3)
    {
        int n = 0;
        check(cnv_batch_autotune(b_, iters, nullptr, 0, &n), "cnv_batch_autotune");
        for (const Item &it : items_)
            it.out->deviceWrote();
    }
    cnv_batch handle() const { return b_; }

private:
    std::vector<Item> items_;
    cnv_batch b_ = nullptr;
};
} // namespace Device

// ------------------------------------------------------------ on-disk format
// IO::CartesianWriteHDF / CartesianReadHDF (IO/CartesianHDF.h:120-136,
// 203-210): the grid's data (host edits uploaded first) to <fname>.h5 +
// <fname>.xmf, and back (host access downloads it).
namespace IO
{
template <typename FileDataType, typename G>
void CartesianWriteHDF(const std::string &fname, const std::string &aname, const G &grid, const double time,
                       const size_t face_dir = 0, const bool create_xdmf = true)
{
    static_assert(sizeof(FileDataType) == 4 || sizeof(FileDataType) == 8, "FP32 or FP64 files");
    const_cast<G &>(grid).pushAll();
    check(cnv_grid_write_hdf(grid.handle(), fname.c_str(), aname.c_str(), sizeof(FileDataType) == 8 ? CNV_F64 : CNV_F32,
                             time, (int)face_dir, create_xdmf ? 1 : 0, nullptr, nullptr, nullptr),
          "CartesianWriteHDF");
}

template <typename FileDataType, typename G>
void CartesianReadHDF(const std::string &fname, G &grid, const size_t face_dir = 0)
{
    static_assert(sizeof(FileDataType) == 4 || sizeof(FileDataType) == 8, "FP32 or FP64 files");
    check(cnv_grid_read_hdf(grid.handle(), fname.c_str(), sizeof(FileDataType) == 8 ? CNV_F64 : CNV_F32,
                            (int)face_dir, nullptr, nullptr, nullptr),
          "CartesianReadHDF");
    grid.deviceWrote();
}
} // namespace IO

} // namespace Cubism

#endif // CUBISM_AMD_H
