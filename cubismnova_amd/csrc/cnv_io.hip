// cnv_io.hip -- the grid's on-disk format: HDF5 container + XDMF wrapper.
//
// The format of IO::CartesianWriteHDF / CartesianReadHDF
// (include/Cubism/IO/CartesianHDF.h:46-215, src/IO/HDFDriver.cpp:26-148,
// IO/XDMFDriver.h:29-160): one dataset "data" in <fname>.h5 of shape
// (Z, Y, X, NComp) over the global index span of the data entity (cells,
// nodes, or the faces of one direction), components interleaved per point
// (IO/FieldAOS.h AOSDriver: buf[c + NComp * i], i the x-fastest flat index
// of the span), chunked as one chunk; and <fname>.xmf describing a
// 3DCoRectMesh for cell and node data.
//
// The block-to-file reordering runs on the GPU: one kernel gathers the
// component slabs of every block into the interleaved file order (with the
// file's element type), one copy moves it to pinned host memory, and the
// HDF5 library writes it.  Reading runs the same way back.  HDF5 is the
// installed C library, loaded at the first call (it is not a link
// dependency of this library); without it these calls fail with
// CNV_ERR_UNSUPPORTED.
#include "cnv_internal.h"

#include <dlfcn.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>

namespace cnv
{

// ------------------------------------------------------------ HDF5 C API
// The few entry points used, with the types of the HDF5 1.10 C API
// (hid_t 64-bit, hsize_t unsigned 64-bit, herr_t int).
namespace h5
{
using hid_t = int64_t;
using hsize_t = unsigned long long;
using herr_t = int;
constexpr hid_t kDefault = 0;       // H5P_DEFAULT, H5E_DEFAULT
constexpr unsigned kAccTrunc = 2u;  // H5F_ACC_TRUNC
constexpr unsigned kAccRdonly = 0u; // H5F_ACC_RDONLY
constexpr int kSelectSet = 0;       // H5S_SELECT_SET

struct Api {
    herr_t (*open)();
    herr_t (*eset_auto2)(hid_t, void *, void *);
    hid_t (*pcreate)(hid_t);
    herr_t (*pclose)(hid_t);
    herr_t (*pset_chunk)(hid_t, int, const hsize_t *);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *, const hsize_t *);
    int (*sget_simple_extent_ndims)(hid_t);
    int (*sget_simple_extent_dims)(hid_t, hsize_t *, hsize_t *);
    hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen2)(hid_t, const char *, hid_t);
    hid_t (*dget_space)(hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    herr_t (*dclose)(hid_t);
    // type and property-class ids, valid after open()
    const hid_t *native_float, *native_double, *cls_file_access, *cls_dataset_create, *cls_dataset_xfer;
};

std::mutex g_mtx; // the serial HDF5 library is not thread-safe
Api g_api;
int g_state = 0; // 0 not tried, 1 loaded, -1 unavailable
std::string g_why;

template <typename F>
bool sym(void *lib, const char *name, F &f)
{
    f = reinterpret_cast<F>(dlsym(lib, name));
    return f != nullptr;
}

// with g_mtx held
const Api *api()
{
    if (g_state == 1)
        return &g_api;
    if (g_state == -1)
        return nullptr;
    g_state = -1;
    const char *env = std::getenv("CNV_HDF5_LIBRARY");
    const char *cands[] = {env, "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103",
                           "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so"};
    void *lib = nullptr;
    for (const char *c : cands)
        if (c && (lib = dlopen(c, RTLD_NOW | RTLD_LOCAL)) != nullptr)
            break;
    if (!lib) {
        g_why = "HDF5 C library not found (set CNV_HDF5_LIBRARY)";
        return nullptr;
    }
    Api &a = g_api;
    bool ok = sym(lib, "H5open", a.open) && sym(lib, "H5Eset_auto2", a.eset_auto2) && sym(lib, "H5Pcreate", a.pcreate) &&
              sym(lib, "H5Pclose", a.pclose) && sym(lib, "H5Pset_chunk", a.pset_chunk) &&
              sym(lib, "H5Fcreate", a.fcreate) && sym(lib, "H5Fopen", a.fopen) && sym(lib, "H5Fclose", a.fclose) &&
              sym(lib, "H5Screate_simple", a.screate_simple) && sym(lib, "H5Sclose", a.sclose) &&
              sym(lib, "H5Sselect_hyperslab", a.sselect_hyperslab) &&
              sym(lib, "H5Sget_simple_extent_ndims", a.sget_simple_extent_ndims) &&
              sym(lib, "H5Sget_simple_extent_dims", a.sget_simple_extent_dims) &&
              sym(lib, "H5Dcreate2", a.dcreate2) && sym(lib, "H5Dopen2", a.dopen2) &&
              sym(lib, "H5Dget_space", a.dget_space) && sym(lib, "H5Dwrite", a.dwrite) &&
              sym(lib, "H5Dread", a.dread) && sym(lib, "H5Dclose", a.dclose) &&
              sym(lib, "H5T_NATIVE_FLOAT_g", a.native_float) && sym(lib, "H5T_NATIVE_DOUBLE_g", a.native_double) &&
              sym(lib, "H5P_CLS_FILE_ACCESS_ID_g", a.cls_file_access) &&
              sym(lib, "H5P_CLS_DATASET_CREATE_ID_g", a.cls_dataset_create) &&
              sym(lib, "H5P_CLS_DATASET_XFER_ID_g", a.cls_dataset_xfer);
    if (!ok || a.open() < 0) {
        g_why = "HDF5 C library lacks the 1.10 API";
        return nullptr;
    }
    a.eset_auto2(kDefault, nullptr, nullptr); // errors come back as status codes
    g_state = 1;
    return &a;
}
} // namespace h5

// ------------------------------------------------------------ layout

// File span of a grid: the global index range of its entity (cells; nodes;
// the faces normal to `dface`), clipped to [lo, hi) when given.
struct Span {
    int lo[3], ext[3];
    int plus;  // bit d: upper-end blocks one larger along d
    int slab0; // first slab of the written components
    long long npts() const { return (long long)ext[0] * ext[1] * ext[2]; }
};

static int make_span(const cnv_grid_s *g, int face_dir, const int *lo, const int *hi, Span &sp)
{
    const int nc = g->desc.ncomp;
    if (g->desc.entity == CNV_FACE) {
        if (face_dir < 0 || face_dir > 2)
            return fail(CNV_ERR_INVALID, "face direction must be 0, 1 or 2");
        sp.slab0 = face_dir * nc;
    } else {
        sp.slab0 = 0;
    }
    sp.plus = plus_mask(g, sp.slab0);
    for (int d = 0; d < 3; ++d) {
        const int full = g->desc.nblocks[d] * g->desc.block_cells[d] + ((sp.plus >> d) & 1);
        int a = lo ? std::max(lo[d], 0) : 0;
        int b = hi ? std::min(hi[d], full) : full;
        if (b <= a)
            return fail(CNV_ERR_INVALID, "empty index range");
        sp.lo[d] = a;
        sp.ext[d] = b - a;
    }
    return CNV_OK;
}

struct IOArgs {
    char *grid;     // grid base
    char *file;     // interleaved file buffer
    long long cstride, bstride;
    int nb[3], n[3], lo[3], ext[3];
    int plus, slab0, nc;
    long long npts;
};

// file point i <-> element of block (b, local) with the block's own extent
// as pitch (node / face blocks at the upper end are one larger)
__device__ inline long long io_elem(const IOArgs &a, long long i)
{
    const long long x = i % a.ext[0], y = (i / a.ext[0]) % a.ext[1], z = i / ((long long)a.ext[0] * a.ext[1]);
    const long long gp[3] = {a.lo[0] + x, a.lo[1] + y, a.lo[2] + z};
    long long blk = 0, loc = 0, bmul = 1, lmul = 1;
    for (int d = 0; d < 3; ++d) {
        const int b = (int)min(gp[d] / a.n[d], (long long)a.nb[d] - 1);
        const int e = a.n[d] + ((((a.plus >> d) & 1) && b == a.nb[d] - 1) ? 1 : 0);
        blk += b * bmul;
        loc += (gp[d] - (long long)b * a.n[d]) * lmul;
        bmul *= a.nb[d];
        lmul *= e;
    }
    return blk * a.bstride + loc;
}

template <typename TG, typename TF, bool WRITE>
__global__ __launch_bounds__(256) void k_io(const IOArgs a)
{
    TG *grid = reinterpret_cast<TG *>(a.grid);
    TF *file = reinterpret_cast<TF *>(a.file);
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < a.npts; i += (long long)gridDim.x * 256) {
        const long long e = io_elem(a, i);
        for (int c = 0; c < a.nc; ++c) {
            const long long ge = (long long)(a.slab0 + c) * a.cstride + e;
            if (WRITE)
                file[i * a.nc + c] = (TF)grid[ge];
            else
                grid[ge] = (TG)file[i * a.nc + c];
        }
    }
}

static int launch_io(const cnv_grid_s *g, const Span &sp, void *file, int file_dtype, bool write, hipStream_t st)
{
    IOArgs a;
    a.grid = static_cast<char *>(g->data);
    a.file = static_cast<char *>(file);
    a.cstride = (long long)(g->component_bytes / g->elem);
    a.bstride = (long long)(g->block_bytes / g->elem);
    for (int d = 0; d < 3; ++d) {
        a.nb[d] = g->desc.nblocks[d];
        a.n[d] = g->desc.block_cells[d];
        a.lo[d] = sp.lo[d];
        a.ext[d] = sp.ext[d];
    }
    a.plus = sp.plus;
    a.slab0 = sp.slab0;
    a.nc = g->desc.ncomp;
    a.npts = sp.npts();
    const long long nwg = std::min<long long>((a.npts + 255) / 256, 65536);
    const bool g64 = g->desc.dtype == CNV_F64, f64 = file_dtype == CNV_F64;
#define CNV_IO(TG, TF)                                                                                          \
    if (write)                                                                                                  \
        hipLaunchKernelGGL((k_io<TG, TF, true>), dim3((unsigned)nwg), dim3(256), 0, st, a);                     \
    else                                                                                                        \
        hipLaunchKernelGGL((k_io<TG, TF, false>), dim3((unsigned)nwg), dim3(256), 0, st, a);
    if (g64 && f64) {
        CNV_IO(double, double)
    } else if (g64) {
        CNV_IO(double, float)
    } else if (f64) {
        CNV_IO(float, double)
    } else {
        CNV_IO(float, float)
    }
#undef CNV_IO
    CNV_HIP(hipGetLastError());
    return CNV_OK;
}

// device staging + pinned host buffer of one call
struct Staging {
    void *dev = nullptr, *host = nullptr;
    ~Staging()
    {
        if (dev)
            (void)hipFree(dev);
        if (host)
            (void)hipHostFree(host);
    }
    int alloc(size_t bytes)
    {
        CNV_HIP(hipMalloc(&dev, bytes));
        CNV_HIP(hipHostMalloc(&host, bytes, hipHostMallocDefault));
        return CNV_OK;
    }
};

// IO/XDMFDriver.h:29-160 (uniform mesh): 3DCoRectMesh over the nodes of the
// written span, origin and spacing in ZYX order, the dataset as an attribute
static int write_xdmf(const cnv_grid_s *g, const Span &sp, const std::string &fname, const std::string &aname,
                      int file_dtype, double time)
{
    const bool node = g->desc.entity == CNV_NODE;
    int mdims[3], ddims[3];
    double orig[3];
    for (int d = 0; d < 3; ++d) {
        ddims[d] = sp.ext[d];
        mdims[d] = node ? sp.ext[d] : sp.ext[d] + 1; // mesh nodes of the span
        orig[d] = g->desc.begin[d] + sp.lo[d] * g->h[d];
    }
    const int nc = g->desc.ncomp;
    const char *attr = nc == 1 ? "Scalar" : (nc == 3 ? "Vector" : "Tensor");
    const std::string base = fname.substr(fname.find_last_of('/') + 1);
    std::FILE *f = std::fopen((fname + ".xmf").c_str(), "w");
    if (!f)
        return fail(CNV_ERR_INVALID, "cannot create " + fname + ".xmf");
    std::fprintf(f, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n<Xdmf Version=\"2.0\">\n");
    std::fprintf(f, "<Domain>\n\t<Grid GridType=\"Uniform\">\n\t\t<Time Value=\"%e\"/>\n\n", time);
    std::fprintf(f, "\t\t<Topology TopologyType=\"3DCoRectMesh\" Dimensions=\"%d %d %d\"/>\n\n", mdims[2], mdims[1],
                 mdims[0]);
    std::fprintf(f, "\t\t<Geometry GeometryType=\"Origin_DxDyDz\">\n");
    std::fprintf(f, "\t\t\t<DataItem Name=\"Origin\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"8\" "
                    "Format=\"XML\">\n\t\t\t\t%.16g %.16g %.16g\n\t\t\t</DataItem>\n",
                 orig[2], orig[1], orig[0]);
    std::fprintf(f, "\t\t\t<DataItem Name=\"Spacing\" Dimensions=\"3\" NumberType=\"Float\" Precision=\"8\" "
                    "Format=\"XML\">\n\t\t\t\t%.16g %.16g %.16g\n\t\t\t</DataItem>\n",
                 g->h[2], g->h[1], g->h[0]);
    std::fprintf(f, "\t\t</Geometry>\n\n");
    std::fprintf(f, "\t\t<Attribute Name=\"%s\" AttributeType=\"%s\" Center=\"%s\">\n", aname.c_str(), attr,
                 node ? "Node" : "Cell");
    std::fprintf(f, "\t\t\t<DataItem Dimensions=\"%d %d %d %d\" NumberType=\"Float\" Precision=\"%d\" Format=\"HDF\">\n",
                 ddims[2], ddims[1], ddims[0], nc, file_dtype == CNV_F64 ? 8 : 4);
    std::fprintf(f, "\t\t\t\t./%s:/data\n\t\t\t</DataItem>\n\t\t</Attribute>\n\t</Grid>\n</Domain>\n</Xdmf>\n",
                 (base + ".h5").c_str());
    std::fclose(f);
    return CNV_OK;
}

static int check_io(const cnv_grid_s *g, const char *fname, int file_dtype)
{
    if (!g || !fname || !*fname)
        return fail(CNV_ERR_INVALID, "grid or file name is NULL");
    if (file_dtype != CNV_F32 && file_dtype != CNV_F64)
        return fail(CNV_ERR_INVALID, "file dtype must be CNV_F32 or CNV_F64");
    return CNV_OK;
}

} // namespace cnv

using namespace cnv;

extern "C" {

int cnv_grid_write_hdf(cnv_grid g, const char *fname, const char *aname, int file_dtype, double time, int face_dir,
                       int create_xdmf, const int begin[3], const int end[3], cnv_stream s)
{
    try {
        int rc = check_io(g, fname, file_dtype);
        if (rc != CNV_OK)
            return rc;
        Span sp;
        if ((rc = make_span(g, face_dir, begin, end, sp)) != CNV_OK)
            return rc;
        std::lock_guard<std::mutex> lk(h5::g_mtx);
        const h5::Api *H = h5::api();
        if (!H)
            return fail(CNV_ERR_UNSUPPORTED, h5::g_why);
        const int nc = g->desc.ncomp;
        const size_t fsz = file_dtype == CNV_F64 ? 8 : 4;
        const size_t bytes = (size_t)sp.npts() * nc * fsz;
        Staging st;
        if ((rc = st.alloc(bytes)) != CNV_OK)
            return rc;
        const hipStream_t hs = (hipStream_t)s;
        if ((rc = launch_io(g, sp, st.dev, file_dtype, true, hs)) != CNV_OK)
            return rc;
        CNV_HIP(hipMemcpyAsync(st.host, st.dev, bytes, hipMemcpyDeviceToHost, hs));
        CNV_HIP(hipStreamSynchronize(hs));

        // HDFDriver::write (src/IO/HDFDriver.cpp:26-103): ZYXC dataset "data"
        const h5::hsize_t dims[4] = {(h5::hsize_t)sp.ext[2], (h5::hsize_t)sp.ext[1], (h5::hsize_t)sp.ext[0],
                                     (h5::hsize_t)nc};
        // one chunk, as the reference writes, unless it exceeds HDF5's 4 GiB
        // chunk limit: then chunks of whole z-planes
        h5::hsize_t chunk[4] = {dims[0], dims[1], dims[2], dims[3]};
        const size_t plane = (size_t)dims[1] * dims[2] * dims[3] * fsz;
        if (bytes >= ((size_t)1 << 32))
            chunk[0] = std::max<size_t>(1, (((size_t)1 << 32) - 1) / plane);
        const h5::hid_t ftype = file_dtype == CNV_F64 ? *H->native_double : *H->native_float;
        const std::string fn = std::string(fname) + ".h5";
        const h5::hid_t fapl = H->pcreate(*H->cls_file_access);
        const h5::hid_t file = H->fcreate(fn.c_str(), h5::kAccTrunc, h5::kDefault, fapl);
        H->pclose(fapl);
        if (file < 0)
            return fail(CNV_ERR_INVALID, "cannot create " + fn);
        bool ok = true;
        const h5::hid_t space = H->screate_simple(4, dims, nullptr);
        const h5::hid_t dcpl = H->pcreate(*H->cls_dataset_create);
        ok = ok && space >= 0 && dcpl >= 0 && H->pset_chunk(dcpl, 4, chunk) >= 0;
        const h5::hid_t dset = ok ? H->dcreate2(file, "data", ftype, space, h5::kDefault, dcpl, h5::kDefault) : -1;
        ok = ok && dset >= 0;
        const h5::hid_t xfer = H->pcreate(*H->cls_dataset_xfer);
        ok = ok && H->dwrite(dset, ftype, space, space, xfer, st.host) >= 0;
        if (xfer >= 0)
            H->pclose(xfer);
        if (dset >= 0)
            H->dclose(dset);
        if (dcpl >= 0)
            H->pclose(dcpl);
        if (space >= 0)
            H->sclose(space);
        ok = (H->fclose(file) >= 0) && ok;
        if (!ok)
            return fail(CNV_ERR_INVALID, "HDF5 write of " + fn + " failed");
        // the reference writes no XDMF for face data (HDFDriver.cpp:52-57)
        if (create_xdmf && g->desc.entity != CNV_FACE)
            return write_xdmf(g, sp, fname, aname ? aname : "data", file_dtype, time);
        return CNV_OK;
    } catch (const std::bad_alloc &) {
        return fail(CNV_ERR_ALLOC, "out of host memory");
    }
}

int cnv_grid_read_hdf(cnv_grid g, const char *fname, int file_dtype, int face_dir, const int begin[3],
                      const int end[3], cnv_stream s)
{
    try {
        int rc = check_io(g, fname, file_dtype);
        if (rc != CNV_OK)
            return rc;
        Span sp;
        if ((rc = make_span(g, face_dir, begin, end, sp)) != CNV_OK)
            return rc;
        const std::string fn = std::string(fname) + ".h5";
        if (std::FILE *t = std::fopen(fn.c_str(), "rb"))
            std::fclose(t);
        else // CartesianHDF.h:175-180
            return fail(CNV_ERR_INVALID, "CartesianReadHDF: File '" + std::string(fname) + "' does not exist");
        std::lock_guard<std::mutex> lk(h5::g_mtx);
        const h5::Api *H = h5::api();
        if (!H)
            return fail(CNV_ERR_UNSUPPORTED, h5::g_why);
        const int nc = g->desc.ncomp;
        const size_t fsz = file_dtype == CNV_F64 ? 8 : 4;
        const size_t bytes = (size_t)sp.npts() * nc * fsz;
        Staging st;
        if ((rc = st.alloc(bytes)) != CNV_OK)
            return rc;

        // HDFDriver::read (src/IO/HDFDriver.cpp:105-148): the leading
        // (Z, Y, X, NComp) hyperslab of "data"
        const h5::hsize_t count[4] = {(h5::hsize_t)sp.ext[2], (h5::hsize_t)sp.ext[1], (h5::hsize_t)sp.ext[0],
                                      (h5::hsize_t)nc};
        const h5::hsize_t offset[4] = {0, 0, 0, 0};
        const h5::hid_t ftype = file_dtype == CNV_F64 ? *H->native_double : *H->native_float;
        const h5::hid_t fapl = H->pcreate(*H->cls_file_access);
        const h5::hid_t file = H->fopen(fn.c_str(), h5::kAccRdonly, fapl);
        H->pclose(fapl);
        if (file < 0)
            return fail(CNV_ERR_INVALID, "cannot open " + fn + " as HDF5");
        bool ok = true;
        const h5::hid_t dset = H->dopen2(file, "data", h5::kDefault);
        const h5::hid_t fspace = dset >= 0 ? H->dget_space(dset) : -1;
        ok = fspace >= 0 && H->sget_simple_extent_ndims(fspace) == 4;
        h5::hsize_t fd[4] = {0, 0, 0, 0};
        ok = ok && H->sget_simple_extent_dims(fspace, fd, nullptr) == 4;
        for (int d = 0; d < 4 && ok; ++d)
            ok = fd[d] >= count[d];
        const h5::hid_t mspace = ok ? H->screate_simple(4, count, nullptr) : -1;
        ok = ok && mspace >= 0 && H->sselect_hyperslab(fspace, h5::kSelectSet, offset, nullptr, count, nullptr) >= 0;
        const h5::hid_t xfer = H->pcreate(*H->cls_dataset_xfer);
        ok = ok && H->dread(dset, ftype, mspace, fspace, xfer, st.host) >= 0;
        if (xfer >= 0)
            H->pclose(xfer);
        if (mspace >= 0)
            H->sclose(mspace);
        if (fspace >= 0)
            H->sclose(fspace);
        if (dset >= 0)
            H->dclose(dset);
        H->fclose(file);
        if (!ok)
            return fail(CNV_ERR_SIZE, "dataset 'data' of " + fn + " is missing or smaller than the grid span");
        const hipStream_t hs = (hipStream_t)s;
        CNV_HIP(hipMemcpyAsync(st.dev, st.host, bytes, hipMemcpyHostToDevice, hs));
        if ((rc = launch_io(g, sp, st.dev, file_dtype, false, hs)) != CNV_OK)
            return rc;
        CNV_HIP(hipStreamSynchronize(hs));
        return CNV_OK;
    } catch (const std::bad_alloc &) {
        return fail(CNV_ERR_ALLOC, "out of host memory");
    }
}

} // extern "C"
