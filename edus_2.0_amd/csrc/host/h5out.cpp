// HDF5 output through the HDF5 1.10 C library, resolved with dlopen/dlsym.
// The file is written the way the reference's HDF5_tree writes it
// (src/core/hdf5/hdf5_tree.hpp): groups created with default property lists;
// every value a one-dimensional dataset (a scalar is a dataset of length 1,
// a string an array of uint8, a complex field its interleaved doubles).
#include "h5out.hpp"

#include <dlfcn.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

namespace edus {
namespace host {

namespace {

// the HDF5 1.10 C ABI types and constants used here (H5Ipublic.h,
// H5public.h, H5Fpublic.h, H5Ppublic.h, H5Spublic.h)
using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kDefault = 0;   // H5P_DEFAULT
constexpr hid_t kAll = 0;       // H5S_ALL
constexpr unsigned kTrunc = 2;  // H5F_ACC_TRUNC
constexpr unsigned kRdwr = 1;   // H5F_ACC_RDWR

}  // namespace

const char* const H5Out::kLibraryNames[] = {"libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103", "libhdf5.so",
                                            "/opt/conda/lib/libhdf5.so", nullptr};

struct H5Out::Impl {
    void* lib = nullptr;
    herr_t (*open)() = nullptr;
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t) = nullptr;
    hid_t (*fopen)(const char*, unsigned, hid_t) = nullptr;
    herr_t (*fclose)(hid_t) = nullptr;
    hid_t (*gcreate)(hid_t, const char*, hid_t, hid_t, hid_t) = nullptr;
    herr_t (*gclose)(hid_t) = nullptr;
    hid_t (*screate)(int, const hsize_t*, const hsize_t*) = nullptr;
    herr_t (*sclose)(hid_t) = nullptr;
    hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*) = nullptr;
    herr_t (*dclose)(hid_t) = nullptr;
    const hid_t* t_double = nullptr;  // H5T_NATIVE_DOUBLE_g
    const hid_t* t_int = nullptr;     // H5T_NATIVE_INT_g
    const hid_t* t_uchar = nullptr;   // H5T_NATIVE_UCHAR_g
    hid_t file = -1;

    template <class F>
    void sym(F& f, const char* name)
    {
        f = reinterpret_cast<F>(dlsym(lib, name));
        if (!f) throw std::runtime_error(std::string("HDF5 library lacks ") + name);
    }

    void load()
    {
        const char* env = std::getenv("EDUS_HDF5_LIB");
        if (env && *env) lib = dlopen(env, RTLD_NOW | RTLD_LOCAL);
        for (int i = 0; !lib && H5Out::kLibraryNames[i]; ++i) lib = dlopen(H5Out::kLibraryNames[i], RTLD_NOW | RTLD_LOCAL);
        if (!lib) throw std::runtime_error("output.h5: no HDF5 1.10 library found (set EDUS_HDF5_LIB)");
        sym(open, "H5open");
        sym(fcreate, "H5Fcreate");
        sym(fopen, "H5Fopen");
        sym(fclose, "H5Fclose");
        sym(gcreate, "H5Gcreate2");
        sym(gclose, "H5Gclose");
        sym(screate, "H5Screate_simple");
        sym(sclose, "H5Sclose");
        sym(dcreate, "H5Dcreate2");
        sym(dwrite, "H5Dwrite");
        sym(dclose, "H5Dclose");
        sym(t_double, "H5T_NATIVE_DOUBLE_g");
        sym(t_int, "H5T_NATIVE_INT_g");
        sym(t_uchar, "H5T_NATIVE_UCHAR_g");
        if (open() < 0) throw std::runtime_error("H5open failed");  // fills the native type ids
    }

    void dataset(const std::string& path, hid_t type, const void* data, size_t n)
    {
        const hsize_t dims[1] = {hsize_t(n)};
        const hid_t space = screate(1, dims, nullptr);
        if (space < 0) throw std::runtime_error("H5Screate_simple failed for " + path);
        const hid_t d = dcreate(file, path.c_str(), type, space, kDefault, kDefault, kDefault);
        if (d < 0) {
            sclose(space);
            throw std::runtime_error("H5Dcreate2 failed for " + path);
        }
        const herr_t rc = n ? dwrite(d, type, kAll, kAll, kDefault, data) : 0;
        dclose(d);
        sclose(space);
        if (rc < 0) throw std::runtime_error("H5Dwrite failed for " + path);
    }
};

H5Out::H5Out(const std::string& path, bool append) : impl_(new Impl)
{
    impl_->load();
    bool exists = false;
    if (append) {
        std::FILE* fp = std::fopen(path.c_str(), "rb");
        exists = fp != nullptr;
        if (fp) std::fclose(fp);
    }
    impl_->file = exists ? impl_->fopen(path.c_str(), kRdwr, kDefault)
                         : impl_->fcreate(path.c_str(), kTrunc, kDefault, kDefault);
    if (impl_->file < 0) throw std::runtime_error("cannot open " + path);
}

H5Out::~H5Out()
{
    if (impl_ && impl_->file >= 0) impl_->fclose(impl_->file);
    // the library stays loaded: HDF5 keeps global state until process exit
}

void H5Out::group(const std::string& path)
{
    const hid_t g = impl_->gcreate(impl_->file, path.c_str(), kDefault, kDefault, kDefault);
    if (g < 0) throw std::runtime_error("H5Gcreate2 failed for " + path);
    impl_->gclose(g);
}

void H5Out::write_doubles(const std::string& path, const double* data, size_t n)
{
    impl_->dataset(path, *impl_->t_double, data, n);
}

void H5Out::write_ints(const std::string& path, const int* data, size_t n)
{
    impl_->dataset(path, *impl_->t_int, data, n);
}

void H5Out::write_bytes(const std::string& path, const std::string& s)
{
    std::vector<unsigned char> b(s.begin(), s.end());
    impl_->dataset(path, *impl_->t_uchar, b.data(), b.size());
}

}  // namespace host
}  // namespace edus
