// Writer of the reference's HDF5 output file (output.h5).  The HDF5 C
// library (1.10) is loaded at run time, so the engine library carries no
// link-time dependency on it; a missing library is reported when a file is
// asked for.
#pragma once
#include <cstddef>
#include <memory>
#include <string>

namespace edus {
namespace host {

class H5Out {
  public:
    // Creates (truncates) the file, or with append opens an existing one for
    // writing (creating it when absent).  Throws when the HDF5 library cannot
    // be loaded or the file cannot be opened.
    explicit H5Out(const std::string& path, bool append = false);
    ~H5Out();
    H5Out(const H5Out&) = delete;
    H5Out& operator=(const H5Out&) = delete;

    // absolute paths ("/json", "/time_au/3"); the parent group must exist
    void group(const std::string& path);
    void write_doubles(const std::string& path, const double* data, size_t n);
    void write_ints(const std::string& path, const int* data, size_t n);
    void write_bytes(const std::string& path, const std::string& s);  // uint8 array, no terminator

    // where the library is looked for: $EDUS_HDF5_LIB, then the names below
    static const char* const kLibraryNames[];

  private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace host
}  // namespace edus
