// Test infrastructure only: C entry points over the reference's own
// MatrixMarket writer (cieutils/packages/io/src/MatrixMarket.cpp, compiled
// from the reference tree by oracle/build.py into oracle/_ref/). The tests
// compare the library's files with the ones this writes, byte for byte.
#include <cstdint>
#include <fstream>

#include "packages/io/inc/MatrixMarket.hpp"

extern "C" {

int ref_mm_write_csr_f64(const char* path, int rows, int columns, int nnz, const int* rowExtents,
                         const int* columnIndices, const double* entries) {
    std::ofstream file(path);
    cie::io::MatrixMarket::Output io(file);
    io(rows, columns, nnz, rowExtents, columnIndices, entries);
    return file.good() ? 0 : 1;
}

int ref_mm_write_csr_f32(const char* path, int rows, int columns, int nnz, const int* rowExtents,
                         const int* columnIndices, const float* entries) {
    std::ofstream file(path);
    cie::io::MatrixMarket::Output io(file);
    io(rows, columns, nnz, rowExtents, columnIndices, entries);
    return file.good() ? 0 : 1;
}

int ref_mm_write_array_f64(const char* path, const double* values, int64_t size) {
    std::ofstream file(path);
    cie::io::MatrixMarket::Output io(file);
    io(values, static_cast<std::size_t>(size));
    return file.good() ? 0 : 1;
}

int ref_mm_write_array_f32(const char* path, const float* values, int64_t size) {
    std::ofstream file(path);
    cie::io::MatrixMarket::Output io(file);
    io(values, static_cast<std::size_t>(size));
    return file.good() ? 0 : 1;
}

}  // extern "C"
