// MatrixMarket files for CSR matrices and vectors (host code).
//
// Writing follows cieutils/packages/io/src/MatrixMarket.cpp:36-78: a
// "coordinate real general" file with one "row column value" line per stored
// entry in CSR order (1-based indices), or an "array real general" file of
// size x 1 for a vector. Values are printed as std::ostream prints them by
// default (%.6g); `digits` > 0 asks for that many significant digits instead
// (17 round-trips a double). Lines are formatted by several threads, chunk by
// chunk, and written in order, so the file is the same for any thread count.
//
// Reading has no counterpart in the reference (MatrixMarket::Input is empty,
// MatrixMarket.hpp:25-26). It accepts coordinate files with real, integer or
// pattern values and general, symmetric or skew-symmetric symmetry, and
// general real/integer arrays. Coordinate entries become CSR rows sorted by
// column; duplicates are summed in file order; the mirror of an off-diagonal
// entry of a symmetric file is added (negated for skew-symmetric).
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cie_hip.h"

namespace ciehip {
int setError(int status, const std::string& message);
}

using ciehip::setError;

struct cie_hip_mm_s {
    int64_t rows = 0, columns = 0;
    bool dense = false;
    std::vector<int32_t> rowExtents, columnIndices;
    std::vector<double> entries;  // CSR values, or the dense array in column-major order
};

namespace {

constexpr int64_t kChunkEntries = 1 << 18;

unsigned workerCount() {
    const unsigned hardware = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hardware ? hardware : 1u));
}

template <typename T>
void appendValue(std::string& out, T value, int digits) {
    char buffer[64];
    const int n = std::snprintf(buffer, sizeof buffer, "%.*g", digits, double(value));
    out.append(buffer, size_t(n));
}

void appendIndex(std::string& out, int64_t value) {
    char buffer[24];
    const auto result = std::to_chars(buffer, buffer + sizeof buffer, value);
    out.append(buffer, result.ptr);
}

struct File {
    std::FILE* handle = nullptr;
    explicit File(const char* path, const char* mode) : handle(std::fopen(path, mode)) {}
    ~File() {
        if (handle) std::fclose(handle);
    }
};

// Formats `count` items in chunks on worker threads and writes the chunks in order.
template <typename Format>
bool writeChunked(std::FILE* file, int64_t count, Format format) {
    const unsigned workers = workerCount();
    std::vector<std::string> buffers(workers);
    for (int64_t first = 0; first < count; first += kChunkEntries * workers) {
        std::vector<std::thread> threads;
        for (unsigned w = 0; w < workers; ++w) {
            const int64_t begin = first + int64_t(w) * kChunkEntries;
            const int64_t end = std::min(count, begin + kChunkEntries);
            buffers[w].clear();
            if (begin >= end) continue;
            threads.emplace_back([&, w, begin, end] { format(buffers[w], begin, end); });
        }
        for (std::thread& t : threads) t.join();
        for (unsigned w = 0; w < workers; ++w)
            if (!buffers[w].empty() && std::fwrite(buffers[w].data(), 1, buffers[w].size(), file) != buffers[w].size())
                return false;
    }
    return true;
}

template <typename T>
int writeCsr(const char* path, int64_t rows, int64_t columns, int64_t nnz, const int32_t* rowExtents,
             const int32_t* columnIndices, const T* entries, int digits, const char* name) {
    if (!path || rows < 0 || columns < 0 || nnz < 0 || !rowExtents || (nnz && (!columnIndices || !entries)) ||
        digits < 0 || digits > 40)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": invalid arguments");
    if (rowExtents[0] != 0 || rowExtents[rows] != nnz)
        return setError(CIE_HIP_ERROR_SHAPE, std::string(name) + ": row extents must run from 0 to nnz");
    std::vector<int32_t> rowOf(size_t(std::max<int64_t>(nnz, 0)));
    for (int64_t r = 0; r < rows; ++r) {
        if (rowExtents[r + 1] < rowExtents[r])
            return setError(CIE_HIP_ERROR_SHAPE, std::string(name) + ": row extents decrease");
        std::fill(rowOf.begin() + rowExtents[r], rowOf.begin() + rowExtents[r + 1], int32_t(r));
    }
    File file(path, "wb");
    if (!file.handle) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": cannot open " + path);
    const int precision = digits ? digits : 6;
    std::string head = "%%MatrixMarket matrix coordinate real general\n";
    appendIndex(head, rows);
    head += ' ';
    appendIndex(head, columns);
    head += ' ';
    appendIndex(head, nnz);
    head += '\n';
    bool ok = std::fwrite(head.data(), 1, head.size(), file.handle) == head.size();
    ok = ok && writeChunked(file.handle, nnz, [&](std::string& out, int64_t begin, int64_t end) {
        out.reserve(size_t(end - begin) * 40);
        for (int64_t k = begin; k < end; ++k) {
            appendIndex(out, int64_t(rowOf[size_t(k)]) + 1);
            out += ' ';
            appendIndex(out, int64_t(columnIndices[k]) + 1);
            out += ' ';
            appendValue(out, entries[k], precision);
            out += '\n';
        }
    });
    if (!ok) return setError(CIE_HIP_ERROR_HIP, std::string(name) + ": write failed");
    return CIE_HIP_OK;
}

template <typename T>
int writeArray(const char* path, const T* values, int64_t size, int digits, const char* name) {
    if (!path || size < 0 || (size && !values) || digits < 0 || digits > 40)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": invalid arguments");
    File file(path, "wb");
    if (!file.handle) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": cannot open " + path);
    const int precision = digits ? digits : 6;
    std::string head = "%%MatrixMarket matrix array real general\n";
    appendIndex(head, size);
    head += " 1\n";
    bool ok = std::fwrite(head.data(), 1, head.size(), file.handle) == head.size();
    ok = ok && writeChunked(file.handle, size, [&](std::string& out, int64_t begin, int64_t end) {
        out.reserve(size_t(end - begin) * 24);
        for (int64_t k = begin; k < end; ++k) {
            appendValue(out, values[k], precision);
            out += '\n';
        }
    });
    if (!ok) return setError(CIE_HIP_ERROR_HIP, std::string(name) + ": write failed");
    return CIE_HIP_OK;
}

// ---- reading --------------------------------------------------------------

std::string lower(std::string s) {
    for (char& c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

const char* skipSpace(const char* p, const char* end) {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    return p;
}

bool parseInt(const char*& p, const char* end, int64_t& value) {
    p = skipSpace(p, end);
    const auto result = std::from_chars(p, end, value);
    if (result.ec != std::errc()) return false;
    p = result.ptr;
    return true;
}

bool parseReal(const char*& p, const char* end, double& value) {
    p = skipSpace(p, end);
    if (p < end && *p == '+') ++p;  // from_chars takes no leading '+'
    const auto result = std::from_chars(p, end, value);
    if (result.ec != std::errc()) return false;
    p = result.ptr;
    return true;
}

// The next line that is neither blank nor a comment; false at the end.
bool nextLine(const char*& p, const char* end, const char*& line, const char*& lineEnd) {
    while (p < end) {
        const char* e = static_cast<const char*>(std::memchr(p, '\n', size_t(end - p)));
        if (!e) e = end;
        const char* s = skipSpace(p, e);
        p = e < end ? e + 1 : end;
        if (s < e && *s != '%') {
            line = s;
            lineEnd = e;
            return true;
        }
    }
    return false;
}

struct Triplet {
    int32_t row, column;
    double value;
};

int readFile(const char* path, std::vector<char>& text) {
    File file(path, "rb");
    if (!file.handle) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string("cie_hip_mm_read: cannot open ") + path);
    if (std::fseek(file.handle, 0, SEEK_END) != 0) return setError(CIE_HIP_ERROR_HIP, "cie_hip_mm_read: cannot seek");
    const long size = std::ftell(file.handle);
    if (size < 0 || std::fseek(file.handle, 0, SEEK_SET) != 0) return setError(CIE_HIP_ERROR_HIP, "cie_hip_mm_read: cannot seek");
    text.resize(size_t(size));
    if (size && std::fread(text.data(), 1, size_t(size), file.handle) != size_t(size))
        return setError(CIE_HIP_ERROR_HIP, "cie_hip_mm_read: read failed");
    return CIE_HIP_OK;
}

}  // namespace

extern "C" {

int cie_hip_mm_write_csr_f64(const char* path, int64_t rows, int64_t columns, int64_t nnz, const int32_t* rowExtents,
                             const int32_t* columnIndices, const double* entries, int digits) {
    return writeCsr(path, rows, columns, nnz, rowExtents, columnIndices, entries, digits, "cie_hip_mm_write_csr_f64");
}

int cie_hip_mm_write_csr_f32(const char* path, int64_t rows, int64_t columns, int64_t nnz, const int32_t* rowExtents,
                             const int32_t* columnIndices, const float* entries, int digits) {
    return writeCsr(path, rows, columns, nnz, rowExtents, columnIndices, entries, digits, "cie_hip_mm_write_csr_f32");
}

int cie_hip_mm_write_array_f64(const char* path, const double* values, int64_t size, int digits) {
    return writeArray(path, values, size, digits, "cie_hip_mm_write_array_f64");
}

int cie_hip_mm_write_array_f32(const char* path, const float* values, int64_t size, int digits) {
    return writeArray(path, values, size, digits, "cie_hip_mm_write_array_f32");
}

int cie_hip_mm_read(const char* path, cie_hip_mm* data) {
    if (!data) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_mm_read: null output");
    *data = nullptr;
    if (!path) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_mm_read: null path");
    std::vector<char> text;
    if (const int status = readFile(path, text)) return status;
    const char* p = text.data();
    const char* end = p + text.size();
    auto shape = [](const std::string& message) { return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_mm_read: " + message); };

    // Banner: %%MatrixMarket matrix <format> <field> <symmetry>
    const char* bannerEnd = static_cast<const char*>(std::memchr(p, '\n', text.size()));
    if (!bannerEnd) bannerEnd = end;
    std::vector<std::string> words;
    for (const char* q = p; q < bannerEnd;) {
        q = skipSpace(q, bannerEnd);
        const char* s = q;
        while (q < bannerEnd && !std::isspace(static_cast<unsigned char>(*q))) ++q;
        if (s < q) words.emplace_back(s, q);
    }
    if (words.size() != 5 || lower(words[0]) != "%%matrixmarket" || lower(words[1]) != "matrix")
        return shape("not a MatrixMarket matrix banner");
    const std::string format = lower(words[2]), field = lower(words[3]), symmetry = lower(words[4]);
    const bool coordinate = format == "coordinate";
    if (!coordinate && format != "array") return shape("unknown format " + format);
    const bool pattern = field == "pattern";
    if (field != "real" && field != "double" && field != "integer" && !(pattern && coordinate))
        return shape("unsupported field " + field);
    const int mirror = symmetry == "general" ? 0 : symmetry == "symmetric" ? 1 : symmetry == "skew-symmetric" ? -1 : 2;
    if (mirror == 2 || (!coordinate && mirror != 0)) return shape("unsupported symmetry " + symmetry);
    p = bannerEnd < end ? bannerEnd + 1 : end;

    const char *line = nullptr, *lineEnd = nullptr;
    if (!nextLine(p, end, line, lineEnd)) return shape("missing size line");
    int64_t rows = 0, columns = 0, count = 0;
    if (!parseInt(line, lineEnd, rows) || !parseInt(line, lineEnd, columns) ||
        (coordinate && !parseInt(line, lineEnd, count)) || rows < 0 || columns < 0 || count < 0)
        return shape("bad size line");
    if (rows > INT32_MAX - 1 || columns > INT32_MAX) return shape("sizes exceed int32 indexing");
    if (!coordinate) count = rows * columns;

    auto result = std::make_unique<cie_hip_mm_s>();
    result->rows = rows;
    result->columns = columns;
    result->dense = !coordinate;

    // Data lines, parsed by workers over newline-aligned slices of the text.
    const unsigned workers = count > 100000 ? workerCount() : 1u;
    std::vector<const char*> cuts(workers + 1);
    cuts[0] = p;
    cuts[workers] = end;
    for (unsigned w = 1; w < workers; ++w) {
        const char* c = p + (end - p) * int64_t(w) / int64_t(workers);
        c = std::max(c, cuts[w - 1]);
        const char* nl = static_cast<const char*>(std::memchr(c, '\n', size_t(end - c)));
        cuts[w] = nl ? nl + 1 : end;
    }
    std::vector<std::vector<Triplet>> parts(workers);
    std::vector<std::vector<double>> values(workers);
    std::vector<std::string> errors(workers);
    auto parse = [&](unsigned w) {
        const char* q = cuts[w];
        const char* stop = cuts[w + 1];
        const char *l = nullptr, *le = nullptr;
        while (nextLine(q, stop, l, le)) {
            if (coordinate) {
                int64_t i = 0, j = 0;
                double v = 1.0;
                if (!parseInt(l, le, i) || !parseInt(l, le, j) || (!pattern && !parseReal(l, le, v))) {
                    errors[w] = "bad entry line";
                    return;
                }
                if (i < 1 || i > rows || j < 1 || j > columns) {
                    errors[w] = "entry index out of range";
                    return;
                }
                parts[w].push_back({int32_t(i - 1), int32_t(j - 1), v});
            } else {
                double v = 0.0;
                if (!parseReal(l, le, v)) {
                    errors[w] = "bad value line";
                    return;
                }
                values[w].push_back(v);
            }
        }
    };
    if (workers == 1) {
        parse(0);
    } else {
        std::vector<std::thread> threads;
        for (unsigned w = 0; w < workers; ++w) threads.emplace_back(parse, w);
        for (std::thread& t : threads) t.join();
    }
    int64_t found = 0;
    for (unsigned w = 0; w < workers; ++w) {
        if (!errors[w].empty()) return shape(errors[w]);
        found += int64_t(coordinate ? parts[w].size() : values[w].size());
    }
    if (found != count) return shape("expected " + std::to_string(count) + " entries, found " + std::to_string(found));

    if (!coordinate) {
        result->entries.reserve(size_t(count));
        for (auto& v : values) result->entries.insert(result->entries.end(), v.begin(), v.end());
        *data = result.release();
        return CIE_HIP_OK;
    }
    // Entries in file order, with the mirrors of symmetric files right after their originals.
    std::vector<Triplet> all;
    all.reserve(size_t(count) * (mirror ? 2 : 1));
    for (auto& part : parts)
        for (const Triplet& t : part) {
            all.push_back(t);
            if (mirror && t.row != t.column) {
                if (t.column >= rows || t.row >= columns) return shape("symmetric entry outside a square matrix");
                all.push_back({t.column, t.row, mirror > 0 ? t.value : -t.value});
            }
        }
    // Stable counting sort by row, then a stable sort by column inside each row, then the duplicates summed.
    std::vector<int64_t> start(size_t(rows) + 1, 0);
    for (const Triplet& t : all) ++start[size_t(t.row) + 1];
    for (int64_t r = 0; r < rows; ++r) start[size_t(r) + 1] += start[size_t(r)];
    std::vector<Triplet> sorted(all.size());
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (const Triplet& t : all) sorted[size_t(fill[size_t(t.row)]++)] = t;
    }
    result->rowExtents.assign(size_t(rows) + 1, 0);
    result->columnIndices.reserve(sorted.size());
    result->entries.reserve(sorted.size());
    for (int64_t r = 0; r < rows; ++r) {
        auto first = sorted.begin() + start[size_t(r)], last = sorted.begin() + start[size_t(r) + 1];
        std::stable_sort(first, last, [](const Triplet& a, const Triplet& b) { return a.column < b.column; });
        for (auto it = first; it != last; ++it) {
            if (!result->columnIndices.empty() && int64_t(result->columnIndices.size()) > result->rowExtents[size_t(r)] &&
                result->columnIndices.back() == it->column)
                result->entries.back() += it->value;
            else {
                result->columnIndices.push_back(it->column);
                result->entries.push_back(it->value);
            }
        }
        if (result->columnIndices.size() > size_t(INT32_MAX)) return shape("more than 2^31 entries");
        result->rowExtents[size_t(r) + 1] = int32_t(result->columnIndices.size());
    }
    *data = result.release();
    return CIE_HIP_OK;
}

int cie_hip_mm_shape(cie_hip_mm data, int64_t* rows, int64_t* columns, int64_t* nnz, int* dense) {
    if (!data) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_mm_shape: null data");
    if (rows) *rows = data->rows;
    if (columns) *columns = data->columns;
    if (nnz) *nnz = int64_t(data->entries.size());
    if (dense) *dense = data->dense ? 1 : 0;
    return CIE_HIP_OK;
}

int cie_hip_mm_arrays(cie_hip_mm data, const int32_t** rowExtents, const int32_t** columnIndices,
                      const double** entries) {
    if (!data) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_mm_arrays: null data");
    if (rowExtents) *rowExtents = data->dense ? nullptr : data->rowExtents.data();
    if (columnIndices) *columnIndices = data->dense ? nullptr : data->columnIndices.data();
    if (entries) *entries = data->entries.data();
    return CIE_HIP_OK;
}

int cie_hip_mm_free(cie_hip_mm data) {
    delete data;
    return CIE_HIP_OK;
}

}  // extern "C"
