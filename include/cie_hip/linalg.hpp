// C++ face of the HIP backend, in the shape of the reference's solver layer.
//
// Header-only over the C ABI (include/cie_hip.h). Names, signatures and
// semantics follow linalg/packages/solvers (paths relative to the reference):
//   LinalgSpaceLike          inc/LinalgSpace.hpp:10-35
//   LinearOperator<TSpace>   inc/LinearOperator.hpp:10-34
//   IterativeSolver          inc/IterativeSolver.hpp:18-56
//   ConjugateGradients       inc/ConjugateGradients.hpp:15-46,
//                            impl/ConjugateGradients_impl.hpp:49-201
//   DiagonalOperator         inc/DiagonalOperator.hpp:16-57, src/DiagonalOperator.cpp
//   CSRView                  utilities/inc/CSRView.hpp:11-86
//   SYCLSpace (counterpart)  inc/SYCLSpace.hpp:17-134
// Inside the reference tree (CIE_HIP_WITH_REFERENCE_HEADERS defined), the
// reference's own LinearOperator, IterativeSolver and ConjugateGradients are
// used and HIPSpace simply becomes one more space for them. Standalone, this
// header declares interface-identical versions of its own.
//
// Errors: every failed C-ABI call throws cie::linalg::HIPError (the
// counterpart of cie::Exception from CIE_CHECK, cieutils/packages/macros/inc/checks.hpp:17-23).
#pragma once

#include "../cie_hip.h"

#include <cmath>
#include <concepts>
#include <cstddef>
#include <iostream>
#include <limits>
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#ifdef CIE_HIP_WITH_REFERENCE_HEADERS
#include "packages/solvers/inc/ConjugateGradients.hpp"
#include "packages/solvers/inc/DiagonalOperator.hpp"
#include "packages/solvers/inc/LinalgSpace.hpp"
#include "packages/solvers/inc/LinearOperator.hpp"
#include "packages/utilities/inc/CSRView.hpp"
#endif

namespace cie::linalg {

class HIPError : public std::runtime_error {
public:
    HIPError(int status, const std::string& message) : std::runtime_error(message), _status(status) {}
    int status() const noexcept { return _status; }

private:
    int _status;
};

namespace hipdetail {
inline void check(int status, const char* what) {
    if (status != CIE_HIP_OK)
        throw HIPError(status, std::string(what) + ": " + cie_hip_last_error());
}

// The report text of IterativeSolver::makeIterationReport (src/IterativeSolver.cpp:17-58) for one solver
// state, formatted by cie_hip_iteration_report.
inline std::string iterationReport(int verbosity, int reportType, std::size_t iteration, double absolute,
                                   double relative, const cie_hip_statistics& settings) {
    const cie_hip_statistics status{iteration, absolute, relative};
    std::size_t length = 0;
    check(cie_hip_iteration_report(verbosity, reportType, &status, &settings, nullptr, 0, &length),
          "cie_hip_iteration_report");
    std::string text(length, '\0');
    check(cie_hip_iteration_report(verbosity, reportType, &status, &settings, text.data(), length + 1, &length),
          "cie_hip_iteration_report");
    return text;
}

// Where the solvers of this header log their reports (the reference logs them into a "ConjugateGradients"
// log block, ConjugateGradients_impl.hpp:72-74): one line per non-empty report, std::clog unless set.
inline std::ostream*& reportStream() {
    static std::ostream* stream = &std::clog;
    return stream;
}

inline void logReport(const std::string& report) {
    if (!report.empty() && reportStream()) *reportStream() << report << '\n';
}
}  // namespace hipdetail

// Sends the solvers' per-iteration reports to `stream` (nullptr: drop them).
inline void setReportStream(std::ostream* stream) { hipdetail::reportStream() = stream; }

#ifndef CIE_HIP_WITH_REFERENCE_HEADERS

template <class T>
concept LinalgSpaceLike = requires(const T space, typename T::Vector vector, const typename T::Vector constVector,
                                   typename T::VectorView view, typename T::ConstVectorView constView) {
    typename T::Value;
    typename T::Vector;
    typename T::VectorView;
    typename T::ConstVectorView;
    { space.view(vector) } -> std::same_as<typename T::VectorView>;
    { space.view(constVector) } -> std::same_as<typename T::ConstVectorView>;
    { space.size(constView) } -> std::same_as<std::size_t>;
    { space.size(view) } -> std::same_as<std::size_t>;
    { space.makeVector(std::size_t()) } -> std::same_as<typename T::Vector>;
    { space.innerProduct(constView, constView) } -> std::same_as<typename T::Value>;
    { space.scale(view, constView, typename T::Value()) } -> std::same_as<void>;
    { space.scale(view, typename T::Value()) } -> std::same_as<void>;
    { space.add(view, constView, typename T::Value()) } -> std::same_as<void>;
    { space.assign(view, constView) } -> std::same_as<void>;
    { space.fill(view, typename T::Value()) } -> std::same_as<void>;
};

template <LinalgSpaceLike TSpace>
class LinearOperator {
public:
    virtual ~LinearOperator() = default;
    /// out = inScale * out + outScale * A * in
    virtual void product(typename TSpace::Value inScale, typename TSpace::ConstVectorView in,
                         typename TSpace::Value outScale, typename TSpace::VectorView out) = 0;
};

template <class TValue, class TIndex = std::conditional_t<std::is_const_v<TValue>, const int, int>>
class CSRView {
public:
    using Value = std::remove_const_t<TValue>;
    using Index = std::remove_const_t<TIndex>;
    constexpr CSRView() noexcept = default;
    constexpr CSRView(std::size_t columnCount, std::span<TIndex> rowExtents, std::span<TIndex> columnIndices,
                      std::span<TValue> entries) noexcept
        : _columnCount(columnCount), _rowExtents(rowExtents), _columnIndices(columnIndices), _entries(entries) {}
    std::size_t rowCount() const noexcept { return _rowExtents.size() - 1; }
    constexpr std::size_t columnCount() const noexcept { return _columnCount; }
    std::span<const Index> rowExtents() const noexcept { return {_rowExtents.data(), _rowExtents.size()}; }
    std::span<const Index> columnIndices() const noexcept { return {_columnIndices.data(), _columnIndices.size()}; }
    std::span<const Value> entries() const noexcept { return {_entries.data(), _entries.size()}; }

private:
    std::size_t _columnCount = 0;
    std::span<TIndex> _rowExtents, _columnIndices;
    std::span<TValue> _entries;
};

template <LinalgSpaceLike TSpace>
class IterativeSolver : public LinearOperator<TSpace> {
public:
    struct Statistics {
        std::size_t iterationCount = 0ul;
        typename TSpace::Value absoluteResidual = static_cast<typename TSpace::Value>(0);
        typename TSpace::Value relativeResidual = static_cast<typename TSpace::Value>(0);
    };
    IterativeSolver() noexcept = default;
    explicit IterativeSolver(Statistics settings) : _settings(settings) {}
    Statistics getConfiguration() const noexcept { return _settings; }
    void configure(Statistics settings) noexcept { _settings = settings; }
    std::optional<Statistics> getStats() const noexcept { return _results; }

protected:
    enum class ReportType { Termination, Iteration };

    // IterativeSolver::makeIterationReport (src/IterativeSolver.cpp:17-58).
    std::string& makeIterationReport(std::string& output, int verbosity, ReportType reportType,
                                     const Statistics& status, const Statistics& settings) const {
        output = hipdetail::iterationReport(
            verbosity, reportType == ReportType::Termination ? CIE_HIP_REPORT_TERMINATION : CIE_HIP_REPORT_ITERATION,
            status.iterationCount, double(status.absoluteResidual), double(status.relativeResidual),
            cie_hip_statistics{settings.iterationCount, double(settings.absoluteResidual),
                               double(settings.relativeResidual)});
        return output;
    }

    void report(Statistics stats) noexcept { _results = stats; }

private:
    Statistics _settings{};
    std::optional<Statistics> _results;
};

#endif  // !CIE_HIP_WITH_REFERENCE_HEADERS

// ---------------------------------------------------------------------------
// Context: one device + one HIP stream (the sycl::queue of SYCLSpace).
// ---------------------------------------------------------------------------

class HIPContext {
public:
    explicit HIPContext(int device = 0, void* stream = nullptr) {
        hipdetail::check(cie_hip_context_create(device, stream, &_handle), "cie_hip_context_create");
    }
    ~HIPContext() { cie_hip_context_destroy(_handle); }
    HIPContext(const HIPContext&) = delete;
    HIPContext& operator=(const HIPContext&) = delete;
    cie_hip_context get() const noexcept { return _handle; }
    void synchronize() const { hipdetail::check(cie_hip_context_synchronize(_handle), "synchronize"); }

private:
    cie_hip_context _handle = nullptr;
};

// ---------------------------------------------------------------------------
// Device vectors and views (SYCLVector / SYCLView, SYCLSpace.hpp:17-71).
// ---------------------------------------------------------------------------

template <class T>
class HIPVector {
public:
    HIPVector() noexcept = default;
    HIPVector(std::shared_ptr<const HIPContext> context, std::size_t size) : _context(std::move(context)), _size(size) {
        void* pointer = nullptr;
        hipdetail::check(cie_hip_malloc(_context->get(), sizeof(T) * (size ? size : 1), &pointer), "cie_hip_malloc");
        _data = static_cast<T*>(pointer);
    }
    ~HIPVector() { release(); }
    HIPVector(HIPVector&& other) noexcept { swap(other); }
    HIPVector& operator=(HIPVector&& other) noexcept {
        if (this != &other) {
            release();
            swap(other);
        }
        return *this;
    }
    HIPVector(const HIPVector&) = delete;
    HIPVector& operator=(const HIPVector&) = delete;
    std::size_t size() const noexcept { return _size; }
    T* get() noexcept { return _data; }
    const T* get() const noexcept { return _data; }

private:
    void release() noexcept {
        if (_data && _context) cie_hip_free(_context->get(), _data);
        _data = nullptr;
        _size = 0;
    }
    void swap(HIPVector& other) noexcept {
        std::swap(_context, other._context);
        std::swap(_data, other._data);
        std::swap(_size, other._size);
    }
    std::shared_ptr<const HIPContext> _context;
    T* _data = nullptr;
    std::size_t _size = 0;
};

template <class T>
class HIPView {
public:
    using Element = std::remove_const_t<T>;
    HIPView() noexcept = default;
    HIPView(T* data, std::size_t size) noexcept : _data(data), _size(size) {}
    HIPView(std::conditional_t<std::is_const_v<T>, const HIPVector<Element>&, HIPVector<Element>&> vector) noexcept
        : _data(vector.get()), _size(vector.size()) {}
    std::size_t size() const noexcept { return _size; }
    T* get() const noexcept { return _data; }
    operator HIPView<const Element>() const noexcept { return {_data, _size}; }

private:
    T* _data = nullptr;
    std::size_t _size = 0;
};

// ---------------------------------------------------------------------------
// HIPSpace<T>: LinalgSpaceLike over device vectors. Every operation is
// synchronous, as SYCLSpace's (each ends in wait_and_throw, SYCLSpace.cpp:166-311).
// ---------------------------------------------------------------------------

template <class T>
class HIPSpace {
    static_assert(std::is_same_v<T, double> || std::is_same_v<T, float>, "HIPSpace holds float or double");

public:
    using Value = T;
    using Vector = HIPVector<T>;
    using VectorView = HIPView<T>;
    using ConstVectorView = HIPView<const T>;

    explicit HIPSpace(std::shared_ptr<const HIPContext> context) : _context(std::move(context)) {}

    static VectorView view(Vector& vector) noexcept { return VectorView(vector); }
    static ConstVectorView view(const Vector& vector) noexcept { return ConstVectorView(vector); }
    static std::size_t size(ConstVectorView view) noexcept { return view.size(); }
    static std::size_t size(VectorView view) noexcept { return view.size(); }
    Vector makeVector(std::size_t size) const { return Vector(_context, size); }

    Value innerProduct(ConstVectorView left, ConstVectorView right) const {
        Value result{};
        if constexpr (std::is_same_v<T, double>)
            hipdetail::check(cie_hip_inner_product_f64(ctx(), left.get(), right.get(), std::int64_t(left.size()), &result),
                             "innerProduct");
        else
            hipdetail::check(cie_hip_inner_product_f32(ctx(), left.get(), right.get(), std::int64_t(left.size()), &result),
                             "innerProduct");
        return result;
    }

    void scale(VectorView target, ConstVectorView source, Value scale) const {
        if constexpr (std::is_same_v<T, double>)
            hipdetail::check(cie_hip_scale_f64(ctx(), target.get(), source.get(), std::int64_t(target.size()), scale), "scale");
        else
            hipdetail::check(cie_hip_scale_f32(ctx(), target.get(), source.get(), std::int64_t(target.size()), scale), "scale");
        sync();
    }

    void scale(VectorView view, Value value) const {
        if constexpr (std::is_same_v<T, double>)
            hipdetail::check(cie_hip_scale_uniform_f64(ctx(), view.get(), std::int64_t(view.size()), value), "scale");
        else
            hipdetail::check(cie_hip_scale_uniform_f32(ctx(), view.get(), std::int64_t(view.size()), value), "scale");
        sync();
    }

    void add(VectorView target, ConstVectorView source, Value scale) const {
        if constexpr (std::is_same_v<T, double>)
            hipdetail::check(cie_hip_add_f64(ctx(), target.get(), source.get(), std::int64_t(target.size()), scale), "add");
        else
            hipdetail::check(cie_hip_add_f32(ctx(), target.get(), source.get(), std::int64_t(target.size()), scale), "add");
        sync();
    }

    void assign(VectorView target, ConstVectorView source) const {
        if constexpr (std::is_same_v<T, double>)
            hipdetail::check(cie_hip_assign_f64(ctx(), target.get(), source.get(), std::int64_t(target.size())), "assign");
        else
            hipdetail::check(cie_hip_assign_f32(ctx(), target.get(), source.get(), std::int64_t(target.size())), "assign");
        sync();
    }

    // host -> device and device -> host (SYCLSpace.hpp:116-122)
    void assign(VectorView target, std::span<const T> source) const {
        hipdetail::check(cie_hip_memcpy(ctx(), target.get(), source.data(), sizeof(T) * source.size(), 0), "assign");
    }
    void assign(std::span<T> target, ConstVectorView source) const {
        hipdetail::check(cie_hip_memcpy(ctx(), target.data(), source.get(), sizeof(T) * target.size(), 1), "assign");
    }

    void fill(VectorView target, Value value) const {
        if constexpr (std::is_same_v<T, double>)
            hipdetail::check(cie_hip_fill_f64(ctx(), target.get(), std::int64_t(target.size()), value), "fill");
        else
            hipdetail::check(cie_hip_fill_f32(ctx(), target.get(), std::int64_t(target.size()), value), "fill");
        sync();
    }

    std::shared_ptr<const HIPContext> getContext() const { return _context; }

private:
    cie_hip_context ctx() const { return _context->get(); }
    void sync() const { _context->synchronize(); }
    std::shared_ptr<const HIPContext> _context;
};

static_assert(LinalgSpaceLike<HIPSpace<double>>);
static_assert(LinalgSpaceLike<HIPSpace<float>>);

// ---------------------------------------------------------------------------
// Device CSR matrix (owning) built from a host CSRView.
// ---------------------------------------------------------------------------

class HIPCSRMatrix {
public:
    template <class TIndex>
    HIPCSRMatrix(std::shared_ptr<const HIPContext> context, CSRView<const double, const TIndex> matrix,
                 int planes = CIE_HIP_PLANE_F64 | CIE_HIP_PLANE_F32)
        : _context(std::move(context)) {
        static_assert(std::is_same_v<std::remove_const_t<TIndex>, int>, "the HIP backend uses int32 indices");
        hipdetail::check(cie_hip_csr_create(_context->get(), std::int64_t(matrix.rowCount()),
                                            std::int64_t(matrix.columnCount()), std::int64_t(matrix.entries().size()),
                                            matrix.rowExtents().data(), matrix.columnIndices().data(),
                                            matrix.entries().data(), CIE_HIP_HOST, planes, &_handle),
                         "cie_hip_csr_create");
    }
    ~HIPCSRMatrix() { cie_hip_csr_destroy(_handle); }
    HIPCSRMatrix(const HIPCSRMatrix&) = delete;
    HIPCSRMatrix& operator=(const HIPCSRMatrix&) = delete;
    cie_hip_csr get() const noexcept { return _handle; }
    std::shared_ptr<const HIPContext> getContext() const { return _context; }
    std::size_t rowCount() const {
        std::int64_t rows = 0;
        cie_hip_csr_shape(_handle, &rows, nullptr, nullptr);
        return std::size_t(rows);
    }
    double autotune(int precision, int repetitions = 10, int kernels = 1 << CIE_HIP_SPMV_STAGED) {
        double seconds = 0.0;
        hipdetail::check(cie_hip_csr_autotune(_handle, precision, kernels, repetitions, &seconds), "autotune");
        return seconds;
    }

private:
    std::shared_ptr<const HIPContext> _context;
    cie_hip_csr _handle = nullptr;
};

// CSROperator / SYCLCSROperator counterpart: out = inScale * out + outScale * A * in.
template <class TIndex, class TValue, class TMatrixValue = TValue>
class HIPCSROperator : public LinearOperator<HIPSpace<TValue>> {
    using Space = HIPSpace<TValue>;

public:
    HIPCSROperator(std::shared_ptr<const HIPCSRMatrix> matrix, std::shared_ptr<Space> space)
        : _matrix(std::move(matrix)), _space(std::move(space)) {}

    void product(typename Space::Value inScale, typename Space::ConstVectorView in, typename Space::Value outScale,
                 typename Space::VectorView out) override {
        const cie_hip_context ctx = _space->getContext()->get();
        if constexpr (std::is_same_v<TValue, double>)
            hipdetail::check(cie_hip_csr_product_f64(ctx, _matrix->get(), inScale, in.get(), std::int64_t(in.size()),
                                                     outScale, out.get(), std::int64_t(out.size())),
                             "HIPCSROperator::product");
        else
            hipdetail::check(cie_hip_csr_product_f32(ctx, _matrix->get(), inScale, in.get(), std::int64_t(in.size()),
                                                     outScale, out.get(), std::int64_t(out.size())),
                             "HIPCSROperator::product");
        _space->getContext()->synchronize();
    }

private:
    std::shared_ptr<const HIPCSRMatrix> _matrix;
    std::shared_ptr<Space> _space;
};

// Jacobi preconditioner over HIPSpace (DiagonalOperator<SYCLSpace<T>> counterpart).
template <class T>
class HIPDiagonalOperator : public LinearOperator<HIPSpace<T>> {
public:
    HIPDiagonalOperator(typename HIPSpace<T>::Vector&& inverseDiagonal, std::shared_ptr<const HIPSpace<T>> space)
        : _inverse(std::move(inverseDiagonal)), _space(std::move(space)) {}

    void product(T inScale, typename HIPSpace<T>::ConstVectorView in, T outScale,
                 typename HIPSpace<T>::VectorView out) override {
        const cie_hip_context ctx = _space->getContext()->get();
        if constexpr (std::is_same_v<T, double>)
            hipdetail::check(cie_hip_diagonal_product_f64(ctx, _inverse.get(), inScale, in.get(), outScale, out.get(),
                                                          std::int64_t(out.size())),
                             "DiagonalOperator::product");
        else
            hipdetail::check(cie_hip_diagonal_product_f32(ctx, _inverse.get(), inScale, in.get(), outScale, out.get(),
                                                          std::int64_t(out.size())),
                             "DiagonalOperator::product");
        _space->getContext()->synchronize();
    }

    const typename HIPSpace<T>::Vector& inverseDiagonal() const noexcept { return _inverse; }

private:
    typename HIPSpace<T>::Vector _inverse;
    std::shared_ptr<const HIPSpace<T>> _space;
};

// makeDiagonalOperator (DiagonalOperator.cpp:56-88): throws on a zero or missing diagonal.
template <class T>
HIPDiagonalOperator<T> makeDiagonalOperator(const HIPCSRMatrix& matrix, std::shared_ptr<const HIPSpace<T>> space) {
    auto inverse = space->makeVector(matrix.rowCount());
    std::int64_t failedRow = -1;
    if constexpr (std::is_same_v<T, double>)
        hipdetail::check(cie_hip_inverse_diagonal_f64(space->getContext()->get(), matrix.get(), inverse.get(), &failedRow),
                         "makeDiagonalOperator");
    else
        hipdetail::check(cie_hip_inverse_diagonal_f32(space->getContext()->get(), matrix.get(), inverse.get(), &failedRow),
                         "makeDiagonalOperator");
    return HIPDiagonalOperator<T>(std::move(inverse), std::move(space));
}

#ifndef CIE_HIP_WITH_REFERENCE_HEADERS

// Generic (preconditioned) conjugate gradients over any LinalgSpaceLike, in the
// step order of ConjugateGradients_impl.hpp:49-201 (see that file for the
// reference; this is an independent implementation of the same algorithm).
template <LinalgSpaceLike TS>
class ConjugateGradients : public IterativeSolver<TS> {
    using Base = IterativeSolver<TS>;

public:
    using typename Base::Statistics;
    using Value = typename TS::Value;

    ConjugateGradients(std::shared_ptr<LinearOperator<TS>> lhs, std::shared_ptr<TS> space,
                       std::shared_ptr<LinearOperator<TS>> preconditioner = nullptr, Statistics settings = {},
                       int verbosity = 0)
        : Base(settings), _lhs(std::move(lhs)), _preconditioner(std::move(preconditioner)), _space(std::move(space)),
          _verbosity(verbosity) {}

    void product(Value inScale, typename TS::ConstVectorView in, Value outScale, typename TS::VectorView out) override {
        TS& space = *_space;
        const std::size_t n = space.size(out);
        if (space.size(in) != n)
            throw HIPError(CIE_HIP_ERROR_SHAPE, "incompatible vector sizes " + std::to_string(space.size(in)) + " != " +
                                                    std::to_string(n));
        const Statistics settings = this->getConfiguration();
        Statistics stats{0, std::numeric_limits<Value>::max(), std::numeric_limits<Value>::max()};
        std::string report;
        auto solution = space.makeVector(n);
        space.assign(space.view(solution), out);
        if (settings.iterationCount) {
            auto search = space.makeVector(n);
            auto searchProduct = space.makeVector(n);
            space.fill(space.view(searchProduct), 0);
            std::optional<typename TS::Vector> preconditioned;
            auto residual = space.makeVector(n);
            space.assign(space.view(residual), in);
            _lhs->product(1, space.view(std::as_const(solution)), -1, space.view(residual));
            Value residualNorm = space.innerProduct(space.view(std::as_const(residual)), space.view(std::as_const(residual)));
            Value preconditionedNorm = 0;
            const Value initialResidualNorm = std::sqrt(residualNorm);
            stats.absoluteResidual = initialResidualNorm;
            stats.relativeResidual = 1;
            if (less(stats.absoluteResidual, settings.absoluteResidual) || less(stats.relativeResidual, settings.relativeResidual)) {
                // :100-111 return before the output step of :190-198: `out` keeps its value.
                hipdetail::logReport(this->makeIterationReport(report, _verbosity, Base::ReportType::Termination,
                                                               stats, settings));
                this->report(stats);
                return;
            }
            if (_preconditioner) {
                preconditioned = space.makeVector(n);
                _preconditioner->product(0, space.view(std::as_const(residual)), 1, space.view(*preconditioned));
                preconditionedNorm = space.innerProduct(space.view(std::as_const(residual)),
                                                        space.view(std::as_const(*preconditioned)));
                space.assign(space.view(search), space.view(std::as_const(*preconditioned)));
            } else {
                preconditionedNorm = residualNorm;
                space.assign(space.view(search), space.view(std::as_const(residual)));
            }
            if (!(Value(0) < preconditionedNorm))
                throw HIPError(CIE_HIP_ERROR_BREAKDOWN, "preconditioned norm vanished in iteration 0 (r^T @ z = 0)");
            for (; stats.iterationCount < settings.iterationCount; ++stats.iterationCount) {
                _lhs->product(0, space.view(std::as_const(search)), 1, space.view(searchProduct));
                const Value denominator = space.innerProduct(space.view(std::as_const(search)),
                                                             space.view(std::as_const(searchProduct)));
                if (!(Value(0) < denominator))
                    throw HIPError(CIE_HIP_ERROR_NOT_SPD,
                                   "p^T @ A @ p vanished in iteration " + std::to_string(stats.iterationCount));
                Value searchScale = preconditionedNorm / denominator;
                space.add(space.view(solution), space.view(std::as_const(search)), searchScale);
                space.add(space.view(residual), space.view(std::as_const(searchProduct)), -searchScale);
                residualNorm = space.innerProduct(space.view(std::as_const(residual)), space.view(std::as_const(residual)));
                const Value previous = preconditionedNorm;
                if (preconditioned) {
                    _preconditioner->product(0, space.view(std::as_const(residual)), 1, space.view(*preconditioned));
                    preconditionedNorm = space.innerProduct(space.view(std::as_const(residual)),
                                                            space.view(std::as_const(*preconditioned)));
                } else {
                    preconditionedNorm = residualNorm;
                }
                searchScale = preconditionedNorm / previous;
                space.scale(space.view(search), searchScale);
                space.add(space.view(search),
                          preconditioned ? space.view(std::as_const(*preconditioned)) : space.view(std::as_const(residual)),
                          1);
                stats.absoluteResidual = std::sqrt(residualNorm);
                stats.relativeResidual = stats.absoluteResidual / initialResidualNorm;
                hipdetail::logReport(this->makeIterationReport(report, _verbosity, Base::ReportType::Iteration, stats,
                                                               settings));
                if (less(stats.absoluteResidual, settings.absoluteResidual) || less(stats.relativeResidual, settings.relativeResidual))
                    break;
            }
        }
        finish(inScale, outScale, out, solution);
        this->report(stats);
    }

private:
    // cieutils/packages/maths/impl/Comparison_impl.hpp:49-63 with the default tolerances.
    static bool less(Value left, Value right) {
        const Value norm = std::min(std::abs(left) + std::abs(right), std::numeric_limits<Value>::max());
        const bool equal = std::abs(left - right) < std::max(std::numeric_limits<Value>::min(),
                                                             std::numeric_limits<Value>::epsilon() * norm);
        return left < right && !equal;
    }

    void finish(Value inScale, Value outScale, typename TS::VectorView out, const typename TS::Vector& solution) {
        TS& space = *_space;
        if (inScale == Value(1)) {
            space.add(out, space.view(solution), outScale);
        } else if (inScale == Value(0)) {
            space.assign(out, space.view(solution));
            space.scale(out, outScale);
        } else {
            space.scale(out, inScale);
            space.add(out, space.view(solution), outScale);
        }
    }

    std::shared_ptr<LinearOperator<TS>> _lhs, _preconditioner;
    std::shared_ptr<TS> _space;
    int _verbosity;
};

#endif  // !CIE_HIP_WITH_REFERENCE_HEADERS

// Fused Jacobi-PCG (cie_hip_pcg_f64) or mixed refined PCG (cie_hip_pcg_mixed)
// behind the same LinearOperator / IterativeSolver face: product(0, b, 1, x)
// solves A x = b with x as the initial guess, getStats() as for
// ConjugateGradients. The preconditioner is always Jacobi. With verbosity >= 2
// the solve records each loop trip's residual on the device and logs the
// reference's per-iteration reports from it afterwards (the start check's
// Termination report, ConjugateGradients_impl.hpp:100-111, or one report per
// trip, :175-182; the last trip's from the solver's statistics).
class FusedConjugateGradients : public IterativeSolver<HIPSpace<double>> {
    using Base = IterativeSolver<HIPSpace<double>>;
    using Space = HIPSpace<double>;

public:
    using typename Base::Statistics;
    static constexpr std::size_t kHistoryCapacity = std::size_t(1) << 20;  // trips recorded for the reports

    FusedConjugateGradients(std::shared_ptr<const HIPCSRMatrix> matrix, std::shared_ptr<Space> space,
                            Statistics settings = {}, bool mixedPrecision = false,
                            double refinementFactor = CIE_HIP_DEFAULT_REFINEMENT_FACTOR, int verbosity = 0)
        : Base(settings), _matrix(std::move(matrix)), _space(std::move(space)), _mixed(mixedPrecision),
          _refinementFactor(refinementFactor), _verbosity(verbosity) {}

    void product(double inScale, Space::ConstVectorView in, double outScale, Space::VectorView out) override {
        const Statistics settings = this->getConfiguration();
        const cie_hip_statistics s{settings.iterationCount, settings.absoluteResidual, settings.relativeResidual};
        const cie_hip_context ctx = _space->getContext()->get();
        std::optional<Space::Vector> history;
        if (_verbosity >= 2 && settings.iterationCount) {
            history = _space->makeVector(std::min<std::size_t>(settings.iterationCount, kHistoryCapacity));
            hipdetail::check(cie_hip_context_set_residual_history(ctx, history->get(), std::int64_t(history->size())),
                             "cie_hip_context_set_residual_history");
        }
        int status;
        if (_mixed) {
            const cie_hip_mixed_settings mixed{s, _refinementFactor, 0, 0};
            status = cie_hip_pcg_mixed(ctx, _matrix->get(), inScale, in.get(), outScale, out.get(),
                                       std::int64_t(out.size()), &mixed, &_info);
        } else {
            status = cie_hip_pcg_f64(ctx, _matrix->get(), inScale, in.get(), outScale, out.get(),
                                     std::int64_t(out.size()), &s, 1, &_info);
        }
        if (history) {
            const std::string message = status == CIE_HIP_OK ? std::string() : cie_hip_last_error();
            cie_hip_context_set_residual_history(ctx, nullptr, 0);
            logReports(*history, s);
            if (status != CIE_HIP_OK) throw HIPError(status, "FusedConjugateGradients::product: " + message);
        }
        hipdetail::check(status, "FusedConjugateGradients::product");
        this->report(Statistics{std::size_t(_info.stats.iterationCount), _info.stats.absoluteResidual,
                                _info.stats.relativeResidual});
    }

    const cie_hip_solve_info& info() const noexcept { return _info; }

private:
    void logReports(const Space::Vector& history, const cie_hip_statistics& settings) const {
        const std::size_t rows = std::size_t(_info.reportedIterations);
        if (rows == 0) {
            if (_info.converged)  // the start check ended the solve
                hipdetail::logReport(hipdetail::iterationReport(_verbosity, CIE_HIP_REPORT_TERMINATION, 0,
                                                                _info.initialResidual, 1.0, settings));
            return;
        }
        std::vector<double> recorded(std::min(rows, history.size()));
        _space->assign(std::span<double>(recorded), _space->view(history));  // the first recorded.size() entries
        for (std::size_t trip = 0; trip < rows; ++trip) {
            if (trip + 1 == rows) {
                hipdetail::logReport(hipdetail::iterationReport(_verbosity, CIE_HIP_REPORT_ITERATION, trip,
                                                                _info.stats.absoluteResidual,
                                                                _info.stats.relativeResidual, settings));
            } else if (trip < recorded.size()) {
                hipdetail::logReport(hipdetail::iterationReport(_verbosity, CIE_HIP_REPORT_ITERATION, trip,
                                                                recorded[trip], recorded[trip] / _info.initialResidual,
                                                                settings));
            }
        }
    }

    std::shared_ptr<const HIPCSRMatrix> _matrix;
    std::shared_ptr<Space> _space;
    bool _mixed;
    double _refinementFactor;
    int _verbosity;
    cie_hip_solve_info _info{};
};

}  // namespace cie::linalg
