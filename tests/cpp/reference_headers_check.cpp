// Compile-only check of the drop-in boundary: include/cie_hip/linalg.hpp built
// inside the reference tree (CIE_HIP_WITH_REFERENCE_HEADERS, include paths
// libraries/linalg and libraries/cieutils of the reference), so that the
// HIP space and operators meet the reference's own concept and base classes:
//   LinalgSpaceLike            linalg/packages/solvers/inc/LinalgSpace.hpp:10-35
//   LinearOperator<TSpace>     linalg/packages/solvers/inc/LinearOperator.hpp:10-34
//   IterativeSolver<TSpace>    linalg/packages/solvers/inc/IterativeSolver.hpp:18-56
//   ConjugateGradients<TSpace> linalg/packages/solvers/inc/ConjugateGradients.hpp:15-46
// HIPCSROperator and HIPDiagonalOperator declare product(...) `override`, so
// a signature that drifted from the reference's virtual would not compile.
// Built with -fsyntax-only by tests/test_reference_boundary.py.
#include "cie_hip/linalg.hpp"

#include <memory>
#include <type_traits>

#ifndef CIE_HIP_WITH_REFERENCE_HEADERS
#error "build with -DCIE_HIP_WITH_REFERENCE_HEADERS against the reference's include paths"
#endif

namespace check {
using namespace cie::linalg;

static_assert(LinalgSpaceLike<HIPSpace<double>>);
static_assert(LinalgSpaceLike<HIPSpace<float>>);
static_assert(std::is_same_v<HIPSpace<double>::Value, double>);
static_assert(std::is_same_v<HIPSpace<float>::Value, float>);

static_assert(std::is_base_of_v<LinearOperator<HIPSpace<double>>, HIPCSROperator<int, double>>);
static_assert(std::is_base_of_v<LinearOperator<HIPSpace<float>>, HIPCSROperator<int, float>>);
static_assert(std::is_base_of_v<LinearOperator<HIPSpace<double>>, HIPDiagonalOperator<double>>);
static_assert(std::is_base_of_v<LinearOperator<HIPSpace<float>>, HIPDiagonalOperator<float>>);

// The reference's own solver template over the HIP space, both instantiations
// the reference ships for its spaces (ConjugateGradients.cpp:10-19).
static_assert(std::is_base_of_v<IterativeSolver<HIPSpace<double>>, ConjugateGradients<HIPSpace<double>>>);
static_assert(std::is_base_of_v<IterativeSolver<HIPSpace<float>>, ConjugateGradients<HIPSpace<float>>>);
static_assert(std::is_base_of_v<LinearOperator<HIPSpace<double>>, FusedConjugateGradients>);

// The driver's construction sequence (poisson2D/solver.cpp:238-288 with the
// SYCL types replaced): space, operator, Jacobi, CG with the driver's settings.
[[maybe_unused]] void driver(std::shared_ptr<const HIPContext> context, const CSRView<const double, const int>& view,
                             typename HIPSpace<double>::ConstVectorView b, typename HIPSpace<double>::VectorView x) {
    auto space = std::make_shared<HIPSpace<double>>(context);
    auto matrix = std::make_shared<const HIPCSRMatrix>(context, view);
    auto lhs = std::make_shared<HIPCSROperator<int, double>>(matrix, space);
    auto jacobi = std::make_shared<HIPDiagonalOperator<double>>(
        makeDiagonalOperator<double>(*matrix, std::shared_ptr<const HIPSpace<double>>(space)));
    using Solver = ConjugateGradients<HIPSpace<double>>;
    Solver solver(lhs, space, jacobi, Solver::Statistics{5000, 1e-6, 1e-6}, 3);
    solver.product(0, b, 1, x);
}
}  // namespace check

int main() { return 0; }
