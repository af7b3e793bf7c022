/* Structured test/benchmark systems for the HIP sparse solver (host only).
 *
 * No counterpart in the reference's solve path: the reference assembles its
 * matrices with the fem library (fem/packages/graph/src/Assembler.cpp:55-175,
 * graph/impl/Assembler_impl.hpp:156-194). These entry points build the same
 * kind of CSR matrix (int32 row extents and sorted column indices, FP64
 * entries) directly for structured meshes, plus the seeded random vectors the
 * workloads are defined with. Every function returns 0 on success and nonzero
 * on failure, with a message in cie_sys_last_error().
 */
#ifndef CIE_HIP_SYSTEMS_H
#define CIE_HIP_SYSTEMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Tensor-product Laplace stiffness on the unit square/cube. */
#define CIE_SYS_BASIS_LAGRANGE 0
#define CIE_SYS_BASIS_LEGENDRE 1
typedef struct cie_sys_lagrange_spec {
    int32_t dim;       /* 2 or 3 */
    int32_t order;     /* polynomial order p */
    int32_t elements;  /* elements per dimension */
    int32_t eliminate; /* 1: drop Dirichlet boundary rows and columns */
    int32_t basis;     /* CIE_SYS_BASIS_LAGRANGE: nodes cos(i pi / p) per element (poisson2D/mesh.cpp:73-77);
                          CIE_SYS_BASIS_LEGENDRE: integrated Legendre polynomials, the hierarchical basis of
                          fem/packages/maths/src/LegendrePolynomial.cpp (poisson2D's default) */
} cie_sys_lagrange_spec;

const char* cie_sys_last_error(void);

/* SplitMix64 raw outputs, U(-1,1) draws and N(0, sigma^2) Box-Muller draws. */
int cie_sys_splitmix64(uint64_t seed, int64_t count, uint64_t* out);
int cie_sys_uniform(uint64_t seed, int64_t count, double* out);
int cie_sys_normal(uint64_t seed, double sigma, int64_t count, double* out);

/* 2D 5-point Laplacian (4 on the diagonal, -1 off it), m x m interior grid. */
int cie_sys_poisson5_size(int m, int64_t* rows, int64_t* nnz);
int cie_sys_poisson5_fill(int m, int32_t* rowExtents, int32_t* columnIndices, double* entries);

/* Lagrange stiffness; conductivity holds one k_e per element (NULL: all 1).
 * threads <= 0 uses all cores, capped by OMP_NUM_THREADS. */
int cie_sys_lagrange_size(const cie_sys_lagrange_spec* spec, int64_t* rows, int64_t* nnz,
                          int64_t* elementCount, int32_t* nodesPerElement);
int cie_sys_lagrange_fill(const cie_sys_lagrange_spec* spec, const double* conductivity,
                          int32_t* rowExtents, int32_t* columnIndices, double* entries, int threads);
/* Element matrix for k_e = 1, nodesPerElement^2 entries, row-major. */
int cie_sys_lagrange_element_matrix(const cie_sys_lagrange_spec* spec, double* elementMatrix);
/* Row index of every element node (elementCount x nodesPerElement), -1 if eliminated. */
int cie_sys_lagrange_element_dofs(const cie_sys_lagrange_spec* spec, int32_t* dofs);
/* Gradient moments of the reference element [-1, 1]^dim: moments[i][j][a][b] =
 * sum_q w_q dN_a/dxi_i dN_b/dxi_j over (order + 1)^dim Gauss-Legendre points,
 * dim^2 x npe^2 values, element nodes in the element matrix's order. With an
 * affine map x = x0 + J xi the element stiffness is
 * |det J| k sum_ij (J^-1 J^-T)_ij moments[i][j] (cie_hip_element_stiffness). */
int cie_sys_lagrange_gradient_moments(const cie_sys_lagrange_spec* spec, double* moments);
/* Ansatz mask of every row for the integrated Legendre basis: the largest
 * polynomial degree among the DOF's 1D factors, at least 1 (makeAnsatzMask,
 * fem/packages/graph/src/connectivity.cpp:62-90 and Assembler.cpp:194-213). */
int cie_sys_lagrange_ansatz_mask(const cie_sys_lagrange_spec* spec, uint16_t* mask);
/* CSR slot of each (row, column) pair, -1 if absent. */
int cie_sys_csr_slots(int64_t rows, const int32_t* rowExtents, const int32_t* columnIndices,
                      int64_t pairCount, const int32_t* pairRows, const int32_t* pairColumns,
                      int64_t* slots);

#ifdef __cplusplus
}
#endif

#endif /* CIE_HIP_SYSTEMS_H */
