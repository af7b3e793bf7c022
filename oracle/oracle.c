/* CPU oracle for the sparse PCG path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's threaded CPU solve path, used by
 * the test-suite as the checker and by the benchmark as the CPU baseline.
 * Nothing in the product library links, loads or calls this file.
 *
 * What each function follows (paths relative to the reference repository):
 *   oracle_csr_product      linalg/packages/solvers/src/CSROperator.cpp:33-111
 *                           (per-row sequential sum in CSR entry order, :46-58;
 *                           out = inScale*out + outScale*A*in, :78-110).  The
 *                           reference's inScale = outScale = -1 branch (:98)
 *                           computes A*in - out; this restatement follows the
 *                           documented contract (LinearOperator.hpp:15-33)
 *                           instead, -A*in - out.  CG never uses that branch.
 *   oracle_inner_product    solvers/src/DefaultSpace.cpp:50-77.  One thread:
 *                           std::inner_product order (:71-75).  Several: one
 *                           contiguous partition per thread, partials summed
 *                           in thread order (the reference sums per-thread
 *                           partials in hash-map order, :66-68, which is not
 *                           reproducible; this one is).
 *   oracle_scale / _uniform / oracle_add / _assign / _fill
 *                           DefaultSpace.cpp:80-187 (same special cases).
 *   oracle_inverse_diagonal DiagonalOperator.cpp:56-88 (CPU: throws on a zero
 *                           or missing diagonal -> status codes 5 / 6).
 *   oracle_diagonal_product DiagonalOperator.cpp:37-53.
 *   oracle_pcg              solvers/impl/ConjugateGradients_impl.hpp:49-201 with
 *                           the Jacobi DiagonalOperator; convergence test of
 *                           cieutils/packages/maths/impl/Comparison_impl.hpp:32-64
 *                           (abs tol DBL_MIN, rel tol DBL_EPSILON).
 *   oracle_assemble         fem/packages/graph/impl/Assembler_impl.hpp:156-194
 *                           (Assembler::addContribution, Serial branch): for
 *                           every element in order, every (local row, local
 *                           column) pair is located by lower_bound in the CSR
 *                           row and added into its slot; a pair missing from
 *                           the pattern is an out-of-range error (status 2).
 *                           The contribution is scale_e * K_e (scale NULL: K_e),
 *                           the product of LinearIsotropicStiffnessIntegrand_impl.hpp:74.
 *   oracle_assemble_vector  Assembler_impl.hpp:197-220 (the vector overload).
 *   oracle_element_stiffness
 *                           LinearIsotropicStiffnessIntegrand_impl.hpp:74
 *                           (K = modulus D^T D) integrated over an affine cell
 *                           x = x0 + J xi (IntegrandProcessor_impl.hpp:519-680):
 *                           |det J| k sum_ij (J^-1 J^-T)_ij M_ij with the
 *                           reference gradient moments M_ij.
 *   oracle_make_reordering  linalg/packages/utilities/src/reorder.cpp:23-173
 *                           (makeReordering, serial path): Cuthill-McKee from
 *                           the first row of least degree (row length); every
 *                           visited row queues its unvisited columns in column
 *                           order, and that batch is then sorted by degree with
 *                           std::sort (:130-135). std::sort is not stable, so
 *                           ties follow the C++ library's algorithm; this is
 *                           GCC libstdc++'s introsort (bits/stl_algo.h
 *                           std::__sort: quicksort to runs of 16 with
 *                           median-of-three pivots, heapsort below 2*lg(n)
 *                           depth, then one insertion-sort pass), restated in
 *                           degree_sort below. When the queue empties, the
 *                           first unvisited row of least degree starts the next
 *                           component (:141-158). Reverse Cuthill-McKee reverses
 *                           the map (:168-170). map[new] = old.
 *   oracle_reorder_csr      reorder.cpp:215-365 (reorder<..., Reverse>): row
 *                           i of the result is row mapView[i] of the input with
 *                           columns renumbered by the inverse and re-sorted;
 *                           reverse = 1 reads the map as old -> new (:273-274).
 *   oracle_reorder_vector   reorder.cpp:368-394: out[i] = in[map[i]], or
 *                           out[map[i]] = in[i] when reverse.
 *   oracle_masked_csr_product
 *                           solvers/src/MaskedCSROperator.cpp:37-118: rows
 *                           and columns with mask >= threshold are left out
 *                           of the row sums (a masked row sums to 0).
 *   oracle_masked_identity  solvers/src/MaskedIdentityOperator.cpp:14-70 (CPU):
 *                           op(out, mask < threshold ? in : 0).
 *   oracle_jacobi           solvers/src/JacobiOperator.cpp:36-62 over a CSR or
 *                           masked CSR lhs with the DiagonalOperator of the
 *                           full matrix (DiagonalOperator.cpp:37-53).
 *   oracle_masked_jacobi    solvers/src/MaskedJacobiOperator.cpp:46-161, the
 *                           threaded path (previous = out before each sweep,
 *                           :103); inScale != 0 with more than one sweep is
 *                           not implemented there (:64-65, status 10).
 *   oracle_approximate_substitution
 *                           solvers/src/ApproximateSubstitution.cpp:27-111,
 *                           threaded path as above (:65).
 *   oracle_pmg              executables/fem_examples/src/poisson2D/solver.cpp:111-232
 *                           (solveMultigrid): the coarsest grid (ansatz mask
 *                           < 2) is Jacobi-PCG on the masked matrix to a
 *                           relative residual of 0.5 (at most 5000 steps); grid
 *                           k = 2..p is 6 Jacobi sweeps (relaxation 2/3) on
 *                           the masked matrix of mask < k + 1 after the mask;
 *                           a cycle visits the grids from the finest down and
 *                           back up from the second, each time restricting
 *                           the residual, adding the grid's update to the
 *                           solution and subtracting the masked product of the
 *                           update from the residual, until
 *                           ||r|| <= 1e-6 ||b|| (the relative tolerance is a
 *                           parameter here, and so is a cycle cap). The
 *                           residual starts as b, as in the driver, so x is
 *                           expected to hold zeros.
 *   oracle_element_gradient dJ/dk_e = -lambda_e^T K_e0 u_e (no reference
 *                           counterpart; K_e = k_e K_e0 as in
 *                           fem/.../LinearIsotropicStiffnessIntegrand_impl.hpp:74).
 *
 * Status codes match include/cie_hip.h: 0 ok, 2 shape, 3 p^T A p <= 0,
 * 4 r^T z <= 0, 5 zero diagonal, 6 missing diagonal, 10 not implemented.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* A persistent thread pool with contiguous static partitions.              */
/* ------------------------------------------------------------------------ */

typedef void (*task_fn)(void* args, int64_t begin, int64_t end, int worker);

typedef struct {
    int size;
    pthread_t* threads;
    pthread_barrier_t start, finish;
    task_fn task;
    void* args;
    int64_t count;
    int stop;
} pool_t;

typedef struct { pool_t* pool; int worker; } worker_arg_t;

static void partition(int64_t count, int parts, int worker, int64_t* begin, int64_t* end) {
    const int64_t chunk = count / parts, extra = count % parts;
    *begin = worker * chunk + (worker < extra ? worker : extra);
    *end = *begin + chunk + (worker < extra ? 1 : 0);
}

static void* worker_main(void* raw) {
    worker_arg_t* arg = (worker_arg_t*)raw;
    pool_t* pool = arg->pool;
    for (;;) {
        pthread_barrier_wait(&pool->start);
        if (pool->stop) break;
        int64_t begin, end;
        partition(pool->count, pool->size, arg->worker, &begin, &end);
        if (begin < end) pool->task(pool->args, begin, end, arg->worker);
        pthread_barrier_wait(&pool->finish);
    }
    free(arg);
    return NULL;
}

static pool_t* pool_create(int size) {
    pool_t* pool = (pool_t*)calloc(1, sizeof(pool_t));
    pool->size = size;
    if (size > 1) {
        pool->threads = (pthread_t*)calloc((size_t)size, sizeof(pthread_t));
        pthread_barrier_init(&pool->start, NULL, (unsigned)size);
        pthread_barrier_init(&pool->finish, NULL, (unsigned)size);
        for (int w = 1; w < size; ++w) {
            worker_arg_t* arg = (worker_arg_t*)malloc(sizeof(worker_arg_t));
            arg->pool = pool;
            arg->worker = w;
            pthread_create(&pool->threads[w], NULL, worker_main, arg);
        }
    }
    return pool;
}

static void pool_run(pool_t* pool, task_fn task, void* args, int64_t count) {
    if (!pool || pool->size <= 1) {
        task(args, 0, count, 0);
        return;
    }
    pool->task = task;
    pool->args = args;
    pool->count = count;
    pthread_barrier_wait(&pool->start);
    int64_t begin, end;
    partition(count, pool->size, 0, &begin, &end);
    if (begin < end) task(args, begin, end, 0);
    pthread_barrier_wait(&pool->finish);
}

static void pool_destroy(pool_t* pool) {
    if (!pool) return;
    if (pool->size > 1) {
        pool->stop = 1;
        pthread_barrier_wait(&pool->start);
        for (int w = 1; w < pool->size; ++w) pthread_join(pool->threads[w], NULL);
        pthread_barrier_destroy(&pool->start);
        pthread_barrier_destroy(&pool->finish);
        free(pool->threads);
    }
    free(pool);
}

/* ------------------------------------------------------------------------ */
/* SpMV: CSROperator::product                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    const int32_t *rowptr, *cols;
    const double* vals;
    double inScale, outScale;
    const double* in;
    double* out;
    const uint16_t* mask; /* NULL: every row and column; else only those with mask < threshold */
    uint16_t threshold;
} spmv_args_t;

/* out = op(out, right) for one (inScale, outScale) class: CSROperator.cpp:78-110
 * and the same table in MaskedCSROperator.cpp:85-118, MaskedIdentityOperator.cpp:40-72,
 * MaskedJacobiOperator.cpp:119-152, ApproximateSubstitution.cpp:78-111. For
 * inScale = outScale = -1 those compute right - out; this follows the
 * documented contract (LinearOperator.hpp:15-33), -right - out. */
static void apply_op(double inScale, double outScale, double* o, double right) {
    if (inScale == 1.0) {
        if (outScale == 1.0) *o += right;
        else if (outScale == -1.0) *o -= right;
        else *o += outScale * right;
    } else if (inScale == 0.0) {
        if (outScale == 1.0) *o = right;
        else if (outScale == -1.0) *o = -right;
        else *o = outScale * right;
    } else if (inScale == -1.0) {
        if (outScale == 1.0) *o = right - *o;
        else if (outScale == -1.0) *o = -right - *o;
        else *o = outScale * right - *o;
    } else {
        if (outScale == 1.0) *o = inScale * *o + right;
        else if (outScale == -1.0) *o = inScale * *o - right;
        else *o = inScale * *o + outScale * right;
    }
}

static void spmv_task(void* raw, int64_t begin, int64_t end, int worker) {
    (void)worker;
    const spmv_args_t* a = (const spmv_args_t*)raw;
    for (int64_t row = begin; row < end; ++row) {
        double contribution = 0.0;
        if (!a->mask) {
            for (int32_t k = a->rowptr[row]; k < a->rowptr[row + 1]; ++k)
                contribution += a->vals[k] * a->in[a->cols[k]];
        } else if (a->mask[row] < a->threshold) {  /* MaskedCSROperator.cpp:53-66: masked entries are skipped */
            for (int32_t k = a->rowptr[row]; k < a->rowptr[row + 1]; ++k)
                if (a->mask[a->cols[k]] < a->threshold) contribution += a->vals[k] * a->in[a->cols[k]];
        }
        apply_op(a->inScale, a->outScale, a->out + row, contribution);
    }
}

static void spmv(pool_t* pool, int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                 double inScale, const double* in, double outScale, double* out) {
    spmv_args_t a = {rowptr, cols, vals, inScale, outScale, in, out, NULL, 0};
    pool_run(pool, spmv_task, &a, rows);
}

/* A CSR operator, optionally masked (MaskedCSROperator). */
typedef struct {
    int64_t rows;
    const int32_t *rowptr, *cols;
    const double* vals;
    const uint16_t* mask;
    uint16_t threshold;
} csr_operator_t;

static void apply_csr(pool_t* pool, const csr_operator_t* A, double inScale, const double* in, double outScale,
                      double* out) {
    spmv_args_t a = {A->rowptr, A->cols, A->vals, inScale, outScale, in, out, A->mask, A->threshold};
    pool_run(pool, spmv_task, &a, A->rows);
}

/* ------------------------------------------------------------------------ */
/* DefaultSpace operations                                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    const double *left, *right;
    double* partials;
} dot_args_t;

static void dot_task(void* raw, int64_t begin, int64_t end, int worker) {
    dot_args_t* a = (dot_args_t*)raw;
    double sum = 0.0;
    for (int64_t i = begin; i < end; ++i) sum += a->left[i] * a->right[i];
    a->partials[worker] = sum;
}

static double dot(pool_t* pool, const double* left, const double* right, int64_t n) {
    if (!pool || pool->size <= 1) {
        double sum = 0.0;
        for (int64_t i = 0; i < n; ++i) sum += left[i] * right[i];
        return sum;
    }
    double partials[256];
    memset(partials, 0, sizeof(partials));
    dot_args_t a = {left, right, partials};
    pool_run(pool, dot_task, &a, n);
    double sum = 0.0;
    for (int w = 0; w < pool->size; ++w) sum += partials[w];
    return sum;
}

enum { OP_SCALE, OP_SCALE_UNIFORM, OP_ADD, OP_ASSIGN, OP_FILL };

typedef struct {
    int op;
    double* target;
    const double* source;
    double value;
} vec_args_t;

static void vec_task(void* raw, int64_t begin, int64_t end, int worker) {
    (void)worker;
    const vec_args_t* a = (const vec_args_t*)raw;
    double* t = a->target;
    const double* s = a->source;
    const double v = a->value;
    switch (a->op) {
    case OP_SCALE: /* DefaultSpace.cpp:100-105 */
        if (v == 1.0) for (int64_t i = begin; i < end; ++i) t[i] = t[i] * s[i];
        else if (v == -1.0) for (int64_t i = begin; i < end; ++i) t[i] = -t[i] * s[i];
        else for (int64_t i = begin; i < end; ++i) t[i] = v * t[i] * s[i];
        break;
    case OP_SCALE_UNIFORM: /* :109-122 */
        for (int64_t i = begin; i < end; ++i) t[i] *= v;
        break;
    case OP_ADD: /* :145-150 */
        if (v == 1.0) for (int64_t i = begin; i < end; ++i) t[i] = t[i] + s[i];
        else if (v == -1.0) for (int64_t i = begin; i < end; ++i) t[i] = t[i] - s[i];
        else for (int64_t i = begin; i < end; ++i) t[i] = t[i] + v * s[i];
        break;
    case OP_ASSIGN: /* :154-169 */
        for (int64_t i = begin; i < end; ++i) t[i] = s[i];
        break;
    case OP_FILL: /* :172-187 */
        for (int64_t i = begin; i < end; ++i) t[i] = v;
        break;
    }
}

static void vec(pool_t* pool, int op, double* target, const double* source, double value, int64_t n) {
    vec_args_t a = {op, target, source, value};
    pool_run(pool, vec_task, &a, n);
}

/* Comparison_impl.hpp:49-63 with the default tolerances (DBL_MIN, DBL_EPSILON). */
static int float_equal(double left, double right) {
    double norm = fabs(left) + fabs(right);
    if (norm > DBL_MAX) norm = DBL_MAX;
    const double tolerance = DBL_EPSILON * norm > DBL_MIN ? DBL_EPSILON * norm : DBL_MIN;
    return fabs(left - right) < tolerance;
}

static int float_less(double left, double right) { return left < right && !float_equal(left, right); }

/* ------------------------------------------------------------------------ */
/* Exported entry points                                                    */
/* ------------------------------------------------------------------------ */

static int resolve_threads(int threads) {
    if (threads <= 0) threads = 1;
    if (threads > 256) threads = 256;
    return threads;
}

int oracle_csr_product(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                       double inScale, const double* in, double outScale, double* out, int threads) {
    pool_t* pool = pool_create(resolve_threads(threads));
    spmv(pool, rows, rowptr, cols, vals, inScale, in, outScale, out);
    pool_destroy(pool);
    return 0;
}

double oracle_inner_product(int64_t n, const double* left, const double* right, int threads) {
    pool_t* pool = pool_create(resolve_threads(threads));
    const double result = dot(pool, left, right, n);
    pool_destroy(pool);
    return result;
}

int oracle_vector_op(int op, double* target, const double* source, double value, int64_t n, int threads) {
    if (op < OP_SCALE || op > OP_FILL) return 1;
    pool_t* pool = pool_create(resolve_threads(threads));
    vec(pool, op, target, source, value, n);
    pool_destroy(pool);
    return 0;
}

int oracle_inverse_diagonal(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                            double* inverse, int64_t* failedRow) {
    for (int64_t row = 0; row < rows; ++row) {
        int found = 0;
        double diagonal = 0.0;
        for (int32_t k = rowptr[row]; k < rowptr[row + 1]; ++k)
            if (cols[k] == row) {
                found = 1;
                diagonal = vals[k];
                break;
            }
        if (!found || diagonal == 0.0) {
            if (failedRow) *failedRow = row;
            return found ? 5 : 6;
        }
        inverse[row] = 1.0 / diagonal;
    }
    return 0;
}

static void diagonal_product(pool_t* pool, const double* inverse, double inScale, const double* in,
                             double outScale, double* out, double* buffer, int64_t n) {
    if (inScale == 0.0) {
        vec(pool, OP_ASSIGN, out, in, 0.0, n);
        vec(pool, OP_SCALE, out, inverse, outScale, n);
    } else {
        vec(pool, OP_ASSIGN, buffer, in, 0.0, n);
        vec(pool, OP_SCALE, buffer, inverse, outScale, n);
        if (inScale != 1.0) vec(pool, OP_SCALE_UNIFORM, out, NULL, inScale, n);
        vec(pool, OP_ADD, out, buffer, 1.0, n);
    }
}

int oracle_diagonal_product(const double* inverse, double inScale, const double* in, double outScale,
                            double* out, int64_t n) {
    double* buffer = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    diagonal_product(NULL, inverse, inScale, in, outScale, out, buffer, n);
    free(buffer);
    return 0;
}

/* Statistics: iterationCount, absoluteResidual, relativeResidual (IterativeSolver.hpp:21-25). */
typedef struct {
    uint64_t iterationCount;
    double absoluteResidual;
    double relativeResidual;
} oracle_statistics_t;

/* ConjugateGradients<DefaultSpace<double>>::product(inScale, b, outScale, x) with
 * a Jacobi (DiagonalOperator) preconditioner when `precondition` is nonzero.
 * history (optional, capacity historyCapacity) receives the absolute residual
 * after every loop trip. */
/* ConjugateGradients<TSpace>::product(inScale, b, outScale, x) over the
 * operator A with the Jacobi DiagonalOperator `inverse` (NULL: no
 * preconditioner); line numbers are ConjugateGradients_impl.hpp's. */
static int pcg_core(pool_t* pool, const csr_operator_t* A, const double* inverse, double inScale, const double* b,
                    double outScale, double* x, uint64_t maxIterations, double absoluteTolerance,
                    double relativeTolerance, oracle_statistics_t* stats, double* history, int64_t historyCapacity) {
    const int64_t n = A->rows;
    const int precondition = inverse != NULL;
    int status = 0;
    double* solution = (double*)malloc((size_t)n * sizeof(double));
    double* search = (double*)malloc((size_t)n * sizeof(double));
    double* product = (double*)malloc((size_t)n * sizeof(double));
    double* residual = (double*)malloc((size_t)n * sizeof(double));
    double* preconditioned = precondition ? (double*)malloc((size_t)n * sizeof(double)) : NULL;
    double* buffer = precondition ? (double*)malloc((size_t)n * sizeof(double)) : NULL;

    oracle_statistics_t s = {0, DBL_MAX, DBL_MAX};

    vec(pool, OP_ASSIGN, solution, x, 0.0, n);                                   /* :74-75 */
    if (maxIterations) {
        vec(pool, OP_FILL, product, NULL, 0.0, n);                               /* :82 */
        vec(pool, OP_ASSIGN, residual, b, 0.0, n);                               /* :86-88 */
        apply_csr(pool, A, 1.0, solution, -1.0, residual);

        double residualNorm = dot(pool, residual, residual, n);                 /* :93 */
        double preconditionedNorm = 0.0;
        const double initialResidualNorm = sqrt(residualNorm);
        s.absoluteResidual = initialResidualNorm;
        s.relativeResidual = 1.0;
        /* :100-111: the early exit returns before the output step of :190-198,
         * so `x` keeps its value whatever inScale and outScale are. */
        if (float_less(s.absoluteResidual, absoluteTolerance) || float_less(s.relativeResidual, relativeTolerance))
            goto done;

        if (precondition) {                                                       /* :114-123 */
            diagonal_product(pool, inverse, 0.0, residual, 1.0, preconditioned, buffer, n);
            preconditionedNorm = dot(pool, residual, preconditioned, n);
            vec(pool, OP_ASSIGN, search, preconditioned, 0.0, n);
        } else {                                                                  /* :124-132 */
            preconditionedNorm = residualNorm;
            vec(pool, OP_ASSIGN, search, residual, 0.0, n);
        }
        if (!(0.0 < preconditionedNorm)) { status = 4; goto done; }

        for (; s.iterationCount < maxIterations; ++s.iterationCount) {           /* :134-186 */
            apply_csr(pool, A, 0.0, search, 1.0, product);
            double searchScale = preconditionedNorm;
            const double denominator = dot(pool, search, product, n);
            if (!(0.0 < denominator)) { status = 3; goto done; }
            searchScale /= denominator;

            vec(pool, OP_ADD, solution, search, searchScale, n);                 /* :154-155 */
            vec(pool, OP_ADD, residual, product, -searchScale, n);

            residualNorm = dot(pool, residual, residual, n);                     /* :158 */
            const double previous = preconditionedNorm;
            if (precondition) {
                diagonal_product(pool, inverse, 0.0, residual, 1.0, preconditioned, buffer, n);
                preconditionedNorm = dot(pool, residual, preconditioned, n);
            } else {
                preconditionedNorm = residualNorm;
            }
            searchScale = preconditionedNorm / previous;                          /* :168-170 */
            vec(pool, OP_SCALE_UNIFORM, search, NULL, searchScale, n);
            vec(pool, OP_ADD, search, precondition ? preconditioned : residual, 1.0, n);

            s.absoluteResidual = sqrt(residualNorm);                              /* :173-174 */
            s.relativeResidual = s.absoluteResidual / initialResidualNorm;
            if (history && (int64_t)s.iterationCount < historyCapacity) history[s.iterationCount] = s.absoluteResidual;
            if (float_less(s.absoluteResidual, absoluteTolerance) || float_less(s.relativeResidual, relativeTolerance))
                break;                                                            /* :184-185 */
        }
    }

    /* :190-198 */
    if (inScale == 1.0) {
        vec(pool, OP_ADD, x, solution, outScale, n);
    } else if (inScale == 0.0) {
        vec(pool, OP_ASSIGN, x, solution, 0.0, n);
        vec(pool, OP_SCALE_UNIFORM, x, NULL, outScale, n);
    } else {
        vec(pool, OP_SCALE_UNIFORM, x, NULL, inScale, n);
        vec(pool, OP_ADD, x, solution, outScale, n);
    }

done:
    if (stats) *stats = s;
    free(solution); free(search); free(product); free(residual);
    free(preconditioned); free(buffer);
    return status;
}

int oracle_pcg(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
               double inScale, const double* b, double outScale, double* x,
               uint64_t maxIterations, double absoluteTolerance, double relativeTolerance,
               int precondition, int threads, oracle_statistics_t* stats,
               double* history, int64_t historyCapacity, double* seconds) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    pool_t* pool = pool_create(resolve_threads(threads));
    int status = 0;
    double* inverse = precondition ? (double*)malloc((size_t)rows * sizeof(double)) : NULL;
    if (precondition) {
        int64_t failed = -1;
        status = oracle_inverse_diagonal(rows, rowptr, cols, vals, inverse, &failed);
    }
    if (status) {
        if (stats) { stats->iterationCount = 0; stats->absoluteResidual = DBL_MAX; stats->relativeResidual = DBL_MAX; }
    } else {
        const csr_operator_t A = {rows, rowptr, cols, vals, NULL, 0};
        status = pcg_core(pool, &A, inverse, inScale, b, outScale, x, maxIterations, absoluteTolerance,
                          relativeTolerance, stats, history, historyCapacity);
    }
    free(inverse);
    pool_destroy(pool);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return status;
}

/* gradient[e] = -sum_ab lambda[dof(e,a)] Ke0[a,b] u[dof(e,b)], eliminated dofs (-1) are zero. */
int oracle_element_gradient(int64_t elements, int32_t nodesPerElement, const int32_t* dofs,
                            const double* elementMatrix, const double* lambda, const double* u,
                            double* gradient) {
    for (int64_t e = 0; e < elements; ++e) {
        const int32_t* d = dofs + e * nodesPerElement;
        double sum = 0.0;
        for (int a = 0; a < nodesPerElement; ++a) {
            if (d[a] < 0) continue;
            double row = 0.0;
            for (int b = 0; b < nodesPerElement; ++b)
                if (d[b] >= 0) row += elementMatrix[(int64_t)a * nodesPerElement + b] * u[d[b]];
            sum += lambda[d[a]] * row;
        }
        gradient[e] = -sum;
    }
    return 0;
}

/* Position of `column` in cols[begin, end) (sorted) by lower bound, or -1. */
static int64_t find_slot(const int32_t* cols, int64_t begin, int64_t end, int32_t column) {
    int64_t lo = begin, hi = end;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (cols[mid] < column) lo = mid + 1;
        else hi = mid;
    }
    return lo < end && cols[lo] == column ? lo : -1;
}

/* entries = sum over elements e (in order) of scale_e * K_e scattered into the
 * pattern; K_e = matrices + e * stride (stride 0: one shared matrix), nodes^2
 * row-major; dofs -1 marks eliminated nodes, which are skipped. entries is
 * zeroed first. Returns 2 if a kept pair is not in the pattern. */
int oracle_assemble(int64_t rows, const int32_t* rowptr, const int32_t* cols, int64_t elements,
                    int32_t nodesPerElement, const int32_t* dofs, const double* matrices, int64_t stride,
                    const double* scale, double* entries) {
    memset(entries, 0, sizeof(double) * (size_t)rowptr[rows]);
    for (int64_t e = 0; e < elements; ++e) {
        const int32_t* d = dofs + e * nodesPerElement;
        const double* k = matrices + e * stride;
        for (int a = 0; a < nodesPerElement; ++a) {
            if (d[a] < 0) continue;
            if (d[a] >= rows) return 2;
            for (int b = 0; b < nodesPerElement; ++b) {
                if (d[b] < 0) continue;
                const int64_t slot = find_slot(cols, rowptr[d[a]], rowptr[d[a] + 1], d[b]);
                if (slot < 0) return 2;
                const double value = k[(int64_t)a * nodesPerElement + b];
                entries[slot] += scale ? scale[e] * value : value;
            }
        }
    }
    return 0;
}

/* out = sum over elements e (in order) of scale_e * v_e scattered by the dof
 * map; v_e = vectors + e * stride (stride 0: one shared vector). */
int oracle_assemble_vector(int64_t rows, int64_t elements, int32_t nodesPerElement, const int32_t* dofs,
                           const double* vectors, int64_t stride, const double* scale, double* out) {
    memset(out, 0, sizeof(double) * (size_t)rows);
    for (int64_t e = 0; e < elements; ++e) {
        const int32_t* d = dofs + e * nodesPerElement;
        const double* v = vectors + e * stride;
        for (int a = 0; a < nodesPerElement; ++a) {
            if (d[a] < 0) continue;
            if (d[a] >= rows) return 2;
            out[d[a]] += scale ? scale[e] * v[a] : v[a];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Reordering                                                               */
/* ------------------------------------------------------------------------ */

/* The order GCC's std::sort leaves items in, for items compared by
 * key[item] (strictly less). Each helper below is one step of that algorithm. */
static int key_less(const int32_t* key, int32_t left, int32_t right) { return key[left] < key[right]; }

static void swap_items(int32_t* a, int64_t i, int64_t j) {
    const int32_t t = a[i];
    a[i] = a[j];
    a[j] = t;
}

/* Sift-down of the hole at `hole` to a leaf, then sift the value back up to `top`. */
static void heap_adjust(int32_t* a, int64_t hole, int64_t len, int32_t value, const int32_t* key) {
    const int64_t top = hole;
    int64_t child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (key_less(key, a[child], a[child - 1])) --child;
        a[hole] = a[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        a[hole] = a[child - 1];
        hole = child - 1;
    }
    int64_t parent = (hole - 1) / 2;
    while (hole > top && key_less(key, a[parent], value)) {
        a[hole] = a[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    a[hole] = value;
}

static void heap_sort_range(int32_t* a, int64_t n, const int32_t* key) {
    if (n >= 2)
        for (int64_t parent = (n - 2) / 2;; --parent) {
            heap_adjust(a, parent, n, a[parent], key);
            if (parent == 0) break;
        }
    for (int64_t last = n; last > 1;) {
        --last;
        const int32_t value = a[last];
        a[last] = a[0];
        heap_adjust(a, 0, last, value, key);
    }
}

/* Moves the median of a[x], a[y], a[z] to a[r]. */
static void median_to_first(int32_t* a, int64_t r, int64_t x, int64_t y, int64_t z, const int32_t* key) {
    if (key_less(key, a[x], a[y])) {
        if (key_less(key, a[y], a[z])) swap_items(a, r, y);
        else if (key_less(key, a[x], a[z])) swap_items(a, r, z);
        else swap_items(a, r, x);
    } else if (key_less(key, a[x], a[z])) swap_items(a, r, x);
    else if (key_less(key, a[y], a[z])) swap_items(a, r, z);
    else swap_items(a, r, y);
}

/* Partition of a[lo, hi) around the pivot a[p] (p < lo); returns the cut. */
static int64_t partition_around(int32_t* a, int64_t lo, int64_t hi, int64_t p, const int32_t* key) {
    for (;;) {
        while (key_less(key, a[lo], a[p])) ++lo;
        --hi;
        while (key_less(key, a[p], a[hi])) --hi;
        if (!(lo < hi)) return lo;
        swap_items(a, lo, hi);
        ++lo;
    }
}

static void introsort_loop(int32_t* a, int64_t lo, int64_t hi, int64_t depth, const int32_t* key) {
    while (hi - lo > 16) {
        if (depth == 0) {
            heap_sort_range(a + lo, hi - lo, key);
            return;
        }
        --depth;
        median_to_first(a, lo, lo + 1, lo + (hi - lo) / 2, hi - 1, key);
        const int64_t cut = partition_around(a, lo + 1, hi, lo, key);
        introsort_loop(a, cut, hi, depth, key);
        hi = cut;
    }
}

/* Insertion of a[i] into the sorted a[lo, i); `guarded` checks against lo first. */
static void insert_item(int32_t* a, int64_t lo, int64_t i, int guarded, const int32_t* key) {
    const int32_t value = a[i];
    if (guarded && key_less(key, value, a[lo])) {
        memmove(a + lo + 1, a + lo, sizeof(int32_t) * (size_t)(i - lo));
        a[lo] = value;
        return;
    }
    int64_t hole = i;
    while (key_less(key, value, a[hole - 1])) {
        a[hole] = a[hole - 1];
        --hole;
    }
    a[hole] = value;
}

static void degree_sort(int32_t* a, int64_t n, const int32_t* key) {
    if (n <= 0) return;
    int64_t lg = 0;
    while (((int64_t)1 << (lg + 1)) <= n) ++lg;
    introsort_loop(a, 0, n, 2 * lg, key);
    const int64_t head = n > 16 ? 16 : n;
    for (int64_t i = 1; i < head; ++i) insert_item(a, 0, i, 1, key);
    for (int64_t i = head; i < n; ++i) insert_item(a, 0, i, 0, key);
}

/* items[0, n) sorted in place by key[item], in the order described above. */
int oracle_degree_sort(int64_t n, int32_t* items, const int32_t* key) {
    degree_sort(items, n, key);
    return 0;
}

/* The heapsort step alone (std::partial_sort over the whole range). */
int oracle_heap_sort(int64_t n, int32_t* items, const int32_t* key) {
    heap_sort_range(items, n, key);
    return 0;
}

/* strategy: 0 none (identity), 1 Cuthill-McKee, 2 reverse Cuthill-McKee. */
int oracle_make_reordering(int64_t rows, const int32_t* rowptr, const int32_t* cols, int strategy, int32_t* map) {
    if (strategy < 0 || strategy > 2) return 1;
    for (int64_t r = 0; r < rows; ++r) map[r] = (int32_t)r;
    if (strategy == 0 || rows == 0 || rowptr[rows] == 0) return 0;
    for (int64_t k = 0; k < rowptr[rows]; ++k)
        if (cols[k] < 0 || cols[k] >= rows) return 2;
    int32_t* degree = (int32_t*)malloc(sizeof(int32_t) * (size_t)rows);
    int32_t* queue = (int32_t*)malloc(sizeof(int32_t) * (size_t)rows);
    char* seen = (char*)calloc((size_t)rows, 1);
    int64_t least = 0;
    for (int64_t r = 0; r < rows; ++r) {
        degree[r] = rowptr[r + 1] - rowptr[r];
        if (degree[r] < degree[least]) least = r;
    }
    int64_t head = 0, tail = 0, placed = 0;
    queue[tail++] = (int32_t)least;
    seen[least] = 1;
    while (head < tail) {
        const int32_t row = queue[head++];
        const int64_t batch = tail;
        for (int64_t k = rowptr[row]; k < rowptr[row + 1]; ++k)
            if (!seen[cols[k]]) {
                seen[cols[k]] = 1;
                queue[tail++] = cols[k];
            }
        degree_sort(queue + batch, tail - batch, degree);
        map[placed++] = row;
        if (head == tail) {
            int64_t next = -1;
            for (int64_t r = 0; r < rows; ++r)
                if (!seen[r] && (next < 0 || degree[r] < degree[next])) next = r;
            if (next >= 0) {
                seen[next] = 1;
                queue[tail++] = (int32_t)next;
            }
        }
    }
    if (strategy == 2)
        for (int64_t i = 0, j = rows - 1; i < j; ++i, --j) {
            const int32_t t = map[i];
            map[i] = map[j];
            map[j] = t;
        }
    free(degree);
    free(queue);
    free(seen);
    return placed == rows ? 0 : 2;
}

/* Row i of the result is row source[i] with columns target[c], re-sorted;
 * source = map, target = inverse(map) (reverse = 0), or the other way round. */
int oracle_reorder_csr(int64_t rows, const int32_t* map, int reverse, const int32_t* rowptr, const int32_t* cols,
                       const double* vals, int32_t* newRowptr, int32_t* newCols, double* newVals) {
    int32_t* inverse = (int32_t*)malloc(sizeof(int32_t) * (size_t)(rows > 0 ? rows : 1));
    for (int64_t r = 0; r < rows; ++r) inverse[r] = -1;
    for (int64_t i = 0; i < rows; ++i) {
        if (map[i] < 0 || map[i] >= rows || inverse[map[i]] >= 0) {
            free(inverse);
            return 2;
        }
        inverse[map[i]] = (int32_t)i;
    }
    const int32_t* source = reverse ? inverse : map;
    const int32_t* target = reverse ? map : inverse;
    newRowptr[0] = 0;
    for (int64_t i = 0; i < rows; ++i) {
        const int32_t old = source[i];
        const int32_t begin = rowptr[old], end = rowptr[old + 1];
        const int32_t out = newRowptr[i];
        newRowptr[i + 1] = out + (end - begin);
        /* insertion sort of the renumbered row (stable) */
        for (int32_t k = begin; k < end; ++k) {
            const int32_t column = target[cols[k]];
            const double value = vals[k];
            int32_t hole = out + (k - begin);
            while (hole > out && newCols[hole - 1] > column) {
                newCols[hole] = newCols[hole - 1];
                newVals[hole] = newVals[hole - 1];
                --hole;
            }
            newCols[hole] = column;
            newVals[hole] = value;
        }
    }
    free(inverse);
    return 0;
}

int oracle_reorder_vector(int64_t n, const int32_t* map, int reverse, const double* in, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        if (map[i] < 0 || map[i] >= n) return 2;
        if (reverse) out[map[i]] = in[i];
        else out[i] = in[map[i]];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Masked operators, Jacobi smoothers and the p-multigrid cycle             */
/* ------------------------------------------------------------------------ */

int oracle_masked_csr_product(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                              const uint16_t* mask, uint16_t threshold, double inScale, const double* in,
                              double outScale, double* out, int threads) {
    pool_t* pool = pool_create(resolve_threads(threads));
    const csr_operator_t A = {rows, rowptr, cols, vals, mask, threshold};
    apply_csr(pool, &A, inScale, in, outScale, out);
    pool_destroy(pool);
    return 0;
}

int oracle_masked_identity(int64_t n, const uint16_t* mask, uint16_t threshold, double inScale, const double* in,
                           double outScale, double* out) {
    for (int64_t i = 0; i < n; ++i) apply_op(inScale, outScale, out + i, mask[i] < threshold ? in[i] : 0.0);
    return 0;
}

/* JacobiOperator::product over A with the inverse diagonal `inverse`; work holds 4 n doubles. */
static void jacobi_core(pool_t* pool, const csr_operator_t* A, const double* inverse, uint64_t iterations,
                        double relaxation, double inScale, const double* in, double outScale, double* out,
                        double* work) {
    const int64_t n = A->rows;
    double *memory = work, *residual = work + n, *buffer = work + 2 * n;
    vec(pool, OP_ASSIGN, memory, out, 0.0, n);                      /* JacobiOperator.cpp:48-49 */
    vec(pool, OP_FILL, out, NULL, 0.0, n);
    for (uint64_t k = 0; k < iterations; ++k) {                      /* :52-57 */
        vec(pool, OP_ASSIGN, residual, in, 0.0, n);
        apply_csr(pool, A, 1.0, out, -1.0, residual);
        diagonal_product(pool, inverse, 1.0, residual, relaxation, out, buffer, n);
    }
    vec(pool, OP_SCALE_UNIFORM, out, NULL, outScale, n);             /* :59-60 */
    vec(pool, OP_ADD, out, memory, inScale, n);
}

int oracle_jacobi(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                  const uint16_t* mask, uint16_t threshold, uint64_t iterations, double relaxation, double inScale,
                  const double* in, double outScale, double* out, int threads) {
    double* inverse = (double*)malloc(sizeof(double) * (size_t)(rows > 0 ? rows : 1));
    int64_t failed = -1;
    int status = oracle_inverse_diagonal(rows, rowptr, cols, vals, inverse, &failed);
    if (!status) {
        pool_t* pool = pool_create(resolve_threads(threads));
        double* work = (double*)malloc(sizeof(double) * 4 * (size_t)(rows > 0 ? rows : 1));
        const csr_operator_t A = {rows, rowptr, cols, vals, mask, threshold};
        jacobi_core(pool, &A, inverse, iterations, relaxation, inScale, in, outScale, out, work);
        free(work);
        pool_destroy(pool);
    }
    free(inverse);
    return status;
}

/* One sweep of MaskedJacobiOperator (masked = 1) or ApproximateSubstitution
 * (masked = 0) from `previous` into out; 6 if a kept row has no diagonal. */
static int substitution_sweep(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                              const uint16_t* mask, uint16_t threshold, int masked, double relaxation,
                              double inScale, const double* in, double outScale, double* out,
                              const double* previous) {
    for (int64_t row = 0; row < rows; ++row) {
        if (masked && !(mask[row] < threshold)) {
            apply_op(inScale, outScale, out + row, 0.0);             /* MaskedJacobiOperator.cpp:89 */
            continue;
        }
        double contribution = 0.0, diagonal = 0.0;
        int found = 0;
        for (int32_t k = rowptr[row]; k < rowptr[row + 1]; ++k) {
            const int32_t column = cols[k];
            if (column != row) {
                if (!masked || mask[column] < threshold) contribution += vals[k] * previous[column];
            } else {
                diagonal = vals[k];
                found = 1;
            }
        }
        if (!found) return 6;  /* std::optional::value() of an empty diagonal throws */
        const double value = masked
            ? relaxation * (in[row] - contribution) / diagonal + (1.0 - relaxation) * previous[row]  /* :87 */
            : previous[row] + (in[row] - contribution) / diagonal;               /* ApproximateSubstitution.cpp:57 */
        apply_op(inScale, outScale, out + row, value);
    }
    return 0;
}

static int substitution(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                        const uint16_t* mask, uint16_t threshold, int masked, uint64_t iterations,
                        double relaxation, double inScale, const double* in, double outScale, double* out) {
    if (inScale != 0.0 && iterations != 1) return 10;                 /* NotImplementedException, :64-65 / :40-41 */
    double* previous = (double*)malloc(sizeof(double) * (size_t)(rows > 0 ? rows : 1));
    int status = 0;
    for (uint64_t k = 0; k < iterations && !status; ++k) {
        memcpy(previous, out, sizeof(double) * (size_t)rows);         /* previous = out before each sweep */
        status = substitution_sweep(rows, rowptr, cols, vals, mask, threshold, masked, relaxation, inScale, in,
                                    outScale, out, previous);
    }
    free(previous);
    return status;
}

int oracle_masked_jacobi(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                         const uint16_t* mask, uint16_t threshold, uint64_t iterations, double relaxation,
                         double inScale, const double* in, double outScale, double* out) {
    return substitution(rows, rowptr, cols, vals, mask, threshold, 1, iterations, relaxation, inScale, in, outScale,
                        out);
}

int oracle_approximate_substitution(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                                    uint64_t iterations, double inScale, const double* in, double outScale,
                                    double* out) {
    return substitution(rows, rowptr, cols, vals, NULL, 0, 0, iterations, 1.0, inScale, in, outScale, out);
}

/* solveMultigrid (poisson2D/solver.cpp:111-232). x accumulates the updates;
 * history[c] is ||r|| after cycle c and coarseHistory[c] the iteration count
 * of its coarse CG; *cycles the cycle count. */
int oracle_pmg(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals, const uint16_t* mask,
               int order, const double* b, double* x, double relativeTolerance, int64_t maxCycles, double* history,
               uint64_t* coarseHistory, int64_t historyCapacity, int64_t* cycles, uint64_t* coarseIterations,
               int threads) {
    if (order < 1) return 1;
    pool_t* pool = pool_create(resolve_threads(threads));
    const int64_t n = rows;
    double* inverse = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* residual = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* gridResidual = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* update = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* buffer = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* work = (double*)malloc(sizeof(double) * 4 * (size_t)(n > 0 ? n : 1));
    int64_t failed = -1;
    int status = oracle_inverse_diagonal(rows, rowptr, cols, vals, inverse, &failed);
    int64_t count = 0;
    uint64_t coarse = 0;
    if (!status) {
        vec(pool, OP_ASSIGN, residual, b, 0.0, n);                                     /* :135-136 */
        const double initial = sqrt(dot(pool, residual, residual, n));
        double norm = initial;
        const int grids = order;  /* grid 0: mask < 2 (CG); grid g >= 1: mask < g + 2 (Jacobi) */
        while (relativeTolerance * initial < norm && count < maxCycles && !status) {    /* :212 */
            for (int pass = 0; pass < 2 && !status; ++pass) {
                const int first = pass == 0 ? grids - 1 : 1, last = pass == 0 ? 0 : grids - 1;
                const int step = pass == 0 ? -1 : 1;
                for (int g = first; pass == 0 ? g >= last : g <= last; g += step) {
                    const uint16_t threshold = (uint16_t)(g + 2);
                    const csr_operator_t A = {rows, rowptr, cols, vals, mask, threshold};
                    oracle_masked_identity(n, mask, threshold, 0.0, residual, 1.0, gridResidual);  /* restriction */
                    vec(pool, OP_FILL, update, NULL, 0.0, n);
                    if (g == 0) {
                        oracle_statistics_t stats;
                        status = pcg_core(pool, &A, inverse, 0.0, gridResidual, 1.0, update, 5000, 0.0, 0.5, &stats,
                                          NULL, 0);                                  /* :148-171 */
                        coarse += stats.iterationCount;
                        if (coarseHistory && count < historyCapacity) coarseHistory[count] = stats.iterationCount;
                        if (status) break;
                    } else {
                        /* NestedProductOperator(smoother, mask): buffer = mask(gridResidual), then Jacobi */
                        vec(pool, OP_FILL, buffer, NULL, 0.0, n);
                        oracle_masked_identity(n, mask, threshold, 0.0, gridResidual, 1.0, buffer);
                        jacobi_core(pool, &A, inverse, 6, 2.0 / 3.0, 0.0, buffer, 1.0, update, work);
                    }
                    vec(pool, OP_ADD, x, update, 1.0, n);                                  /* :217, :225 */
                    apply_csr(pool, &A, 1.0, update, -1.0, residual);                      /* :218, :226 */
                }
            }
            norm = sqrt(dot(pool, residual, residual, n));                                 /* :229 */
            if (history && count < historyCapacity) history[count] = norm;
            ++count;
        }
    }
    if (cycles) *cycles = count;
    if (coarseIterations) *coarseIterations = coarse;
    free(inverse); free(residual); free(gridResidual); free(update); free(buffer); free(work);
    pool_destroy(pool);
    return status;
}

/* out[e] = |det J_e| k_e sum_ij G_ij moments[i][j], G = J^-1 J^-T by cofactors. */
int oracle_element_stiffness(int64_t elements, int dim, int npe, const double* moments, const double* jacobians,
                             const double* modulus, double* out) {
    if (dim != 2 && dim != 3) return 2;
    const int64_t entries = (int64_t)npe * npe;
    for (int64_t e = 0; e < elements; ++e) {
        const double* J = jacobians + e * dim * dim;
        double inverse[9], G[9], det;
        if (dim == 2) {
            det = J[0] * J[3] - J[1] * J[2];
            inverse[0] = J[3] / det;
            inverse[1] = -J[1] / det;
            inverse[2] = -J[2] / det;
            inverse[3] = J[0] / det;
        } else {
            const double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
            const double c10 = J[2] * J[7] - J[1] * J[8], c11 = J[0] * J[8] - J[2] * J[6], c12 = J[1] * J[6] - J[0] * J[7];
            const double c20 = J[1] * J[5] - J[2] * J[4], c21 = J[2] * J[3] - J[0] * J[5], c22 = J[0] * J[4] - J[1] * J[3];
            det = J[0] * c00 + J[1] * c01 + J[2] * c02;
            inverse[0] = c00 / det; inverse[1] = c10 / det; inverse[2] = c20 / det;
            inverse[3] = c01 / det; inverse[4] = c11 / det; inverse[5] = c21 / det;
            inverse[6] = c02 / det; inverse[7] = c12 / det; inverse[8] = c22 / det;
        }
        for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j) {
                double g = 0.0;
                for (int k = 0; k < dim; ++k) g += inverse[i * dim + k] * inverse[j * dim + k];
                G[i * dim + j] = g;
            }
        const double factor = fabs(det) * (modulus ? modulus[e] : 1.0);
        for (int64_t entry = 0; entry < entries; ++entry) {
            double v = 0.0;
            for (int ij = 0; ij < dim * dim; ++ij) v += G[ij] * moments[ij * entries + entry];
            out[e * entries + entry] = factor * v;
        }
    }
    return 0;
}
