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
 *   oracle_element_gradient dJ/dk_e = -lambda_e^T K_e0 u_e (no reference
 *                           counterpart; K_e = k_e K_e0 as in
 *                           fem/.../LinearIsotropicStiffnessIntegrand_impl.hpp:74).
 *
 * Status codes match include/cie_hip.h: 0 ok, 2 shape, 3 p^T A p <= 0,
 * 4 r^T z <= 0, 5 zero diagonal, 6 missing diagonal.
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
} spmv_args_t;

static void spmv_task(void* raw, int64_t begin, int64_t end, int worker) {
    (void)worker;
    const spmv_args_t* a = (const spmv_args_t*)raw;
    for (int64_t row = begin; row < end; ++row) {
        double contribution = 0.0;
        for (int32_t k = a->rowptr[row]; k < a->rowptr[row + 1]; ++k)
            contribution += a->vals[k] * a->in[a->cols[k]];
        double* o = a->out + row;
        /* CSROperator.cpp:78-110, one case per (inScale, outScale) class. */
        if (a->inScale == 1.0) {
            if (a->outScale == 1.0) *o += contribution;
            else if (a->outScale == -1.0) *o -= contribution;
            else *o += a->outScale * contribution;
        } else if (a->inScale == 0.0) {
            if (a->outScale == 1.0) *o = contribution;
            else if (a->outScale == -1.0) *o = -contribution;
            else *o = a->outScale * contribution;
        } else if (a->inScale == -1.0) {
            if (a->outScale == 1.0) *o = contribution - *o;
            else if (a->outScale == -1.0) *o = -contribution - *o;
            else *o = a->outScale * contribution - *o;
        } else {
            if (a->outScale == 1.0) *o = a->inScale * *o + contribution;
            else if (a->outScale == -1.0) *o = a->inScale * *o - contribution;
            else *o = a->inScale * *o + a->outScale * contribution;
        }
    }
}

static void spmv(pool_t* pool, int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
                 double inScale, const double* in, double outScale, double* out) {
    spmv_args_t a = {rowptr, cols, vals, inScale, outScale, in, out};
    pool_run(pool, spmv_task, &a, rows);
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
int oracle_pcg(int64_t rows, const int32_t* rowptr, const int32_t* cols, const double* vals,
               double inScale, const double* b, double outScale, double* x,
               uint64_t maxIterations, double absoluteTolerance, double relativeTolerance,
               int precondition, int threads, oracle_statistics_t* stats,
               double* history, int64_t historyCapacity, double* seconds) {
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    pool_t* pool = pool_create(resolve_threads(threads));
    const int64_t n = rows;
    int status = 0;
    double* solution = (double*)malloc((size_t)n * sizeof(double));
    double* search = (double*)malloc((size_t)n * sizeof(double));
    double* product = (double*)malloc((size_t)n * sizeof(double));
    double* residual = (double*)malloc((size_t)n * sizeof(double));
    double* preconditioned = precondition ? (double*)malloc((size_t)n * sizeof(double)) : NULL;
    double* inverse = precondition ? (double*)malloc((size_t)n * sizeof(double)) : NULL;
    double* buffer = precondition ? (double*)malloc((size_t)n * sizeof(double)) : NULL;

    oracle_statistics_t s = {0, DBL_MAX, DBL_MAX};

    if (precondition) {
        int64_t failed = -1;
        status = oracle_inverse_diagonal(rows, rowptr, cols, vals, inverse, &failed);
        if (status) goto done;
    }

    vec(pool, OP_ASSIGN, solution, x, 0.0, n);                                   /* :74-75 */
    if (maxIterations) {
        vec(pool, OP_FILL, product, NULL, 0.0, n);                               /* :82 */
        vec(pool, OP_ASSIGN, residual, b, 0.0, n);                               /* :86-88 */
        spmv(pool, rows, rowptr, cols, vals, 1.0, solution, -1.0, residual);

        double residualNorm = dot(pool, residual, residual, n);                 /* :93 */
        double preconditionedNorm = 0.0;
        const double initialResidualNorm = sqrt(residualNorm);
        s.absoluteResidual = initialResidualNorm;
        s.relativeResidual = 1.0;
        if (float_less(s.absoluteResidual, absoluteTolerance) || float_less(s.relativeResidual, relativeTolerance))
            goto finish;                                                          /* :100-111 */

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
            spmv(pool, rows, rowptr, cols, vals, 0.0, search, 1.0, product);
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

finish:
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
    free(preconditioned); free(inverse); free(buffer);
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
