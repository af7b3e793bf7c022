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
