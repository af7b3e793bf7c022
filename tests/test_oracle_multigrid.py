"""CPU checks of the oracle's masked operators, Jacobi smoothers and p-multigrid
(solvers/src/Masked*.cpp, JacobiOperator.cpp, ApproximateSubstitution.cpp,
poisson2D/solver.cpp:111-232) against dense numpy, closed forms and fixed points."""
import numpy as np
import pytest
import scipy.sparse

import oracle
from cie_amd import systems

SCALES = [(0.0, 1.0), (0.0, -1.0), (0.0, 2.5), (1.0, 1.0), (1.0, -1.0), (1.0, 0.5), (-1.0, 1.0), (-1.0, -1.0),
          (-1.0, 3.0), (2.0, 1.0), (2.0, -1.0), (0.5, -2.0)]


def integer_csr(seed, rows=40, density=0.2):
    rng = np.random.default_rng(seed)
    a = scipy.sparse.random(rows, rows, density=density, random_state=seed, format="csr")
    a = (a + scipy.sparse.eye(rows)).tocsr()
    a.sort_indices()
    a.data = rng.integers(-9, 10, a.nnz).astype(np.float64)
    a.data[a.indices == np.repeat(np.arange(rows), np.diff(a.indptr))] = 20.0 + rng.integers(0, 5, rows)
    return systems.CSR(rows, rows, a.indptr.astype(np.int32), a.indices.astype(np.int32), a.data)


def contract(in_scale, out_scale, out, right):
    return in_scale * out + out_scale * right  # exact for the integer data used below


@pytest.mark.parametrize("in_scale,out_scale", SCALES)
def test_masked_product_against_dense(in_scale, out_scale):
    csr = integer_csr(1)
    rng = np.random.default_rng(2)
    mask = rng.integers(0, 4, csr.rows).astype(np.uint16)
    x = rng.integers(-5, 6, csr.rows).astype(np.float64)
    y0 = rng.integers(-5, 6, csr.rows).astype(np.float64)
    for threshold in (0, 1, 2, 4):
        keep = mask < threshold
        dense = csr.to_scipy().toarray() * keep[:, None] * keep[None, :]
        expected = contract(in_scale, out_scale, y0, dense @ x)
        got = oracle.masked_csr_product(csr, mask, threshold, in_scale, x, out_scale, y0.copy())
        assert np.array_equal(got, expected)
        ident = oracle.masked_identity(mask, threshold, in_scale, x, out_scale, y0.copy())
        assert np.array_equal(ident, contract(in_scale, out_scale, y0, np.where(keep, x, 0.0)))


def test_jacobi_closed_forms_and_convergence():
    # diagonal A, relaxation 1: one sweep is the exact solve; out = outScale x + inScale out_old
    d = np.array([2.0, 4.0, 8.0, 0.5])
    diag = systems.CSR(4, 4, np.arange(5, dtype=np.int32), np.arange(4, dtype=np.int32), d)
    b = np.array([1.0, 2.0, 3.0, 4.0])
    out = np.array([1.0, 1.0, 1.0, 1.0])
    oracle.jacobi(diag, 1, 1.0, 2.0, b, 3.0, out)
    assert np.array_equal(out, 3.0 * b / d + 2.0)
    # damped Jacobi on a diagonally dominant SPD matrix converges to the solution
    csr = systems.poisson5(6)
    b = systems.uniform(3, csr.rows)
    x = oracle.jacobi(csr, 400, 2.0 / 3.0, 0.0, b, 1.0, np.zeros(csr.rows))
    assert np.allclose(x, np.linalg.solve(csr.to_scipy().toarray(), b), rtol=1e-8, atol=1e-10)
    # k sweeps = the dense recurrence x <- x + (w r) / d
    a = csr.to_scipy().toarray()
    ref = np.zeros(csr.rows)
    for _ in range(5):
        ref = ref + (2.0 / 3.0 * (b - a @ ref)) * (1.0 / np.diag(a))
    got = oracle.jacobi(csr, 5, 2.0 / 3.0, 0.0, b, 1.0, np.zeros(csr.rows))
    assert np.allclose(got, ref, rtol=1e-14, atol=1e-15)


def test_jacobi_on_a_masked_matrix_is_jacobi_on_the_kept_rows():
    order, n = 3, 4
    csr = systems.lagrange(2, order, n, basis="legendre")
    mask = systems.ansatz_mask(2, order, n)
    b = systems.uniform(4, csr.rows)
    a = csr.to_scipy().toarray()
    for threshold in (2, 3):
        keep = mask < threshold
        rhs = np.where(keep, b, 0.0)
        got = oracle.jacobi(csr, 6, 2.0 / 3.0, 0.0, rhs, 1.0, np.zeros(csr.rows), mask=mask, threshold=threshold)
        sub = a[np.ix_(keep, keep)]
        ref = np.zeros(keep.sum())
        for _ in range(6):
            ref = ref + (2.0 / 3.0 * (rhs[keep] - sub @ ref)) * (1.0 / np.diag(sub))
        assert np.all(got[~keep] == 0.0)
        assert np.allclose(got[keep], ref, rtol=1e-13, atol=1e-15)


def test_masked_jacobi_and_approximate_substitution():
    csr = systems.CSR(2, 2, np.array([0, 2, 4], dtype=np.int32), np.array([0, 1, 0, 1], dtype=np.int32),
                      np.array([2.0, 1.0, 1.0, 3.0]))
    b = np.array([1.0, 1.0])
    # ApproximateSubstitution: out = previous + (b - offdiag(previous)) / diag, previous = out before each sweep
    once = oracle.approximate_substitution(csr, 1, 0.0, b, 1.0, np.zeros(2))
    assert np.array_equal(once, [0.5, 1.0 / 3.0])
    twice = oracle.approximate_substitution(csr, 2, 0.0, b, 1.0, np.zeros(2))
    assert np.allclose(twice, [0.5 + (1.0 - 1.0 / 3.0) / 2.0, 1.0 / 3.0 + (1.0 - 0.5) / 3.0], rtol=1e-15)
    # MaskedJacobiOperator: a masked row gives 0; the solution of the kept system is a fixed point
    big = systems.lagrange(2, 2, 5, basis="legendre")
    mask = systems.ansatz_mask(2, 2, 5)
    keep = mask < 2
    a = big.to_scipy().toarray()
    rhs = systems.uniform(6, big.rows)
    solution = np.zeros(big.rows)
    solution[keep] = np.linalg.solve(a[np.ix_(keep, keep)], rhs[keep])
    out = oracle.masked_jacobi(big, mask, 2, 3, 0.7, 0.0, rhs, 1.0, solution.copy())
    assert np.allclose(out, solution, rtol=1e-12, atol=1e-14)
    first = oracle.masked_jacobi(big, mask, 2, 1, 0.7, 0.0, rhs, 1.0, np.zeros(big.rows))
    assert np.allclose(first, np.where(keep, 0.7 * rhs / np.diag(a), 0.0), rtol=1e-15)
    with pytest.raises(oracle.OracleError) as error:
        oracle.masked_jacobi(big, mask, 2, 2, 0.7, 1.0, rhs, 1.0, np.zeros(big.rows))
    assert error.value.status == 10
    missing = systems.CSR(2, 2, np.array([0, 1, 2], dtype=np.int32), np.array([1, 0], dtype=np.int32),
                          np.array([1.0, 1.0]))
    with pytest.raises(oracle.OracleError):
        oracle.approximate_substitution(missing, 1, 0.0, b, 1.0, np.zeros(2))


def dense_pmg(a, mask, order, b, tolerance=1e-6, max_cycles=500):
    """The same cycle written with dense numpy (an independent statement of solveMultigrid)."""
    d = np.diag(a)

    def cg(m, rhs):
        x = np.zeros_like(rhs)
        r = rhs - m @ x
        init = np.sqrt(r @ r)
        z = r / d
        rz = r @ z
        p = z.copy()
        for it in range(5000):
            q = m @ p
            alpha = rz / (p @ q)
            x = x + alpha * p
            r = r - alpha * q
            z = r / d
            rz, old = r @ z, rz
            p = (rz / old) * p + z
            if np.sqrt(r @ r) / init < 0.5:
                return x, it
        return x, 5000

    masked = {t: a * (mask < t)[:, None] * (mask < t)[None, :] for t in range(2, order + 2)}
    x = np.zeros_like(b)
    r = b.copy()
    initial = norm = np.sqrt(r @ r)
    counts = []
    while tolerance * initial < norm and len(counts) < max_cycles:
        for g in list(range(order - 1, -1, -1)) + list(range(1, order)):
            t = g + 2
            rg = np.where(mask < t, r, 0.0)
            if g == 0:
                update, it = cg(masked[t], rg)
                counts.append(it)
            else:
                update = np.zeros_like(b)
                for _ in range(6):
                    update = update + (2.0 / 3.0 * (rg - masked[t] @ update)) / d
            x = x + update
            r = r - masked[t] @ update
        norm = np.sqrt(r @ r)
    return x, counts


@pytest.mark.parametrize("order,n", [(1, 9), (2, 7), (3, 5), (4, 5)])
def test_pmg_against_a_dense_statement(order, n):
    csr = systems.lagrange(2, order, n, basis="legendre")
    mask = systems.ansatz_mask(2, order, n)
    b = systems.uniform(1, csr.rows)
    x, history, coarse = oracle.pmg(csr, mask, order, b)
    ref, ref_counts = dense_pmg(csr.to_scipy().toarray(), mask, order, b)
    assert list(coarse) == ref_counts
    assert np.allclose(x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    assert history[-1] <= 1e-6 * np.linalg.norm(b) < history[-2]
    assert np.all(np.diff(np.log(history)) < 0)  # the tracked residual falls every cycle


def test_pmg_order_one_solves_exactly_and_higher_orders_track_only_the_masked_residual():
    # order 1: one grid, the tracked residual is the true one
    csr = systems.lagrange(2, 1, 12, basis="legendre")
    b = systems.uniform(2, csr.rows)
    x, history, _ = oracle.pmg(csr, systems.ansatz_mask(2, 1, 12), 1, b)
    a = csr.to_scipy()
    assert np.linalg.norm(b - a @ x) <= 1e-6 * np.linalg.norm(b) * (1 + 1e-9)
    # order >= 2: the coarse update's residual correction uses the masked matrix, so the coupling of vertex
    # updates into bubble rows is never subtracted; the tracked residual converges, the true one stalls
    csr = systems.lagrange(2, 3, 6, basis="legendre")
    b = systems.uniform(2, csr.rows)
    x, history, _ = oracle.pmg(csr, systems.ansatz_mask(2, 3, 6), 3, b)
    true = np.linalg.norm(b - csr.to_scipy() @ x) / np.linalg.norm(b)
    assert history[-1] / np.linalg.norm(b) <= 1e-6 and true > 1e-3


def test_pmg_order_one_cycles_are_restarted_cg_to_one_half():
    """Order 1 has one grid, the coarse CG (solver.cpp:146-170) with relative tolerance 0.5 from a zero update:
    the first cycle is the oracle's PCG to 0.5 on b, and the next starts it again on the tracked residual
    r = b - A u (product(1, update, -1, residual), solver.cpp:219)."""
    csr = systems.lagrange(2, 1, 10, basis="legendre")
    mask = systems.ansatz_mask(2, 1, 10)
    b = systems.uniform(7, csr.rows)
    x1, history1, coarse1 = oracle.pmg(csr, mask, 1, b, max_cycles=1)
    u1, s1, _, _ = oracle.pcg(csr, b, relative_tolerance=0.5, max_iterations=5000)
    assert np.array_equal(x1, u1) and list(coarse1) == [s1.iterationCount]
    r1 = oracle.csr_product(csr, 1.0, u1, -1.0, b.copy())
    assert history1[0] == pytest.approx(np.linalg.norm(r1), rel=1e-14)
    x2, _, coarse2 = oracle.pmg(csr, mask, 1, b, max_cycles=2)
    u2, s2, _, _ = oracle.pcg(csr, r1, relative_tolerance=0.5, max_iterations=5000)
    assert np.allclose(x2, u1 + u2, rtol=0, atol=1e-15 * np.abs(u1 + u2).max())
    assert list(coarse2) == [s1.iterationCount, s2.iterationCount]


@pytest.mark.parametrize("order", [1, 2, 3])
def test_pmg_is_scale_invariant_and_linear_in_b(order):
    """Power-of-two scalings are exact: the cycle on (2^k A, 2^k b) returns the same x bit for bit (Jacobi
    divides by the scaled diagonal, CG's alpha and beta are ratios), and on (A, 2^k b) it returns 2^k x; the
    coarse iteration counts do not change."""
    csr = systems.lagrange(2, order, 5, basis="legendre")
    mask = systems.ansatz_mask(2, order, 5)
    b = systems.uniform(3, csr.rows)
    x, history, coarse = oracle.pmg(csr, mask, order, b)
    scaled = systems.CSR(csr.rows, csr.columns, csr.row_extents, csr.column_indices, 8.0 * csr.entries)
    xs, hs, cs = oracle.pmg(scaled, mask, order, 8.0 * b)
    assert np.array_equal(xs, x) and np.array_equal(hs, 8.0 * history) and np.array_equal(cs, coarse)
    xb, hb, cb = oracle.pmg(csr, mask, order, 0.25 * b)
    assert np.array_equal(xb, 0.25 * x) and np.array_equal(hb, 0.25 * history) and np.array_equal(cb, coarse)


def test_pmg_jacobi_grid_on_an_uncoupled_diagonal_block():
    """Order 2 on A = blockdiag(A_c, D_f) with the coarse rows masked 1 and the fine rows masked 2: the fine rows
    couple to nothing, so the coarse CG never touches them and each visit of the Jacobi grid (6 sweeps of
    relaxation 2/3 from zero, solver.cpp:188-195) multiplies their residual by (1 - 2/3)^6 exactly in exact
    arithmetic; a cycle visits that grid twice (down and up, solver.cpp:210-226)."""
    coarse = systems.poisson5(6)
    nc = coarse.rows
    d_f = np.array([1.0, 2.5, 4.0, 0.75, 3.0, 8.0])
    a = scipy.sparse.block_diag([coarse.to_scipy(), scipy.sparse.diags(d_f)]).tocsr()
    a.sort_indices()
    csr = systems.CSR(a.shape[0], a.shape[1], a.indptr.astype(np.int32), a.indices.astype(np.int32), a.data)
    mask = np.concatenate([np.ones(nc), 2 * np.ones(len(d_f))]).astype(np.uint16)
    b = systems.uniform(11, csr.rows)
    for cycles in (1, 2):
        x, history, _ = oracle.pmg(csr, mask, 2, b, max_cycles=cycles)
        # r = b - A x cancels to about eps |b| here, so the ratio is known to about 1e-4 of (1/3)^12
        ratio = (b[nc:] - d_f * x[nc:]) / b[nc:]
        assert np.allclose(ratio, (1.0 / 3.0) ** (12 * cycles), rtol=1e-3, atol=0)
