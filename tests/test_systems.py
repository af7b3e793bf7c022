"""The structured system generator: closed-form sizes, element matrices and
assembly invariants (symmetry, zero row sums before elimination, linearity in
the element conductivities), and the seeded RNG against its fixture."""
import json
from pathlib import Path

import numpy as np
import pytest

from cie_amd import systems

RNG = json.loads((Path(__file__).parent / "golden" / "rng.json").read_text())


def test_rng_fixture():
    for seed, data in RNG.items():
        raw = systems.splitmix64(int(seed), 16)
        assert [str(v) for v in raw] == data["raw"]
        assert systems.uniform(int(seed), 16).tolist() == data["uniform"]


def test_normal_moments():
    g = systems.normal(2, 1.0, 200000)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    g = systems.normal(4, 0.5, 200000)
    assert abs(g.std() - 0.5) < 0.005


@pytest.mark.parametrize("n", [1, 2, 5, 16])
@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_closed_form_sizes(n, order):
    # per-dimension couplings: full 2*order*n + 1 - ... ; eliminated Q2: 8n-9, Q4: 24n-17 (3D cube root)
    rows, nnz, count, npe = systems.lagrange_size(3, order, n, eliminate=False)
    assert rows == (order * n + 1) ** 3 and count == n ** 3 and npe == (order + 1) ** 3
    csr = systems.lagrange(2, order, n, eliminate=False)
    per_dim = np.diff(csr.row_extents).reshape(order * n + 1, order * n + 1)
    assert csr.rows == (order * n + 1) ** 2
    if order == 2:
        assert systems.lagrange_size(3, 2, n, eliminate=False)[1] == (8 * n + 1) ** 3
        if n > 1:
            assert systems.lagrange_size(3, 2, n)[1] == (8 * n - 9) ** 3
    if order == 4:
        assert systems.lagrange_size(3, 4, n, eliminate=False)[1] == (24 * n + 1) ** 3
        if n > 1:
            assert systems.lagrange_size(3, 4, n)[1] == (24 * n - 17) ** 3
    assert per_dim.sum() == csr.nnz


def test_benchmark_sizes():
    assert systems.lagrange_size(3, 2, 64) == (2048383, 127263527, 262144, 27)
    assert systems.lagrange_size(3, 4, 32)[:2] == (2048383, 423564751)
    assert systems.lagrange_size(3, 2, 48)[:2] == (857375, 52734375)
    assert systems.poisson5(256).nnz == 326656


def test_q2_element_matrix_closed_form():
    h = 1.0 / 3
    k1 = np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * h)
    m1 = np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) * h / 30
    expected = np.kron(np.kron(m1, m1), k1) + np.kron(np.kron(m1, k1), m1) + np.kron(np.kron(k1, m1), m1)
    assert np.allclose(systems.element_matrix(3, 2, 3), expected, rtol=0, atol=1e-14)
    expected2 = np.kron(m1 / h, k1 * h) + np.kron(k1 * h, m1 / h)
    assert np.allclose(systems.element_matrix(2, 2, 3), expected2, atol=1e-14)


@pytest.mark.parametrize("dim,order", [(2, 2), (2, 4), (3, 2), (3, 4)])
def test_element_matrix_properties(dim, order):
    ke = systems.element_matrix(dim, order, 2)
    assert np.allclose(ke, ke.T, atol=1e-14)
    assert np.allclose(ke.sum(axis=1), 0.0, atol=1e-12)
    eig = np.linalg.eigvalsh(ke)
    assert eig[0] > -1e-12 and eig[1] > 1e-8  # positive semidefinite, rank npe - 1


@pytest.mark.parametrize("dim,order,n", [(2, 2, 4), (2, 4, 3), (3, 2, 3), (3, 4, 2)])
def test_assembly_symmetry_row_sums_and_dense(dim, order, n):
    rng = np.random.default_rng(1)
    count = n ** dim
    k = np.exp(rng.standard_normal(count))
    full = systems.lagrange(dim, order, n, conductivity=k, eliminate=False).to_scipy().toarray()
    assert np.allclose(full, full.T, atol=1e-12)
    assert np.allclose(full.sum(axis=1), 0.0, atol=1e-10)
    # dense element-by-element assembly from the element matrix and dof map
    ke = systems.element_matrix(dim, order, n)
    dofs = systems.element_dofs(dim, order, n, eliminate=False)
    dense = np.zeros_like(full)
    for e in range(count):
        d = dofs[e]
        dense[np.ix_(d, d)] += k[e] * ke
    assert np.allclose(full, dense, atol=1e-12)
    # elimination = deleting boundary rows and columns
    elim = systems.lagrange(dim, order, n, conductivity=k).to_scipy().toarray()
    m = order * n + 1
    grid = np.indices((m,) * dim).reshape(dim, -1)
    interior = np.all((grid > 0) & (grid < m - 1), axis=0)
    assert np.allclose(elim, full[np.ix_(interior, interior)], atol=1e-12)
    assert np.all(np.linalg.eigvalsh(elim) > 0)


def test_rows_sorted_and_unique():
    csr = systems.lagrange(3, 2, 5)
    for row in range(csr.rows):
        c = csr.column_indices[csr.row_extents[row]:csr.row_extents[row + 1]]
        assert np.all(np.diff(c) > 0)
        assert row in c


def test_thread_count_independence():
    k = np.exp(systems.normal(4, 0.7, 4 ** 3))
    a = systems.lagrange(3, 4, 4, conductivity=k, threads=1)
    b = systems.lagrange(3, 4, 4, conductivity=k, threads=7)
    assert np.array_equal(a.entries, b.entries) and np.array_equal(a.column_indices, b.column_indices)


def test_five_point_structure():
    csr = systems.poisson5(5)
    a = csr.to_scipy().toarray()
    t = 2 * np.eye(5) - np.eye(5, k=1) - np.eye(5, k=-1)
    assert np.array_equal(a, np.kron(np.eye(5), t) + np.kron(t, np.eye(5)))


def test_csr_slots():
    csr = systems.poisson5(4)
    slots = systems.csr_slots(csr, [0, 5, 5, 15], [0, 4, 7, 14])
    assert csr.entries[slots[0]] == 4.0 and csr.entries[slots[1]] == -1.0
    assert slots[2] == -1
    assert csr.column_indices[slots[3]] == 14


def test_batch_workload_layout():
    w = systems.q2_quad_batch(count=6, seed=3, smallest=2, largest=9)
    sizes = w.extra["sizes"]
    raw = systems.splitmix64(3, 6)
    unit = (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    assert sizes.tolist() == (2 + np.floor(unit * 8)).astype(int).tolist()
    offsets = w.extra["offsets"]
    assert np.array_equal(np.diff(offsets), (2 * sizes - 1) ** 2)
    a = w.matrix.to_scipy()
    for lo, hi in zip(offsets[:-1], offsets[1:]):  # block diagonal: nothing outside the diagonal blocks
        assert a[lo:hi, :lo].nnz == 0 and a[lo:hi, hi:].nnz == 0
    assert w.matrix.nnz == sum((8 * n - 9) ** 2 for n in sizes)


def _legendre_1d(order, t):
    """Integrated Legendre basis in position order at t in [0, 1] (x = 2t - 1), evaluated with scipy."""
    from scipy.special import eval_legendre

    x = 2.0 * np.asarray(t) - 1.0
    out = np.zeros((order + 1,) + x.shape)
    out[0] = 0.5 * (1.0 - x)
    out[order] = 0.5 * (1.0 + x)
    for pos in range(1, order):
        k = pos + 1
        out[pos] = (eval_legendre(k, x) - eval_legendre(k - 2, x)) / (2 * k - 1)
    return out


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_legendre_1d_stiffness_closed_form(order):
    # 1D factor of the 2D element matrix on one element (h = 1): K = K1 (x) M1 + M1 (x) K1. With the
    # vertex block [[1,-1],[-1,1]] and bubbles 4/(2k-1) on the diagonal K1 is known exactly; M1 from scipy.
    from scipy.special import roots_legendre

    points, weights = roots_legendre(order + 2)
    t, w = 0.5 * (points + 1.0), 0.5 * weights
    values = _legendre_1d(order, t)
    m1 = (values * w) @ values.T
    k1 = np.zeros((order + 1, order + 1))
    k1[np.ix_([0, order], [0, order])] = [[1.0, -1.0], [-1.0, 1.0]]
    for pos in range(1, order):
        k1[pos, pos] = 4.0 / (2 * (pos + 1) - 1)
    expected = np.kron(m1, k1) + np.kron(k1, m1)  # local index x fastest
    got = systems.element_matrix(2, order, 2, basis="legendre")  # in 2D the element matrix does not depend on h
    assert np.allclose(got, expected, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("dim,order,n", [(2, 2, 3), (2, 3, 2), (2, 4, 2), (3, 2, 2), (3, 3, 2)])
def test_legendre_system_is_the_lagrange_system_in_another_basis(dim, order, n):
    # Both bases span Q_p with u = 0 on the boundary, so K_leg = T^T K_lag T with T[i, j] the value of
    # Legendre function j at Lagrange node i (interior positions only).
    positions = order * n + 1
    lagrange_points = np.concatenate([[0.0]] + [
        (e + 0.5 * (1.0 - np.cos(np.arange(1, order + 1) * np.pi / order))) / n for e in range(n)])
    t1 = np.zeros((positions, positions))
    for e in range(n):
        local = np.arange(order + 1)
        for i in range(order * e, order * e + order + 1):
            t = lagrange_points[i] * n - e
            t1[i, order * e + local] = _legendre_1d(order, np.array([t]))[:, 0]
    interior = t1[1:-1, 1:-1]
    t_full = interior
    for _ in range(dim - 1):
        t_full = np.kron(interior, t_full)  # x fastest
    k_lag = systems.lagrange(dim, order, n).to_scipy().toarray()
    k_leg = systems.lagrange(dim, order, n, basis="legendre").to_scipy().toarray()
    assert np.allclose(k_leg, t_full.T @ k_lag @ t_full, rtol=1e-11, atol=1e-11 * np.abs(k_leg).max())
    # same pattern as the Lagrange system, symmetric, positive definite
    assert np.array_equal(systems.lagrange(dim, order, n, basis="legendre").row_extents,
                          systems.lagrange(dim, order, n).row_extents)
    assert np.allclose(k_leg, k_leg.T, rtol=0, atol=1e-14 * np.abs(k_leg).max())
    assert np.linalg.eigvalsh(k_leg).min() > 0


def test_ansatz_mask_degrees():
    order, n = 4, 3
    mask = systems.ansatz_mask(2, order, n)
    positions = np.arange(1, order * n)  # interior positions along one dimension
    degree = np.where(positions % order == 0, 1, positions % order + 1)
    expected = np.maximum(degree[None, :], degree[:, None]).ravel()  # row = x + (n_pos) * y, x fastest
    assert np.array_equal(mask, expected.astype(np.uint16))
    assert np.count_nonzero(mask == 1) == (n - 1) ** 2  # interior vertex DOFs
    assert mask.max() == order
    # the mask is only defined for the hierarchical basis
    from ctypes import byref

    from cie_amd._lib import LagrangeSpec, check_sys, library

    out = np.empty(mask.size, dtype=np.uint16)
    with pytest.raises(RuntimeError):
        check_sys(library().cie_sys_lagrange_ansatz_mask(byref(LagrangeSpec(2, order, n, 1, 0)),
                                                         out.ctypes.data_as(__import__("ctypes").c_void_p)))
