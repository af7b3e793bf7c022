"""HIP SpMV and space operations against the oracle and the reference's KATs."""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from cie_amd import systems

pytestmark = pytest.mark.gpu
KAT = json.loads((Path(__file__).parent / "golden" / "kat.json").read_text())
EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


def to_dev(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0", dtype=dtype or torch.float64)


def random_csr(rng, rows, cols, density, empty_rows=()):
    a = rng.standard_normal((rows, cols)) * (rng.random((rows, cols)) < density)
    for r in empty_rows:
        a[r, :] = 0.0
    rr, cc = np.nonzero(a)
    ext = np.zeros(rows + 1, dtype=np.int64)
    np.add.at(ext, rr + 1, 1)
    return systems.CSR(rows, cols, np.cumsum(ext).astype(np.int32), cc.astype(np.int32), a[rr, cc])


def spmv_bound(csr, x, y0, in_scale, out_scale, eps):
    absolute = np.abs(csr.entries) * np.abs(x[csr.column_indices])
    lengths = np.diff(csr.row_extents).astype(np.int64)
    sums = np.bincount(np.repeat(np.arange(csr.rows), lengths), weights=absolute, minlength=csr.rows)
    return 2 * (lengths + 3) * eps * (abs(out_scale) * sums + abs(in_scale) * np.abs(y0)) + 1e-300


def test_spmv_kat(hip, torch_gpu):
    k = KAT["spmv"]                        # SYCLSpace_test.cpp:140-186 -> [233, 462]
    csr = systems.CSR(2, k["columns"], np.array(k["row_extents"], np.int32), np.array(k["column_indices"], np.int32),
                      np.array(k["entries"]))
    matrix = hip.CSRMatrix(csr)
    for dtype in (torch_gpu.float64, torch_gpu.float32):
        space = hip.HIPSpace(dtype)
        op = hip.HIPCSROperator(matrix, space)
        x = to_dev(torch_gpu, k["in"], dtype)
        y = to_dev(torch_gpu, k["out"], dtype)
        op.product(k["in_scale"], x, k["out_scale"], y)
        assert y.cpu().tolist() == k["expected"]


SCALES = [(0, 1), (0, -1), (0, 2.5), (1, 1), (1, -1), (1, 0.5), (-1, 1), (-1, 3.0), (0.25, 1), (0.25, -2.0)]


LAUNCHES = [("staged", 512, 4), ("staged", 512, 8), ("staged", 1024, 4), ("staged", 1024, 8), ("vector", 1, 1),
            ("vector", 2, 4), ("vector", 4, 2), ("vector", 8, 8), ("vector", 16, 4), ("vector", 32, 1),
            ("vector", 64, 2), ("tiled", 256, 4), ("tiled", 256, 8), ("tiled", 256, 16), ("blocks", 2, 8),
            ("blocks", 3, 4), ("blocks", 3, 16), ("blocks", 4, 8)]


@pytest.mark.parametrize("kernel,width,unroll", LAUNCHES)
def test_spmv_random_against_oracle(hip, torch_gpu, kernel, width, unroll):
    rng = np.random.default_rng(width * 10 + unroll)
    csr = random_csr(rng, 3001, 2500, 0.02, empty_rows=(0, 17, 3000))
    matrix = hip.CSRMatrix(csr)
    matrix.set_launch(64, kernel, width, unroll)
    matrix.set_launch(32, kernel, width, unroll)
    x = rng.standard_normal(2500)
    y0 = rng.standard_normal(3001)
    for in_scale, out_scale in SCALES:
        expected = oracle.csr_product(csr, in_scale, x, out_scale, y0.copy())
        space = hip.HIPSpace(torch_gpu.float64)
        y = to_dev(torch_gpu, y0)
        hip.HIPCSROperator(matrix, space).product(in_scale, to_dev(torch_gpu, x), out_scale, y)
        got = y.cpu().numpy()
        if kernel == "staged":
            assert np.array_equal(got, expected)  # same rounding sequence as the CPU row loop
        else:
            assert np.all(np.abs(got - expected) <= spmv_bound(csr, x, y0, in_scale, out_scale, EPS64))
        space32 = hip.HIPSpace(torch_gpu.float32)
        y32 = to_dev(torch_gpu, y0, torch_gpu.float32)
        hip.HIPCSROperator(matrix, space32).product(in_scale, to_dev(torch_gpu, x, torch_gpu.float32), out_scale, y32)
        bound32 = spmv_bound(csr, x, y0, in_scale, out_scale, EPS32) * 4
        assert np.all(np.abs(y32.cpu().numpy() - expected) <= bound32)


@pytest.mark.parametrize("unroll", [4, 8, 16])
@pytest.mark.parametrize("cols", [16000, 200000])
def test_spmv_tiled_long_short_and_empty_rows(hip, torch_gpu, unroll, cols):
    """Block tiles: rows longer than a tile (summed in passes), tiles of exactly 256 short rows, empty rows;
    with 16000 columns the tiled kernel reads the narrow 16-bit column plane, with 200000 the int32 one."""
    rng = np.random.default_rng(unroll)
    rows = 2000
    lengths = rng.integers(1, 4, rows)
    lengths[rng.choice(rows, 40, replace=False)] = 0
    lengths[[0, 700, 701, 1999]] = [5000, 256 * unroll + 1, 3 * 256 * unroll, 4000]
    ext = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    columns = np.concatenate([np.sort(rng.choice(cols, n, replace=False)) for n in lengths]).astype(np.int32)
    csr = systems.CSR(rows, cols, ext, columns, rng.standard_normal(int(ext[-1])))
    matrix = hip.CSRMatrix(csr)
    matrix.set_launch(64, "tiled", 256, unroll)
    matrix.set_launch(32, "tiled", 256, unroll)
    assert matrix.launch(64) == ("tiled", 256, unroll)
    assert matrix.column_bytes() == (2 if cols < 65536 else 4)
    x = rng.standard_normal(cols)
    y0 = rng.standard_normal(rows)
    for in_scale, out_scale in [(0, 1), (1, -1), (0.5, 2.0)]:
        expected = oracle.csr_product(csr, in_scale, x, out_scale, y0.copy())
        space = hip.HIPSpace(torch_gpu.float64)
        results = []
        for _ in range(2):
            y = to_dev(torch_gpu, y0)
            hip.HIPCSROperator(matrix, space).product(in_scale, to_dev(torch_gpu, x), out_scale, y)
            results.append(y.cpu().numpy())
        assert np.array_equal(results[0], results[1])  # bitwise run to run
        assert np.all(np.abs(results[0] - expected) <= spmv_bound(csr, x, y0, in_scale, out_scale, EPS64))
        space32 = hip.HIPSpace(torch_gpu.float32)
        y32 = to_dev(torch_gpu, y0, torch_gpu.float32)
        hip.HIPCSROperator(matrix, space32).product(in_scale, to_dev(torch_gpu, x, torch_gpu.float32), out_scale, y32)
        assert np.all(np.abs(y32.cpu().numpy() - expected) <= 4 * spmv_bound(csr, x, y0, in_scale, out_scale, EPS32))


@pytest.mark.parametrize("width,unroll", [(2, 4), (3, 8), (4, 16)])
@pytest.mark.parametrize("cols", [16000, 200000])
def test_spmv_blocks_long_short_and_empty_rows(hip, torch_gpu, width, unroll, cols):
    """Column blocks: rows with more blocks than a tile (summed in passes), runs of adjacent columns cut at
    every offset, a lone entry in the last column (its block starts at columns - width), empty rows."""
    rng = np.random.default_rng(width * 100 + unroll)
    rows = 2000
    lengths = rng.integers(1, 9, rows)
    lengths[rng.choice(rows, 40, replace=False)] = 0
    lengths[[0, 700, 701, 1999]] = [5000, 256 * unroll + 1, 3 * 256 * unroll, 4000]
    rows_columns = []
    for n in lengths:
        if n >= 256:  # scattered: one block per entry
            rows_columns.append(np.sort(rng.choice(cols, n, replace=False)))
        else:  # runs of adjacent columns of random length
            starts = np.sort(rng.choice(cols // 16, max(1, n // 3), replace=False)) * 16
            run = np.unique(np.concatenate([s + np.arange(rng.integers(1, 6)) for s in starts]))
            rows_columns.append(run[:n])
    rows_columns[5] = np.array([cols - 1])
    rows_columns[6] = np.array([0, cols - 2, cols - 1])
    lengths = np.array([len(c) for c in rows_columns])
    ext = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    csr = systems.CSR(rows, cols, ext, np.concatenate(rows_columns).astype(np.int32),
                      rng.standard_normal(int(ext[-1])))
    matrix = hip.CSRMatrix(csr)
    matrix.set_launch(64, "blocks", width, unroll)
    matrix.set_launch(32, "blocks", width, unroll)
    assert matrix.launch(32) == ("blocks", width, unroll)
    x = rng.standard_normal(cols)
    y0 = rng.standard_normal(rows)
    for in_scale, out_scale in [(0, 1), (1, -1), (0.5, 2.0)]:
        expected = oracle.csr_product(csr, in_scale, x, out_scale, y0.copy())
        space = hip.HIPSpace(torch_gpu.float64)
        results = []
        for _ in range(2):
            y = to_dev(torch_gpu, y0)
            hip.HIPCSROperator(matrix, space).product(in_scale, to_dev(torch_gpu, x), out_scale, y)
            results.append(y.cpu().numpy())
        assert np.array_equal(results[0], results[1])  # bitwise run to run
        assert np.all(np.abs(results[0] - expected) <= spmv_bound(csr, x, y0, in_scale, out_scale, EPS64))
        space32 = hip.HIPSpace(torch_gpu.float32)
        y32 = to_dev(torch_gpu, y0, torch_gpu.float32)
        hip.HIPCSROperator(matrix, space32).product(in_scale, to_dev(torch_gpu, x, torch_gpu.float32), out_scale, y32)
        assert np.all(np.abs(y32.cpu().numpy() - expected) <= 4 * spmv_bound(csr, x, y0, in_scale, out_scale, EPS32))


def test_spmv_deterministic_and_autotuned(hip, torch_gpu):
    w = systems.q2_laplace_3d(16)
    matrix = hip.CSRMatrix(w.matrix)
    best = matrix.autotune(64, repetitions=5, kernels=("staged",))
    assert best > 0
    assert matrix.launch(64)[0] == "staged"
    space = hip.HIPSpace()
    op = hip.HIPCSROperator(matrix, space)
    x = to_dev(torch_gpu, w.rhs)
    outs = []
    for _ in range(3):
        y = space.makeVector(w.matrix.rows)
        op.product(0, x, 1, y)
        outs.append(y.cpu().numpy())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])  # bitwise run to run
    expected = oracle.csr_product(w.matrix, 0.0, w.rhs, 1.0, np.zeros(w.matrix.rows))
    assert np.array_equal(outs[0], expected)
    # vector rows: deterministic too, within the rounding bound of the CPU order
    matrix.autotune(64, repetitions=3, kernels=("vector",))
    assert matrix.launch(64)[0] == "vector"
    first = space.makeVector(w.matrix.rows)
    op.product(0, x, 1, first)
    second = space.makeVector(w.matrix.rows)
    op.product(0, x, 1, second)
    assert torch_gpu.equal(first, second)
    bound = spmv_bound(w.matrix, w.rhs, np.zeros(w.matrix.rows), 0, 1, EPS64)
    assert np.all(np.abs(first.cpu().numpy() - expected) <= bound)
    # block tiles: the same
    matrix.autotune(64, repetitions=3, kernels=("tiled",))
    assert matrix.launch(64)[0] == "tiled"
    first = space.makeVector(w.matrix.rows)
    op.product(0, x, 1, first)
    second = space.makeVector(w.matrix.rows)
    op.product(0, x, 1, second)
    assert torch_gpu.equal(first, second)
    assert np.all(np.abs(first.cpu().numpy() - expected) <= bound)


def test_spmv_full_size_q2_64(hip, torch_gpu):
    """3D Q2, 64^3 hexahedra (2,048,383 rows, 127,263,527 nonzeros): FP64 and FP32 planes vs the oracle."""
    w = systems.q2_laplace_3d(64)
    assert w.matrix.nnz == 127263527
    matrix = hip.CSRMatrix(w.matrix)
    assert matrix.launch(64) == ("tiled", 256, 8)  # the default launch
    matrix.set_launch(64, "staged", 1024, 8)
    matrix.set_launch(32, "staged", 1024, 8)
    expected = oracle.csr_product(w.matrix, 0.0, w.rhs, 1.0, np.zeros(w.matrix.rows), threads=8)
    space = hip.HIPSpace()
    y = space.makeVector(w.matrix.rows)
    hip.HIPCSROperator(matrix, space).product(0, to_dev(torch_gpu, w.rhs), 1, y)
    got = y.cpu().numpy()
    assert np.array_equal(got, expected)  # bit-identical to the sequential CPU rows
    space32 = hip.HIPSpace(torch_gpu.float32)
    y32 = space32.makeVector(w.matrix.rows)
    hip.HIPCSROperator(matrix, space32).product(0, to_dev(torch_gpu, w.rhs, torch_gpu.float32), 1, y32)
    # FP32 plane: the inputs' and the result's FP32 roundings are inside the summation bound's (len + 3) terms
    bound32 = spmv_bound(w.matrix, w.rhs, np.zeros(w.matrix.rows), 0, 1, EPS32)
    assert np.all(np.abs(y32.cpu().numpy() - expected) <= bound32)
    # block tiles over the narrow 16-bit column plane (every 4096-entry run spans <= 65,160 columns)
    assert matrix.column_bytes() == 2
    bound = spmv_bound(w.matrix, w.rhs, np.zeros(w.matrix.rows), 0, 1, EPS64)
    for unroll in (8, 16):
        matrix.set_launch(64, "tiled", 256, unroll)
        y = space.makeVector(w.matrix.rows)
        hip.HIPCSROperator(matrix, space).product(0, to_dev(torch_gpu, w.rhs), 1, y)
        assert np.all(np.abs(y.cpu().numpy() - expected) <= bound)
    # column blocks (runs of 3 and 5 adjacent columns: 1 x 3 blocks are 13% padding), and the same blocks in
    # column-sorted tiles, whose products are bitwise the plain blocks'
    for width, unroll in ((3, 16), (2, 8), (4, 4), (3, 8)):
        plain = {}
        for kernel in ("blocks", "sorted"):
            matrix.set_launch(64, kernel, width, unroll)
            matrix.set_launch(32, kernel, width, unroll)
            y = space.makeVector(w.matrix.rows)
            hip.HIPCSROperator(matrix, space).product(0, to_dev(torch_gpu, w.rhs), 1, y)
            got64 = y.cpu().numpy()
            assert np.all(np.abs(got64 - expected) <= bound), (kernel, width, unroll)
            y32 = space32.makeVector(w.matrix.rows)
            hip.HIPCSROperator(matrix, space32).product(0, to_dev(torch_gpu, w.rhs, torch_gpu.float32), 1, y32)
            got32 = y32.cpu().numpy()
            assert np.all(np.abs(got32 - expected) <= bound32), (kernel, width, unroll)
            if kernel == "blocks":
                plain = {64: got64, 32: got32}
            else:
                assert np.array_equal(got64, plain[64]) and np.array_equal(got32, plain[32]), (width, unroll)
    # paired 1 x 5 / 1 x 3 tiles (the runs of 5 and 3 without padding) and 1 x 3 / 1 x 2 tiles
    for width, unroll in ((53, 2), (53, 4), (32, 4)):
        matrix.set_launch(64, "paired", width, unroll)
        matrix.set_launch(32, "paired", width, unroll)
        y = space.makeVector(w.matrix.rows)
        hip.HIPCSROperator(matrix, space).product(0, to_dev(torch_gpu, w.rhs), 1, y)
        assert np.all(np.abs(y.cpu().numpy() - expected) <= bound), ("paired", width, unroll)
        y32 = space32.makeVector(w.matrix.rows)
        hip.HIPCSROperator(matrix, space32).product(0, to_dev(torch_gpu, w.rhs, torch_gpu.float32), 1, y32)
        assert np.all(np.abs(y32.cpu().numpy() - expected) <= bound32), ("paired", width, unroll)


def test_spmv_shape_error(hip, torch_gpu):
    matrix = hip.CSRMatrix(systems.poisson5(4))
    space = hip.HIPSpace()
    with pytest.raises(hip._lib.ShapeError):
        hip.HIPCSROperator(matrix, space).product(0, space.makeVector(15), 1, space.makeVector(16))


def test_space_kats(hip, torch_gpu):
    for dtype in (torch_gpu.float64, torch_gpu.float32):
        space = hip.HIPSpace(dtype)
        for n in KAT["inner_product_sizes"]:     # SYCLSpace_test.cpp:27-47
            left = torch_gpu.arange(n, dtype=torch_gpu.float64, device="cuda:0").to(dtype)
            right = torch_gpu.full((n,), -1.0, dtype=dtype, device="cuda:0")
            if dtype == torch_gpu.float64:
                assert space.innerProduct(left, right) == -n * (n - 1) / 2
            else:
                assert space.innerProduct(left, right) == pytest.approx(-n * (n - 1) / 2, rel=1e-5)
        for n in KAT["assign_add_scale_sizes"]:
            source_host = np.arange(n, dtype=np.float64)
            source = space.makeVector(n)
            space.assign(source, source_host)            # :49-68
            copy = space.makeVector(n)
            space.assign(copy, source)
            back = np.empty(n)
            space.assign(back, copy)
            assert np.array_equal(back, source_host)
            target = space.makeVector(n)
            space.assign(target, 10.0 * source_host)
            space.add(target, source, 100.0)             # :70-98
            assert np.allclose(target.cpu().numpy(), 110.0 * source_host, rtol=1e-6, atol=0)
            space.add(target, source, 1.0)
            assert np.allclose(target.cpu().numpy(), 111.0 * source_host, rtol=1e-6, atol=0)
            v = space.makeVector(n)
            space.assign(v, source)
            space.scale(v, 2.0)                          # :100-115
            assert np.array_equal(v.cpu().numpy(), 2.0 * source_host)
            t = space.makeVector(n)
            space.assign(t, 10.0 * source_host)
            space.scale(t, source, 100.0)                # :117-138
            assert np.allclose(t.cpu().numpy(), 1e3 * source_host ** 2, rtol=1e-6, atol=0)
            space.fill(t, -2.5)
            assert np.all(t.cpu().numpy() == -2.5)


def test_space_ops_match_oracle_bitwise(hip, torch_gpu):
    rng = np.random.default_rng(11)
    n = 100003
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    space = hip.HIPSpace()
    for op, scale in [(oracle.OP_SCALE, 1.0), (oracle.OP_SCALE, -1.0), (oracle.OP_SCALE, 0.3), (oracle.OP_ADD, 1.0),
                      (oracle.OP_ADD, -1.0), (oracle.OP_ADD, 0.7), (oracle.OP_SCALE_UNIFORM, 1.7)]:
        expected = oracle.vector_op(op, a.copy(), b, scale)
        t = to_dev(torch_gpu, a)
        if op == oracle.OP_SCALE:
            space.scale(t, to_dev(torch_gpu, b), scale)
        elif op == oracle.OP_ADD:
            space.add(t, to_dev(torch_gpu, b), scale)
        else:
            space.scale(t, scale)
        got = t.cpu().numpy()
        # elementwise ops: the same operations in the same order, no contraction (-ffp-contract=off)
        assert np.array_equal(got, expected), (op, scale)
    dot = space.innerProduct(to_dev(torch_gpu, a), to_dev(torch_gpu, b))
    ref = oracle.inner_product(a, b)
    assert abs(dot - ref) <= 2 * n * EPS64 * np.sum(np.abs(a * b))
    again = space.innerProduct(to_dev(torch_gpu, a), to_dev(torch_gpu, b))
    assert dot == again


def test_diagonal_operator(hip, torch_gpu):
    csr = systems.poisson5(32)
    matrix = hip.CSRMatrix(csr)
    space = hip.HIPSpace()
    diag = hip.makeDiagonalOperator(matrix, space)
    assert np.all(diag.inverseDiagonal.cpu().numpy() == 0.25)
    rng = np.random.default_rng(5)
    x, y0 = rng.standard_normal(csr.rows), rng.standard_normal(csr.rows)
    for in_scale, out_scale in [(0, 1), (0, 2.0), (1, 1), (0.5, -3.0)]:
        expected = oracle.diagonal_product(np.full(csr.rows, 0.25), in_scale, x, out_scale, y0.copy())
        y = to_dev(torch_gpu, y0)
        diag.product(in_scale, to_dev(torch_gpu, x), out_scale, y)
        assert np.allclose(y.cpu().numpy(), expected, rtol=4 * EPS64, atol=1e-300)


def test_diagonal_errors(hip, torch_gpu):
    space = hip.HIPSpace()
    missing = systems.CSR(3, 3, np.array([0, 2, 3, 5], np.int32), np.array([0, 1, 0, 0, 2], np.int32),
                          np.array([2.0, 1.0, 1.0, 1.0, 3.0]))
    with pytest.raises(hip._lib.MissingDiagonal, match="row 1"):
        hip.makeDiagonalOperator(hip.CSRMatrix(missing), space)
    zero = systems.CSR(3, 3, np.array([0, 1, 2, 3], np.int32), np.array([0, 1, 2], np.int32),
                       np.array([1.0, 2.0, 0.0]))
    with pytest.raises(hip._lib.ZeroDiagonal, match="row 2"):
        hip.makeDiagonalOperator(hip.CSRMatrix(zero), space)
