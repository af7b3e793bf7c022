"""Batched PCG: many uneven systems in one launch per step, against the oracle."""
import numpy as np
import pytest

import oracle
from cie_amd import systems
from cie_amd._lib import Statistics

pytestmark = pytest.mark.gpu


def small_batch(count=12, seed=7):
    return systems.q2_quad_batch(count=count, seed=seed, smallest=3, largest=24)


def set_format(batch, blocks):
    """0 / 3: entry or 1 x 3 column-block tiles; "sorted" / "paired": those tile kernels."""
    if blocks in ("sorted", "paired"):
        batch.set_tile_kernel(blocks)
        assert batch.tile_kernel() == blocks
    else:
        batch.set_column_blocks(blocks)
        assert batch.column_blocks() == blocks


def per_system_residuals(w, x):
    r = oracle.csr_product(w.matrix, 1.0, x, -1.0, w.rhs.copy(), threads=8)
    offsets = w.extra["offsets"]
    return np.array([np.linalg.norm(r[a:b]) / np.linalg.norm(w.rhs[a:b]) for a, b in zip(offsets[:-1], offsets[1:])])


@pytest.mark.parametrize("blocks", [0, 3, "paired"])
def test_batched_fp64_matches_oracle_per_system(hip, torch_gpu, blocks):
    w = small_batch()
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    set_format(batch, blocks)
    b = torch_gpu.from_numpy(w.rhs).cuda()
    x = torch_gpu.zeros_like(b)
    stats, status, _, summary = batch.solve(b, x, Statistics(20000, 0.0, 1e-10), mixed=False)
    assert np.all(status == 1) and summary.converged == 1
    got = x.cpu().numpy()
    offsets = w.extra["offsets"]
    for i, (a, c) in enumerate(zip(offsets[:-1], offsets[1:])):
        sub = systems.CSR(int(c - a), int(c - a), w.matrix.row_extents[a:c + 1] - w.matrix.row_extents[a],
                          w.matrix.column_indices[w.matrix.row_extents[a]:w.matrix.row_extents[c]] - a,
                          w.matrix.entries[w.matrix.row_extents[a]:w.matrix.row_extents[c]])
        x_ref, s_ref, _, _ = oracle.pcg(sub, w.rhs[a:c], relative_tolerance=1e-10)
        assert abs(int(stats[i].iterationCount) - int(s_ref.iterationCount)) <= 1
        assert np.linalg.norm(got[a:c] - x_ref) <= 1e-9 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("restart,blocks,correction", [(False, 0, True), (True, 0, True), (False, 3, True),
                                                      (False, 3, False), (False, "paired", True)])
def test_batched_mixed_reaches_fp64_tolerance(hip, torch_gpu, restart, blocks, correction):
    w = small_batch(seed=8)
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    set_format(batch, blocks)
    b = torch_gpu.from_numpy(w.rhs).cuda()
    x = torch_gpu.zeros_like(b)
    stats, status, refinements, _ = batch.solve(b, x, Statistics(20000, 0.0, 1e-10), mixed=True, restart=restart,
                                                row_sum_correction=correction)
    assert np.all(status == 1) and np.all(refinements >= 1)
    assert np.all(per_system_residuals(w, x.cpu().numpy()) < 1e-10)
    assert all(s.relativeResidual < 1e-10 for s in stats)


@pytest.mark.parametrize("blocks", [None, 0, 3, "sorted", "paired"])
def test_batched_graph_replays_match_eager(hip, torch_gpu, blocks):
    w = small_batch(seed=9)
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    batch.autotune(32, 2)
    assert set(batch.tile_rows().tolist()) <= {32, 64, 128, 256}
    assert batch.column_blocks() in (0, 3)
    assert batch.tile_kernel() in ("tiled", "blocks", "sorted", "paired")
    if blocks is not None:
        set_format(batch, blocks)
        if blocks == "sorted":
            assert batch.column_blocks() == 3
    b = torch_gpu.from_numpy(w.rhs).cuda()
    results = []
    for graph in (False, True, True, True):   # the graph is captured once and replayed on later calls
        x = torch_gpu.zeros_like(b)
        stats, status, _, _ = batch.solve(b, x, Statistics(20000, 0.0, 1e-10), mixed=True, graph=graph)
        torch_gpu.cuda.synchronize()
        _ = (b * 2).sum().item()                # eager work between replays
        results.append((x.cpu().numpy(), [s.iterationCount for s in stats]))
    for x, iterations in results[1:]:
        assert np.array_equal(x, results[0][0]) and iterations == results[0][1]


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_graph_follows_new_settings(hip, torch_gpu, mixed):
    """A captured trip takes the tolerances, the iteration cap and the refinement threshold by value: a
    second graph solve of the same b and x with other settings must capture again, reach its own tolerance
    and equal the eager solve with those settings bit for bit."""
    w = small_batch(seed=13)
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    b = torch_gpu.from_numpy(w.rhs).cuda()
    x = torch_gpu.zeros_like(b)
    loose, _, _, _ = batch.solve(b, x, Statistics(20000, 0.0, 1e-6), mixed=mixed, graph=True)
    assert all(s.relativeResidual < 1e-6 for s in loose)
    runs = []
    for graph in (True, False):
        x = torch_gpu.zeros_like(b)
        stats, status, _, _ = batch.solve(b, x, Statistics(20000, 0.0, 1e-12), mixed=mixed, graph=graph)
        torch_gpu.cuda.synchronize()
        assert np.all(status == 1)
        assert all(s.relativeResidual < 1e-12 for s in stats)
        runs.append((x.cpu().numpy(), [s.iterationCount for s in stats]))
    assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]
    assert sum(runs[0][1]) > sum(s.iterationCount for s in loose)
    # and a cap: the graph captured at 20000 must not keep that cap
    x = torch_gpu.zeros_like(b)
    capped, status, _, _ = batch.solve(b, x, Statistics(5, 0.0, 1e-12), mixed=mixed, graph=True)
    assert all(s.iterationCount <= 5 for s in capped) and not np.all(status == 1)


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_sorted_tiles_equal_column_blocks(hip, torch_gpu, mixed):
    """Column-sorted tiles stage every product at its block's position: the batch solves bitwise as it does on
    the plain column-block tiles (tile partials are per tile, so the grids do not matter)."""
    w = small_batch(seed=11)
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    batch.set_tile_rows(np.full(batch.systems, 64, dtype=np.int32))
    b = torch_gpu.from_numpy(w.rhs).cuda()
    results = []
    for kernel in ("blocks", "sorted", "blocks"):
        batch.set_tile_kernel(kernel)
        assert batch.tile_kernel() == kernel
        x = torch_gpu.zeros_like(b)
        stats, status, _, _ = batch.solve(b, x, Statistics(20000, 0.0, 1e-10), mixed=mixed)
        assert np.all(status == 1)
        results.append((x.cpu().numpy(), [s.iterationCount for s in stats]))
    for x, iterations in results[1:]:
        assert np.array_equal(x, results[0][0]) and iterations == results[0][1]
    assert np.all(per_system_residuals(w, results[1][0]) < 1e-10)
    with pytest.raises(hip._lib.InvalidArgument):
        batch.set_tile_kernel("vector")


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_paired_tiles_switch_and_agree(hip, torch_gpu, mixed):
    """Paired 1 x 5 / 1 x 3 tiles, cut per system like the other formats: switching to them and back leaves
    every other format's results bitwise as they were; the paired solves are bitwise reproducible, reach the
    tolerance per system and agree with the column-block solves to the tolerance's accuracy."""
    w = small_batch(seed=12)
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    batch.set_tile_rows(np.full(batch.systems, 64, dtype=np.int32))
    b = torch_gpu.from_numpy(w.rhs).cuda()
    results = {}
    for kernel in ("blocks", "paired", "tiled", "paired", "blocks"):
        batch.set_tile_kernel(kernel)
        assert batch.tile_kernel() == kernel
        assert batch.tile_rows().tolist() == [64] * batch.systems
        x = torch_gpu.zeros_like(b)
        stats, status, _, _ = batch.solve(b, x, Statistics(20000, 0.0, 1e-10), mixed=mixed)
        assert np.all(status == 1)
        got = (x.cpu().numpy(), [s.iterationCount for s in stats])
        if kernel in results:
            assert np.array_equal(got[0], results[kernel][0]) and got[1] == results[kernel][1], kernel
        results[kernel] = got
    assert np.all(per_system_residuals(w, results["paired"][0]) < 1e-10)
    reference = results["blocks"][0]
    assert np.linalg.norm(results["paired"][0] - reference) <= 1e-8 * np.linalg.norm(reference)


def test_batched_rejects_coupled_systems(hip, torch_gpu):
    csr = systems.poisson5(8)
    matrix = hip.CSRMatrix(csr)
    with pytest.raises(hip._lib.ShapeError):
        hip.Batch(matrix, [0, 32, 64])   # the 5-point stencil couples rows 31 and 32


def test_full_batch_256_uneven_systems(hip, torch_gpu):
    """256 2D Q2 systems, n_i in [16, 256] (seed 3), lognormal(0, 0.5) conductivity, mixed refinement to 1e-10."""
    w = systems.q2_quad_batch(256)
    assert len(w.extra["sizes"]) == 256
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    batch.autotune(32, 2)
    b = torch_gpu.from_numpy(w.rhs).cuda()
    x = torch_gpu.zeros_like(b)
    stats, status, refinements, summary = batch.solve(b, x, Statistics(100000, 0.0, 1e-10), mixed=True)
    assert np.all(status == 1)
    residuals = per_system_residuals(w, x.cpu().numpy())
    assert np.all(residuals < 1e-10)
    print(f"batch: {w.matrix.rows} rows, {w.matrix.nnz} nnz, {summary.seconds:.3f} s, max iterations "
          f"{summary.stats.iterationCount}, refinements {int(refinements.sum())}, tile format "
          f"{batch.tile_kernel()}, tile rows "
          f"{np.bincount(batch.tile_rows(), minlength=257)[[32, 64, 128, 256]].tolist()}")


def test_batched_tile_rows(hip, torch_gpu):
    """Rows per block tile change the sum order only: same iteration counts (within one), same solutions."""
    w = small_batch(seed=10)
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    b = torch_gpu.from_numpy(w.rhs).cuda()
    results = []
    for rows in (32, 256):
        batch.set_tile_rows(np.full(batch.systems, rows, dtype=np.int32))
        assert set(batch.tile_rows().tolist()) == {rows}
        x = torch_gpu.zeros_like(b)
        stats, status, _, _ = batch.solve(b, x, Statistics(20000, 0.0, 1e-10), mixed=False)
        assert np.all(status == 1)
        results.append((x.cpu().numpy(), np.array([s.iterationCount for s in stats])))
    assert np.all(np.abs(results[0][1].astype(int) - results[1][1].astype(int)) <= 1)
    assert np.linalg.norm(results[0][0] - results[1][0]) <= 1e-9 * np.linalg.norm(results[1][0])
    with pytest.raises(hip._lib.CieHipError):
        batch.set_tile_rows(np.full(batch.systems, 48, dtype=np.int32))
    with pytest.raises(hip._lib.InvalidArgument):
        batch.set_column_blocks(2)


def test_batch_keeps_its_column_blocks_through_the_matrix_autotuner(hip, torch_gpu):
    """A batch walking column blocks pins the block form: the matrix's own autotuner, which frees the block forms
    its launches do not use, keeps it, and later value updates reach it."""
    w = small_batch(seed=11)
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    batch.set_column_blocks(3)
    b = torch_gpu.from_numpy(w.rhs).cuda()
    x0 = torch_gpu.zeros_like(b)
    batch.solve(b, x0, Statistics(20000, 0.0, 1e-10), mixed=False)
    matrix.autotune(64, 2, kernels=("tiled",))
    matrix.update_entries(torch_gpu.from_numpy(2.0 * w.matrix.entries).cuda())
    x1 = torch_gpu.zeros_like(b)
    stats, status, _, _ = batch.solve(b, x1, Statistics(20000, 0.0, 1e-10), mixed=False)
    assert np.all(status == 1)
    # (2A) x1 = b: x1 = x0 / 2 to the solve's tolerance
    assert np.linalg.norm(2.0 * x1.cpu().numpy() - x0.cpu().numpy()) <= 1e-8 * np.linalg.norm(x0.cpu().numpy())


@pytest.mark.parametrize("systems_count", [1024, 1100])
def test_batched_selection_of_many_systems(hip, torch_gpu, systems_count):
    """The running and refining systems are listed by one kernel up to 1024 systems and by a compaction and an
    expansion kernel above; both lists give each system the oracle's solution (FP64) and the FP64 tolerance
    (mixed)."""
    w = systems.q2_quad_batch(count=systems_count, seed=5, smallest=1, largest=4)
    matrix = hip.CSRMatrix(w.matrix)
    batch = hip.Batch(matrix, w.extra["offsets"])
    b = torch_gpu.from_numpy(w.rhs).cuda()
    x = torch_gpu.zeros_like(b)
    stats, status, _, _ = batch.solve(b, x, Statistics(20000, 0.0, 1e-10), mixed=False)
    assert np.all(status == 1)
    got = x.cpu().numpy()
    offsets = w.extra["offsets"]
    for i in range(0, systems_count, 37):  # a sample of the systems against the oracle
        a, c = int(offsets[i]), int(offsets[i + 1])
        sub = systems.CSR(c - a, c - a, w.matrix.row_extents[a:c + 1] - w.matrix.row_extents[a],
                          w.matrix.column_indices[w.matrix.row_extents[a]:w.matrix.row_extents[c]] - a,
                          w.matrix.entries[w.matrix.row_extents[a]:w.matrix.row_extents[c]])
        x_ref, s_ref, _, _ = oracle.pcg(sub, w.rhs[a:c], relative_tolerance=1e-10)
        assert abs(int(stats[i].iterationCount) - int(s_ref.iterationCount)) <= 1
        assert np.linalg.norm(got[a:c] - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    x.zero_()
    stats, status, refinements, _ = batch.solve(b, x, Statistics(20000, 0.0, 1e-10), mixed=True)
    assert np.all(status == 1)
    assert np.all(per_system_residuals(w, x.cpu().numpy()) < 1e-10)
