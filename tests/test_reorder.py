"""CPU checks of the oracle's reordering (reorder.cpp): hand-worked orderings,
the tie order of std::sort against GCC's own, and the CSR / vector
permutations against scipy and numpy."""
import subprocess

import numpy as np
import pytest
import scipy.sparse

import oracle
from cie_amd import systems

PROBE = __import__("pathlib").Path(__file__).resolve().parent / "cpp" / "degree_sort_probe.cpp"


def pattern(rows, edges, diagonal=True):
    """Symmetric CSR pattern (values = 1) with the given undirected edges."""
    r = [i for i in range(rows)] if diagonal else []
    c = list(r)
    for a, b in edges:
        r += [a, b]
        c += [b, a]
    m = scipy.sparse.csr_matrix((np.ones(len(r)), (r, c)), shape=(rows, rows))
    m.sum_duplicates()
    m.sort_indices()
    return systems.CSR(rows, rows, m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.astype(np.float64))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    binary = tmp_path_factory.mktemp("probe") / "degree_sort_probe"
    subprocess.run(["g++", "-O2", "-std=c++17", str(PROBE), "-o", str(binary)], check=True)
    return binary


def run_probe(binary, cases):
    text = [str(len(cases))]
    for mode, key, items in cases:
        text.append(f"{mode} {len(items)}")
        text.append(" ".join(map(str, key)))
        text.append(" ".join(map(str, items)))
    out = subprocess.run([str(binary)], input="\n".join(text) + "\n", capture_output=True, text=True, check=True)
    lines = out.stdout.split("\n")
    return [np.array(line.split(), dtype=np.int32) for line in lines[:len(cases)]]


def test_degree_sort_matches_gcc_std_sort(probe):
    rng = np.random.default_rng(0)
    cases = []
    for n in list(range(0, 40)) + [63, 64, 65, 100, 127, 128, 200, 513, 1000]:
        for distinct in (1, 2, 3, 7, 1000):
            key = rng.integers(0, distinct, n).astype(np.int32)
            cases.append((0, key, rng.permutation(n).astype(np.int32)))
    n = 300  # presorted, reversed and organ-pipe keys
    for key in (np.arange(n), np.arange(n)[::-1], np.minimum(np.arange(n), n - np.arange(n))):
        cases.append((0, key.astype(np.int32), np.arange(n, dtype=np.int32)))
    expected = run_probe(probe, cases)
    for (mode, key, items), want in zip(cases, expected):
        got = oracle.degree_sort(items, key)
        assert np.array_equal(got, want), (len(items), key[:8])


def test_heap_step_matches_gcc_partial_sort(probe):
    rng = np.random.default_rng(1)
    cases = [(1, rng.integers(0, d, n).astype(np.int32), rng.permutation(n).astype(np.int32))
             for n in (0, 1, 2, 3, 4, 5, 16, 17, 31, 100, 257) for d in (1, 3, 50)]
    for (mode, key, items), want in zip(cases, run_probe(probe, cases)):
        assert np.array_equal(oracle.heap_sort(items, key), want)


def test_cuthill_mckee_hand_worked_cases():
    path = pattern(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    assert list(oracle.make_reordering(path, "cuthill-mckee")) == [0, 1, 2, 3, 4]
    assert list(oracle.make_reordering(path, "reverse-cuthill-mckee")) == [4, 3, 2, 1, 0]
    assert list(oracle.make_reordering(path, "none")) == [0, 1, 2, 3, 4]
    # the path 3-0-4-1-2 numbered out of order: the first least-degree end (row 2) starts
    scrambled = pattern(5, [(3, 0), (0, 4), (4, 1), (1, 2)])
    assert list(oracle.make_reordering(scrambled, "cuthill-mckee")) == [2, 1, 4, 0, 3]
    # star: centre 0 (degree 6), leaves of degree 2; start at leaf 1, then the centre, then the other leaves
    star = pattern(6, [(0, k) for k in range(1, 6)])
    assert list(oracle.make_reordering(star, "cuthill-mckee")) == [1, 0, 2, 3, 4, 5]
    # components {0,1}, {2} (diagonal only), {3,4}: the isolated row (degree 1) first, then by least degree
    parts = pattern(5, [(0, 1), (3, 4)])
    assert list(oracle.make_reordering(parts, "cuthill-mckee")) == [2, 0, 1, 3, 4]
    # 3x3 five-point grid: corners have degree 3, edges 4, the centre 5
    grid = pattern(9, [(0, 1), (1, 2), (3, 4), (4, 5), (6, 7), (7, 8), (0, 3), (3, 6), (1, 4), (4, 7), (2, 5), (5, 8)])
    assert list(oracle.make_reordering(grid, "cuthill-mckee")) == [0, 1, 3, 2, 4, 6, 5, 7, 8]
    # the queued batch is sorted by degree: from row 0, neighbours 1 (degree 2) and 2 (degree 4) -> 1 first;
    # from row 1 (batch of 3 with degrees 3, 2), 3 goes after 4
    small = pattern(6, [(0, 2), (0, 1), (2, 3), (2, 5), (1, 4), (4, 3)])
    assert list(oracle.make_reordering(small, "cuthill-mckee"))[:2] == [5, 2]


def bandwidth(csr):
    rows = np.repeat(np.arange(csr.rows), np.diff(csr.row_extents))
    return int(np.max(np.abs(rows - csr.column_indices))) if csr.nnz else 0


@pytest.mark.parametrize("dim,order,n", [(2, 1, 20), (2, 2, 9), (3, 1, 7), (3, 2, 6)])
def test_rcm_restores_a_narrow_band_after_a_random_numbering(dim, order, n):
    natural = systems.lagrange(dim, order, n)
    shuffle = np.random.default_rng(n).permutation(natural.rows).astype(np.int32)
    shuffled = systems.CSR(natural.rows, natural.rows, *oracle.reorder_csr(natural, shuffle))
    mapping = oracle.make_reordering(shuffled, "reverse-cuthill-mckee")
    assert np.array_equal(np.sort(mapping), np.arange(natural.rows))
    assert np.array_equal(oracle.make_reordering(shuffled, "cuthill-mckee")[::-1], mapping)
    reordered = systems.CSR(natural.rows, natural.rows, *oracle.reorder_csr(shuffled, mapping))
    assert bandwidth(shuffled) > 2 * bandwidth(reordered)
    # level sets from a corner run diagonally through the mesh: within 2.5 times the lexicographic band
    assert bandwidth(reordered) <= 2.5 * bandwidth(natural)
    # scipy's reverse Cuthill-McKee (its own start and tie rules) reaches the same band width
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    perm = reverse_cuthill_mckee(shuffled.to_scipy(), symmetric_mode=True)
    reference = shuffled.to_scipy()[perm][:, perm]
    ref_rows = np.repeat(np.arange(natural.rows), np.diff(reference.indptr))
    assert bandwidth(reordered) <= 1.1 * int(np.max(np.abs(ref_rows - reference.indices)))


@pytest.mark.parametrize("reverse", [False, True])
def test_reorder_csr_against_scipy_permutation(reverse):
    rng = np.random.default_rng(3)
    rows = 300
    a = scipy.sparse.random(rows, rows, density=0.03, random_state=4, format="csr") + scipy.sparse.eye(rows)
    a = a.tocsr()
    a.sort_indices()
    a.data = rng.standard_normal(a.nnz)
    csr = systems.CSR(rows, rows, a.indptr.astype(np.int32), a.indices.astype(np.int32), a.data)
    mapping = rng.permutation(rows).astype(np.int32)
    extents, columns, entries = oracle.reorder_csr(csr, mapping, reverse=reverse)
    source = np.argsort(mapping) if reverse else mapping  # new row i is old row source[i]
    expected = a[source][:, source].tocsr()
    expected.sort_indices()
    assert np.array_equal(extents, expected.indptr)
    assert np.array_equal(columns, expected.indices)
    assert np.array_equal(entries, expected.data)
    with pytest.raises(oracle.OracleError):
        oracle.reorder_csr(csr, np.zeros(rows, dtype=np.int32))


def test_reorder_vector_gather_and_scatter():
    rng = np.random.default_rng(5)
    values = rng.standard_normal(50)
    mapping = rng.permutation(50).astype(np.int32)
    gathered = oracle.reorder_vector(mapping, values)
    assert np.array_equal(gathered, values[mapping])
    assert np.array_equal(oracle.reorder_vector(mapping, gathered, reverse=True), values)
