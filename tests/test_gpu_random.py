"""Seeded random configurations against the oracle (bit-exact after the
canonical sort), to reach corners the fixed cases may not:

* 1 and 2 atoms, coincident atoms (d = 0 pairs), atoms several periods
  outside the cell (unwrapped input);
* small and strongly skewed cells where the cutoff exceeds the cell (many
  images, self images), mixed periodicity flags;
* one to four cutoffs, unsorted and repeated, cumulative and disjoint;
* single systems and ragged batches mixing all of the above;
* full lists (the oracle's half list mirrored).

Each case prints its seed before it runs."""
import numpy as np
import pytest

from conftest import as_tuple, assert_same_edges

pytestmark = pytest.mark.gpu


def _mirror(ei, ej, s):
    keep = ei != ej
    return (np.concatenate([ei, ej[keep]]), np.concatenate([ej, ei[keep]]),
            np.concatenate([s, -s[keep]]))


def _random_system(rng):
    """(positions, H with lattice vectors as columns or None, pbc)"""
    n = int(rng.choice([1, 2, 3, 17, 64, 150, 400, 900, 1200]))
    kind = rng.integers(0, 3)
    if kind == 0:
        span = rng.uniform(2.0, 25.0)
        pos = rng.uniform(0.0, span, (n, 3))
        H, pbc = None, None
    else:
        L = rng.uniform(3.0, 20.0, 3)
        H = np.diag(L)
        if kind == 2:  # triclinic: skew the upper triangle
            H[0, 1] = rng.uniform(-0.5, 0.5) * L[1]
            H[0, 2] = rng.uniform(-0.5, 0.5) * L[2]
            H[1, 2] = rng.uniform(-0.5, 0.5) * L[2]
        # at most 0.15 atoms per cubic angstrom (liquid-like), so small cells
        # with many images stay small problems
        n = max(1, min(n, int(0.15 * abs(np.linalg.det(H)))))
        f = rng.uniform(-2.0, 3.0, (n, 3))  # unwrapped: up to two periods outside
        pos = f @ H.T
        pbc = [bool(x) for x in rng.integers(0, 2, 3)]
        if not any(pbc):
            pbc[int(rng.integers(0, 3))] = True
    if n >= 3 and rng.random() < 0.3:
        pos[1] = pos[0]  # coincident atoms
    return np.ascontiguousarray(pos), H, pbc


def _random_cutoffs(rng, H):
    k = int(rng.integers(1, 5))
    top = 12.0 if H is None else 1.2 * float(np.min(np.diag(H))) + 2.0
    cut = [float(rng.uniform(0.5, top)) for _ in range(k)]
    if k > 1 and rng.random() < 0.3:
        cut[-1] = cut[0]  # repeated cutoff
    return cut


def _cell(nr, H, pbc):
    return None if H is None else nr.PyCell(H.tolist(), pbc)


@pytest.mark.parametrize("seed", range(48))
def test_random_single_vs_oracle(nr, oracle, seed):
    rng = np.random.default_rng(1000 + seed)
    pos, H, pbc = _random_system(rng)
    cut = _random_cutoffs(rng, H)
    disjoint = bool(rng.integers(0, 2))
    print("[case] random single seed", seed, "n", pos.shape[0], "cell", H is not None, "pbc", pbc,
          "cut", [round(c, 3) for c in cut], "disjoint", disjoint, flush=True)
    cell = _cell(nr, H, pbc)
    want = oracle.search_single_multi(pos, H, pbc, cut, disjoint)
    got = nr.build_neighborlists_multi(cell, pos, cut, disjoint=disjoint)
    for q in range(len(cut)):
        assert_same_edges(as_tuple(got[q]), want[q], f"seed {seed} list {q}")
    one = nr.build_neighborlists(cell, pos, cut[0])
    assert_same_edges(as_tuple(one), oracle.search_single(pos, H, pbc, cut[0]),
                      f"seed {seed} single cutoff")


@pytest.mark.parametrize("seed", range(16))
def test_random_batch_vs_oracle(nr, oracle, seed):
    rng = np.random.default_rng(2000 + seed)
    systems = [_random_system(rng) for _ in range(int(rng.integers(1, 9)))]
    cut = _random_cutoffs(rng, None)
    disjoint = bool(rng.integers(0, 2))
    sizes = [p.shape[0] for p, _, _ in systems]
    pos = np.ascontiguousarray(np.concatenate([p for p, _, _ in systems]))
    batch = np.repeat(np.arange(len(systems), dtype=np.int32), sizes)
    # batch cells are ASE rows with periodicity all true (zero rows: isolated)
    rows = np.zeros((len(systems), 3, 3))
    ocells = []
    for s, (_, H, _) in enumerate(systems):
        if H is not None:
            rows[s] = H.T
            ocells.append((H, [True] * 3))
        else:
            ocells.append(None)
    print("[case] random batch seed", seed, "sizes", sizes, "cut", [round(c, 3) for c in cut],
          "disjoint", disjoint, flush=True)
    want = oracle.search_batch_multi(pos, batch, ocells, cut, disjoint)
    got = nr.build_neighborlists_batch_multi(pos, batch, rows, cut, disjoint=disjoint)
    for q in range(len(cut)):
        assert_same_edges(as_tuple(got[q]), want[q], f"seed {seed} list {q}")


@pytest.mark.parametrize("seed", range(16))
def test_random_full_lists_vs_oracle(nr, oracle, seed):
    import torch
    rng = np.random.default_rng(3000 + seed)
    pos, H, pbc = _random_system(rng)
    rc = _random_cutoffs(rng, H)[0]
    print("[case] random full seed", seed, "n", pos.shape[0], "cell", H is not None, "pbc", pbc,
          "rc", round(rc, 3), flush=True)
    got = nr.build_neighborlists_torch(_cell(nr, H, pbc), torch.from_numpy(pos).cuda(), rc,
                                       full=True)
    c = lambda x: x.detach().cpu().numpy()
    assert_same_edges((c(got["edge_index"][0]), c(got["edge_index"][1]), c(got["shift"])),
                      _mirror(*oracle.search_single(pos, H, pbc, rc)), f"seed {seed} full")
