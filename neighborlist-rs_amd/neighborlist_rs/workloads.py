"""Synthetic inputs of the benchmark configurations (deterministic, seeded).

  c1_isolated      N=1000 uniform in a 22 A cube, no cell (auto-box), rc 6
  c2_ethanol       112 rigid ethanol-like 9-atom units on a jittered 5^3
                   lattice in a cubic box L=22.15 A (N=1008), rc 6 / 14
  c3_triclinic     1112 units (N=10008) in the triclinic cell
                   a=(47.6,0,0) b=(14.3,45.4,0) c=(9.5,11.2,44.8), rc 6 / 14
  c4_ragged_batch  256 systems, N log-uniform in [8, 2048], half isolated,
                   half periodic (orthorhombic / triclinic) at liquid density,
                   cutoffs {5, 6, 14}
  c5_serving_trace Poisson arrivals of single systems, N log-uniform in
                   [16, 50000], isolated / orthorhombic / triclinic, rc U[4, 12]
  gnn_molecule_batch  the reference's GNN batch: 128 isolated molecules,
                   42 atoms on average (5,376 atoms), rc 6

Cells are returned as H (lattice vectors as columns), the convention of
PyCell and of the C ABI.
"""
import numpy as np

DENSITY = 0.0928  # atoms / A^3 of the ethanol-like liquid at L = 22.15 A


def ethanol_unit():
    """9-atom CH3-CH2-OH geometry: C-C 1.54, C-O 1.43, C-H 1.09, O-H 0.96 A,
    tetrahedral angles (109.47 deg), centred on its centroid."""
    tet = np.arccos(-1.0 / 3.0)
    c1 = np.zeros(3)
    c2 = np.array([1.54, 0.0, 0.0])

    def tetra_dirs(axis):
        # three directions at the tetrahedral angle from -axis, 120 deg apart
        axis = axis / np.linalg.norm(axis)
        perp = np.cross(axis, [0.0, 0.0, 1.0])
        if np.linalg.norm(perp) < 1e-8:
            perp = np.cross(axis, [0.0, 1.0, 0.0])
        perp /= np.linalg.norm(perp)
        perp2 = np.cross(axis, perp)
        out = []
        for k in range(3):
            phi = 2.0 * np.pi * k / 3.0
            v = (-np.cos(tet) * -axis + np.sin(tet) * (np.cos(phi) * perp + np.sin(phi) * perp2))
            out.append(-v if np.dot(v, axis) > 0 else v)
        return out

    d1 = tetra_dirs(c2 - c1)          # around C1, pointing away from C2
    d2 = tetra_dirs(c1 - c2)          # around C2, pointing away from C1
    o = c2 + 1.43 * d2[0]
    hs = [c1 + 1.09 * v for v in d1] + [c2 + 1.09 * d2[1], c2 + 1.09 * d2[2]]
    od = tetra_dirs(c2 - o)
    ho = o + 0.96 * od[0]
    unit = np.array([c1, c2, o, *hs, ho])
    return unit - unit.mean(axis=0)


def _rotations(rng, m):
    q = rng.normal(size=(m, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], axis=1)
    return R


def molecular_liquid(H, n_units, seed, jitter=0.25):
    """Rigid ethanol-like units with random rotations on a jittered lattice
    of fractional sites inside the cell H (columns = lattice vectors)."""
    rng = np.random.default_rng(seed)
    g = int(np.ceil(n_units ** (1.0 / 3.0)))
    sites = np.array([(i, j, k) for i in range(g) for j in range(g) for k in range(g)],
                     dtype=np.float64)
    sites = sites[rng.permutation(len(sites))[:n_units]]
    frac = (sites + 0.5 + rng.uniform(-jitter, jitter, size=sites.shape)) / g
    centres = frac @ np.asarray(H, dtype=np.float64).T
    unit = ethanol_unit()
    R = _rotations(rng, n_units)
    pos = centres[:, None, :] + np.einsum("mij,aj->mai", R, unit)
    return np.ascontiguousarray(pos.reshape(-1, 3))


def c1_isolated(seed=1):
    rng = np.random.default_rng(seed)
    return np.ascontiguousarray(rng.uniform(0.0, 22.0, size=(1000, 3))), None, 6.0


def c2_ethanol(seed=2):
    H = np.eye(3) * 22.15
    return molecular_liquid(H, 112, seed), H, (6.0, 14.0)


C3_ROWS = np.array([[47.6, 0.0, 0.0], [14.3, 45.4, 0.0], [9.5, 11.2, 44.8]])


def c3_triclinic(seed=3):
    H = C3_ROWS.T.copy()
    return molecular_liquid(H, 1112, seed), H, (6.0, 14.0)


def random_system(rng, n, kind, density=0.1):
    """One system of n atoms: kind 0 isolated cluster, 1 orthorhombic PBC,
    2 triclinic PBC, at the given number density."""
    vol = n / density
    if kind == 0:
        L = vol ** (1.0 / 3.0)
        return rng.uniform(0.0, L, size=(n, 3)), None
    if kind == 1:
        a = rng.uniform(0.8, 1.25, size=3)
        a *= (vol / np.prod(a)) ** (1.0 / 3.0)
        H = np.diag(a)
    else:
        a = vol ** (1.0 / 3.0)
        rows = np.array([[a, 0, 0],
                         [rng.uniform(-0.3, 0.3) * a, a, 0],
                         [rng.uniform(-0.3, 0.3) * a, rng.uniform(-0.3, 0.3) * a, a]])
        H = rows.T.copy()
    frac = rng.uniform(0.0, 1.0, size=(n, 3))
    return frac @ H.T, H


def c4_ragged_batch(seed=4, n_systems=256, nmin=8, nmax=2048):
    """Returns positions (N,3), sizes (S,), cells list (H or None), cutoffs."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=n_systems)).astype(np.int64)
    sizes = np.clip(sizes, nmin, nmax)
    pos, cells = [], []
    for s in range(n_systems):
        kind = 0 if s % 2 == 0 else 1 + (s // 2) % 2
        p, H = random_system(rng, int(sizes[s]), kind)
        pos.append(p)
        cells.append(H)
    return np.ascontiguousarray(np.concatenate(pos)), sizes, cells, (5.0, 6.0, 14.0)


def c5_serving_trace(seed=5, n_requests=200, qps=200.0, nmin=16, nmax=50000):
    """Poisson arrivals: list of (arrival_s, positions, H or None, cutoff)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / qps, size=n_requests))
    out = []
    for r in range(n_requests):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        kind = int(rng.integers(0, 3))
        p, H = random_system(rng, n, kind)
        out.append((float(t[r]), np.ascontiguousarray(p), H, float(rng.uniform(4.0, 12.0))))
    return out


def benzene_unit():
    """12-atom C6H6 ring: C-C 1.39 A, C-H 1.09 A, planar, centred."""
    ang = np.arange(6) * (np.pi / 3.0)
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(6)], -1)
    return np.concatenate([1.39 * ring, (1.39 + 1.09) * ring])


def gnn_molecule_batch(seed=6, n_molecules=128):
    """The molecular batch of the reference's GNN benchmark
    (benchmarks/batch_radius_benchmark.py:15-51): isolated molecules of 30 to
    50 atoms, each a small molecule replicated along x at a 10 A period
    (ASE's m * (2, 1, 1) in a 10 A cell) until it has at least 30 atoms,
    128 molecules averaging 42 atoms (5,376 in all, batch_results.json:2-4).
    Without ASE: alternately four copies of the ethanol-like unit (36 atoms)
    and of a benzene ring (48 atoms), each molecule randomly rotated.
    Returns positions (N, 3) and the int32 batch ids."""
    rng = np.random.default_rng(seed)
    units = [ethanol_unit(), benzene_unit()]
    R = _rotations(rng, n_molecules)
    pos, ids = [], []
    for m in range(n_molecules):
        u = units[m % 2]
        mol = np.concatenate([u + np.array([10.0 * c, 0.0, 0.0]) for c in range(4)])
        mol = mol @ R[m].T + rng.uniform(-5.0, 5.0, size=3)
        pos.append(mol)
        ids.append(np.full(mol.shape[0], m, np.int32))
    return np.ascontiguousarray(np.concatenate(pos)), np.concatenate(ids)


def batch_arrays(sizes, cells):
    """(batch ids int32, cells (S,3,3) ASE rows with zeros for isolated)."""
    batch = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    rows = np.zeros((len(sizes), 3, 3))
    for s, H in enumerate(cells):
        if H is not None:
            rows[s] = np.asarray(H).T
    return batch, rows
