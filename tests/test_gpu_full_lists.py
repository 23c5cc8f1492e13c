"""GPU full (directed) lists and device-resident torch outputs.

A full list is defined from the half list the reference emits: every edge
(i, j, S) with i != j also appears as (j, i, -S); a self image (i, i, S)
appears once per S (the half list already holds both signs).  Checks:

* against the oracle's half list mirrored, bit-exact after the canonical sort;
* the layout: slot 2e holds an edge and slot 2e + 1 its mirror, and without
  self images the even slots are the half list itself, order included;
* an independent ordered-pair search over lattice images on small cells;
* the reference-logged count of ethanol_10000 at rc 6 (2 x logged + self images);
* the torch functions against the numpy ones, DLPack lifetime and stream order.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import as_tuple, assert_same_edges
from xyzio import read_xyz

pytestmark = pytest.mark.gpu

NL_FULL, NL_DIST = 16, 8


def _print_case(*a):
    print("[case]", *a, flush=True)


def _ex(pos, sizes, cells, cutoffs, disjoint=False, multi=False, opts=0):
    """nl_search_ex through the C ABI with host positions; per list
    (ei, ej, shifts, dist or None)."""
    from neighborlist_rs import _lib
    lib = _lib.lib
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    sizes = np.ascontiguousarray(sizes, dtype=np.int64)
    S = sizes.size
    H = np.zeros((S, 9))
    has = np.zeros(S, np.uint8)
    pb = np.ones((S, 3), np.uint8)
    for s, c in enumerate(cells):
        if c is not None:
            H[s] = np.asarray(c[0], dtype=np.float64).reshape(9)
            has[s] = 1
            pb[s] = [1 if v else 0 for v in c[1]]
    cut = np.ascontiguousarray(cutoffs, dtype=np.float64)
    r = C.c_void_p()
    rc = lib.nl_search_ex(pos.ctypes.data_as(C.c_void_p), pos.shape[0],
                          sizes.ctypes.data_as(_lib.i64p), S, H.ctypes.data_as(_lib.dp),
                          has.ctypes.data_as(_lib.u8p), pb.ctypes.data_as(_lib.u8p),
                          cut.ctypes.data_as(_lib.dp), cut.size, int(disjoint), int(multi), opts,
                          None, C.byref(r))
    assert rc == 0, _lib.last_error()
    out = []
    try:
        for q in range(cut.size):
            m = int(lib.nl_result_num_edges(r, q))
            ei, ej = np.empty(m, np.int64), np.empty(m, np.int64)
            sh = np.empty((m, 3), np.int32)
            d = np.empty(m) if opts & NL_DIST else None
            if m:
                assert lib.nl_result_copy(r, q, ei.ctypes.data_as(_lib.i64p),
                                          ej.ctypes.data_as(_lib.i64p),
                                          sh.ctypes.data_as(_lib.i32p)) == 0
                if d is not None:
                    assert lib.nl_result_copy_dist(r, q, d.ctypes.data_as(_lib.dp)) == 0
            out.append((ei, ej, sh, d))
    finally:
        lib.nl_result_free(r)
    return out


def _mirror(ei, ej, s):
    keep = ei != ej
    return (np.concatenate([ei, ej[keep]]), np.concatenate([ej, ei[keep]]),
            np.concatenate([s, -s[keep]]))


def _check_layout(full, half, what):
    fi, fj, fs, fd = full
    hi, hj, hs, hd = half
    assert fi.size % 2 == 0, what
    np.testing.assert_array_equal(fi[0::2], fj[1::2], err_msg=what)
    np.testing.assert_array_equal(fj[0::2], fi[1::2], err_msg=what)
    np.testing.assert_array_equal(fs[0::2], -fs[1::2], err_msg=what)
    n_self = int((hi == hj).sum())
    assert fi.size == 2 * hi.size - n_self, what
    if n_self == 0:
        # the even slots are the half list, order included
        np.testing.assert_array_equal(fi[0::2], hi, err_msg=what)
        np.testing.assert_array_equal(fj[0::2], hj, err_msg=what)
        np.testing.assert_array_equal(fs[0::2], hs, err_msg=what)
        if fd is not None and hd is not None:
            np.testing.assert_array_equal(fd[0::2], hd, err_msg=what)
    if fd is not None:
        np.testing.assert_array_equal(fd[0::2], fd[1::2], err_msg=what)


def _cases():
    rng = np.random.default_rng(21)
    out = []
    # isolated, N = 1000 (cell-list path)
    out.append(("isolated_1000_rc6", rng.uniform(0, 22, (1000, 3)), None, 6.0))
    # isolated brute force
    out.append(("isolated_bf_200_rc5", rng.uniform(0, 15, (200, 3)), None, 5.0))
    # periodic brute force (minimum image: 2 rc < min width)
    out.append(("pbc_bf_60_rc4", rng.uniform(0, 20, (60, 3)), (np.diag([20.0, 21.0, 22.0]),
                                                               [True] * 3), 4.0))
    # triclinic cell list with many images and self images (rc > box)
    H = np.array([[9.0, 2.0, 1.0], [0.0, 8.5, 1.5], [0.0, 0.0, 9.5]])
    f = rng.uniform(0, 1, (300, 3))
    out.append(("triclinic_300_rc11", f @ H.T, (H, [True] * 3), 11.0))
    # mixed periodicity
    out.append(("mixed_pbc_400_rc6", rng.uniform(0, 18, (400, 3)),
                (np.diag([18.0, 18.0, 18.0]), [True, False, True]), 6.0))
    return out


@pytest.mark.parametrize("case", range(5))
def test_full_vs_mirrored_oracle(oracle, case):
    name, pos, cell, rc = _cases()[case]
    _print_case("full", name)
    H, pbc = (None, None) if cell is None else cell
    want = _mirror(*oracle.search_single(pos, H, pbc, rc))
    full = _ex(pos, [pos.shape[0]], [cell], [rc], opts=NL_FULL | NL_DIST)[0]
    half = _ex(pos, [pos.shape[0]], [cell], [rc], opts=NL_DIST)[0]
    assert_same_edges(full[:3], want, name)
    _check_layout(full, half, name)


def test_full_ragged_batch_multi(oracle):
    from neighborlist_rs.workloads import c4_ragged_batch
    pos, sizes, cells_list, cutoffs = c4_ragged_batch(seed=4, n_systems=24, nmax=700)
    cells = [None if c is None else (np.asarray(c), [True] * 3) for c in cells_list]
    batch = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    for disjoint in (False, True):
        _print_case("full batch", len(sizes), disjoint)
        full = _ex(pos, sizes, cells, cutoffs, disjoint, True, NL_FULL)
        half = _ex(pos, sizes, cells, cutoffs, disjoint, True, 0)
        want = oracle.search_batch_multi(pos, batch, cells, cutoffs, disjoint)
        for q in range(len(cutoffs)):
            what = f"list {q} disjoint={disjoint}"
            _check_layout(full[q], half[q], what)
            assert_same_edges(full[q][:3], _mirror(*want[q]), what)


def _ordered_pairs(pos, H, rc):
    """Every ordered (i, j, S), S != 0 when i == j, with |r_j + H S - r_i| < rc,
    by direct enumeration of lattice images (independent of the half list)."""
    Hinv = np.linalg.inv(H)
    w = 1.0 / np.linalg.norm(Hinv, axis=1)
    span = np.ptp(pos @ Hinv.T, axis=0)
    k = [int(np.ceil(rc / w[a] + span[a])) + 1 for a in range(3)]
    rng = [np.arange(-k[a], k[a] + 1) for a in range(3)]
    S = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 3)
    off = S @ H.T
    ei, ej, es = [], [], []
    for i in range(pos.shape[0]):
        d = (pos[None, :, :] - pos[i][None, None, :]) + off[:, None, :]
        r2 = np.einsum("abk,abk->ab", d, d)
        a, j = np.nonzero(r2 < rc * rc)
        keep = (j != i) | np.any(S[a] != 0, axis=1)
        ei.append(np.full(keep.sum(), i))
        ej.append(j[keep])
        es.append(S[a][keep])
        assert np.all(np.abs(np.sqrt(r2[a, j]) - rc) > 1e-9)
    return np.concatenate(ei), np.concatenate(ej), np.concatenate(es).astype(np.int32)


def test_full_vs_ordered_pair_enumeration():
    rng = np.random.default_rng(5)
    H = np.array([[6.0, 1.0, 0.5], [0.0, 5.5, 1.0], [0.0, 0.0, 6.5]])
    for n, rc in ((24, 7.0), (40, 4.0)):
        f = rng.uniform(-0.3, 1.3, (n, 3))  # unwrapped positions too
        pos = f @ H.T
        _print_case("full vs enumeration", n, rc)
        got = _ex(pos, [n], [(H, [True] * 3)], [rc], opts=NL_FULL)[0]
        assert_same_edges(got[:3], _ordered_pairs(pos, H, rc), f"n={n} rc={rc}")


def test_full_logged_count_ethanol_10000():
    pos, rows, pbc = read_xyz("ethanol_10000.xyz")
    _print_case("full ethanol_10000 rc 6")
    ei, ej, s, _ = _ex(pos, [pos.shape[0]], [(rows.T, pbc)], [6.0], opts=NL_FULL)[0]
    # benchmarks/benchmark_EPYC7R13.txt: 568606 pairs without self images
    assert int((ei != ej).sum()) == 2 * 568606
    assert int((ei == ej).sum()) % 2 == 0


# ---------------------------------------------------------------------------
# torch tensors (DLPack views of the engine's output blocks)

def _np(t):
    return t.detach().cpu().numpy()


def test_torch_single_and_multi_match_numpy(nr):
    import torch
    rng = np.random.default_rng(7)
    # H: lattice vectors as columns (the reference's convention, PyCell's argument)
    H = np.array([[14.0, 2.0, 1.0], [0.0, 13.0, 1.5], [0.0, 0.0, 12.5]])
    pos = rng.uniform(0, 1, (1500, 3)) @ H.T
    cell = nr.PyCell(H.tolist(), [True] * 3)
    pt = torch.from_numpy(pos).cuda()
    _print_case("torch single")
    got = nr.build_neighborlists_torch(cell, pt, 5.5)
    assert got["edge_index"].device.type == "cuda" and got["edge_index"].dtype == torch.int64
    assert got["shift"].dtype == torch.int32 and got["shift"].shape[1] == 3
    want = nr.build_neighborlists(cell, pos, 5.5)
    assert_same_edges((_np(got["edge_index"][0]), _np(got["edge_index"][1]), _np(got["shift"])),
                      as_tuple(want), "torch single")
    _print_case("torch multi full")
    gm = nr.build_neighborlists_multi_torch(cell, pt, [3.0, 5.5], labels=["a", "b"], full=True,
                                            distances=True)
    wm = nr.build_neighborlists_multi(cell, pos, [3.0, 5.5], labels=["a", "b"])
    for key in ("a", "b"):
        e, s, d = gm[key]["edge_index"], gm[key]["shift"], gm[key]["distance"]
        assert_same_edges((_np(e[0]), _np(e[1]), _np(s)), _mirror(*as_tuple(wm[key])), key)
        # distances: |r_j + S H^T - r_i| from the raw positions
        v = pos[_np(e[1])] + _np(s) @ H.T - pos[_np(e[0])]
        np.testing.assert_allclose(_np(d), np.linalg.norm(v, axis=1), rtol=0, atol=1e-12)


def test_torch_batch_matches_numpy(nr):
    import torch
    from neighborlist_rs.workloads import c4_ragged_batch
    pos, sizes, cells_list, cutoffs = c4_ragged_batch(seed=9, n_systems=16, nmax=900)
    batch = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    # Python batch convention: (B, 3, 3) rows, zero = isolated, periodic in all directions
    cells = np.zeros((len(sizes), 3, 3))
    for s, c in enumerate(cells_list):
        if c is not None:
            cells[s] = np.asarray(c).T
    _print_case("torch batch")
    pt = torch.from_numpy(pos).cuda()
    got = nr.build_neighborlists_batch_torch(pt, torch.from_numpy(batch).cuda(),
                                             torch.from_numpy(cells), cutoff=6.0)
    want = nr.build_neighborlists_batch(pos, batch, cells, 6.0)
    e, s = got["edge_index"], got["shift"]
    assert_same_edges((_np(e[0]), _np(e[1]), _np(s)), as_tuple(want), "torch batch")
    gm = nr.build_neighborlists_batch_multi_torch(pt, batch, cells, [4.0, 6.0], disjoint=True,
                                                  full=True)
    wm = nr.build_neighborlists_batch_multi(pos, batch, cells, [4.0, 6.0], disjoint=True)
    for q in (0, 1):
        e, s = gm[q]["edge_index"], gm[q]["shift"]
        assert_same_edges((_np(e[0]), _np(e[1]), _np(s)), _mirror(*as_tuple(wm[q])), f"q={q}")


def test_torch_views_outlive_result_and_other_searches(nr):
    import torch
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, 30, (3000, 3))
    pt = torch.from_numpy(pos).cuda()
    _print_case("torch lifetime")
    keep = nr.build_neighborlists_torch(None, pt, 6.0, full=True)
    saved = {k: _np(v).copy() for k, v in keep.items()}
    # other searches of several sizes reuse pooled blocks; the views must hold theirs
    for n in (500, 3000, 8000, 3000):
        tmp = nr.build_neighborlists_torch(None, torch.from_numpy(rng.uniform(0, 30, (n, 3))).cuda(),
                                           6.0, full=True)
        del tmp
    for k, v in keep.items():
        np.testing.assert_array_equal(_np(v), saved[k], err_msg=k)
    # a view's consumer work queued on the stream, then dropped: the block's
    # next user sees the finished result of that work
    e = keep["edge_index"]
    total = int(e.sum().item())
    big = e.to(torch.float64) * 2.0  # queued on the current stream
    del keep, e
    again = nr.build_neighborlists_torch(None, pt, 6.0, full=True)
    assert int((big.sum() / 2.0).item()) == total
    np.testing.assert_array_equal(_np(again["edge_index"]), saved["edge_index"])


def test_torch_reads_positions_after_queued_writes(nr):
    import torch
    rng = np.random.default_rng(10)
    base = rng.uniform(0, 25, (4000, 3))
    want = nr.build_neighborlists(None, base, 5.0)
    _print_case("torch stream order")
    src = torch.from_numpy(base).cuda()
    pt = torch.empty_like(src)
    for _ in range(3):
        pt.zero_()
        pt.copy_(src * 1.0)  # queued on the current stream right before the search
        got = nr.build_neighborlists_torch(None, pt, 5.0)
        e, s = got["edge_index"], got["shift"]
        assert_same_edges((_np(e[0]), _np(e[1]), _np(s)), as_tuple(want), "stream order")


def test_torch_errors_and_empty(nr):
    import torch
    _print_case("torch errors")
    with pytest.raises(TypeError, match="float64"):
        nr.build_neighborlists_torch(None, torch.zeros((4, 3), dtype=torch.float32).cuda(), 2.0)
    with pytest.raises(ValueError, match="device"):
        nr.build_neighborlists_torch(None, torch.zeros((4, 3), dtype=torch.float64), 2.0)
    with pytest.raises(ValueError, match="Cutoff must be positive and finite"):
        nr.build_neighborlists_torch(None, torch.zeros((4, 3), dtype=torch.float64).cuda(), -1.0)
    r = nr.build_neighborlists_torch(None, torch.zeros((0, 3), dtype=torch.float64).cuda(), 2.0,
                                     distances=True)
    assert tuple(r["edge_index"].shape) == (2, 0) and tuple(r["distance"].shape) == (0,)
    r = nr.build_neighborlists_torch(None, torch.zeros((1, 3), dtype=torch.float64).cuda(), 2.0,
                                     full=True)
    assert tuple(r["edge_index"].shape) == (2, 0)


# ---------------------------------------------------------------------------
# displacements r_ij = r_j + H S - r_i (NL_DISP), half and full lists

NL_DISP = 32


def _check_disp(pos, cell, got, what):
    ei, ej, s, d, disp = got
    H = np.zeros((3, 3)) if cell is None else np.asarray(cell[0])
    want = pos[ej] + s.astype(np.float64) @ H.T - pos[ei]
    np.testing.assert_allclose(disp, want, rtol=0, atol=1e-12, err_msg=what)
    # the vector whose norm the search compared: its norm is the distance
    n = np.sqrt((disp[:, 0] * disp[:, 0] + disp[:, 1] * disp[:, 1]) + disp[:, 2] * disp[:, 2])
    assert np.all(np.abs(n - d) <= 2 * np.spacing(np.maximum(n, d))), what


def _ex_disp(pos, sizes, cells, cutoffs, disjoint=False, multi=False, opts=0):
    from neighborlist_rs import _lib
    out = _ex(pos, sizes, cells, cutoffs, disjoint, multi, opts | NL_DIST)
    # second call for the displacements (the helper copies edges and distances)
    lib = _lib.lib
    p = np.ascontiguousarray(pos, dtype=np.float64)
    sz = np.ascontiguousarray(sizes, dtype=np.int64)
    S = sz.size
    Hs = np.zeros((S, 9))
    has = np.zeros(S, np.uint8)
    pb = np.ones((S, 3), np.uint8)
    for k, c in enumerate(cells):
        if c is not None:
            Hs[k] = np.asarray(c[0], dtype=np.float64).reshape(9)
            has[k] = 1
            pb[k] = [1 if v else 0 for v in c[1]]
    cut = np.ascontiguousarray(cutoffs, dtype=np.float64)
    r = C.c_void_p()
    assert lib.nl_search_ex(p.ctypes.data_as(C.c_void_p), p.shape[0], sz.ctypes.data_as(_lib.i64p),
                            S, Hs.ctypes.data_as(_lib.dp), has.ctypes.data_as(_lib.u8p),
                            pb.ctypes.data_as(_lib.u8p), cut.ctypes.data_as(_lib.dp), cut.size,
                            int(disjoint), int(multi), opts | NL_DIST | NL_DISP, None,
                            C.byref(r)) == 0, _lib.last_error()
    res = []
    try:
        for q in range(cut.size):
            m = int(lib.nl_result_num_edges(r, q))
            ei, ej = np.empty(m, np.int64), np.empty(m, np.int64)
            sh = np.empty((m, 3), np.int32)
            d, v = np.empty(m), np.empty((m, 3))
            if m:
                assert lib.nl_result_copy(r, q, ei.ctypes.data_as(_lib.i64p),
                                          ej.ctypes.data_as(_lib.i64p),
                                          sh.ctypes.data_as(_lib.i32p)) == 0
                assert lib.nl_result_copy_dist(r, q, d.ctypes.data_as(_lib.dp)) == 0
                assert lib.nl_result_copy_disp(r, q, v.ctypes.data_as(_lib.dp)) == 0
            # the lists themselves do not change when displacements are asked for
            np.testing.assert_array_equal(ei, out[q][0])
            np.testing.assert_array_equal(ej, out[q][1])
            np.testing.assert_array_equal(sh, out[q][2])
            np.testing.assert_array_equal(d, out[q][3])
            res.append((ei, ej, sh, d, v))
    finally:
        lib.nl_result_free(r)
    return res


@pytest.mark.parametrize("case", range(5))
def test_displacements_half_and_full(case):
    name, pos, cell, rc = _cases()[case]
    _print_case("displacements", name)
    half = _ex_disp(pos, [pos.shape[0]], [cell], [rc])[0]
    full = _ex_disp(pos, [pos.shape[0]], [cell], [rc], opts=NL_FULL)[0]
    _check_disp(pos, cell, half, name + " half")
    _check_disp(pos, cell, full, name + " full")
    np.testing.assert_array_equal(full[4][1::2], -full[4][0::2])
    if int((half[0] == half[1]).sum()) == 0:
        np.testing.assert_array_equal(full[4][0::2], half[4])


def test_displacements_fp32_filter_and_batch():
    from neighborlist_rs import _lib
    from neighborlist_rs.workloads import c4_ragged_batch
    lib = _lib.lib
    pos, sizes, cells_list, cutoffs = c4_ragged_batch(seed=6, n_systems=20, nmax=800)
    cells = [None if c is None else (np.asarray(c), [True] * 3) for c in cells_list]
    starts = np.concatenate([[0], np.cumsum(sizes)])
    was = lib.nl_get_autotune()
    lib.nl_set_autotune(0)
    try:
        got = {}
        for f32 in (0, 1):
            lib.nl_set_fp32_filter(f32)
            _print_case("displacements batch fp32", f32)
            got[f32] = _ex_disp(pos, sizes, cells, cutoffs, False, True)
    finally:
        lib.nl_set_fp32_filter(0)
        lib.nl_set_autotune(was)
    for q in range(len(cutoffs)):
        # FP32 filter with FP64 recheck: the same lists and displacements, bit for bit
        for a, b in zip(got[0][q], got[1][q]):
            np.testing.assert_array_equal(a, b)
        ei, ej, s, d, v = got[0][q]
        for k in range(len(sizes)):
            sel = (ei >= starts[k]) & (ei < starts[k + 1])
            _check_disp(pos, cells[k], (ei[sel], ej[sel], s[sel], d[sel], v[sel]),
                        f"list {q} system {k}")


def test_torch_displacements(nr):
    import torch
    rng = np.random.default_rng(12)
    H = np.array([[12.0, 1.0, 0.5], [0.0, 11.0, 1.0], [0.0, 0.0, 13.0]])
    pos = rng.uniform(0, 1, (900, 3)) @ H.T
    cell = nr.PyCell(H.tolist(), [True] * 3)
    _print_case("torch displacements")
    g = nr.build_neighborlists_torch(cell, torch.from_numpy(pos).cuda(), 5.0, full=True,
                                     distances=True, displacements=True)
    v = g["displacement"]
    assert v.dtype == torch.float64 and tuple(v.shape) == (g["shift"].shape[0], 3)
    e, s = _np(g["edge_index"]), _np(g["shift"])
    _check_disp(pos, (H, [True] * 3), (e[0], e[1], s, _np(g["distance"]), _np(v)), "torch")
    r = nr.build_neighborlists_torch(None, torch.zeros((1, 3), dtype=torch.float64).cuda(), 2.0,
                                     displacements=True)
    assert tuple(r["displacement"].shape) == (0, 3)
