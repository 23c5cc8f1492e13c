"""Device-resident form of the search functions for torch callers.

The four search functions of the reference's Python module
(src/python_api.rs:139-546: build_neighborlists, build_neighborlists_multi,
build_neighborlists_batch, build_neighborlists_batch_multi) with the same
arguments, taking positions as a float64 (N, 3) torch tensor on a ROCm
device and returning the lists as torch tensors on that device, with no copy
through host memory: the tensors are DLPack views of the engine's own output
blocks (nl_result_dlpack), which stay alive as long as any view does.

Three options go beyond the reference (all keyword-only):

* ``full=True``: full (directed) lists, every pair of the half list in both
  directions, (i, j, S) at an even position and (j, i, -S) right after it; a
  self image (i, i, S) appears once per S.  This is the edge set a
  message-passing model consumes.
* ``distances=True``: adds ``"distance"``, the float64 length
  |r_j + S H^T - r_i| of every edge, from the same expression the search
  compares with the cutoff.
* ``displacements=True``: adds ``"displacement"``, the float64 vector
  r_j + S H^T - r_i of every edge (M, 3), the one whose norm is the distance.

Import torch before this package: torch carries its own HIP runtime, and
the engine library binds to it only if it is loaded first (otherwise the
functions below raise, naming both runtimes).

Ordering: the search runs after the work queued on torch's current stream
(positions may still be being written there), and output blocks dropped by
the last view are reused only after the work queued on that stream up to
then.  Use the tensors on that stream, or synchronise before handing them to
another one.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import lib
from . import PyCell, _batch_cells, _check, _cutoff_list, _keyed, _log

NL_DEVICE_POSITIONS = 1
NL_LOCAL_INDEX = 2
NL_RANGES = 4
NL_DIST = 8
NL_FULL = 16
NL_DISP = 32

_DLTENSOR = b"dltensor"  # capsule name; kept alive by the module
_capsule_new = C.pythonapi.PyCapsule_New
_capsule_new.restype = C.py_object
_capsule_new.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]


_runtime_checked = False


def _torch():
    """torch, after checking that the process holds one HIP runtime.

    torch ships its own HIP runtime, and the engine library binds to it
    whichever package is imported first (the loader maps torch's runtime
    before the library).  A second runtime can only appear when something
    else mapped another copy first; torch then finds no device, so that is
    reported here rather than as a failed device call."""
    global _runtime_checked
    import torch
    if not _runtime_checked:
        try:
            with open("/proc/self/maps") as f:
                libs = {os.path.realpath(l.split()[-1]) for l in f if "libamdhip64" in l}
        except OSError:
            libs = set()
        if len(libs) > 1:
            raise RuntimeError(
                "two HIP runtimes are loaded (" + ", ".join(sorted(libs)) + "): a runtime "
                "other than torch's was mapped before neighborlist_rs was imported")
        _runtime_checked = True
    return torch


def _positions_t(positions):
    torch = _torch()
    if not isinstance(positions, torch.Tensor):
        raise TypeError("argument 'positions': expected a torch.Tensor")
    if positions.dtype != torch.float64:
        raise TypeError("argument 'positions': tensor must be float64")
    if positions.ndim != 2:
        raise TypeError(f"argument 'positions': expected a 2-d tensor, got {positions.ndim}-d")
    if positions.shape[1] != 3:
        raise ValueError("Positions must be (N, 3)")
    if positions.device.type != "cuda":
        raise ValueError("argument 'positions': tensor must be on a ROCm (cuda) device")
    return positions.contiguous()


def _view(res, q, what):
    """Zero-copy torch view of one array of list q (0 edge_index, 1 shift,
    2 distance, 3 displacement)."""
    torch = _torch()
    p = C.c_void_p()
    _check(lib.nl_result_dlpack(res.ptr, q, what, C.byref(p)))
    return torch.utils.dlpack.from_dlpack(_capsule_new(p, _DLTENSOR, None))


def _empty_t(device, distances, displacements=False):
    torch = _torch()
    out = {"edge_index": torch.zeros((2, 0), dtype=torch.int64, device=device),
           "shift": torch.zeros((0, 3), dtype=torch.int32, device=device)}
    if distances:
        out["distance"] = torch.zeros((0,), dtype=torch.float64, device=device)
    if displacements:
        out["displacement"] = torch.zeros((0, 3), dtype=torch.float64, device=device)
    return out


def _lists(res, k, device, distances, displacements):
    out = []
    for q in range(k):
        m = int(lib.nl_result_num_edges(res.ptr, q))
        if m == 0:
            out.append(_empty_t(device, distances, displacements))
            continue
        d = {"edge_index": _view(res, q, 0), "shift": _view(res, q, 1)}
        if distances:
            d["distance"] = _view(res, q, 2)
        if displacements:
            d["displacement"] = _view(res, q, 3)
        out.append(d)
    return out


def _search(pos, sizes, H, has, pb, cutoffs, disjoint, multi, full, distances, displacements):
    from . import _Result
    torch = _torch()
    dev = pos.device
    _check(lib.nl_set_device(dev.index if dev.index is not None else 0))
    stream = torch.cuda.current_stream(dev).cuda_stream
    sz = np.ascontiguousarray(np.asarray(sizes, dtype=np.int64))
    c = np.ascontiguousarray(np.asarray(cutoffs, dtype=np.float64))
    opts = (NL_DEVICE_POSITIONS | (NL_FULL if full else 0) | (NL_DIST if distances else 0) |
            (NL_DISP if displacements else 0))
    res = _Result()
    _check(lib.nl_search_ex(
        C.c_void_p(pos.data_ptr()), pos.shape[0], sz.ctypes.data_as(_lib.i64p), sz.size,
        None if H is None else H.ctypes.data_as(_lib.dp),
        None if has is None else has.ctypes.data_as(_lib.u8p),
        None if pb is None else pb.ctypes.data_as(_lib.u8p), c.ctypes.data_as(_lib.dp), c.size,
        int(bool(disjoint)), int(bool(multi)), opts, C.c_void_p(stream), C.byref(res.ptr)))
    return _lists(res, c.size, dev, distances, displacements)


def _single_cell(cell):
    if cell is None:
        return None, None, None
    if not isinstance(cell, PyCell):
        raise TypeError("argument 'cell': expected PyCell or None")
    H = np.ascontiguousarray(np.asarray(cell._h, dtype=np.float64).reshape(1, 9))
    has = np.ones(1, np.uint8)
    pb = np.ascontiguousarray(np.array([[1 if v else 0 for v in cell._pbc]], np.uint8))
    return H, has, pb


def _batch_sizes(batch, n):
    """System sizes of a batch-id vector: maximal runs of equal ids, ids
    non-decreasing (src/python_api.rs:360-372, batch.rs:25-46)."""
    torch = _torch()
    b = batch.detach().to("cpu").numpy() if isinstance(batch, torch.Tensor) else np.asarray(batch)
    if b.ndim != 1:
        raise TypeError("argument 'batch': expected a 1-d array")
    if not np.issubdtype(b.dtype, np.integer):
        raise TypeError("argument 'batch': array must be integer")
    if b.size != n:
        raise ValueError("positions and batch must have the same length")
    b = b.astype(np.int64)
    d = np.diff(b)
    if np.any(d < 0):
        raise ValueError("Batch IDs must be monotonic (non-decreasing)")
    starts = np.concatenate([[0], np.nonzero(d)[0] + 1, [n]])
    return np.diff(starts)


def _batch_cells_t(cells, ns):
    torch = _torch()
    if isinstance(cells, torch.Tensor):
        cells = cells.detach().to("cpu").numpy()
    return _batch_cells(cells, ns)


def _check_cutoffs(cutoffs, labels):
    cutoffs = _cutoff_list(cutoffs)
    if any((not np.isfinite(c)) or c <= 0.0 for c in cutoffs):
        raise ValueError("All cutoffs must be positive and finite")
    if labels is not None:
        labels = [str(l) for l in labels]
        if len(labels) != len(cutoffs):
            raise ValueError("Length of labels must match length of cutoffs")
    return cutoffs, labels


def build_neighborlists_torch(cell, positions, cutoff, *, full=False, distances=False,
                              displacements=False):
    """build_neighborlists (src/python_api.rs:139-186) on device tensors."""
    cutoff = float(cutoff)
    if not np.isfinite(cutoff) or cutoff <= 0.0:
        raise ValueError("Cutoff must be positive and finite")
    pos = _positions_t(positions)
    n = pos.shape[0]
    if n == 0:
        return _empty_t(pos.device, distances, displacements)
    H, has, pb = _single_cell(cell)
    out = _search(pos, [n], H, has, pb, [cutoff], False, False, full, distances,
                  displacements)
    _log("build_neighborlists_torch", n)
    return out[0]


def build_neighborlists_multi_torch(cell, positions, cutoffs, labels=None, disjoint=False, *,
                                    full=False, distances=False, displacements=False):
    """build_neighborlists_multi (src/python_api.rs:235-309) on device tensors."""
    pos = _positions_t(positions)
    n = pos.shape[0]
    cutoffs, labels = _check_cutoffs(cutoffs, labels)
    if len(cutoffs) == 0:
        return {}
    if n == 0:
        return _keyed([_empty_t(pos.device, distances, displacements) for _ in cutoffs], labels,
                      len(cutoffs))
    H, has, pb = _single_cell(cell)
    out = _search(pos, [n], H, has, pb, cutoffs, disjoint, True, full, distances,
                  displacements)
    _log("build_neighborlists_multi_torch", n)
    return _keyed(out, labels, len(cutoffs))


def build_neighborlists_batch_torch(positions, batch, cells=None, cutoff=5.0, *, full=False,
                                    distances=False, displacements=False):
    """build_neighborlists_batch (src/python_api.rs:315-423) on device
    tensors; indices are offset by each system's start, as there."""
    cutoff = float(cutoff)
    if not np.isfinite(cutoff) or cutoff <= 0.0:
        raise ValueError("Cutoff must be positive and finite")
    pos = _positions_t(positions)
    n = pos.shape[0]
    sizes = _batch_sizes(batch, n)
    if n == 0:
        return _empty_t(pos.device, distances, displacements)
    _, H, has, pb = _batch_cells_t(cells, sizes.size)
    out = _search(pos, sizes, H, has, pb, [cutoff], False, False, full, distances,
                  displacements)
    _log("build_neighborlists_batch_torch", n)
    return out[0]


def build_neighborlists_batch_multi_torch(positions, batch, cells=None, cutoffs=[5.0],
                                          labels=None, disjoint=False, *, full=False,
                                          distances=False, displacements=False):
    """build_neighborlists_batch_multi (src/python_api.rs:425-546) on device
    tensors."""
    pos = _positions_t(positions)
    n = pos.shape[0]
    sizes = _batch_sizes(batch, n)
    cutoffs, labels = _check_cutoffs(cutoffs, labels)
    if n == 0 or len(cutoffs) == 0:
        return {}
    _, H, has, pb = _batch_cells_t(cells, sizes.size)
    out = _search(pos, sizes, H, has, pb, cutoffs, disjoint, True, full, distances,
                  displacements)
    _log("build_neighborlists_batch_multi_torch", n)
    return _keyed(out, labels, len(cutoffs))
