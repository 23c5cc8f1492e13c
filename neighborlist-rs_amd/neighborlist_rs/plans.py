"""Graph-captured searches for torch callers (repeated searches of one shape).

A plan fixes a batch shape (atom count, system sizes, cells, cutoffs) whose
positions live in one float64 (N, 3) torch tensor on the GPU and change
between runs, as in an MD loop or a model's input pipeline.  The whole search
(uploads, binning, sort, count, scan, fill, and the optional full lists,
distances, displacements and per-system ranges) is captured once in a HIP
graph; ``run()`` replays it with one launch (``nl_plan_*`` in
include/nl_engine_ext.h).  The lists equal those of the search functions for
the same positions: same arguments and options as
``build_neighborlists_batch_multi_torch`` (tensors.py), the reference's
src/python_api.rs:425-546.

``run()`` returns torch tensors owned by torch (copies of the plan's output
arrays, made on the device), so they stay valid after later runs.  Callers
that need the arrays without a copy use the eager search functions
(zero-copy DLPack views) or the plan's device addresses through the C ABI.
Import torch before this package.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib
from . import PyCell, _check, _keyed
from .tensors import (NL_DIST, NL_DISP, NL_FULL, NL_LOCAL_INDEX, NL_RANGES, _batch_cells_t,
                      _batch_sizes, _check_cutoffs, _positions_t, _single_cell, _torch)


class NeighborPlan:
    """Graph-captured search over the positions tensor ``positions``.

    One system (``cell``: PyCell or None) or a ragged batch (``batch``: ids,
    ``cells``: (B, 3, 3) rows as in build_neighborlists_batch).  ``cutoffs``
    is a number (``run()`` returns one dict) or a sequence (``run()`` returns
    a dict keyed by ``labels`` or list index).  ``full``, ``distances`` and
    ``displacements`` are those of the torch search functions; ``ranges=True``
    records per-system edge ranges (``system_ranges()``).  Batch indices are
    offset by each system's start unless ``local_index=True``.  ``capacity``
    sizes the output (edges over all lists; <= 0: from one count plus 1/8);
    a run whose lists outgrow it grows the arrays and re-captures."""

    def __init__(self, positions, cutoffs, cell=None, *, batch=None, cells=None, labels=None,
                 disjoint=False, full=False, distances=False, displacements=False,
                 ranges=False, local_index=False, capacity=0):
        torch = _torch()
        pos = _positions_t(positions)
        if pos.data_ptr() != positions.data_ptr():
            raise ValueError("a plan needs a contiguous positions tensor (it reads that memory "
                             "on every run)")
        n = pos.shape[0]
        if n == 0:
            raise ValueError("a plan needs at least one atom")
        self._single = np.ndim(cutoffs) == 0
        cut, self._labels = _check_cutoffs([cutoffs] if self._single else cutoffs, labels)
        if len(cut) == 0:
            raise ValueError("a plan needs at least one cutoff")
        if batch is None:
            if cells is not None:
                raise ValueError("cells belongs to batches; use cell for one system")
            sizes = np.array([n], np.int64)
            H, has, pb = _single_cell(cell)
        else:
            if cell is not None:
                raise ValueError("cell belongs to one system; use cells for batches")
            sizes = np.asarray(_batch_sizes(batch, n), np.int64)
            _, H, has, pb = _batch_cells_t(cells, sizes.size)
        self._pos = pos  # the plan reads this memory on every run
        self._device = pos.device
        self._k = len(cut)
        self._sizes = sizes
        self._dist, self._disp, self._ranges = bool(distances), bool(displacements), bool(ranges)
        opts = ((NL_FULL if full else 0) | (NL_DIST if distances else 0) |
                (NL_DISP if displacements else 0) | (NL_RANGES if ranges else 0) |
                (NL_LOCAL_INDEX if local_index else 0))
        c = np.ascontiguousarray(cut, dtype=np.float64)
        self._ptr = C.c_void_p()
        _check(lib.nl_set_device(self._device.index or 0))
        # positions written on torch's stream must be complete before the
        # plan's own stream reads them
        torch.cuda.current_stream(self._device).synchronize()
        _check(lib.nl_plan_create_ex(
            C.c_void_p(pos.data_ptr()), n, sizes.ctypes.data_as(_lib.i64p), sizes.size,
            None if H is None else H.ctypes.data_as(_lib.dp),
            None if has is None else has.ctypes.data_as(_lib.u8p),
            None if pb is None else pb.ctypes.data_as(_lib.u8p), c.ctypes.data_as(_lib.dp),
            c.size, int(bool(disjoint)), int(not self._single), opts, int(capacity),
            C.byref(self._ptr)))

    def close(self):
        if self._ptr:
            lib.nl_plan_free(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _array(self, q, what, shape, dtype):
        torch = _torch()
        t = torch.empty(shape, dtype=dtype, device=self._device)
        if t.numel():
            p = C.c_void_p()
            _check(lib.nl_plan_array(self._ptr, q, what, C.byref(p)))
            _check(lib.nl_memcpy_d2d(C.c_void_p(t.data_ptr()), p, t.numel() * t.element_size()))
        return t

    def run(self):
        """Replays the search on the current contents of the positions tensor
        and returns the lists (torch tensors on the positions' device)."""
        torch = _torch()
        if not self._ptr:
            raise ValueError("the plan is closed")
        torch.cuda.current_stream(self._device).synchronize()
        _check(lib.nl_plan_run(self._ptr))
        out = []
        for q in range(self._k):
            m = int(lib.nl_plan_num_edges(self._ptr, q))
            d = {"edge_index": self._array(q, 0, (2, m), torch.int64),
                 "shift": self._array(q, 1, (m, 3), torch.int32)}
            if self._dist:
                d["distance"] = self._array(q, 2, (m,), torch.float64)
            if self._disp:
                d["displacement"] = self._array(q, 3, (m, 3), torch.float64)
            out.append(d)
        return out[0] if self._single else _keyed(out, self._labels, self._k)

    def system_ranges(self):
        """(ranges, modes) of the last run: ranges[s, q] = [first, end) edge of
        system s in list q; modes[s] = 0 brute force on raw positions, 1 brute
        force minimum image, 2 cell list (as nl_result_system_ranges)."""
        S = self._sizes.size
        r = np.zeros((S, self._k, 2), np.int64)
        m = np.zeros(S, np.int32)
        _check(lib.nl_plan_system_ranges(self._ptr, r.ctypes.data_as(_lib.i64p),
                                         m.ctypes.data_as(_lib.i32p)))
        return r, m

    def stats(self):
        rp, rc, cap = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib.nl_plan_stats(self._ptr, C.byref(rp), C.byref(rc), C.byref(cap)))
        return {"replays": rp.value, "recaptures": rc.value, "capacity": cap.value}
