"""Neighbour lists with a skin for molecular dynamics (torch tensors).

The dynamic manager the reference specifies for its MD phase (spec.md:41-48
and the ``DynamicNeighborManager`` sketch of spec.md:209-222: ``new(config,
skins)``, ``update(positions, cell)``, ``needs_rebuild(positions)``,
``get()``), which the reference never implemented.  The lists are built at
r_list = rc + skin from the positions of the last build and kept while no
atom has moved more than half the smallest skin since then; every pair now
within rc is then in the list, and the consumer applies rc itself.  The
largest displacement is measured on the GPU; a rebuild replays a
graph-captured build (``nl_skin_*`` in include/nl_engine_ext.h).

Positions are float64 (N, 3) torch tensors on the GPU, one continuous
trajectory (re-wrapping atoms into the cell between updates counts as a
large displacement and forces a rebuild).  Import torch before this package.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib
from . import PyCell, _check, _keyed
from .tensors import (NL_FULL, _batch_cells_t, _batch_sizes, _check_cutoffs,
                      _positions_t, _single_cell, _torch)


# Skin lists hold every pair within rc_q + skin_q at the last build; a pair
# that drifts across a shell boundary between rebuilds would sit in the wrong
# disjoint shell, so only cumulative lists keep the coverage guarantee.
_DISJOINT_MSG = ("skin lists are cumulative (disjoint shells cannot keep every pair "
                 "within rc in its shell between rebuilds); split the cumulative lists "
                 "by distance instead")


def _as_list(x, n, what):
    if np.ndim(x) == 0:
        return [float(x)] * n
    v = [float(t) for t in x]
    if len(v) != n:
        raise ValueError(f"one {what} per cutoff is needed")
    return v


class DynamicNeighborManager:
    """Skin lists for one system (``cell``: PyCell or None) or a ragged batch
    (``batch``: ids, ``cells``: (B, 3, 3) rows as in build_neighborlists_batch).

    ``cutoffs`` is a number (one list, ``get()`` returns a dict) or a sequence
    (``get()`` returns a dict keyed by ``labels`` or list index); ``skins`` is
    a number or one value per cutoff.  ``full=True`` gives full lists."""

    def __init__(self, positions, cutoffs, skins, cell=None, *, batch=None, cells=None,
                 labels=None, disjoint=False, full=False):
        if disjoint:
            raise ValueError(_DISJOINT_MSG)
        torch = _torch()
        pos = _positions_t(positions)
        n = pos.shape[0]
        if n == 0:
            raise ValueError("the skin manager needs at least one atom")
        self._single = np.ndim(cutoffs) == 0
        cut, self._labels = _check_cutoffs([cutoffs] if self._single else cutoffs, labels)
        if len(cut) == 0:
            raise ValueError("the skin manager needs at least one cutoff")
        sk = _as_list(skins, len(cut), "skin")
        if any((not np.isfinite(x)) or x < 0.0 for x in sk):
            raise ValueError("Skins must be non-negative and finite")
        self._device = pos.device
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
        self._n, self._sizes, self._has = n, sizes, has
        self._k = len(cut)
        self.cutoffs, self.skins = cut, sk
        opts = NL_FULL if full else 0
        c = np.ascontiguousarray(cut, dtype=np.float64)
        s = np.ascontiguousarray(sk, dtype=np.float64)
        self._ptr = C.c_void_p()
        _check(lib.nl_set_device(self._device.index or 0))
        _check(lib.nl_skin_create(
            C.c_void_p(pos.data_ptr()), n, sizes.ctypes.data_as(_lib.i64p), sizes.size,
            None if H is None else H.ctypes.data_as(_lib.dp),
            None if has is None else has.ctypes.data_as(_lib.u8p),
            None if pb is None else pb.ctypes.data_as(_lib.u8p), c.ctypes.data_as(_lib.dp),
            s.ctypes.data_as(_lib.dp), c.size, int(bool(disjoint)), int(not self._single), opts,
            C.c_void_p(torch.cuda.current_stream(self._device).cuda_stream),
            C.byref(self._ptr)))
        self._cache = None

    def close(self):
        if getattr(self, "_ptr", None):
            lib.nl_skin_free(self._ptr)
            self._ptr = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _pos(self, positions):
        pos = _positions_t(positions)
        if pos.shape[0] != self._n or pos.device != self._device:
            raise ValueError("positions must keep the atom count and device of the manager")
        return pos

    def _stream(self):
        return C.c_void_p(_torch().cuda.current_stream(self._device).cuda_stream)

    def _cells(self, cell):
        """New cells as row-major H per system, or None."""
        if cell is None:
            return None
        if len(self._sizes) == 1 and isinstance(cell, PyCell):
            H, _, _ = _single_cell(cell)
            return H
        _, H, _, _ = _batch_cells_t(cell, len(self._sizes))
        return H

    def update(self, positions, cell=None):
        """New positions (and optionally a new cell, or (B, 3, 3) cells for a
        batch); returns True when the lists were rebuilt."""
        pos = self._pos(positions)
        H = self._cells(cell)
        r = C.c_int(0)
        _check(lib.nl_skin_update(self._ptr, C.c_void_p(pos.data_ptr()),
                                  None if H is None else H.ctypes.data_as(_lib.dp),
                                  self._stream(), C.byref(r)))
        if r.value:
            self._cache = None
        return bool(r.value)

    def max_displacement(self, positions):
        """Largest displacement from the positions of the last build."""
        pos = self._pos(positions)
        d = C.c_double(0.0)
        _check(lib.nl_skin_check(self._ptr, C.c_void_p(pos.data_ptr()), self._stream(),
                                 C.byref(d), None))
        return d.value

    def needs_rebuild(self, positions):
        pos = self._pos(positions)
        r = C.c_int(0)
        _check(lib.nl_skin_check(self._ptr, C.c_void_p(pos.data_ptr()), self._stream(), None,
                                 C.byref(r)))
        return bool(r.value)

    def _list(self, q):
        torch = _torch()
        m = int(lib.nl_skin_num_edges(self._ptr, q))
        ei = torch.empty((2, m), dtype=torch.int64, device=self._device)
        sh = torch.empty((m, 3), dtype=torch.int32, device=self._device)
        if m:
            pe, ps = C.c_void_p(), C.c_void_p()
            _check(lib.nl_skin_device(self._ptr, q, C.byref(pe), C.byref(ps)))
            # torch-owned copies: they outlive the next rebuild
            _check(lib.nl_memcpy_d2d(C.c_void_p(ei.data_ptr()), pe, 16 * m))
            _check(lib.nl_memcpy_d2d(C.c_void_p(sh.data_ptr()), ps, 12 * m))
        return {"edge_index": ei, "shift": sh}

    def get(self):
        """The lists at rc + skin (torch tensors, copied once per build)."""
        if self._cache is None:
            lists = [self._list(q) for q in range(self._k)]
            self._cache = lists[0] if self._single else _keyed(lists, self._labels, self._k)
        return self._cache

    def stats(self):
        u, b, d = C.c_int64(), C.c_int64(), C.c_double()
        _check(lib.nl_skin_stats(self._ptr, C.byref(u), C.byref(b), C.byref(d)))
        return {"updates": u.value, "builds": b.value, "last_displacement": d.value}
