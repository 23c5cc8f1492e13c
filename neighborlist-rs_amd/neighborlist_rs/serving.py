"""Continuous-batching request server over libnlmi355.

The reference answers one request per synchronous call
(src/python_api.rs:139-186, build_neighborlists).  `Server` takes the same
inputs (a PyCell or None, an (N, 3) float64 array, one cutoff) from any
number of threads, batches whatever is queued into ragged batches on the GPU
(several in flight on separate HIP streams) and hands each request back the
same dict build_neighborlists would return for it.

Options beyond the reference, the same for every request of a server:
``full=True`` (full lists: every pair in both directions, (i, j, S) then
(j, i, -S)), ``distances=True`` (adds "distance", the FP64 |r_ij|) and
``displacements=True`` (adds "displacement", the FP64 r_ij = r_j + S H^T -
r_i), as in the torch search functions.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib


class Server:
    """submit() returns a request id at once; result(id) blocks for its lists.

    slots        batches in flight at once (each has its own stream and buffers;
                 four measured best on a trace of 600 requests of 16-50,000
                 atoms: 77,300-77,800 systems/s against 73,600 with three
                 and 72,200-74,200 with five)
    max_atoms    atoms per batch
    max_systems  requests per batch
    edge_budget  estimated edges of all in-flight batches (admission control;
                 a request alone is always admitted)
    full, distances, displacements   outputs of every request (see above)

    Each slot's stream should get a hardware queue of its own: HIP's default
    of four (GPU_MAX_HW_QUEUES, read when the runtime initialises) is shared
    with the engine's stream and the null stream, and two slots on one queue
    cost 6 % of the capacity with three slots.  Set GPU_MAX_HW_QUEUES to at
    least slots + 2 before the process first touches the GPU.
    """

    def __init__(self, slots=4, max_atoms=1 << 20, max_systems=1024, edge_budget=1 << 28, *,
                 full=False, distances=False, displacements=False):
        self._dist, self._disp = bool(distances), bool(displacements)
        self._borrowed = {}  # request id -> positions the server reads until it is done
        opts = (16 if full else 0) | (8 if distances else 0) | (32 if displacements else 0)
        self._p = lib.nl_server_create_ex(int(slots), int(max_atoms), int(max_systems),
                                          int(edge_budget), opts)
        if not self._p:
            raise RuntimeError(_lib.last_error())

    def close(self):
        if self._p:
            lib.nl_server_free(self._p)
            self._p = None
        self._borrowed.clear()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, cell, positions, cutoff, *, copy=True):
        """Queues one request; returns its id.  copy=False lets the server
        read `positions` when the request's batch is staged instead of copying
        them now (the server keeps a reference until the request is done):
        the array must not be modified in place before then."""
        from . import _cell_args, _positions

        cutoff = float(cutoff)
        if not np.isfinite(cutoff) or cutoff <= 0.0:
            raise ValueError("Cutoff must be positive and finite")
        pos = _positions(positions)
        h, pb = _cell_args(cell)
        rid = int(lib.nl_server_submit_ex(self._p, pos.ctypes.data_as(_lib.dp), pos.shape[0], h,
                                          pb, cutoff, 0 if copy else 1))
        if rid < 0:
            raise ValueError(_lib.last_error())
        if not copy:
            self._borrowed[rid] = pos
        return rid

    def wait(self, rid, timeout=None):
        rc = lib.nl_server_wait(self._p, int(rid), -1.0 if timeout is None else float(timeout))
        if rc == 1:
            return False
        self._borrowed.pop(int(rid), None)  # done or failed: no longer read
        if rc != 0:
            raise RuntimeError(_lib.last_error())
        return True

    def info(self, rid):
        """(edges, latency in seconds from submit to completion, batch size)"""
        m = C.c_int64()
        lat = C.c_double()
        bs = C.c_int()
        if lib.nl_server_result_info(self._p, int(rid), C.byref(m), C.byref(lat), C.byref(bs)):
            raise RuntimeError(_lib.last_error())
        return int(m.value), float(lat.value), int(bs.value)

    def result(self, rid, timeout=None):
        """Blocks until done; returns {"edge_index": (2, M) int64, "shift": (M, 3) int32}
        (plus "distance" (M,) and "displacement" (M, 3) float64 when the server
        emits them) and forgets the request."""
        if not self.wait(rid, timeout):
            raise TimeoutError(f"request {rid} not done after {timeout} s")
        from . import _host_block

        m, _, _ = self.info(rid)
        specs = [((2, m), np.int64)]
        specs += [((m,), np.float64)] if self._dist else []
        specs += [((m, 3), np.float64)] if self._disp else []
        arrs = _host_block(specs + [((m, 3), np.int32)])
        ei, sh = arrs[0], arrs[-1]
        d = arrs[1] if self._dist else None
        v = arrs[1 + int(self._dist)] if self._disp else None
        if lib.nl_server_take_ex(self._p, int(rid), ei[0].ctypes.data_as(_lib.i64p),
                                 ei[1].ctypes.data_as(_lib.i64p), sh.ctypes.data_as(_lib.i32p),
                                 None if d is None else d.ctypes.data_as(_lib.dp),
                                 None if v is None else v.ctypes.data_as(_lib.dp)):
            raise RuntimeError(_lib.last_error())
        out = {"edge_index": ei, "shift": sh}
        if d is not None:
            out["distance"] = d
        if v is not None:
            out["displacement"] = v
        return out

    def result_device(self, rid, timeout=None):
        """Blocks until done; returns (edges, addr_i, addr_j, addr_shift): device
        addresses of the request's lists (M int64, M int64, M x 3 int32), valid
        until release(rid)."""
        if not self.wait(rid, timeout):
            raise TimeoutError(f"request {rid} not done after {timeout} s")
        m, _, _ = self.info(rid)
        pi, pj, ps = C.c_void_p(), C.c_void_p(), C.c_void_p()
        if lib.nl_server_result_device(self._p, int(rid), C.byref(pi), C.byref(pj),
                                       C.byref(ps)):
            raise RuntimeError(_lib.last_error())
        return m, pi.value or 0, pj.value or 0, ps.value or 0

    def release(self, rid):
        if lib.nl_server_release(self._p, int(rid)):
            raise RuntimeError(_lib.last_error())

    def stats(self):
        v = [C.c_int64() for _ in range(6)]
        lib.nl_server_stats(self._p, *[C.byref(x) for x in v[:5]])
        lib.nl_server_allocations(self._p, C.byref(v[5]))
        return dict(zip(("completed", "batches", "max_inflight", "max_batch", "refills",
                         "allocations"), (int(x.value) for x in v)))
