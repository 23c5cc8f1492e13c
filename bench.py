"""Benchmark of the MI355X neighbour-list engine (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Headline workload (one "step"): the 256-system ragged batch (N log-uniform in
[8, 2048], half isolated, half periodic orthorhombic/triclinic at liquid
density) built in ONE pass with three lists at cutoffs {5, 6, 14} A, inputs
resident in HBM, outputs left on the device.  `value` = edges/s summed over
the three lists.  Beside it: systems/s of the same batch through the batched
path and through the plain path (one call per system), p50/p99 latency of a
batch build, the 10k-atom triclinic system at rc 6 against the CPU oracle,
the serving trace (continuous batching: capacity, p50/p99 latency at 70% load,
against the plain path), the roofline of the dominant kernel (pair fill) and
the CPU baseline.

Prints ONE JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

# Hardware queues of the HIP runtime (read when it initialises, so set before
# anything touches the GPU).  The serving measurement runs one HIP stream per
# slot beside the engine's stream and the null stream; with HIP's default of
# four queues two of them share one, measured 6 % of the serving capacity
# (68,400-69,000 against 72,000-74,800 systems/s on one box).  An environment
# that already asks for four (HIP's default) is raised to eight.
try:
    _hw_queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0)
except ValueError:
    _hw_queues = 0
if _hw_queues < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neighborlist-rs_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


FILL_PROFILE = os.path.join(ROOT, "profiles", "k_pairs_fill_c4_ragged_batch.md")


def profiled_traffic(edges):
    """HBM bytes per fill launch measured by rocprofv3 --pmc on the same
    workload (FETCH_SIZE and WRITE_SIZE, each scaled by a copy-kernel
    calibration), read from the committed profile summary; None when the
    profile is missing or was taken on a different edge count."""
    vals = {}
    try:
        with open(FILL_PROFILE) as f:
            for line in f:
                if ":" in line and line.split(":")[0].strip() in ("traffic_bytes_per_launch",
                                                                  "edges_per_launch"):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = int(v.strip().strip("`"))
    except OSError:
        return None
    if vals.get("edges_per_launch") != edges:
        return None
    return vals.get("traffic_bytes_per_launch")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class DevArray:
    def __init__(self, lib, host):
        self.lib = lib
        self.nbytes = host.nbytes
        self.ptr = lib.nl_device_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise RuntimeError("device allocation failed")
        if lib.nl_memcpy_h2d(C.c_void_p(self.ptr), host.ctypes.data_as(C.c_void_p),
                             self.nbytes) != 0:
            raise RuntimeError("H2D copy failed")

    def free(self):
        self.lib.nl_device_free(C.c_void_p(self.ptr))


def run_systems(lib, L, dpos, n, sizes, H9, has, pbc, cut, multi):
    r = C.c_void_p()
    rc = lib.nl_search_batch_device(
        C.c_void_p(dpos), n, sizes.ctypes.data_as(L.i64p), sizes.size,
        H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p), pbc.ctypes.data_as(L.u8p),
        cut.ctypes.data_as(L.dp), cut.size, 0, int(multi), 1, 0, 0, C.byref(r))
    if rc != 0:
        raise RuntimeError(L.last_error())
    edges = sum(int(lib.nl_result_num_edges(r, q)) for q in range(cut.size))
    lib.nl_result_free(r)
    return edges


def run_ex(lib, L, dpos, n, sizes, H9, has, pbc, cut, multi, opts):
    """nl_search_ex on device positions (options: NL_* bits of nl_engine_ext.h)."""
    r = C.c_void_p()
    rc = lib.nl_search_ex(
        C.c_void_p(dpos), n, sizes.ctypes.data_as(L.i64p), sizes.size,
        H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p), pbc.ctypes.data_as(L.u8p),
        cut.ctypes.data_as(L.dp), cut.size, 0, int(multi), opts, None, C.byref(r))
    if rc != 0:
        raise RuntimeError(L.last_error())
    edges = sum(int(lib.nl_result_num_edges(r, q)) for q in range(cut.size))
    lib.nl_result_free(r)
    return edges


def small_single_systems(lib, L, calls=200):
    """Single small isolated systems, one call each, through the small-system
    kernels (nl_set_small_path 1, the default) and through the pipeline (0).
    Engine timing events are off here: per-call latency as a caller sees it."""
    lib.nl_set_timing(0)
    small_calls = {"workload": "isolated clusters at 0.1 atoms/A^3, rc 6.0, one call each through "
                               "nl_search_batch_device (positions on the device, lists left there); "
                               f"median host-clock time per call over {calls} calls, engine "
                               "timing events off",
                   "timing": "small_path_us = nl_set_small_path(1), the default (one launch "
                             "up to 128 atoms, two launches above), pipeline_us = "
                             "nl_set_small_path(0)"}
    for n_s in (9, 100, 500):
        ps = np.random.default_rng(n_s).uniform(0.0, (n_s / 0.1) ** (1 / 3), (n_s, 3))
        dps = DevArray(lib, ps)
        ss = np.array([n_s], np.int64)
        row = {}
        for mode, key in ((1, "small_path_us"), (0, "pipeline_us")):
            lib.nl_set_small_path(mode)
            run_systems(lib, L, dps.ptr, n_s, ss, np.zeros((1, 9)), np.zeros(1, np.uint8),
                        np.zeros(3, np.uint8), np.array([6.0]), False)
            tt = []
            for _ in range(calls):
                t0 = time.perf_counter()
                es = run_systems(lib, L, dps.ptr, n_s, ss, np.zeros((1, 9)),
                                 np.zeros(1, np.uint8), np.zeros(3, np.uint8), np.array([6.0]),
                                 False)
                tt.append(time.perf_counter() - t0)
            row[key] = round(float(np.median(tt)) * 1e6, 1)
            row["edges"] = es
        lib.nl_set_small_path(1)
        dps.free()
        small_calls[f"{n_s}_atoms"] = row
    lib.nl_set_timing(1)
    return small_calls


def md_skin_trajectory(lib, L, p3, H3, rc=6.0, skin=1.0, steps=200, sigma=0.01, seed=3):
    """Skin lists on the 10k-atom triclinic system along a random walk
    (sigma per coordinate per step): time per update against a full
    graph-replayed build every step."""
    rng = np.random.default_rng(seed)
    n = p3.shape[0]
    pos = np.ascontiguousarray(p3.copy())
    # the trajectory is generated up front and resident on the device
    traj = np.ascontiguousarray(pos[None] + np.cumsum(rng.normal(0.0, sigma, (steps, n, 3)), 0))
    dtraj = DevArray(lib, traj)
    sz = np.array([n], np.int64)
    H = np.ascontiguousarray(H3.reshape(1, 9))
    has = np.ones(1, np.uint8)
    pb = np.ones((1, 3), np.uint8)
    cut = np.array([rc])
    sk = np.array([skin])
    k = C.c_void_p()
    if lib.nl_skin_create(C.c_void_p(dtraj.ptr), n, sz.ctypes.data_as(L.i64p), 1,
                          H.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                          pb.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp),
                          sk.ctypes.data_as(L.dp), 1, 0, 0, 0, None, C.byref(k)) != 0:
        raise RuntimeError(L.last_error())
    r = C.c_int(0)
    lib.nl_device_synchronize()
    t0 = time.perf_counter()
    for t in range(steps):
        if lib.nl_skin_update(k, C.c_void_p(dtraj.ptr + 24 * n * t), None, None, C.byref(r)) != 0:
            raise RuntimeError(L.last_error())
    skin_ms = 1e3 * (time.perf_counter() - t0) / steps
    u, b, d = C.c_int64(), C.c_int64(), C.c_double()
    lib.nl_skin_stats(k, C.byref(u), C.byref(b), C.byref(d))
    m_skin = int(lib.nl_skin_num_edges(k, 0))
    lib.nl_skin_free(k)
    # reference: a graph-captured build at rc every step (positions copied in)
    dpos = DevArray(lib, pos)
    plan = C.c_void_p()
    if lib.nl_plan_create(C.c_void_p(dpos.ptr), n, sz.ctypes.data_as(L.i64p), 1,
                          H.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                          pb.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp), 1, 0, 0, 1, 0,
                          C.byref(plan)) != 0:
        raise RuntimeError(L.last_error())
    lib.nl_plan_run(plan)
    lib.nl_device_synchronize()
    t0 = time.perf_counter()
    for t in range(steps):
        lib.nl_memcpy_d2d(C.c_void_p(dpos.ptr), C.c_void_p(dtraj.ptr + 24 * n * t), 24 * n)
        if lib.nl_plan_run(plan) != 0:
            raise RuntimeError(L.last_error())
    plan_ms = 1e3 * (time.perf_counter() - t0) / steps
    m_rc = int(lib.nl_plan_num_edges(plan, 0))
    lib.nl_plan_free(plan)
    dpos.free()
    dtraj.free()
    return {"workload": f"c3_triclinic 10,008 atoms, rc {rc}, skin {skin}, {steps} random-walk "
                        f"steps of sigma {sigma} A per coordinate",
            "ms_per_step": round(skin_ms, 4), "builds": int(b.value),
            "rebuild_every_step_ms": round(plan_ms, 4),
            "speedup": round(plan_ms / skin_ms, 2),
            "edges_at_rc_plus_skin": m_skin, "edges_at_rc": m_rc,
            "timing": "host clock around each update (displacement check on the device, "
                      "graph replay when rebuilding)"}


def timing(lib):
    b, c, f, t = C.c_float(), C.c_float(), C.c_float(), C.c_float()
    e, nb = C.c_int64(), C.c_int64()
    lib.nl_last_timing(C.byref(b), C.byref(c), C.byref(f), C.byref(t), C.byref(e), C.byref(nb))
    return b.value, c.value, f.value, t.value, e.value


def cell_arrays(cells):
    S = len(cells)
    H9 = np.zeros((S, 9))
    has = np.zeros(S, np.uint8)
    pbc = np.ones((S, 3), np.uint8)
    for s, H in enumerate(cells):
        if H is not None:
            H9[s] = np.asarray(H).reshape(9)
            has[s] = 1
    return H9, has, pbc


def serve_trace(L, n_req, slots, load):
    """Serving trace: Poisson arrivals of single systems (N log-uniform in
    [16, 50000], isolated / orthorhombic / triclinic, rc uniform in [4, 12]).

    Capacity: all requests queued at once (closed loop, median of five runs
    after two warm-up runs), results left on the device (the consumer of a
    served list is a GPU model) and, separately, copied into numpy arrays.  Latency: open loop, Poisson arrivals at `load`
    times the measured capacity; latency = submit -> lists complete on the
    device.  Plain: the same requests one call each through the plain path
    (host input, lists left on the device).  Requests are submitted with
    copy=False: the server reads the trace's arrays when it stages a batch
    instead of copying each one at submit."""
    import threading

    import neighborlist_rs as nr
    from neighborlist_rs.workloads import c5_serving_trace

    lib = L.lib
    trace = c5_serving_trace(seed=5, n_requests=n_req)
    cells = [None if H is None else nr.PyCell(np.asarray(H).tolist(), [True] * 3)
             for (_, _, H, _) in trace]
    atoms = sum(p.shape[0] for (_, p, _, _) in trace)

    def plain_pass():
        tot = 0
        for (_, p, H, rc) in trace:
            sz = np.array([p.shape[0]], np.int64)
            H9, has, pbc = cell_arrays([H])
            c = np.array([rc])
            r = C.c_void_p()
            if lib.nl_search_systems(p.ctypes.data_as(L.dp), p.shape[0], sz.ctypes.data_as(L.i64p),
                                     1, H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                                     pbc.ctypes.data_as(L.u8p), c.ctypes.data_as(L.dp), 1, 0, 0,
                                     0, 0, 0, C.byref(r)) != 0:
                raise RuntimeError(L.last_error())
            tot += int(lib.nl_result_num_edges(r, 0))
            lib.nl_result_free(r)
        return tot

    plain_pass()
    t0 = time.perf_counter()
    plain_edges = plain_pass()
    plain_s = time.perf_counter() - t0

    srv = nr.Server(slots=slots, max_atoms=1 << 18, max_systems=256, edge_budget=1 << 27)

    def closed(host):
        t0 = time.perf_counter()
        # the trace's arrays are not modified: the server may read them in place
        ids = [srv.submit(c, p, rc, copy=False) for c, (_, p, _, rc) in zip(cells, trace)]
        tot = 0
        for rid in ids:
            if host:
                tot += srv.result(rid)["shift"].shape[0]
            else:
                m, _, _, _ = srv.result_device(rid)
                srv.release(rid)
                tot += m
        return time.perf_counter() - t0, tot

    # warm-up: two loops grow the slot buffers and the output pool (a warmed
    # server allocates nothing more; a freed device buffer stalls the device)
    closed(False)
    closed(False)
    allocs_warm = srv.stats()["allocations"]
    runs = [closed(False) for _ in range(5)]  # median of five closed loops
    allocs_timed = srv.stats()["allocations"] - allocs_warm
    cap_s, served_edges = sorted(runs)[2]
    assert all(e == plain_edges for _, e in runs), (runs, plain_edges)
    host_s, host_edges = closed(True)
    assert host_edges == plain_edges
    capacity = n_req / cap_s
    log(f"serving: plain {n_req / plain_s:.1f} sys/s, served capacity {capacity:.1f} sys/s "
        f"(device lists), {n_req / host_s:.1f} sys/s (numpy lists)")

    # open loop at `load` x capacity
    qps = load * capacity
    rng = np.random.default_rng(55)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, size=n_req))
    ids = [None] * n_req
    ready = threading.Semaphore(0)

    def producer():
        t0 = time.perf_counter()
        for k, (c, (_, p, _, rc)) in enumerate(zip(cells, trace)):
            dt = arrivals[k] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids[k] = srv.submit(c, p, rc, copy=False)
            ready.release()

    st0 = srv.stats()
    th = threading.Thread(target=producer)
    t_start = time.perf_counter()
    th.start()
    lat, bsz = [], []
    for k in range(n_req):
        ready.acquire()
        srv.wait(ids[k])
        _, l, b = srv.info(ids[k])
        lat.append(l)
        bsz.append(b)
        srv.release(ids[k])
    t_end = time.perf_counter()
    th.join()
    st = srv.stats()
    srv.close()
    lat = np.array(lat)
    return {
        "workload": f"c5_serving_trace: {n_req} requests, {atoms} atoms, {plain_edges} edges; "
                    "N log-uniform [16,50000], rc U[4,12], 1/3 isolated, 1/3 orthorhombic, "
                    "1/3 triclinic",
        "slots": slots,
        "hip_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "capacity_systems_per_s": round(capacity, 1),
        "capacity_edges_per_s": round(plain_edges / cap_s, 1),
        "numpy_delivery_systems_per_s": round(n_req / host_s, 1),
        "plain_gpu_systems_per_s": round(n_req / plain_s, 1),
        "served_over_plain": round(capacity / (n_req / plain_s), 2),
        "open_loop": {"offered_qps": round(qps, 1),
                      "sustained_systems_per_s": round(n_req / (t_end - t_start), 1),
                      "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                                     "p99": round(1e3 * float(np.percentile(lat, 99)), 3),
                                     "mean": round(1e3 * float(lat.mean()), 3)},
                      "mean_batch": round(float(np.mean(bsz)), 2),
                      "batches": st["batches"] - st0["batches"],
                      "max_inflight": st["max_inflight"]},
        "refills": st["refills"],
        "allocations": {"warm_up": allocs_warm, "timed_closed_loops": allocs_timed,
                        "open_loop": st["allocations"] - st0["allocations"]},
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def ethanol_calibration(lib, L, O, threads):
    """The reference's own published case, ethanol_10000 (tests/data of the
    reference, PBC cube) at rc 6: the oracle at `threads` threads and at one
    thread beside the reference's logged times, and the GPU on the same input
    (device-resident positions and lists; and numpy in, numpy out through the
    Python API, host copies included)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from xyzio import read_xyz
    import neighborlist_rs as nr
    pos, rows, pbc = read_xyz("ethanol_10000.xyz")
    H = np.ascontiguousarray(rows.T)  # lattice vectors as columns
    n = pos.shape[0]
    d = DevArray(lib, pos)
    sz = np.array([n], np.int64)
    H9 = np.ascontiguousarray(H.reshape(1, 9))
    has = np.ones(1, np.uint8)
    pb = np.ascontiguousarray(np.array(pbc, np.uint8).reshape(1, 3))
    cut = np.array([6.0])
    for _ in range(3):
        run_systems(lib, L, d.ptr, n, sz, H9, has, pb, cut, False)
    t0 = time.perf_counter()
    for _ in range(20):
        e = run_systems(lib, L, d.ptr, n, sz, H9, has, pb, cut, False)
    gpu_ms = 1e3 * (time.perf_counter() - t0) / 20
    d.free()
    cell = nr.PyCell(H.tolist(), [bool(x) for x in pbc])
    for _ in range(3):
        nr.build_neighborlists(cell, pos, 6.0)
    t0 = time.perf_counter()
    for _ in range(10):
        r = nr.build_neighborlists(cell, pos, 6.0)
    api_ms = 1e3 * (time.perf_counter() - t0) / 10
    assert r["edge_index"].shape[1] == e
    O.set_threads(threads)
    O.search_single(pos, H, pbc, 6.0)
    t0 = time.perf_counter()
    for _ in range(10):
        ce = len(O.search_single(pos, H, pbc, 6.0)[0])
    cpu_ms = 1e3 * (time.perf_counter() - t0) / 10
    O.set_threads(1)
    t0 = time.perf_counter()
    for _ in range(3):
        O.search_single(pos, H, pbc, 6.0)
    cpu1_ms = 1e3 * (time.perf_counter() - t0) / 3
    O.set_threads(threads)
    assert ce == e, (ce, e)
    return {"fixture": f"ethanol_10000.xyz ({n} atoms, periodic cube), rc 6", "edges": int(e),
            "oracle_ms": round(cpu_ms, 3), "oracle_threads": threads,
            "oracle_1_thread_ms": round(cpu1_ms, 3), "cpu_model": cpu_model(),
            "reference_published_ms": {
                "grace_20_threads": 7.32, "grace_1_thread": 27.98, "epyc7r13_64_threads": 11.46,
                "source": "benchmarks/benchmark_grace.txt:39-40, "
                          "benchmarks/benchmark_EPYC7R13.txt:40 (Python API, numpy out)"},
            "gpu_device_ms": round(gpu_ms, 4),
            "gpu_python_api_ms": round(api_ms, 4),
            "gpu_python_api_note": "build_neighborlists, numpy in and out (H2D and D2H "
                                   "included), as the published times"}


def _median_ms(fn, reps, warm=2):
    for _ in range(warm):
        fn()
    tt = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        tt.append(time.perf_counter() - t0)
    return 1e3 * float(np.median(tt))


def config_lines(lib, L, O, threads, only=None):
    """The benchmark configurations one by one, and the reference's own
    published cases, each with: GPU time with positions and lists on the
    device (host clock per call, median), the device stages alone (hipEvents),
    numpy in / numpy out through the Python API (host copies included, as
    the reference's published times), the roofline fraction of the device
    stages (24 B per atom + 72 B per cell + 28 B per edge against 8 TB/s), and
    the oracle on `threads` host cores (None: no CPU timing).  only: the
    name of the one line to run (all when None)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from xyzio import read_xyz
    import neighborlist_rs as nr
    from neighborlist_rs import workloads as W
    out = {}

    def single(name, pos, H, pbc, cutoffs, published=None, reps=30):
        if only is not None and name != only:
            return
        n = pos.shape[0]
        multi = len(cutoffs) > 1
        d = DevArray(lib, np.ascontiguousarray(pos))
        sz = np.array([n], np.int64)
        H9, has, pb = cell_arrays([H])
        if H is not None:
            pb = np.ascontiguousarray(np.array(pbc, np.uint8).reshape(1, 3))
        cut = np.array(cutoffs, np.float64)
        edges = [0]

        def dev():
            edges[0] = run_systems(lib, L, d.ptr, n, sz, H9, has, pb, cut, multi)
        lib.nl_set_timing(0)
        dev_ms = _median_ms(dev, reps)
        stage = []
        lib.nl_set_timing(1)
        for _ in range(5):
            dev()
            stage.append(timing(lib)[3])
        lib.nl_set_timing(0)
        d.free()
        cell = None if H is None else nr.PyCell(np.asarray(H).tolist(), [bool(x) for x in pbc])
        if multi:
            api = lambda: nr.build_neighborlists_multi(cell, pos, list(cutoffs))
        else:
            api = lambda: nr.build_neighborlists(cell, pos, float(cutoffs[0]))
        api_ms = _median_ms(api, reps)
        dev_stage_ms = float(np.median(stage))
        alg = 24.0 * n + (72.0 if H is not None else 0.0) + 28.0 * edges[0]
        row = {"atoms": int(n), "cutoffs": list(cutoffs), "edges": int(edges[0]),
               "gpu_device_ms": round(dev_ms, 4), "gpu_device_stages_ms": round(dev_stage_ms, 4),
               "gpu_numpy_ms": round(api_ms, 4),
               "roofline_frac": round(alg / (dev_stage_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)}
        if O is not None:
            Hc = None if H is None else np.asarray(H)
            pc = None if H is None else list(pbc)
            if multi:
                cpu = lambda: O.search_single_multi(pos, Hc, pc, list(cutoffs), False)
            else:
                cpu = lambda: O.search_single(pos, Hc, pc, float(cutoffs[0]))
            ocpu = _median_ms(cpu, 5, warm=1)
            row["oracle_ms"] = round(ocpu, 3)
            row["oracle_threads"] = threads
            row["gpu_numpy_speedup_vs_oracle"] = round(ocpu / api_ms, 2)
        if published:
            row["reference_published"] = published
        log(f"config {name}: {row}")
        out[name] = row

    p1, _, rc1 = W.c1_isolated()
    single("c1_isolated_1000_rc6", p1, None, None, (rc1,),
           {"ms": 2.97, "threads": 8, "note": "isolated_1000.xyz rc 6 (not the same positions)",
            "source": "README.md:109-110"})
    p2, H2, _ = W.c2_ethanol()
    for rc, pub in ((6.0, {"ms": 5.35, "threads": 8, "note": "ethanol_1000.xyz",
                           "source": "README.md:122-123"}),
                    (14.0, {"ms": 37.88, "threads": 8, "note": "ethanol_1000.xyz",
                            "source": "README.md:165-166"})):
        single(f"c2_ethanol_1008_rc{rc:g}", p2, H2, [1, 1, 1], (rc,), pub)
    p3, H3, _ = W.c3_triclinic()
    single("c3_triclinic_10008_rc14", p3, H3, [1, 1, 1], (14.0,))
    pe, rows, pbc = read_xyz("ethanol_10000.xyz")
    He = np.ascontiguousarray(rows.T)
    single("ethanol_10000_rc14", pe, He, pbc, (14.0,),
           {"ms_grace_20_threads": 51.67, "ms_epyc_64_threads": 62.54,
            "ms_grace_1_thread": 247.26,
            "source": "benchmarks/benchmark_grace.txt:90-91, benchmark_EPYC7R13.txt:91"})
    single("ethanol_10000_multi_6_14", pe, He, pbc, (6.0, 14.0),
           {"ms": 313.0, "threads": 8, "note": "multi-cutoff [6, 14] in one pass",
            "source": "README.md:188-193"})

    # the reference's GNN batch (batch_radius_benchmark.py:109-133,
    # batch_results.json:2-12): numpy in, numpy out, no cells
    if only is not None and only != "gnn_molecule_batch_rc6":
        return out
    pg, bg = W.gnn_molecule_batch()
    res = [None]

    def gnn():
        res[0] = nr.build_neighborlists_batch(pg, bg, None, 6.0)
    lib.nl_set_timing(0)
    api_ms = _median_ms(gnn, 50, warm=5)
    m = int(res[0]["edge_index"].shape[1])
    sizes = np.bincount(bg).astype(np.int64)
    d = DevArray(lib, pg)
    H9, has, pb = cell_arrays([None] * sizes.size)
    cut = np.array([6.0])
    dev_ms = _median_ms(lambda: run_systems(lib, L, d.ptr, pg.shape[0], sizes, H9, has, pb, cut,
                                            False), 50, warm=5)
    lib.nl_set_timing(1)
    run_systems(lib, L, d.ptr, pg.shape[0], sizes, H9, has, pb, cut, False)
    gnn_stage_ms = timing(lib)[3]
    lib.nl_set_timing(0)
    d.free()
    row = {"workload": "128 isolated molecules, 42 atoms on average (5,376 atoms), rc 6, "
                       "build_neighborlists_batch(positions, batch, None, 6.0)",
           "gpu_device_stages_ms": round(gnn_stage_ms, 4),
           "edges": m, "gpu_numpy_ms": round(api_ms, 4), "gpu_device_ms": round(dev_ms, 4),
           "reference_published": {"ms_8_cpu": 0.24, "ms_1_cpu": 0.26,
                                   "source": "benchmarks/batch_results.json:2-12"}}
    if O is not None:
        ocpu = _median_ms(lambda: O.search_batch(pg, bg, None, 6.0), 20, warm=2)
        row["oracle_ms"] = round(ocpu, 3)
        row["oracle_threads"] = threads
    log(f"config gnn_batch: {row}")
    out["gnn_molecule_batch_rc6"] = row
    return out


def precision_report(L, pos, H, rc):
    """FP64 |r_ij| emitted with the edges (cell-list path) against a float64
    recomputation from the raw positions, |p_j + H S - p_i|, for the FP32
    filter on and off (the edge sets must be identical)."""
    lib = L.lib
    out = {}
    sets = []
    old = lib.nl_get_fp32_filter()
    old_tune = lib.nl_get_autotune()
    lib.nl_set_autotune(0)  # the filter setting must reach the kernels
    try:
        for f32 in (0, 1):
            lib.nl_set_fp32_filter(f32)
            sz = np.array([pos.shape[0]], np.int64)
            H9 = np.ascontiguousarray(np.asarray(H).reshape(9))
            has = np.ones(1, np.uint8)
            pb = np.ones(3, np.uint8)
            c = np.array([rc])
            r = C.c_void_p()
            if lib.nl_search_systems(pos.ctypes.data_as(L.dp), pos.shape[0],
                                     sz.ctypes.data_as(L.i64p), 1, H9.ctypes.data_as(L.dp),
                                     has.ctypes.data_as(L.u8p), pb.ctypes.data_as(L.u8p),
                                     c.ctypes.data_as(L.dp), 1, 0, 0, 1, 0, 1, C.byref(r)) != 0:
                raise RuntimeError(L.last_error())
            m = int(lib.nl_result_num_edges(r, 0))
            ei = np.empty(m, np.int64)
            ej = np.empty(m, np.int64)
            s = np.empty((m, 3), np.int32)
            d = np.empty(m)
            lib.nl_result_copy(r, 0, ei.ctypes.data_as(L.i64p), ej.ctypes.data_as(L.i64p),
                               s.ctypes.data_as(L.i32p))
            lib.nl_result_copy_dist(r, 0, d.ctypes.data_as(L.dp))
            lib.nl_result_free(r)
            rij = pos[ej] + s @ np.asarray(H).T - pos[ei]
            ref = np.sqrt((rij * rij).sum(1))
            err = np.abs(d - ref)
            key = np.lexsort((s[:, 2], s[:, 1], s[:, 0], ej, ei))
            sets.append((ei[key], ej[key], s[key]))
            out["fp32_filter" if f32 else "fp64"] = {
                "edges": m, "max_abs_err_A": float(err.max()) if m else 0.0,
                "max_rel_err": float((err / ref).max()) if m else 0.0,
                "max_dist_A": float(d.max()) if m else 0.0}
    finally:
        lib.nl_set_fp32_filter(old)
        lib.nl_set_autotune(old_tune)
    same = all(np.array_equal(a, b) for a, b in zip(sets[0], sets[1]))
    out["fp32_and_fp64_edge_sets_identical"] = bool(same)
    out["reference"] = "float64 numpy |p_j + H S - p_i| from the raw positions"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--tuning", default="auto",
                    help="'auto' (autotuner) or cell_div,group,parts,fp32 pinned")
    ap.add_argument("--only-serving", action="store_true",
                    help="run only the serving trace (for rocprofv3)")
    ap.add_argument("--only-small", action="store_true",
                    help="run only the single-small-system calls (for rocprofv3)")
    ap.add_argument("--only-headline", action="store_true",
                    help="run only the timed batch steps (for rocprofv3)")
    ap.add_argument("--only-configs", action="store_true",
                    help="run only the per-configuration lines")
    ap.add_argument("--only-config", default=None,
                    help="with --only-configs: the one line to run (e.g. c3_triclinic_10008_rc14)")
    ap.add_argument("--bin-order", type=int, default=-1,
                    help="1: Morton-ranked bins, 0: linear bin order (the engine default)")
    ap.add_argument("--serve-requests", type=int, default=600,
                    help="requests in the serving trace (0: skip serving)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="open-loop arrival rate as a fraction of the served capacity")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run a device-to-device copy of this size (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path "
              "(batches and serving traces would shard by system, one process per GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from neighborlist_rs import _lib as L
    from neighborlist_rs import workloads as W
    lib = L.lib
    lib.nl_set_timing(1)
    if args.bin_order >= 0:
        lib.nl_set_bin_order(args.bin_order)
    if args.tuning != "auto":
        dv, g, pp, f32 = (int(x) for x in args.tuning.split(","))
        lib.nl_set_autotune(0)
        lib.nl_set_cell_div(dv)
        lib.nl_set_work_split(g, pp)
        lib.nl_set_fp32_filter(f32)
    if args.calibrate_copy_mib > 0:
        nb = args.calibrate_copy_mib << 20
        a = lib.nl_device_malloc(nb)
        b = lib.nl_device_malloc(nb)
        for _ in range(3):
            lib.nl_memcpy_d2d(C.c_void_p(b), C.c_void_p(a), nb)
        lib.nl_device_synchronize()
        lib.nl_device_free(C.c_void_p(a))
        lib.nl_device_free(C.c_void_p(b))
        log(f"calibration copy: 3 x {args.calibrate_copy_mib} MiB device-to-device")

    if args.only_small:
        print(json.dumps({"small_single_systems": small_single_systems(lib, L)}), flush=True)
        return

    if args.only_configs:
        O = None
        if not args.no_cpu:
            from oracle import oracle as O
        threads = O.max_threads() if O is not None else None
        print(json.dumps({"configs": config_lines(lib, L, O, threads, args.only_config)}),
              flush=True)
        return

    if args.only_serving:
        print(json.dumps({"serving": serve_trace(L, args.serve_requests, args.serve_slots,
                                                 args.serve_load)}), flush=True)
        return

    # ---- headline: ragged batch, three cutoffs, one pass ------------------
    pos, sizes, cells, cut = W.c4_ragged_batch()
    cut = np.array(cut, dtype=np.float64)
    H9, has, pbc = cell_arrays(cells)
    n = pos.shape[0]
    S = sizes.size
    dpos = DevArray(lib, pos)
    log(f"c4: {S} systems, {n} atoms, cutoffs {cut.tolist()}")
    # eager batched path (first call autotunes this workload class)
    for _ in range(args.warmup):
        run_systems(lib, L, dpos.ptr, n, sizes, H9, has, pbc, cut, True)
    buf = C.create_string_buffer(4096)
    lib.nl_tune_report(buf, 4096)
    tuned = buf.value.decode().strip()
    log("tuned:", tuned)
    fills, counts = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        edges = run_systems(lib, L, dpos.ptr, n, sizes, H9, has, pbc, cut, True)
        b, c, f, t, _ = timing(lib)
        counts.append(c)
        fills.append(f)
    eager_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    log(f"eager batched: {eager_ms:.3f} ms/step, {edges} edges, build+count "
        f"{np.mean(counts):.3f} fill {np.mean(fills):.3f} ms")

    # steady state: the same batch as a graph-captured plan (one launch/step)
    plan = C.c_void_p()
    if lib.nl_plan_create(C.c_void_p(dpos.ptr), n, sizes.ctypes.data_as(L.i64p), S,
                          H9.ctypes.data_as(L.dp), has.ctypes.data_as(L.u8p),
                          pbc.ctypes.data_as(L.u8p), cut.ctypes.data_as(L.dp), cut.size, 0, 1,
                          1, 0, C.byref(plan)) != 0:
        raise RuntimeError(L.last_error())
    for _ in range(args.warmup):
        lib.nl_plan_run(plan)
    lat = []
    lib.nl_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        if lib.nl_plan_run(plan) != 0:
            raise RuntimeError(L.last_error())
        lat.append(time.perf_counter() - ts)
    lib.nl_device_synchronize()
    elapsed = time.perf_counter() - t0
    plan_edges = sum(int(lib.nl_plan_num_edges(plan, q)) for q in range(cut.size))
    assert plan_edges == edges, (plan_edges, edges)
    lib.nl_plan_free(plan)
    ms_per_step = 1e3 * elapsed / args.steps
    value = edges * args.steps / elapsed
    sys_per_s = S * args.steps / elapsed
    log(f"graph batched: {ms_per_step:.3f} ms/step")

    # roofline of the dominant kernel: the fill pass writes 28 B per edge
    # (2 x int64 indices + 3 x int32 shift) and reads each position once
    fill_ms = float(np.mean(fills))
    alg_bytes = 28.0 * edges + 24.0 * n
    achieved = alg_bytes / (fill_ms * 1e-3) / 1e9
    traffic = profiled_traffic(edges)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "traffic_source": None if traffic is None else
                "profiles/k_pairs_fill_c4_ragged_batch.md (rocprofv3 --pmc FETCH_SIZE, "
                "WRITE_SIZE; copy-calibrated)",
                "kernel": ("k_pairs<fill, fp32 filter>" if "fp32=1" in tuned
                           else "k_pairs_w6<fill> (FP64 path)"),
                "kernel_ms": round(fill_ms, 4),
                "alg_bytes_per_launch": int(alg_bytes),
                # the whole graph-replayed step (build, count, scan, fill)
                # against the same algorithmic bytes
                "whole_step": {"ms": round(ms_per_step, 4),
                               "achieved": round(alg_bytes / (ms_per_step * 1e-3) / 1e9, 2),
                               "frac": round(alg_bytes / (ms_per_step * 1e-3) / 1e9 /
                                             HBM_PEAK_GBPS, 4)}}

    if args.only_headline:
        print(json.dumps({"metric": "edges/s", "value": round(value, 1), "unit": "edges/s",
                          "ms_per_step": round(ms_per_step, 4), "roofline": roofline,
                          "tuning": tuned, "eager_batched_ms_per_step": round(eager_ms, 4)}),
              flush=True)
        dpos.free()
        return

    # full (directed) lists of the same batch: the pair kernels write both
    # directions of every pair (56 B per half-list pair), no second pass
    NL_DEVICE_POSITIONS, NL_FULL = 1, 16
    for _ in range(2):
        run_ex(lib, L, dpos.ptr, n, sizes, H9, has, pbc, cut, True, NL_DEVICE_POSITIONS | NL_FULL)
    ffill, fcount, ftot = [], [], []
    for _ in range(max(5, args.steps // 2)):
        full_edges = run_ex(lib, L, dpos.ptr, n, sizes, H9, has, pbc, cut, True,
                            NL_DEVICE_POSITIONS | NL_FULL)
        b, c, f, t, _ = timing(lib)
        fcount.append(c)
        ffill.append(f)
        ftot.append(t)
    full_fill = float(np.mean(ffill))
    full_lists = {"edges_per_step": int(full_edges),
                  "build_and_count_ms": round(float(np.mean(fcount)), 4),
                  "fill_ms": round(full_fill, 4),
                  "device_ms": round(float(np.mean(ftot)), 4),
                  "edges_per_s": round(full_edges / (float(np.mean(ftot)) * 1e-3), 1),
                  "fill_alg_GBps": round((28.0 * full_edges + 24.0 * n) / (full_fill * 1e-3) / 1e9,
                                         2),
                  "timing": "hipEvents around the device stages of the eager path"}
    log(f"full lists: {full_edges} edges, fill {full_fill:.3f} ms")

    # half lists with FP64 distances and displacements (28 + 8 + 24 B per edge)
    NL_DIST, NL_DISP = 8, 32
    geo = NL_DEVICE_POSITIONS | NL_DIST | NL_DISP
    for _ in range(2):
        run_ex(lib, L, dpos.ptr, n, sizes, H9, has, pbc, cut, True, geo)
    gfill, gtot = [], []
    for _ in range(max(5, args.steps // 2)):
        geo_edges = run_ex(lib, L, dpos.ptr, n, sizes, H9, has, pbc, cut, True, geo)
        b, c, f, t, _ = timing(lib)
        gfill.append(f)
        gtot.append(t)
    geo_fill = float(np.mean(gfill))
    geometry = {"edges_per_step": int(geo_edges),
                "fill_ms": round(geo_fill, 4),
                "device_ms": round(float(np.mean(gtot)), 4),
                "fill_alg_GBps": round((60.0 * geo_edges + 24.0 * n) / (geo_fill * 1e-3) / 1e9, 2),
                "bytes_per_edge": 60,
                "timing": "hipEvents around the device stages (fill includes the square-root pass)"}
    log(f"lists + distances + displacements: fill {geo_fill:.3f} ms")

    # ---- plain path: one call per system (same inputs, device-resident) --
    # per-call timings from here on run with the engine's timing events off
    # (they add event records and a query to every call); the stage timings
    # of the per-configuration lines switch them on where they are read
    lib.nl_set_timing(0)
    starts = np.concatenate([[0], np.cumsum(sizes)])
    one = np.ones(1, np.int64)
    def plain_pass():
        tot = 0
        for s in range(S):
            sz = np.array([sizes[s]], np.int64)
            tot += run_systems(lib, L, dpos.ptr + 24 * int(starts[s]), int(sizes[s]), sz,
                               np.ascontiguousarray(H9[s:s + 1]), has[s:s + 1].copy(),
                               pbc[s:s + 1].copy(), cut, True)
        return tot
    plain_pass()
    reps = max(1, min(5, args.steps // 4))
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_edges = plain_pass()
    plain_elapsed = time.perf_counter() - t0
    plain_sys_per_s = S * reps / plain_elapsed
    assert plain_edges == edges, (plain_edges, edges)
    log(f"plain: {plain_sys_per_s:.1f} systems/s vs batched {sys_per_s:.1f}")

    small_calls = small_single_systems(lib, L)
    log(f"small single systems: {small_calls}")

    # ---- 10k-atom triclinic PBC at rc 6 (single system) ------------------
    p3, H3, _ = W.c3_triclinic()
    d3 = DevArray(lib, p3)
    s3 = np.array([p3.shape[0]], np.int64)
    c3 = np.array([6.0])
    H3a = np.ascontiguousarray(H3.reshape(1, 9))
    h1 = np.ones(1, np.uint8)
    p1 = np.ones((1, 3), np.uint8)
    for _ in range(3):
        run_systems(lib, L, d3.ptr, p3.shape[0], s3, H3a, h1, p1, c3, False)
    t0 = time.perf_counter()
    k3 = 20
    for _ in range(k3):
        e3 = run_systems(lib, L, d3.ptr, p3.shape[0], s3, H3a, h1, p1, c3, False)
    c3_ms = 1e3 * (time.perf_counter() - t0) / k3
    log(f"c3 rc6: {c3_ms:.3f} ms, {e3} edges")
    precision = precision_report(L, p3, H3, 6.0)
    log(f"precision: {precision}")
    md_skin = md_skin_trajectory(lib, L, p3, H3)
    log(f"md skin: {md_skin}")

    serving = None
    if args.serve_requests > 0:
        serving = serve_trace(L, args.serve_requests, args.serve_slots, args.serve_load)

    cpu = None
    c3_cpu_ms = None
    calibration = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = O.max_threads()
        ocells = [None if H is None else (H, [1, 1, 1]) for H in cells]
        batch = np.repeat(np.arange(S, dtype=np.int32), sizes)
        t0 = time.perf_counter()
        passes = 0
        cpu_edges = 0
        while True:
            res = O.search_batch_multi(pos, batch, ocells, cut.tolist(), False)
            passes += 1
            cpu_edges = sum(len(r[0]) for r in res)
            if time.perf_counter() - t0 > args.cpu_seconds or passes >= 40:
                break
        cpu_elapsed = time.perf_counter() - t0
        assert cpu_edges == edges, (cpu_edges, edges)
        cpu = {"value": round(cpu_edges * passes / cpu_elapsed, 1), "unit": "edges/s",
               "cores": threads, "kind": "port",
               "sample": f"{passes} full passes of the same 256-system batch, cutoffs "
                         f"{{5,6,14}} ({cpu_elapsed:.1f} s), oracle/nl_oracle.c (OpenMP over "
                         f"systems and atoms)",
               "systems_per_s": round(S * passes / cpu_elapsed, 1)}
        t0 = time.perf_counter()
        for _ in range(10):
            O.search_single(p3, H3, [1, 1, 1], 6.0)
        c3_cpu_ms = 1e3 * (time.perf_counter() - t0) / 10
        cpu["cpu_model"] = cpu_model()
        log(f"cpu: {cpu['value']:.3e} edges/s on {threads} threads; c3 rc6 {c3_cpu_ms:.2f} ms")
        calibration = ethanol_calibration(lib, L, O, threads)
        log(f"ethanol_10000 rc 6: {calibration}")

    Oc = None
    if not args.no_cpu:
        from oracle import oracle as Oc
    configs = config_lines(lib, L, Oc, Oc.max_threads() if Oc is not None else None)

    out = {
        "metric": "edges/s and systems/s (batched/served vs plain GPU); p50/p99 build latency",
        "value": round(value, 1),
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generators in neighborlist_rs/workloads.py)",
        "config": {"workload": "c4_ragged_batch: 256 systems, N log-uniform [8,2048], half "
                               "isolated / half periodic, cutoffs {5,6,14} in one pass; "
                               "graph-captured steady state",
                   "atoms": int(n), "systems": int(S), "edges_per_step": int(edges)},
        "systems_per_s": round(sys_per_s, 1),
        "plain_gpu_systems_per_s": round(plain_sys_per_s, 1),
        "batched_over_plain": round(sys_per_s / plain_sys_per_s, 2),
        "small_single_systems": small_calls,
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 4),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 4)},
        "stage_ms": {"build_and_count": round(float(np.mean(counts)), 4),
                     "fill": round(fill_ms, 4)},
        "eager_batched_ms_per_step": round(eager_ms, 4),
        "tuning": tuned,
        "c3_triclinic_rc6": {"gpu_ms": round(c3_ms, 4), "edges": int(e3),
                             "edges_per_s": round(e3 / (c3_ms * 1e-3), 1),
                             "cpu_ms": None if c3_cpu_ms is None else round(c3_cpu_ms, 3),
                             "speedup_vs_cpu": None if c3_cpu_ms is None
                             else round(c3_cpu_ms / c3_ms, 2)},
        "precision_c3_rc6": precision,
        "md_skin_c3": md_skin,
        "full_lists_c4": full_lists,
        "distances_displacements_c4": geometry,
        "serving": serving,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "ethanol_10000_rc6_calibration": calibration,
        "configs": configs,
    }
    dpos.free()
    d3.free()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
