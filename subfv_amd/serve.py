"""Served mixed-size explicit solves on one GPU.

A request carries a raw mesh (quad or prism grid of a seeded size, as a
msh file is read: vertices, elements, boundary faces), a seeded two-state
initial condition (init_1drp, ns_euler_module.F90:36-43) and a step count.
The server builds the connectivity after the request arrives (build_mesh,
inside the request's latency: on the device, or on the host), admits
arriving requests while the HBM budget, the batch capacity and the cap on
cells in flight allow, keeps every admitted state resident, and advances
all of them together with the multi-instance launches of subfv_batch_run
(continuous batching: requests that arrive while others run are admitted
between chunks of steps, in arrival order, and finished ones leave).  Each
instance's arithmetic is that of a solve alone, so results are
bit-identical to subfv_run on the same request.

The cap on cells in flight keeps the batch at the size that fills the GPU:
beyond it, more instances share the same throughput and every one of them
finishes later.

`serve_plain` is the baseline: the same requests one at a time, in arrival
order, each with its own connectivity build and launches (subfv_run).
`autotune` picks the chunk length, the launch form (per-step launches,
graph replay, or the persistent kernel and its workgroups per CU), the
cells cap and where connectivity is built by timing candidates on a
calibration sample.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor
from concurrent.futures import wait as wait_futures
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import meshgen as mg
from .api import HostMesh, Solver, SubfvError, make_config

SEED = 20261202


@dataclass
class Request:
    rid: int
    kind: str           # "quad" or "tri"
    n: int              # grid size (tri: 2 n^2 prisms)
    arrival: float      # seconds after the start
    wl: np.ndarray      # primitive left state
    wr: np.ndarray      # primitive right state
    x_split: float
    steps: int
    host: HostMesh | None = None   # connectivity built beforehand (else built on admission)
    mesh: mg.Mesh | None = None    # the raw mesh
    boundary_2d: bool = True
    cells: int = 0
    mesh_path: str | None = None  # a msh 2.2 file instead of the generated mesh
    bc_names: tuple = ()


def make_requests(n_req: int = 4096, rate: float = 200.0, steps: int = 300, seed: int = SEED,
                  cmin: int = 32 ** 2, cmax: int = 512 ** 2) -> list[Request]:
    """Seeded request stream: Poisson arrivals at `rate` per second, quad or
    tri 50/50, cell count log-uniform in [cmin, cmax], left/right states with
    rho, p in [0.1, 1] and |u| <= 0.5, split position in [0.3, 0.7]."""
    ids = np.arange(n_req, dtype=np.uint64)
    u = [mg.uniform(seed, 100 + k, ids) for k in range(12)]
    t = np.cumsum(-np.log1p(-u[0]) / rate)
    out = []
    for i in range(n_req):
        kind = "quad" if u[1][i] < 0.5 else "tri"
        cells = math.exp(math.log(cmin) + u[2][i] * (math.log(cmax) - math.log(cmin)))
        n = max(4, int(round(math.sqrt(cells if kind == "quad" else cells / 2))))
        sp_l, an_l = 0.5 * math.sqrt(u[5][i]), 2 * math.pi * u[6][i]
        sp_r, an_r = 0.5 * math.sqrt(u[9][i]), 2 * math.pi * u[10][i]
        wl = np.array([0.1 + 0.9 * u[3][i], sp_l * math.cos(an_l), sp_l * math.sin(an_l), 0.0, 0.1 + 0.9 * u[4][i]])
        wr = np.array([0.1 + 0.9 * u[7][i], sp_r * math.cos(an_r), sp_r * math.sin(an_r), 0.0, 0.1 + 0.9 * u[8][i]])
        out.append(Request(i, kind, n, float(t[i]), wl, wr, 0.3 + 0.4 * float(u[11][i]), steps))
    return out


def prepare(reqs: list[Request], workers: int = 16, connectivity: bool = False) -> list[Request]:
    """Client side: generate each mesh (or read its msh 2.2 file); with
    `connectivity`, also build it here on the host (threads; the C-ABI call
    releases the interpreter lock), outside the server."""
    def one(r: Request):
        if r.mesh_path:
            mesh = mg.read_msh(r.mesh_path, r.bc_names, r.boundary_2d)
        elif r.kind == "quad":
            mesh = mg.quad_mesh(r.n, seed=SEED + r.rid)
        else:
            mesh = mg.tri_mesh(r.n, seed=SEED + r.rid)
        r.mesh = mesh
        if connectivity:
            r.host = HostMesh(mesh)
        r.boundary_2d = mesh.boundary_2d
        r.cells = mesh.n_elems
        return r
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, reqs))


def config_for(r: Request):
    return make_config("multi_point", bc_style=2, boundary_2d=r.boundary_2d, local_time_step=False,
                       second_order=False, cfl=0.4, tmax=1e30)


def admit(s: Solver, r: Request, connectivity: str = "device"):
    """Build the connectivity (on the device of `s`, or on the host) unless
    the request brings it, upload the mesh, set the configuration and the
    init_1drp state."""
    if r.host is not None:
        s.upload_mesh(r.host, r.boundary_2d)
    elif connectivity == "device":
        s.build_mesh(r.mesh)
    elif connectivity == "host":
        s.upload_mesh(HostMesh(r.mesh), r.boundary_2d)
    else:
        raise ValueError(f"connectivity {connectivity!r}: 'device' or 'host'")
    s.set_config(config_for(r))
    s.init_1drp(r.x_split, r.wl, r.wr)
    s.set_time(0.0)


def device_bytes(r: Request) -> int:
    """HBM an admitted instance holds (mesh, partials, states, work arrays)."""
    return 3200 * r.cells + (4 << 20)


class Batch:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = _lib.lib().subfv_batch_create(C.byref(h), device)
        if st:
            raise SubfvError(st, "subfv_batch_create failed")
        self.h = h

    def set_grid(self, wgs_per_cu: int):
        """Workgroups per CU of the persistent launch (0: as many as fit)."""
        st = _lib.lib().subfv_batch_set_grid(self.h, wgs_per_cu)
        if st:
            raise SubfvError(st, "subfv_batch_set_grid failed")

    def run(self, solvers, n_steps: int, mode: int = 1):
        """mode 0: one multi-instance launch per kernel of a step, 1: replayed
        3-step hipGraph, 2: one persistent cooperative launch."""
        arr = (C.c_void_p * len(solvers))(*[s.h.value for s in solvers])
        st = _lib.lib().subfv_batch_run(self.h, arr, len(solvers), n_steps, int(mode))
        if st:
            raise SubfvError(st, _lib.lib().subfv_batch_error(self.h).decode())

    def wait(self):
        st = _lib.lib().subfv_batch_wait(self.h)
        if st:
            raise SubfvError(st, _lib.lib().subfv_batch_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().subfv_batch_destroy(self.h)
            self.h = None


@dataclass
class Result:
    rid: int
    latency: float
    finish: float
    sol: np.ndarray | None = None


@dataclass
class Params:
    chunk: int = 30          # steps per batch call (a multiple of 3 keeps the graph cycle aligned)
    max_batch: int = 256     # instances resident at once
    use_graph: bool = True
    persistent: bool = False  # one persistent cooperative launch per chunk (overrides use_graph)
    wgs_per_cu: int = 0       # its workgroups per CU (0: as many as fit)
    budget: int = 96 << 30   # HBM bytes for admitted instances
    admit_workers: int = 8   # host threads uploading admitted requests
    prefetch: int = 64       # instances uploaded ahead of free batch capacity
    first: int = 64          # uploads to wait for before the first chunk of an idle server
    max_cells: int = 1 << 40  # cells of the instances in flight (one instance always fits)
    connectivity: str = "device"  # where admission builds a request's connectivity


def serve(reqs: list[Request], params: Params = Params(), device: int = 0, keep: bool = False,
          realtime: bool = True, pool: list | None = None, on_finish=None) -> tuple[list[Result], float]:
    """Continuous batching.  Returns the results (latency = finish - arrival)
    and the makespan in seconds.  Arrived requests are uploaded into idle
    contexts by a pool of host threads in the background, up to the free
    batch capacity plus `params.prefetch` ahead, and join the batch at the
    next chunk boundary; an idle server waits for `params.first` uploads
    (or all that are under way) before it starts.  `pool`: idle contexts to
    admit into (a server keeps them between calls; they are returned to it);
    without one, contexts are created as needed and closed at the end.
    `on_finish(request, solver)`: called as each solve finishes, before its
    context is reused (e.g. to hand a snapshot to output.AsyncWriter)."""
    if params.persistent and params.connectivity == "device":
        # one run of this combination (96 small requests, the admission
        # threads building connectivity on the device while the cooperative
        # kernel ran) stopped making progress; the cause is not established
        raise ValueError("the persistent kernel is served with connectivity built on the host only")
    batch = None    # created with the first chunk
    own = pool is None
    pool = [] if own else pool
    pending = sorted(reqs, key=lambda r: r.arrival)
    pi = 0
    active = []     # [request, solver, remaining], in the batch
    staged = []     # uploaded, waiting for batch capacity
    uploading = {}  # future -> entry
    used = 0
    results = []
    t0 = time.perf_counter()

    def now():
        return time.perf_counter() - t0 if realtime else float("inf")

    def start_uploads(limit):
        nonlocal pi, used
        t = now()
        n = 0
        while (pi < len(pending) and pending[pi].arrival <= t and n < limit
               and used + device_bytes(pending[pi]) <= params.budget):
            r = pending[pi]
            pi += 1
            s = pool.pop() if pool else Solver(device)
            used += device_bytes(r)
            # each upload works on its own context and stream; the C-ABI calls
            # release the interpreter lock, so uploads overlap the batch
            uploading[ex.submit(admit, s, r, params.connectivity)] = [r, s, r.steps]
            n += 1

    def collect(wait_for=0):
        if wait_for > 0 and uploading:
            done = set()
            while len(done) < min(wait_for, len(uploading)):
                d, _ = wait_futures(list(set(uploading) - done), return_when=FIRST_COMPLETED)
                done |= d
        for f in [f for f in uploading if f.done()]:
            entry = uploading.pop(f)
            try:
                f.result()  # an upload error ends the serve
            except Exception:
                pool.append(entry[1])
                raise
            staged.append(entry)

    ex = ThreadPoolExecutor(max_workers=params.admit_workers)
    try:
        while pi < len(pending) or active or staged or uploading:
            collect()
            # staged instances join in arrival order while the batch capacity
            # and the cells cap allow
            staged.sort(key=lambda e: e[0].arrival)
            cells = sum(a[0].cells for a in active)
            k = 0
            while (k < len(staged) and len(active) < params.max_batch
                   and (not active or cells + staged[k][0].cells <= params.max_cells)):
                cells += staged[k][0].cells
                active.append(staged[k])
                k += 1
            del staged[:k]
            if not active:
                start_uploads(params.max_batch + params.prefetch - len(uploading))
                if uploading:
                    collect(wait_for=params.first)
                elif pending[pi].arrival <= now():
                    raise SubfvError(2, f"request {pending[pi].rid} needs {device_bytes(pending[pi])} bytes, "
                                        f"more than the HBM budget of {params.budget}")
                else:
                    time.sleep(max(0.0, pending[pi].arrival - (time.perf_counter() - t0)))
                continue
            steps = min(params.chunk, min(a[2] for a in active))
            if batch is None:
                batch = Batch(device)
                batch.set_grid(params.wgs_per_cu)
            batch.run([a[1] for a in active], steps, 2 if params.persistent else int(params.use_graph))
            start_uploads(params.max_batch - len(active) + params.prefetch - len(staged) - len(uploading))
            batch.wait()
            done_t = time.perf_counter() - t0
            still = []
            for a in active:
                a[2] -= steps
                if a[2] == 0:
                    r, s = a[0], a[1]
                    s.run(0, 0)  # raises the instance's error, if any
                    sol = s.get_sol() if keep else None
                    results.append(Result(r.rid, done_t - r.arrival, done_t, sol))
                    if on_finish is not None:
                        on_finish(r, s)
                    pool.append(s)
                    used -= device_bytes(r)
                else:
                    still.append(a)
            active = still
    finally:
        ex.shutdown(wait=True)
        if batch is not None:
            batch.close()
        makespan = time.perf_counter() - t0
        for a in active + staged + list(uploading.values()):
            pool.append(a[1])
        if own:
            for s in pool:
                s.close()
    return results, makespan


def serve_plain(reqs: list[Request], device: int = 0, keep: bool = False,
                realtime: bool = True, solver: Solver | None = None,
                connectivity: str = "host", on_finish=None) -> tuple[list[Result], float]:
    """One solve at a time in arrival order, each with its own connectivity
    build and launches (on `solver` if given, kept open).  `on_finish` as
    for serve."""
    s = solver if solver is not None else Solver(device)
    results = []
    t0 = time.perf_counter()
    try:
        for r in sorted(reqs, key=lambda q: q.arrival):
            if realtime:
                wait = r.arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(wait)
            admit(s, r, connectivity)
            s.run(r.steps, 0)
            sol = s.get_sol() if keep else None
            t = time.perf_counter() - t0
            results.append(Result(r.rid, t - (r.arrival if realtime else 0.0), t, sol))
            if on_finish is not None:
                on_finish(r, s)
    finally:
        makespan = time.perf_counter() - t0
        if solver is None:
            s.close()
    return results, makespan


def summarize(results: list[Result], makespan: float) -> dict:
    lat = np.array([r.latency for r in results])
    return {"solves": len(results), "makespan_s": makespan, "solves_per_s": len(results) / makespan,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99))}


def autotune(sample: list[Request], device: int = 0, candidates=None, slack: float = 0.95,
             log=None) -> tuple[Params, list]:
    """Times candidates on the sample (all requests available at once:
    latency = finish time) and returns, among the candidates within `slack`
    of the best throughput, the one with the lowest median latency, with the
    table of trials.  Without `candidates`, two rounds: the launch form
    (per-step launches or a replayed graph) x where connectivity is built,
    then the chunk length x the cells cap for the best of those.
    `log(trial)` is called after each trial."""
    cap = max(p.max_batch + p.prefetch for p in candidates) if candidates is not None else 256 + Params().prefetch
    pool = [Solver(device) for _ in range(cap)]
    trials = []

    def trial(p):
        res, ms = serve(sample, p, device, realtime=False, pool=pool)
        lat = np.array([r.latency for r in res])
        t = {"chunk": p.chunk, "max_batch": p.max_batch, "use_graph": p.use_graph,
             "persistent": p.persistent, "wgs_per_cu": p.wgs_per_cu,
             "max_cells": p.max_cells, "connectivity": p.connectivity,
             "solves_per_s": len(sample) / ms, "p50_s": float(np.percentile(lat, 50))}
        trials.append(t)
        if log is not None:
            log(t)
        return t

    def pick(cands, ts):
        top = max(t["solves_per_s"] for t in ts)
        ok = [i for i, t in enumerate(ts) if t["solves_per_s"] >= slack * top]
        return cands[min(ok, key=lambda i: (ts[i]["p50_s"], -ts[i]["solves_per_s"]))]

    try:
        if candidates is not None:
            best = pick(candidates, [trial(p) for p in candidates])
        else:
            # the persistent kernel is tried through `candidates` only: it
            # never won these rounds, and it is served with host-built
            # connectivity only
            forms = [dict(use_graph=False), dict(use_graph=True)]
            first = [Params(chunk=30, max_batch=256, connectivity=cn, **f) for f in forms for cn in ("device", "host")]
            b = pick(first, [trial(p) for p in first])
            second = [Params(chunk=c, max_batch=256, max_cells=mc, connectivity=b.connectivity, use_graph=b.use_graph,
                             persistent=b.persistent, wgs_per_cu=b.wgs_per_cu)
                      for c in (30, 60) for mc in (1 << 20, 4 << 20, 1 << 40)]
            best = pick(second, [trial(p) for p in second])
    finally:
        for s in pool:
            s.close()
    return best, trials
