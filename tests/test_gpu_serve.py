"""Served batches: every request's final state, advanced in a batch of
mixed-size instances that join and leave between chunks of steps, equals
the state of the same request solved alone (tolerance zero: each
instance's arithmetic is that of subfv_run)."""
import numpy as np
import pytest

from subfv_amd import meshgen as mg
from subfv_amd import serve as S
from subfv_amd.api import Solver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def requests():
    reqs = S.make_requests(n_req=7, rate=1e6, steps=21, cmin=6 ** 2, cmax=22 ** 2)
    reqs[2].steps = 12   # leaves early
    reqs[5].steps = 33   # stays longer
    return S.prepare(reqs, workers=4)


@pytest.mark.parametrize("mode", ["launches", "graph", "persistent", "persistent1"])
def test_batched_results_equal_solves_alone(requests, mode):
    """Per-step launches, replayed graphs, and the persistent cooperative
    kernel (all workgroups that fit, or one per CU so that each takes many
    queue items)."""
    print(f"[gpu] serve equality mode={mode}", flush=True)
    alone, _ = S.serve_plain(requests, keep=True, realtime=False)
    # capacity 3 and 6-step chunks: instances are admitted while others run
    pers = mode.startswith("persistent")
    p = S.Params(chunk=6, max_batch=3, use_graph=mode == "graph", persistent=pers,
                 wgs_per_cu=1 if mode == "persistent1" else 0, connectivity="host" if pers else "device")
    mixed, _ = S.serve(requests, p, keep=True, realtime=False)
    a = {r.rid: r.sol for r in alone}
    m = {r.rid: r.sol for r in mixed}
    assert set(a) == set(m) == {r.rid for r in requests}
    for rid in a:
        assert np.array_equal(a[rid], m[rid]), rid


def test_cells_cap_and_connectivity_on_admission_keep_results(requests):
    """Admission under a cap on cells in flight, with the connectivity built
    on the host at admission, against solves alone whose connectivity was
    built on the device."""
    print("[gpu] serve cells cap, connectivity on admission", flush=True)
    alone, _ = S.serve_plain(requests, keep=True, realtime=False, connectivity="device")
    capped, _ = S.serve(requests, S.Params(chunk=6, max_batch=8, max_cells=600, connectivity="host"), keep=True,
                        realtime=False)
    a = {r.rid: r.sol for r in alone}
    m = {r.rid: r.sol for r in capped}
    assert set(a) == set(m) == {r.rid for r in requests}
    for rid in a:
        assert np.array_equal(a[rid], m[rid]), rid


def test_autotuner_picks_a_candidate(requests):
    print("[gpu] serve autotune", flush=True)
    best, trials = S.autotune(requests[:4], candidates=[S.Params(chunk=3, max_batch=2),
                                                         S.Params(chunk=9, max_batch=4, use_graph=True)])
    assert len(trials) == 2
    assert best.chunk in (3, 9)


def test_init_1drp_on_the_device_equals_the_host_state():
    print("[gpu] init_1drp", flush=True)
    mesh = mg.tri_mesh(9)
    s = Solver(0)
    s.build_mesh(mesh)
    wl, wr = np.array([1.0, 0.1, -0.2, 0.0, 1.0]), np.array([0.125, 0.0, 0.3, 0.0, 0.1])
    s.init_1drp(0.45, wl, wr)
    got = s.get_sol()
    ref = mg.init_1drp(s.elem_coord, 0.45, wl, wr)
    s.close()
    assert np.array_equal(got, ref)


def test_periodic_instances_in_a_batch_equal_solves_alone(requests):
    """Instances on periodic meshes (extended vertex lists) batched with
    ordinary ones."""
    print("[gpu] serve periodic", flush=True)
    from subfv_amd.api import HostMesh
    per = []
    for i, (kind, n) in enumerate((("quad", 10), ("tri", 8))):
        mesh = (mg.quad_mesh if kind == "quad" else mg.tri_mesh)(n, x0=-0.5, y0=-0.5, seed=40 + i)
        r = S.Request(100 + i, kind, n, 0.0, np.array([1.0, 0.2, 0.1, 0.0, 1.0]),
                      np.array([0.3, -0.1, 0.0, 0.0, 0.4]), 0.1 * i, 15)
        r.host = HostMesh(mesh, periodic=True)
        r.boundary_2d = True
        r.cells = mesh.n_elems
        per.append(r)
    reqs = per + list(requests[:3])
    alone, _ = S.serve_plain(reqs, keep=True, realtime=False)
    mixed, _ = S.serve(reqs, S.Params(chunk=6, max_batch=4, use_graph=True), keep=True, realtime=False)
    a = {r.rid: r.sol for r in alone}
    m = {r.rid: r.sol for r in mixed}
    for rid in a:
        assert np.array_equal(a[rid], m[rid]), rid


def test_finished_solves_are_written_in_the_background(requests, tmp_path):
    """on_finish hands each finished solve's snapshot to the asynchronous
    writer; the written tables hold the solves' final states."""
    print("[gpu] serve async output", flush=True)
    from subfv_amd import output as OUT
    wr = OUT.AsyncWriter(workers=2)
    futs = {}

    def on_finish(r, s):
        futs[r.rid] = wr.dat(str(tmp_path / f"req{r.rid}.dat"), OUT.solver_snapshot(s))

    res, _ = S.serve(requests, S.Params(chunk=6, max_batch=3), keep=True, realtime=False, on_finish=on_finish)
    wr.close()
    for r in res:
        assert futs[r.rid].exception() is None
        rows = np.loadtxt(tmp_path / f"req{r.rid}.dat").reshape(-1, 13)
        assert np.array_equal(rows[:, 3:8], r.sol)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_pooled_context_reused_for_a_different_mesh(requests, mode):
    """One context serves every request in turn (capacity 1, no prefetch):
    each new request uploads a mesh of another size into the context the
    batch used before, so the batch must rebuild its tables and graph."""
    print(f"[gpu] serve pooled context reuse mode={mode}", flush=True)
    alone, _ = S.serve_plain(requests, keep=True, realtime=False)
    p = S.Params(chunk=6, max_batch=1, prefetch=0, first=1, use_graph=mode == 1, persistent=mode == 2,
                 connectivity="host" if mode == 2 else "device")
    one, _ = S.serve(requests, p, keep=True, realtime=False)
    a = {r.rid: r.sol for r in alone}
    m = {r.rid: r.sol for r in one}
    assert set(a) == set(m)
    for rid in a:
        assert np.array_equal(a[rid], m[rid]), rid

