"""The served-request stream (config-5 shape) on the CPU: seeded and
deterministic, quad/tri about 50/50, cell counts log-uniform in
[32^2, 512^2], states inside the stated ranges, Poisson arrivals at the
requested rate; client-side preparation generates each mesh (and, when
asked, builds its connectivity)."""
import numpy as np
import pytest

from subfv_amd import serve as S


def test_stream_is_seeded_and_in_range():
    a = S.make_requests(2000, rate=50.0)
    b = S.make_requests(2000, rate=50.0)
    assert [(r.kind, r.n, r.arrival, r.x_split) for r in a] == [(r.kind, r.n, r.arrival, r.x_split) for r in b]
    kinds = np.array([r.kind == "quad" for r in a])
    assert 0.45 < kinds.mean() < 0.55
    cells = np.array([r.n * r.n * (1 if r.kind == "quad" else 2) for r in a], float)
    assert cells.min() >= 0.9 * 32 ** 2 and cells.max() <= 1.1 * 512 ** 2
    lc = np.log(cells)
    # log-uniform: the quartiles of log(cells) sit near the quartiles of the interval
    lo, hi = np.log(32 ** 2), np.log(512 ** 2)
    q = np.percentile(lc, [25, 50, 75])
    assert np.allclose(q, lo + (hi - lo) * np.array([0.25, 0.5, 0.75]), atol=0.35)
    for r in a:
        for w in (r.wl, r.wr):
            assert 0.1 <= w[0] <= 1.0 and 0.1 <= w[4] <= 1.0
            assert np.hypot(w[1], w[2]) <= 0.5 + 1e-12
        assert 0.3 <= r.x_split <= 0.7
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0


def test_prepare_builds_connectivity():
    reqs = S.prepare(S.make_requests(6, rate=10.0, cmin=8 ** 2, cmax=20 ** 2), workers=3, connectivity=True)
    for r in reqs:
        assert r.cells == r.n * r.n * (1 if r.kind == "quad" else 2)
        assert r.host.geti("elem_vert").size > 0
    raw = S.prepare(S.make_requests(2, rate=10.0, cmin=8 ** 2, cmax=20 ** 2), workers=2)
    for r in raw:
        assert r.host is None and r.mesh.n_elems == r.cells


def test_requests_from_mesh_files(tmp_path):
    """A request may name a msh 2.2 file; its connectivity equals that of
    the generated mesh the file was written from."""
    from subfv_amd import meshgen as mg
    reqs = S.make_requests(2, rate=10.0, cmin=8 ** 2, cmax=12 ** 2)
    gen = S.prepare([S.Request(**{**r.__dict__}) for r in reqs], workers=2, connectivity=True)
    for r in reqs:
        mesh = mg.quad_mesh(r.n, seed=S.SEED + r.rid) if r.kind == "quad" else mg.tri_mesh(r.n, seed=S.SEED + r.rid)
        r.mesh_path = str(tmp_path / f"r{r.rid}.msh")
        mg.write_msh(mesh, r.mesh_path)
    from_files = S.prepare(reqs, workers=2, connectivity=True)
    for a, b in zip(gen, from_files):
        assert a.cells == b.cells
        for name in ("elem_vert", "sub_face_left_elem", "face_right", "vert_sub_face"):
            assert np.array_equal(a.host.geti(name), b.host.geti(name)), name


def test_a_request_beyond_the_budget_is_refused():
    """Admission control: a request that cannot fit the HBM budget even on
    an idle server is an error, not a wait (no device call is made)."""
    import pytest
    reqs = S.make_requests(1, rate=10.0, cmin=8 ** 2, cmax=8 ** 2)
    reqs[0].cells = 64
    with pytest.raises(S.SubfvError, match="HBM budget"):
        S.serve(reqs, S.Params(budget=1024), realtime=False, pool=[])


def test_persistent_kernel_refuses_device_connectivity():
    """The persistent kernel is served with host-built connectivity only:
    refused before any device work."""
    with pytest.raises(ValueError, match="host only"):
        S.serve([], S.Params(persistent=True, connectivity="device"), realtime=False)
