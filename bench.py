"""Benchmark of the explicit residual path on one MI355X.

Default workload: 1024 x 1024 jittered quads split along seeded diagonals
(2,097,152 one-layer prisms), FP64, second-order reconstruction with
limited gradients (method 4), multi_point nodal-solver fluxes, viscous terms
(mu = 1e-3, Pr = 0.72, point-implicit nodal momentum and heat systems),
SSP-RK3, a seeded vortex initial state.  A "step" is one SSP-RK3 step
(3 stages).
``value`` = cell updates per RK stage per second over the timed steps, with
the state resident in HBM.  One JSON line is printed.

The roofline block times the dominant kernel with HIP events recorded on
the context's stream during the timed region and divides its algorithmic
bytes per launch (mesh/state bytes it must read, intermediates excluded) by
the average launch time: the viscous apply k_diff_apply when the viscous
terms are on (``roofline_node_sweep`` then gives the node sweep's), the node
sweep k_node_flux otherwise.  The CPU baseline times the oracle on the host
cores (one process per core of the GPU's CPU share) on bounded samples of the
same workload.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s per RK stage (% HBM roofline); served solves/s + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(C, V, SF, SE, second_order):
    """Bytes a node sweep must move: state of every cell, sub-face normal,
    area and left/right cell ids, node-sorted sub-element -> cell ids and the
    vertex offsets; the second-order sweep also reads the cell gradients,
    cell centroids and vertex coordinates."""
    b = C * 40 + SF * 40 + SE * 4 + V * 4
    if second_order:
        b += C * 120 + C * 24 + V * 24
    return b


def diff_apply_bytes(C, V, SF, SE, block_doubles):
    """Bytes the viscous apply must move per launch: the packed per-vertex
    blocks (N^-1 S, S~^T N^-1 S, their right-hand sides and the heat-system
    counterparts, built once per mesh), each sub-face's unit normal, area and
    left cell, each slot's cell, sub-element id, sub-element volume, offset
    and sub-face list (three entries for a prism corner), each cell's state
    once, the vertex offsets, and the 18-double slot rows it writes."""
    return block_doubles * 8 + SF * 36 + SE * (4 + 4 + 8 + 4 + 12) + C * 40 + V * 16 + SE * 144


def stage_bytes_m1(C, V, SF, SE):
    """First-order residual + update per stage: U in and out, volume, the
    sub-face data, sub-element -> cell ids and vertex offsets."""
    return C * 88 + SF * 40 + SE * 4 + V * 4


def read_value(summary, key):
    """A figure recorded in a committed rocprofv3 summary under profiles/."""
    path = os.path.join(ROOT, "profiles", f"{summary}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def read_traffic(kernel_tag):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel factor), if a summary exists."""
    return read_value(f"pmc_{kernel_tag}", "traffic_bytes_per_launch")


def _cpu_sample(job):
    """One process of the CPU baseline: the oracle on its own n x n tri mesh;
    returns (cells, wall-clock start, end, status) of the timed run."""
    n, steps, viscous, first_order = job
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    from subfv_amd import meshgen as mg

    om = O.OracleMesh(mg.tri_mesh(n))
    sol = mg.vortex_state(om.elem_coord)
    cfg = O.make_config("multi_point", second_order=not first_order, method=4, cfl=0.4,
                        diffusion=viscous and not first_order, mu=1e-3, prandtl=0.72)
    t0 = time.time()
    st, _, _ = om.run(cfg, sol, steps, 0 if first_order else 1)
    return om.n_elems, t0, time.time(), st


def cpu_baseline(n=192, steps=1, viscous=True, first_order=False, procs=None):
    """Oracle on the host cores: `procs` processes (the CPU share of one GPU,
    16, or fewer if the process may use fewer cores), each on its own n x n
    tri mesh with the timed workload's scheme (second order, SSP-RK3 and the
    same viscous terms, or the inviscid first-order forward-Euler stage);
    value = all processes' cell updates / the wall time from the first start
    to the last end.  Run before anything touches the GPU (the processes are
    started fresh)."""
    import multiprocessing

    if procs is None:
        procs = max(1, min(16, len(os.sched_getaffinity(0))))
    stages = 1 if first_order else 3
    with multiprocessing.get_context("spawn").Pool(procs) as pool:
        res = pool.map(_cpu_sample, [(n, steps, viscous, first_order)] * procs)
    if any(r[3] for r in res):
        raise RuntimeError(f"oracle status {[r[3] for r in res]}")
    wall = max(r[2] for r in res) - min(r[1] for r in res)
    cells = sum(r[0] for r in res)
    one = sorted(r[0] * stages * steps / (r[2] - r[1]) for r in res)[len(res) // 2]
    if first_order:
        what = f"{steps} forward-Euler step(s), first order, multi_point, inviscid"
    else:
        what = (f"{steps} SSP-RK3 step(s), second order, multi_point"
                + (", viscous mu=1e-3 Pr=0.72" if viscous else ", inviscid"))
    return {"value": cells * stages * steps / wall, "unit": "cell-updates/s", "cores": procs, "kind": "port",
            "per_core_value": one,
            "sample": f"{procs} processes, each tri {n}x{n} ({res[0][0]} prisms), {what}; wall {wall:.1f} s"}


def bcsr_bench(n=512, restart=30):
    """The implicit low-Mach system (512^2 periodic quads, 4x4 blocks, Mach
    0.01 vortex, backward Euler at acoustic CFL 5): bcsr_gmres_diag to
    ||r|| <= 1e-10 ||b|| from x = 0, in FP64 and with the SpMV and
    block-Jacobi of each cycle in FP32 (FP64 restarts).  Wall time of the
    device-resident solve (one call: rhs in, x out), SpMV count, the SpMV's
    algorithmic bytes (blocks, column ids, row offsets, x read once, y
    written) and the solution difference against FP64."""
    import numpy as np

    from subfv_amd import lowmach as lm
    from subfv_amd.api import Bcsr, Solver

    sysm = lm.build_system(n)
    m, nnzb = sysm["m"], len(sysm["col"])
    rhs = sysm["rhs"]
    tol = 1e-10 * float(np.linalg.norm(rhs))
    s = Solver(0)
    A = Bcsr(s, m, 4, sysm["row"], sysm["col"], sysm["v"])
    out = {"system": f"periodic quad {n}x{n}, 4x4 blocks, m={m}, nnzb={nnzb}", "restart": restart,
           "tol_rel": 1e-10}
    xs = {}
    for prec, name in ((0, "fp64"), (1, "fp32_inner")):
        A.gmres_diag(rhs, tol, restart, prec)  # warm-up (factors, work arrays)
        t0 = time.perf_counter()
        x, conv, st = A.gmres_diag(rhs, tol, restart, prec)
        el = time.perf_counter() - t0
        xs[name] = x
        vbytes = 4 if prec else 8
        spmv_bytes = nnzb * (16 * vbytes + 4) + (m + 1) * 4 + 2 * m * 32
        out[name] = {"seconds": el, "converged": conv, "restarts": st["restarts"], "spmv": st["spmv"],
                     "rel_residual": st["rho"] / float(np.linalg.norm(rhs)),
                     "spmv_bytes": spmv_bytes,
                     "effective_spmv_GBs_incl_orthogonalisation": st["spmv"] * spmv_bytes / el / 1e9}
    out["fp32_inner_rel_diff_vs_fp64"] = float(np.linalg.norm(xs["fp32_inner"] - xs["fp64"]) /
                                               np.linalg.norm(xs["fp64"]))
    A.close()
    s.close()
    return out


def served_bench(n_req=192, steps=300, rate=1000.0, calib_n=24, cmin=32 ** 2, cmax=512 ** 2):
    """Served mixed-size first-order solves (quad/tri 50/50, 32^2..512^2
    cells log-uniform, seeded discontinuities, `steps` steps each, Poisson
    arrivals at `rate`/s): continuous batching with autotuned parameters
    against the same requests solved one at a time.  The raw meshes are
    generated on the client before the arrivals (as mesh files are read);
    the connectivity build, upload, geometry, initial state, the steps and
    the read-back are inside each request's latency (plain: connectivity
    built on the host, which the autotuner also picks for these sizes)."""
    from subfv_amd import serve as S
    from subfv_amd.api import Solver

    def note(msg):
        print(f"[served] {msg}", file=sys.stderr, flush=True)

    count = [0]

    def progress(r, s):
        count[0] += 1
        if count[0] % 256 == 0:
            note(f"{count[0]} solves done")

    reqs = S.prepare(S.make_requests(n_req, rate=rate, steps=steps, cmin=cmin, cmax=cmax))
    calib = S.prepare(S.make_requests(calib_n, rate=1e9, steps=30, seed=S.SEED + 1, cmin=cmin, cmax=cmax))
    note(f"{n_req} requests generated; autotuning on {calib_n}")
    best, trials = S.autotune(calib, log=lambda t: note(f"trial {t}"))
    note(f"autotuned: {best}")
    # contexts are created before the timed runs, as a server holds them
    one = Solver(0)
    plain_res, plain_ms = S.serve_plain(reqs, solver=one, keep=True, connectivity="host", on_finish=progress)
    plain = S.summarize(plain_res, plain_ms)
    one.close()
    for r in plain_res:  # final states kept as digests (the comparison below is bitwise)
        r.sol = hashlib.sha256(r.sol.tobytes()).hexdigest()
    note(f"plain: {plain}")
    pool = [Solver(0) for _ in range(best.max_batch + best.prefetch)]
    count[0] = 0
    served_res, served_ms = S.serve(reqs, best, pool=pool, keep=True, on_finish=progress)
    served = S.summarize(served_res, served_ms)
    for r in served_res:
        r.sol = hashlib.sha256(r.sol.tobytes()).hexdigest()
    note(f"served: {served}")
    for s in pool:
        s.close()
    # every served solve must equal the same request solved alone, bit for bit
    ref = {r.rid: r.sol for r in plain_res}
    mismatches = sum(int(r.sol != ref[r.rid]) for r in served_res)
    cells = sum(r.cells for r in reqs)
    return {
        "requests": n_req, "steps_per_request": steps, "arrival_rate_per_s": rate,
        "cells_range": [cmin, cmax],
        "mean_cells": cells / n_req, "cell_steps": cells * steps,
        "served": served, "plain": plain,
        "throughput_vs_plain": served["solves_per_s"] / plain["solves_per_s"],
        "served_vs_alone_mismatches": mismatches,
        "autotuned": {"chunk": best.chunk, "max_batch": best.max_batch, "use_graph": best.use_graph,
                      "persistent": best.persistent, "wgs_per_cu": best.wgs_per_cu,
                      "max_cells": best.max_cells, "connectivity": best.connectivity},
        "autotune_trials": trials,
        "note": ("same requests and arrival times for both; connectivity built inside each latency"
                 + ("" if n_req >= 4096 else "; subsample of the 4096-request stream")),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024, help="quads per side before the diagonal split")
    ap.add_argument("--first-order", action="store_true",
                    help="inviscid first-order forward-Euler stage (the M1 configuration)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=192)
    ap.add_argument("--inviscid", action="store_true", help="drop the viscous terms (Euler only)")
    ap.add_argument("--arith", choices=("exact", "fast"), default="exact",
                    help="node-sweep arithmetic: exact (bit-identical to the reference, the default) or fast "
                         "(fused multiply-adds and reciprocal products, within 1e-12)")
    ap.add_argument("--served-requests", type=int, default=192,
                    help="requests in the served-solves measurement (0 skips it)")
    ap.add_argument("--no-bcsr", dest="bcsr", action="store_false",
                    help="skip the implicit low-Mach GMRES measurement")
    ap.add_argument("--bcsr-only", action="store_true",
                    help="only the implicit low-Mach GMRES measurement (one JSON object) and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (calibrates the "
                         "FETCH_SIZE/WRITE_SIZE counters against known bytes) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path is single-GPU; multi-GPU sharding is design-only).", file=sys.stderr)
        sys.exit(2)
    if args.bcsr_only:
        print(json.dumps(bcsr_bench()), flush=True)
        return
    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n}), flush=True)
        return

    cpu = None
    if not args.no_cpu_baseline:
        # before anything touches the GPU: the baseline's processes start fresh
        print("# CPU baseline (oracle on the host cores)", file=sys.stderr, flush=True)
        if args.first_order:
            cpu = cpu_baseline(args.cpu_n, steps=40, viscous=False, first_order=True)
        else:
            cpu = cpu_baseline(args.cpu_n, steps=2, viscous=not args.inviscid)

    from subfv_amd import meshgen as mg
    from subfv_amd.api import Solver, make_config

    t0 = time.perf_counter()
    mesh = mg.tri_mesh(args.n)
    s = Solver(0)
    s.build_mesh(mesh)
    C, V = s.n_elems, s.n_vert
    SF, SE = s.n_sub_faces, s.n_sub_elems
    o2 = not args.first_order
    s.set_config(make_config("multi_point", second_order=o2, method=4, cfl=0.4))
    s.set_arith(args.arith == "fast")
    viscous = not (args.inviscid or args.first_order)  # M1 is the inviscid first-order stage
    if viscous:
        s.set_diffusion(True, mu=1e-3, prandtl=0.72)
    s.set_time(0.0)
    sol0 = mg.vortex_state(s.elem_coord)
    s.put_sol(sol0)
    vol = s.elem_volume
    setup_s = time.perf_counter() - t0
    print(f"# mesh {C} cells, {V} vertices, {SF} sub-faces, {SE} sub-elements; setup {setup_s:.1f}s",
          file=sys.stderr, flush=True)

    scheme = 0 if args.first_order else 1
    stages = 1 if args.first_order else 3
    if args.warmup:
        s.run(args.warmup, scheme)
    s.profile(True)
    t1 = time.perf_counter()
    tf, dtf = s.run(args.steps, scheme)  # ends with a stream synchronisation
    elapsed = time.perf_counter() - t1
    phases = {}
    for ph in ("node_flux", "node_flux_o2", "grad", "gather", "diffusion", "diff_apply", "update"):
        ms, cnt = s.profile_read(ph)
        if cnt:
            phases[ph] = {"ms_total": ms, "launches": cnt, "us_avg": 1e3 * ms / cnt}
    s.profile(False)
    # sanity of the timed run: finite states and mass conserved to round-off
    # (all-wall mesh: sum of volume * rho is invariant)
    import numpy as np

    sol1 = s.get_sol()
    mass0 = float(np.dot(vol, sol0[:, 0]))
    mass1 = float(np.dot(vol, sol1[:, 0]))
    checks = {"finite": bool(np.isfinite(sol1).all()), "mass_drift_rel": abs(mass1 - mass0) / mass0,
              "mass_drift_bound": 1e-13}
    checks["ok"] = checks["finite"] and checks["mass_drift_rel"] <= checks["mass_drift_bound"]

    value = C * stages * args.steps / elapsed
    # the same steps again in the other node-sweep arithmetic (exact <-> fast),
    # continuing from the timed state: reported beside `value`, not in it
    other = "fast" if args.arith == "exact" else "exact"
    s.set_arith(other == "fast")
    s.run(1, scheme)
    s.profile(True)
    t2 = time.perf_counter()
    s.run(args.steps, scheme)
    el2 = time.perf_counter() - t2
    sw2 = [s.profile_read(p) for p in ("node_flux", "node_flux_o2")]
    s.profile(False)
    s.set_arith(args.arith == "fast")
    other_arith = {"node_sweep_arith": other, "value": C * stages * args.steps / el2,
                   "us_per_stage": 1e6 * el2 / (stages * args.steps),
                   "node_sweep_us_per_launch": 1e3 * sum(m for m, _ in sw2) / max(1, sum(c for _, c in sw2))}
    # node sweep k_node_flux, launched once per stage for the first-order
    # residual and, in second order, once more for the reconstructed one; its
    # roofline averages over both launch kinds
    sweeps = [(p, p == "node_flux_o2") for p in ("node_flux", "node_flux_o2") if p in phases]
    n_launch = sum(phases[p]["launches"] for p, _ in sweeps)
    sweep_us = 1e3 * sum(phases[p]["ms_total"] for p, _ in sweeps) / n_launch
    abytes = sum(phases[p]["launches"] * algorithmic_bytes(C, V, SF, SE, o2f) for p, o2f in sweeps) / n_launch
    achieved = abytes / (sweep_us * 1e-6) / 1e9
    tag = f"tri{args.n}_" + ("o1_fe" if args.first_order else "o2_rk3")
    traffic = read_traffic(f"k_node_flux_{tag}")
    node_roof = {
        "bound": "hbm", "kernel": "k_node_flux<multi_point>", "achieved": achieved, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        "algorithmic_bytes_per_launch": abytes, "us_per_launch": sweep_us, "launches": n_launch,
        "valu_issue_frac": read_value(f"valu_k_node_flux_{tag}", "valu_issue_frac"),
    }
    roofline, extra = node_roof, {}
    if "diff_apply" in phases:
        # with the viscous terms the dominant kernel is the viscous apply
        # (longest launch, and more time per stage than both node sweeps)
        da = phases["diff_apply"]
        dbytes = diff_apply_bytes(C, V, SF, SE, s.count("diff_block_doubles"))
        dach = dbytes / (da["us_avg"] * 1e-6) / 1e9
        roofline = {
            "bound": "hbm", "kernel": "k_diff_apply", "achieved": dach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dach / HBM_PEAK_GBS, "traffic": read_traffic(f"k_diff_apply_{tag}"),
            "algorithmic_bytes_per_launch": dbytes, "us_per_launch": da["us_avg"], "launches": da["launches"],
        }
        extra = {"roofline_node_sweep": node_roof}
    stage_us = 1e6 * elapsed / (args.steps * stages)
    m1 = stage_bytes_m1(C, V, SF, SE)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"tri {args.n}x{args.n} jittered, {C} prisms, multi_point, "
                         + ("first order, forward Euler" if args.first_order else
                            "second order (method 4 limited gradients), SSP-RK3")
                         + (", viscous mu=1e-3 Pr=0.72" if viscous else ", inviscid")),
            "cells": C, "stages_per_step": stages, "viscous": viscous, "mu": 1e-3 if viscous else 0.0,
            "prandtl": 0.72 if viscous else None, "node_sweep_arith": args.arith,
        },
        "roofline": roofline,
        **extra,
        "stage": {
            "us_per_stage": stage_us,
            "m1_bytes": m1,
            "m1_frac_of_peak": m1 / (stage_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
            "phases_us_avg": {k: v["us_avg"] for k, v in phases.items()},
        },
        "other_arith": other_arith,
        "t_final": tf,
        "checks": checks,
    }
    s.close()
    print(f"# timed stage run: {1e6 * elapsed / (stages * args.steps):.1f} us per stage", file=sys.stderr, flush=True)
    if args.served_requests > 0 and not args.first_order:
        out["served"] = served_bench(args.served_requests)
        # the same loop on small meshes (32^2 .. 96^2 cells), where one solve
        # alone cannot fill the GPU and launches dominate the plain path
        out["served_small"] = served_bench(2 * args.served_requests, rate=4000.0, calib_n=96, cmax=96 ** 2)
    if args.bcsr and not args.first_order:
        print("# implicit low-Mach GMRES", file=sys.stderr, flush=True)
        out["bcsr_gmres"] = bcsr_bench()
    if cpu is not None:
        # samples of about 10-15 s of work per process
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    if not checks["ok"]:
        print(f"bench: timed run failed its checks: {checks}", file=sys.stderr, flush=True)
        sys.exit(1)


if __name__ == "__main__":
    main()
