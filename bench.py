#!/usr/bin/env python3
"""Benchmark of the electrostatic PIC step on one MI355X.

Workload (the 2-D configuration the headline metric is quoted on): a 2-D plasma
slab, 256 x 256 cells (257^2 nodes), Dirichlet phi = 0 walls, 2^24 electrons +
2^24 ions (m_i = 1836 m_e), Te = 2 eV / Ti = 0.03 eV Maxwellians, absorbing
walls, Debye length = cell size, omega_pe dt = 0.1.  One step = gather + push +
boundary + CIC deposit of both species, deposit merge, Poisson solve to
relative residual 1e-8 (FP64 residual; by default the direct sine-transform
solve, multigrid and PCG timed beside it on the same charge density), E-field,
and the counting sort every `sort_every` steps.  Inputs are generated on the
device (Philox loader), so they are resident in HBM before the timed region.

After the timed step, a reduced learned leg (513^2 nodes, the 4-level
32-256-channel network, 1000 two-stream snapshots instead of 20,000 and
1000 steps and 4-8 rounds of iterate pairs, as at full size; the full size is
`python -m iskra_amd.warmstart`) trains the network on the device as the
coarse solver of the two-level cycle that preconditions a flexible CG
(iskra_ws_solve), then compares, on held-out snapshots, Jacobi-PCG from
zero, from the previous phi and from the network's prediction, multigrid
likewise, and the network-preconditioned solve, all to the same FP64
relative residual 1e-8: the "learned vs plain" part of the metric;
--no-warmstart skips it.  Then the single-context 1-D configurations are
timed (runs_1d: the two-stream instability, 64 cells and 20,000 electrons
over 500 steps, and the RF discharge with MCC, 512 cells and 1 M electrons +
1 M ions over 2000 steps); --no-runs-1d skips them.

Prints one JSON line (see the keys below).  --gpus above 1 exits: this build
runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("particle-steps/s + % peak HBM; Poisson solve ms & iters at 1e-8 residual, "
          "learned vs plain")
PEAK_HBM_GBS = 8000.0
# push bytes per particle-step: read cell, ox, oy, vx, vy and write ox, oy,
# vx, vy (36 B), plus 4 B for the cell of a particle that changed cell
BYTES_PER_PSTEP_2D = 36
BYTES_CELL = 4


def cell_change_fraction(ctx, sids, dt, everies):
    """Share of the pushed particles whose cell changes in one step (the push
    stores a cell only then).  Measured after the timed region on one step
    that sorts no species: the device order, and with it every particle's
    slot, is then the same before and after the step."""
    n = ctx.stats()["steps"]
    while any(e > 0 and (n + 1) % e == 0 for e in everies):
        ctx.step(dt, 1)
        n += 1
    before = [ctx.get_state(s)["cell"] for s in sids]
    ctx.step(dt, 1)
    moved = live = 0
    for s, c0 in zip(sids, before):
        c1 = ctx.get_state(s)["cell"]
        ok = c0 >= 0
        live += int(ok.sum())
        moved += int((c1[ok] != c0[ok]).sum())
    return moved / max(live, 1)


def cpu_baseline(cells, budget_s=10.0):
    """The oracle's FP64 restatement of the reference's per-step particle work
    (grid_to_particle, push_in_cartesian! with B = 0, discard!, particle_to_grid)
    for a bounded sample of electrons on the same grid: on all the host cores
    this process may use (OpenMP, ref_particles_omp), and on one thread (the
    reference is single-threaded)."""
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    allc = cpu_baseline_omp(cells, threads, budget_s)
    one = cpu_baseline_1thread(cells, budget_s)
    allc["single_thread"] = one
    return allc


def cpu_baseline_omp(cells, threads, budget_s, sample=1 << 23, max_steps=100):
    import oracle as O
    from iskra_amd import problems as P
    p = P.slab_2d(cells=cells)
    nx = cells + 1
    dx = p["dx"]
    rng = np.random.default_rng(0)
    L = cells * dx
    x = np.ascontiguousarray(np.concatenate([rng.random(sample) * L, rng.random(sample) * L]))
    v = np.ascontiguousarray(rng.standard_normal(3 * sample) * p["ve"])
    E = np.ascontiguousarray(rng.standard_normal(3 * nx * nx) * 10.0)
    E[2 * nx * nx:] = 0.0
    wg = np.full(sample, p["w0"])
    u = np.zeros(nx * nx)
    qm = -P.QE / P.ME
    lib = O.lib()
    n, steps, psteps = sample, 0, 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < max_steps:
        m = lib.ref_particles_omp(n, O.ptr(x), O.ptr(v), O.ptr(wg), nx, nx, dx, dx, O.ptr(E), qm,
                                  p["dt"], O.ptr(u), threads)
        psteps += n
        steps += 1
        if m != n:  # keep leading dimension = count for the next step
            x = np.ascontiguousarray(np.concatenate([x[:m], x[n:n + m]]))
            v = np.ascontiguousarray(np.concatenate([v[:m], v[n:n + m], v[2 * n:2 * n + m]]))
            wg = wg[:m].copy()
            n = m
    dt = time.perf_counter() - t0
    return dict(value=psteps / dt, unit="particle-steps/s", cores=threads, kind="port",
                sample=(f"{sample} electrons on the {cells}x{cells}-cell grid, {steps} steps of "
                        f"the FP64 reference algorithm (gather, push B=0, discard, CIC deposit) "
                        f"on {threads} OpenMP threads, {dt:.1f} s; Poisson excluded (the "
                        f"reference's dense A\\b at 257^2 nodes is infeasible)"))


def cpu_baseline_1thread(cells, budget_s, sample=1 << 21, max_steps=100):
    import oracle as O
    from iskra_amd import problems as P
    p = P.slab_2d(cells=cells)
    nx = cells + 1
    dx = p["dx"]
    rng = np.random.default_rng(0)
    L = cells * dx
    x = np.ascontiguousarray(np.concatenate([rng.random(sample) * L, rng.random(sample) * L]))
    v = np.ascontiguousarray((rng.standard_normal(3 * sample) * p["ve"]))
    E = np.ascontiguousarray(rng.standard_normal(3 * nx * nx) * 10.0)
    E[2 * nx * nx:] = 0.0
    B = np.zeros(3 * sample)
    wg = np.full(sample, p["w0"])
    ids = np.arange(sample, dtype=np.uint32)
    n = np.array([nx, nx], dtype=np.int64)
    dh = np.array([dx, dx])
    org = np.zeros(2)
    dims = np.array([0, 1], dtype=np.int32)
    lib = O.lib()
    pE = np.zeros(3 * sample)
    u = np.zeros(nx * nx)
    qm = -P.QE / P.ME
    steps = 0
    psteps = 0
    npart = np.array([sample], dtype=np.int64)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < max_steps:
        m = int(npart[0])
        lib.ref_grid_to_particle(2, m, O.ptr(x), nx, nx, dx, dx, O.ptr(E), O.ptr(pE))
        lib.ref_push(m, 2, O.ptr(x), O.ptr(v), O.ptr(pE), O.ptr(B), qm, p["dt"])
        lib.ref_discard(O.ptr(npart), m, 2, O.ptr(x), O.ptr(v), O.ptr(wg), O.ptr(ids), p["w0"],
                        O.ptr(n), O.ptr(dh), O.ptr(org), 2, O.ptr(dims))
        # remaining particles keep leading dimension m; repack for the next step
        k = int(npart[0])
        if k != m:
            x = np.ascontiguousarray(np.concatenate([x[:k], x[m:m + k]]))
            v = np.ascontiguousarray(np.concatenate([v[:k], v[m:m + k], v[2 * m:2 * m + k]]))
            wg = wg[:k].copy()
            ids = ids[:k].copy()
            pE = np.zeros(3 * k)
            B = np.zeros(3 * k)
        lib.ref_particle_to_grid(2, k, O.ptr(x), O.ptr(wg), nx, nx, dx, dx, O.ptr(u))
        steps += 1
        psteps += m
    dt = time.perf_counter() - t0
    return dict(value=psteps / dt, unit="particle-steps/s", cores=1, kind="port",
                sample=(f"{sample} electrons on the {cells}x{cells}-cell grid, {steps} steps of "
                        f"the FP64 reference algorithm (gather, push B=0, discard, CIC deposit), "
                        f"1 host thread, {dt:.1f} s; Poisson excluded (the reference's dense "
                        f"A\\b at 257^2 nodes is infeasible)"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--cells", type=int, default=256)
    ap.add_argument("--ppc", type=int, default=256)
    ap.add_argument("--solver", default="spectral", choices=["spectral", "mg", "mgpcg", "pcg"])
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    # electrons re-sorted about every 12 steps with a 3-cell window margin:
    # 2.6 % faster per step than every 8 steps with 2 cells (the sort costs
    # less than the slightly larger deposit windows add).  The interval used
    # is the largest divisor of --steps not above this, so the timed steps
    # always hold whole sort cycles (--steps 20: every 10); 0 / -1 = library
    # default
    ap.add_argument("--margin", type=int, default=3)
    ap.add_argument("--sort-every", type=int, default=12)
    # ions at 0.03 eV drift ~4e-4 cells per step here (0.04 cells over the
    # configuration's 100 steps), so they keep the order of the initial sort
    # for the whole run (0: not re-sorted; those crossing a tile edge deposit
    # through the fallback grid)
    ap.add_argument("--ion-sort-every", type=int, default=0)
    # and their deposit windows need no margin (-1: the electrons' margin)
    ap.add_argument("--ion-margin", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-warmstart", action="store_true")
    ap.add_argument("--no-runs-1d", action="store_true",
                    help="skip the 1-D two-stream and RF-discharge timings")
    ap.add_argument("--ws-runs", type=int, default=4)
    ap.add_argument("--ws-train-steps", type=int, default=1000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU decomposition is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from iskra_amd import Context, _lib as L, problems as P

    p = P.slab_2d(cells=args.cells, ppc=args.ppc)
    nodes = p["nodes"]
    n_species = p["n_part"]
    # whole electron sort cycles in the timed steps: the largest divisor of
    # --steps not above --sort-every (any window of --steps steps then holds
    # exactly --steps / interval sorts, whatever the warm-up)
    sort_every = args.sort_every
    if sort_every > 0 and args.steps > 0:
        sort_every = max(d for d in range(1, sort_every + 1) if args.steps % d == 0)
    ion_every = args.ion_sort_every if args.ion_sort_every > 0 else 1 << 30
    t_setup = time.perf_counter()
    ctx = Context(0)
    ctx.grid(2, (nodes, nodes), (p["dx"], p["dx"]))
    ctx.tuning(tile=args.tile, chunk=args.chunk, margin=args.margin, sort_every=sort_every)
    tags = np.zeros((nodes, nodes), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
    kinds = dict(spectral=L.POISSON_SPECTRAL, mg=L.POISSON_MG, mgpcg=L.POISSON_MGPCG,
                 pcg=L.POISSON_PCG)
    kind = kinds[args.solver]
    ctx.poisson(kind, p["eps0"], tags.ravel(), rtol=1e-8)
    ctx.tag_value(1, 0.0)
    cap = n_species + 1024
    e = ctx.species(-p["q"], p["me"], p["w0"], cap, track_ids=False)
    i = ctx.species(p["q"], p["mi"], p["w0"], cap, track_ids=False)
    ctx.boundary(e, [L.BND_DISCARD, L.BND_DISCARD])
    ctx.boundary(i, [L.BND_DISCARD, L.BND_DISCARD])
    ctx.species_sort_every(i, ion_every)
    ctx.species_margin(i, args.ion_margin)
    Lx = p["L"]
    ctx.sample_maxwellian(e, n_species, [Lx, Lx], [p["ve"]] * 3, seed=1, stream=0)
    ctx.sample_maxwellian(i, n_species, [Lx, Lx], [p["vi"]] * 3, seed=1, stream=1)
    ctx.sync()
    t_setup = time.perf_counter() - t_setup

    dt = p["dt"]
    if args.warmup:
        ctx.step(dt, args.warmup)
    # timed region: exactly K steps, synchronised on both sides
    ctx.sync()
    s0 = ctx.stats()
    t0 = time.perf_counter()
    ctx.step(dt, args.steps)
    ctx.sync()
    t1 = time.perf_counter()
    s1 = ctx.stats()
    elapsed = t1 - t0
    psteps = s1["particle_steps"] - s0["particle_steps"]
    value = psteps / elapsed
    solves = s1["poisson_solves"] - s0["poisson_solves"]
    iters = s1["poisson_iters"] - s0["poisson_iters"]

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device Philox Maxwellian loader)",
        "config": {
            "workload": f"2-D ES-PIC slab {args.cells}x{args.cells} cells, "
                        f"{n_species} e- + {n_species} ions, Dirichlet walls, "
                        f"{args.solver} Poisson to 1e-8",
            "cells": args.cells,
            "particles_per_species": n_species,
            "poisson": args.solver,
            "tuning": dict(ctx.get_tuning(),
                           ion_sort_every=args.ion_sort_every if args.ion_sort_every > 0
                           else "not re-sorted in the run", ion_margin=args.ion_margin),
            "sorts_in_timed_steps": {"electrons": args.steps // sort_every if sort_every > 0
                                     else 0, "ions": args.steps // ion_every},
        },
        "poisson": {"solver": args.solver, "iters_per_solve": iters / max(solves, 1),
                    "rtol": 1e-8},
        "setup_s": t_setup,
    }

    if not args.no_profile:
        # profiled pass of the same K steps: per-kernel HIP-event durations
        ctx.profile(True)
        ctx.step(dt, args.steps)
        sp = ctx.stats()
        ctx.profile(False)
        push_ms = sp["kernel_ms"]["push"]
        push_calls = sp["kernel_calls"]["push"]
        pp = sp["push_particles"]
        moved = cell_change_fraction(ctx, (e, i), dt, (sort_every, ion_every))
        bytes_pstep = BYTES_PER_PSTEP_2D + BYTES_CELL * moved
        bytes_total = bytes_pstep * pp
        achieved = bytes_total / (push_ms * 1e-3) / 1e9
        traffic = None
        tf = os.path.join(ROOT, "profiles", "traffic_push2d.json")
        if os.path.exists(tf):
            with open(tf) as f:
                t = json.load(f)
            if t.get("cells") == args.cells and t.get("particles_per_species") == n_species:
                traffic = t.get("hbm_bytes_per_launch")
        out["roofline"] = {
            "bound": "hbm",
            "kernel": "k_push_multi<2> (fused gather+push+boundary+deposit, both species per launch)",
            "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
            "bytes_per_particle_step": bytes_pstep,
            "cell_change_fraction": moved,
            "avg_launch_ms": push_ms / max(push_calls, 1),
            "algorithmic_bytes_per_launch": bytes_total / max(push_calls, 1),
        }
        out["kernel_ms_per_step"] = {k: v / args.steps for k, v in sp["kernel_ms"].items()}
        out["poisson"]["ms_per_solve"] = sp["kernel_ms"]["poisson"] / max(
            sp["kernel_calls"]["poisson"], 1)
        nbytes = 1 << 30
        cms = ctx.bench_copy(nbytes, 20, nontemporal=True)
        pms = ctx.bench_copy(nbytes, 5, nontemporal=False)
        out["hbm_copy_gbps"] = 2 * nbytes / (cms * 1e-3) / 1e9
        out["hbm_copy_plain_gbps"] = 2 * nbytes / (pms * 1e-3) / 1e9
        out["roofline"]["frac_of_copy"] = achieved / out["hbm_copy_gbps"]
        # the other solvers on this step's charge density (zero and previous-phi starts)
        phi_prev = ctx.field(L.F_PHI)
        ctx.step(dt, 1)
        rho = ctx.field(L.F_RHO)
        alt = {}
        for name in ("spectral", "mg", "pcg"):
            ctx.poisson(kinds[name], p["eps0"], tags.ravel(), rtol=1e-8, max_iter=100000)
            ctx.tag_value(1, 0.0)
            ctx.set_field(L.F_RHO, rho)
            ctx.solve(L.WARM_ZERO)  # first use
            row = {}
            for start in ("zero", "prev"):
                if start == "prev":
                    ctx.set_field(L.F_PHI, phi_prev)
                ctx.sync()
                t0 = time.perf_counter()
                it, rr = ctx.solve(L.WARM_ZERO if start == "zero" else L.WARM_PREV)
                ctx.sync()
                row[start] = {"iters": it, "ms": 1e3 * (time.perf_counter() - t0), "relres": rr}
            alt[name] = row
        out["poisson"]["solvers_same_rho"] = alt

    ctx.close()
    if not args.no_warmstart:
        from iskra_amd import warmstart as W
        ws = W.experiment(side=512, levels=4, base=32, runs=args.ws_runs, steps_per_run=250,
                          train_steps=args.ws_train_steps, held_steps=41, every=10, log=None)
        pcg = ws["solves"]["pcg"]
        out["learned_warm_start"] = {
            "grid": "513x513 nodes, phi = 0 walls, 2-D two-stream snapshots",
            "reduced": f"{ws['snapshots']} of 20,000 snapshots, {ws['train_steps']} training "
                       "steps (full size: python -m iskra_amd.warmstart)",
            "network": f"{ws['levels']}-level conv encoder-decoder, {ws['base']}-"
                       f"{min(ws['base'] << (ws['levels'] - 1), ws['base'] * 8)} channels, "
                       f"{ws['params']} parameters",
            "objective": ws["objective"], "rounds": ws["rounds"],
            "iterate_pairs": ws["iterate_pairs"],
            "snapshots": ws["snapshots"], "train_steps": ws["train_steps"],
            "train_ms_per_step": ws["train_ms_per_step"],
            "inference_ms": ws["inference_ms"], "heldout_loss": ws["heldout_loss_after"],
            "solves": ws["solves"],
            "learned_converged": pcg["fcg"]["converged"],
            "pcg_iters_plain_over_learned": (pcg["zero"]["iters"] / pcg["fcg"]["iters"]
                                             if pcg["fcg"]["converged"] else None),
            "pcg_ms_plain_over_learned": (pcg["zero"]["ms"] / pcg["fcg"]["ms"]
                                          if pcg["fcg"]["converged"] else None),
            "pcg_iters_plain_over_warm_start": pcg["zero"]["iters"] / pcg["learned"]["iters"],
        }
    if not args.no_runs_1d:
        # the single-context 1-D configurations (not the headline workload):
        # the periodic two-stream instability and the RF discharge with MCC,
        # each timed over whole device calls after a warm-up
        from iskra_amd import runs
        from iskra_amd.ensemble import ccp_context
        out["runs_1d"] = [
            runs.run("two-stream 1-D, 64 cells, 20,000 electrons", runs.two_stream_context, 500,
                     batch=False),
            runs.run("RF discharge 1-D with MCC, 512 cells, 1M e- + 1M ions",
                     lambda: ccp_context(dict(cells=512, particles=1_000_000, volts=200.0,
                                              torr=1.0, seed=1)), 2000, batch=False)]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cells)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
