"""Ensembles of independent 1-D RF discharges advanced together (iskra_batch_*).

Each instance is an ordinary context set up like the 1-D capacitively
coupled discharge (problem/11_rf_discharge.jl structure: electrons and ions
in a gap between a driven and a grounded electrode, discard walls, electron
and ion Monte Carlo collisions with a neutral gas, direct Poisson solve) with
its own grid size, particle counts, voltage, pressure and seed.  The batch
advances all of them with one launch per kernel type per step, in replayed
hipGraphs; every instance computes exactly what it would compute alone.

The tile and chunk shapes of the instances come from `autotune`, which
measures candidate shapes per grid-size class on a throw-away copy of the
ensemble.
"""
import time

import numpy as np

from . import _lib as L
from . import problems as P
from ._lib import Batch, Context


def draw_instances(count=256, seed=0, cells=(128, 1024), particles=(1e4, 5e5),
                   volts=(100.0, 400.0), torr=(0.1, 2.0)):
    """Instance parameters: cells uniform in [cells], particles per species
    log-uniform in [particles], RF amplitude and pressure uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        out.append(dict(
            cells=int(rng.integers(cells[0], cells[1] + 1)),
            particles=int(np.exp(rng.uniform(np.log(particles[0]), np.log(particles[1])))),
            volts=float(rng.uniform(*volts)),
            torr=float(rng.uniform(*torr)),
            seed=int(seed * 100003 + k + 1)))
    return out


def ccp_context(inst, tile=0, chunk=0, sort_every=16, gap=0.025, freq=13.56e6, Te_eV=2.5,
                Tg=300.0, n0=1e15, cap_factor=2.0, device=0):
    """One instance as an ordinary context.  Returns (ctx, dt).  Both species
    are re-sorted every `sort_every` steps (16: 3 % faster per batched step
    than 8 on the 256-member ensemble, the same as 32)."""
    cells = inst["cells"]
    npart = inst["particles"]
    p = P.ccp_1d(cells=cells, gap=gap, electrons=npart, ions=npart, Te=Te_eV,
                 pressure_torr=inst["torr"], Tg=Tg, volts=inst["volts"], freq=freq, n0=n0)
    ctx = Context(device)
    ctx.grid(1, (cells + 1,), (p["dx"],))
    ctx.tuning(tile=tile, chunk=chunk, sort_every=sort_every)
    ctx.rng(inst["seed"])
    cap = int(cap_factor * npart) + 1024
    e = ctx.species(-p["q"], p["me"], p["w0"], cap, track_ids=True)
    i = ctx.species(p["q"], p["mi"], p["w0"], cap, track_ids=True)
    ctx.boundary(e, [L.BND_DISCARD])
    ctx.boundary(i, [L.BND_DISCARD])
    ctx.sample_maxwellian(e, npart, [gap], [p["ve"]] * 3, seed=inst["seed"], stream=0)
    ctx.sample_maxwellian(i, npart, [gap], [p["vi"]] * 3, seed=inst["seed"], stream=1)
    xs = p["xs"]
    me = ctx.mcc(e, [L.COLL_ELASTIC_ISOTROPIC, L.COLL_EXCITATION, L.COLL_IONIZATION],
                 xs["energies"], xs["electron"], ion=i, target_m=p["mi"], target_q=0.0,
                 target_T=Tg)
    mi = ctx.mcc(i, [L.COLL_ELASTIC_BACKWARD], [0.0], xs["ion"], ion=-1, target_m=p["mi"],
                 target_q=0.0, target_T=Tg)
    tn = np.full(cells + 1, p["ng"])
    ctx.set_field(L.F_TARGET, tn, index=me)
    ctx.set_field(L.F_TARGET, tn, index=mi)
    tags = np.zeros(cells + 1, np.int8)
    tags[0], tags[-1] = 1, 2
    ctx.poisson(L.POISSON_DIRECT, p["eps0"], tags)
    ctx.tag_waveform(1, 0.0, inst["volts"], freq, 0.0)
    ctx.sync()
    return ctx, p["dt"]


def size_class(cells, bounds=(384, 640)):
    return int(np.searchsorted(bounds, cells, side="right"))


def build(instances, tunings=None, **kw):
    """Contexts of all instances (tunings: list of (tile, chunk) or None)."""
    ctxs, dt = [], None
    for k, inst in enumerate(instances):
        t, ch = (0, 0) if tunings is None else tunings[k]
        c, dt = ccp_context(inst, tile=t, chunk=ch, **kw)
        ctxs.append(c)
    return ctxs, dt


def close_all(ctxs, batch=None):
    if batch is not None:
        batch.close()
    for c in ctxs:
        c.close()


def time_batch(ctxs, dt, steps, warmup=0):
    """Wall time of `steps` batched steps (after `warmup`); returns
    (seconds, particle-steps)."""
    b = Batch(ctxs)
    try:
        if warmup:
            b.step(dt, warmup)
        ps0 = sum(c.stats()["particle_steps"] for c in ctxs)
        t0 = time.perf_counter()
        b.step(dt, steps)
        t = time.perf_counter() - t0
        ps1 = sum(c.stats()["particle_steps"] for c in ctxs)
    finally:
        b.close()
    return t, ps1 - ps0


def autotune(instances, candidates=((32, 1024), (64, 2048), (64, 4096), (128, 4096)),
             steps=16, log=None, **kw):
    """Per grid-size class, the (tile, chunk) with the shortest batched step:
    classes are tuned one after the other (coordinate descent), the other
    classes keeping their current best; each measurement builds a fresh copy
    of the ensemble and times `steps` batched steps after one sort cycle of
    warm-up.  Returns the per-instance tunings and the timing table."""
    classes = sorted({size_class(i["cells"]) for i in instances})
    best = {c: candidates[0] for c in classes}
    table = []
    for cl in classes:
        timings = {}
        for cand in candidates:
            trial = {**best, cl: cand}
            tun = [trial[size_class(i["cells"])] for i in instances]
            ctxs, dt = build(instances, tun, **kw)
            try:
                t, ps = time_batch(ctxs, dt, steps, warmup=16)
            finally:
                close_all(ctxs)
            timings[cand] = t
            table.append(dict(cls=cl, tile=cand[0], chunk=cand[1], s=t, psteps=ps))
            if log:
                log(f"class {cl} tile {cand[0]} chunk {cand[1]}: {1e3 * t / steps:.3f} ms/step")
        best[cl] = min(timings, key=timings.get)
    return [best[size_class(i["cells"])] for i in instances], table


def experiment(count=256, steps=500, seed=0, tune_subset=48, single_steps=48, check=4,
               log=print):
    """Autotune, then advance `count` instances `steps` steps as one batch;
    time the same instances stepped one by one (iskra_step per context) for
    `single_steps` steps; check `check` members bit for bit against single
    runs.  Returns the measured figures."""
    import json as _json
    t0 = time.perf_counter()

    def say(m):
        if log is not None:
            log(f"[{time.perf_counter() - t0:7.1f} s] {m}")

    insts = draw_instances(count, seed=seed)
    sub = insts[:tune_subset]
    tun_sub, table = autotune(sub, candidates=((32, 1024), (64, 2048), (128, 4096)), steps=16,
                              log=say)
    best = {size_class(i["cells"]): t for i, t in zip(sub, tun_sub)}
    tunings = [best.get(size_class(i["cells"]), (64, 2048)) for i in insts]
    say(f"autotuned shapes per size class: {best}")
    t = time.perf_counter()
    ctxs, dt = build(insts, tunings)
    setup_s = time.perf_counter() - t
    np0 = sum(c.count(0) + c.count(1) for c in ctxs)
    say(f"built {count} instances ({np0} particles) in {setup_s:.1f} s")
    b = Batch(ctxs)
    res = {}
    try:
        b.step(dt, 16)
        ps0 = sum(c.stats()["particle_steps"] for c in ctxs)
        t = time.perf_counter()
        b.step(dt, steps)
        tb = time.perf_counter() - t
        ps1 = sum(c.stats()["particle_steps"] for c in ctxs)
        res.update(batch_s=tb, batch_steps=steps, batch_psteps=ps1 - ps0,
                   batch_rate=(ps1 - ps0) / tb, batch_ms_per_step=1e3 * tb / steps,
                   graphs=b.graphs_launched())
        say(f"batch: {steps} steps in {tb:.2f} s, {(ps1 - ps0) / tb:.3e} particle-steps/s")
        # the same contexts stepped one by one
        ps0 = sum(c.stats()["particle_steps"] for c in ctxs)
        t = time.perf_counter()
        for c in ctxs:
            c.step(dt, single_steps)
        ts = time.perf_counter() - t
        ps1 = sum(c.stats()["particle_steps"] for c in ctxs)
        res.update(single_s=ts, single_steps=single_steps, single_psteps=ps1 - ps0,
                   single_rate=(ps1 - ps0) / ts)
        say(f"one by one: {single_steps} steps each in {ts:.2f} s, "
            f"{(ps1 - ps0) / ts:.3e} particle-steps/s")
    finally:
        b.close()
    # bitwise spot check: a fresh batch of the first 8 instances against
    # fresh single runs of its first `check` members
    ok = []
    m = min(8, count)
    fresh, _ = build(insts[:m], tunings[:m])
    bb = Batch(fresh)
    bb.step(dt, 48)
    bb.close()
    for k in range(min(check, m)):
        cs, _ = ccp_context(insts[k], tile=tunings[k][0], chunk=tunings[k][1])
        cs.step(dt, 48)
        cb = fresh[k]
        same = all(np.array_equal(cb.get_state(s)[f].view(np.uint32),
                                  cs.get_state(s)[f].view(np.uint32))
                   for s in (0, 1) for f in ("cell", "ox", "vx", "vy", "vz", "id"))
        same = same and np.array_equal(cb.field(L.F_PHI), cs.field(L.F_PHI))
        ok.append(bool(same))
        cs.close()
    close_all(fresh)
    say(f"bitwise spot checks: {ok}")
    res.update(instances=count, particles_initial=int(np0), setup_s=setup_s,
               tunings_by_class={str(k): list(v) for k, v in best.items()}, autotune=table,
               bitwise_equal=ok,
               cells_range=[min(i["cells"] for i in insts), max(i["cells"] for i in insts)],
               particles_range=[min(i["particles"] for i in insts),
                                max(i["particles"] for i in insts)])
    close_all(ctxs)
    say(_json.dumps({k: v for k, v in res.items() if k != "autotune"}))
    return res


def main(argv=None):
    import argparse
    import json as _json
    ap = argparse.ArgumentParser(description="batched ensemble of 1-D RF discharges")
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    res = experiment(a.count, a.steps, a.seed, log=lambda m: print(m, flush=True))
    if a.out:
        with open(a.out, "w") as f:
            f.write(_json.dumps(res) + "\n")
    print("closed", flush=True)


if __name__ == "__main__":
    main()
