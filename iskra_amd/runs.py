"""Timed runs of the single-context 1-D configurations.

- The periodic two-stream instability (ts_unstable.jl / 10_two_streams.jl
  structure): 64 cells, 20,000 electrons in two interleaved beams, immobile
  ion background, direct periodic Poisson solve, 500 steps.
- The capacitively coupled RF discharge with MCC (11_rf_discharge.jl
  structure): 512 cells, 2.5 cm gap, 1 M electrons + 1 M ions, 200 V at
  13.56 MHz, 1 Torr, electron elastic / excitation / ionisation and ion
  elastic collisions, 2000 steps.

Each runs eagerly through iskra_step, and (when the batch path supports the
configuration) on a second, identical context as a one-member batch, whose
sort cycles are replayed hipGraphs.  Both contexts compute the same bits, so
the final particle counts must agree.

    python -m iskra_amd.runs [--out FILE]
"""
import argparse
import json
import time

import numpy as np

from . import _lib as L
from . import problems as P
from ._lib import Batch, Context
from .ensemble import ccp_context


def two_stream_context():
    p = P.two_stream_1d()
    ctx = Context()
    ctx.grid(1, (p["nodes"],), (p["dx"],))
    n = len(p["x"])
    e = ctx.species(p["q"], p["m"], p["w0"], n, track_ids=True)
    ctx.boundary(e, [L.BND_WRAP])
    ctx.upload(e, p["x"], p["v"])
    ctx.set_field(L.F_BG, np.full(p["nodes"], p["bg"]))
    ctx.poisson(L.POISSON_DIRECT, p["eps0"], None, periodic_x=True)
    ctx.sync()
    return ctx, p["dt"]


def _timed(step, sync, nsteps):
    sync()
    t0 = time.perf_counter()
    step(nsteps)
    sync()
    return time.perf_counter() - t0


def run(name, make, nsteps, warmup=8, batch=True):
    out = {"config": name, "steps": nsteps}
    ctx, dt = make()
    ctx.step(dt, warmup)
    s0 = ctx.stats()["particle_steps"]
    t = _timed(lambda k: ctx.step(dt, k), ctx.sync, nsteps)
    ps = ctx.stats()["particle_steps"] - s0
    out["eager"] = {"s": t, "ms_per_step": 1e3 * t / nsteps, "particle_steps_per_s": ps / t}
    counts = [ctx.count(s) for s in range(2) if s < _nspecies(ctx)]
    if not batch:
        out["particles_final"] = counts
        ctx.close()
        return out
    ctx2, _ = make()
    try:
        b = Batch([ctx2])
    except L.IskraError as ex:
        out["batch"] = {"unsupported": str(ex)}
    else:
        b.step(dt, warmup)
        s0 = ctx2.stats()["particle_steps"]
        t = _timed(lambda k: b.step(dt, k), ctx2.sync, nsteps)
        ps = ctx2.stats()["particle_steps"] - s0
        out["batch"] = {"s": t, "ms_per_step": 1e3 * t / nsteps, "particle_steps_per_s": ps / t,
                        "graphs": b.graphs_launched()}
        counts2 = [ctx2.count(s) for s in range(len(counts))]
        out["same_counts"] = counts == counts2
        b.close()
    out["particles_final"] = counts
    ctx.close()
    ctx2.close()
    return out


def _nspecies(ctx):
    n = 0
    while True:
        try:
            ctx.count(n)
        except L.IskraError:
            return n
        n += 1


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=None)
    ap.add_argument("--no-batch", action="store_true",
                    help="time the single-context path only (no one-member batch)")
    ap.add_argument("--steps", type=int, default=0, help="steps of each run (0: 500 / 2000)")
    args = ap.parse_args(argv)
    b = not args.no_batch
    res = [run("two-stream 1-D, 64 cells, 20,000 electrons", two_stream_context,
               args.steps or 500, batch=b),
           run("RF discharge 1-D with MCC, 512 cells, 1M e- + 1M ions",
               lambda: ccp_context(dict(cells=512, particles=1_000_000, volts=200.0, torr=1.0,
                                        seed=1)), args.steps or 2000, batch=b)]
    for r in res:
        print(json.dumps(r), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
