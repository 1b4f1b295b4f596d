"""Learned warm start for the 2-D Poisson solve: host side.

The network, its training and its inference run in libiskra_hip.so
(iskra_ws_*, hand-written BF16 MFMA kernels); this module initialises
parameters, generates (rho, phi) snapshots with the device simulator, trains,
and measures what the trained network gains: Poisson iterations and time to
relative residual rtol from a zero start, from the previous step's phi and
from the network's prediction, with the same solver and the same stopping
rule (the solver always iterates to rtol, so the warm start cannot change the
answer beyond that tolerance).  Trained with the energy objective, the same
network also preconditions every iteration of a flexible CG
(iskra_ws_solve), measured beside Jacobi-PCG to the same FP64 residual.

Snapshot source: a 2-D two-stream plasma (the 1-D two-stream instability of
the 64-cell configuration, set on a side x side grid with phi = 0 walls):
electrons in two counter-streaming beams along a random direction with small
random multi-mode density perturbations, an immobile neutralising
background, normalised units (eps0 = 1, omega_p = 1, dx = 1).
"""
import json
import time

import numpy as np

from . import _lib as L
from ._lib import Context, WarmStart

WS_IN, WS_OUT = 8, 16


def _ch(base, l):
    """Channels of level l: base * 2^l, held at base * 8 from level 3 on."""
    return base << min(l, 3)


def layer_shapes(levels, base):
    """(cin, cout) of each convolution, in parameter order."""
    out = []
    for l in range(levels):
        out.append((WS_IN if l == 0 else _ch(base, l - 1), _ch(base, l)))
        out.append((_ch(base, l), _ch(base, l)))
    for l in range(levels - 2, -1, -1):
        out.append((_ch(base, l + 1), _ch(base, l)))
        out.append((_ch(base, l), _ch(base, l)))
    out.append((base, WS_OUT))
    return out


def layer_slices(levels, base):
    """(offset, count, cin, cout) of each layer in the parameter vector."""
    res, off = [], 0
    for cin, cout in layer_shapes(levels, base):
        n = 9 * cin * cout
        res.append((off, n, cin, cout))
        off += n
    return res


def he_init(levels, base, seed=0):
    """He-normal initialisation ([cout][3x3][cin] per layer).  Padded
    channels (inputs 1..7 of the first layer, outputs 1..15 of the last) are
    zero; the output layer starts small."""
    rng = np.random.default_rng(seed)
    parts = []
    shapes = layer_shapes(levels, base)
    for k, (cin, cout) in enumerate(shapes):
        w = np.zeros((cout, 9, cin), dtype=np.float32)
        if k == 0:
            w[:, :, 0] = rng.standard_normal((cout, 9)) * np.sqrt(2.0 / 9.0)
        elif k == len(shapes) - 1:
            w[0] = rng.standard_normal((9, cin)) * np.sqrt(1.0 / (9.0 * cin))
        else:
            w[:] = rng.standard_normal((cout, 9, cin)) * np.sqrt(2.0 / (9.0 * cin))
        parts.append(w.ravel())
    return np.concatenate(parts).astype(np.float32)


# ---------------------------------------------------------------------------
# snapshots

def two_stream_setup(ctx, side, ppc=16, seed=0, v0=None, amp=0.02, modes=4, solver=None,
                     rtol=1e-10):
    """Grid, solver and one electron species on `ctx` for a 2-D two-stream
    run; returns the species id and the time step."""
    rng = np.random.default_rng(seed)
    n = side + 1
    ctx.grid(2, (n, n), (1.0, 1.0))
    tags = np.zeros((n, n), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
    ctx.poisson(L.POISSON_MG if solver is None else solver, 1.0, tags.ravel(), rtol=rtol)
    ctx.tag_value(1, 0.0)
    npart = ppc * side * side
    w0 = float(side * side) / npart          # n0 = 1
    sid = ctx.species(-1.0, 1.0, w0, npart + 16, track_ids=False)
    ctx.boundary(sid, [L.BND_DISCARD, L.BND_DISCARD])
    if v0 is None:
        v0 = 0.9 * (side / 8.0) / (2 * np.pi)  # most unstable wavelength ~ side / 8
    x = rng.random((npart, 2)) * side
    # multi-mode density perturbation by displacement along random directions
    for _ in range(modes):
        k = 2 * np.pi * rng.integers(1, 9, size=2) / side
        ph = rng.random() * 2 * np.pi
        a = amp * rng.random() / np.linalg.norm(k)
        arg = x @ k + ph
        x += (a * np.sin(arg))[:, None] * (k / np.linalg.norm(k))[None, :]
    x = np.clip(x, 1e-6, side - 1e-6)
    ang = rng.random() * 2 * np.pi
    d = np.array([np.cos(ang), np.sin(ang), 0.0])
    v = np.zeros((npart, 3))
    sgn = np.where(np.arange(npart) % 2 == 0, 1.0, -1.0)
    v += (sgn * v0)[:, None] * d[None, :]
    v += rng.standard_normal((npart, 3)) * (0.05 * v0)
    ctx.upload(sid, x, v)
    ctx.set_field(L.F_BG, np.full(n * n, 1.0))
    return sid, 0.1


def capture_two_stream(ws, runs, steps, skip=20, ppc=16, seed=0, record=None, record_every=1):
    """Run `runs` two-stream simulations on the network's context, `steps`
    steps each, capturing (f, phi) after every step beyond `skip`.  `record`
    (list, optional) receives (rho, phi_prev, phi) of every record_every-th
    captured step."""
    ctx = ws.ctx
    for r in range(runs):
        sid, dt = two_stream_setup(ctx, ws.side, ppc=ppc, seed=seed * 1000 + r)
        ctx.step(dt, skip)
        prev = ctx.field(L.F_PHI) if record is not None else None
        for s in range(steps):
            ctx.step(dt, 1)
            ws.capture()
            if record is not None and (record_every == 1 or (s + 1) % record_every == 0):
                phi = ctx.field(L.F_PHI)
                record.append((ctx.field(L.F_RHO), prev, phi))
                prev = phi
            elif record is not None and (s + 2) % record_every == 0:
                prev = ctx.field(L.F_PHI)


def capture_iterate_pairs(ws, rhos, iters=(1, 2, 4, 8, 16), omega=0.8, sweeps=2):
    """For each rho of `rhos`: the exact phi by the spectral direct solve,
    then the pairs (r_k, phi - x_k) of iskra_ws_solve's iterates after each
    count of `iters` appended to the dataset (the residuals the network meets
    inside the iteration, whose distribution depends on the network itself).
    Returns the number of pairs appended."""
    ctx = ws.ctx
    n = ws.side + 1
    tags = np.zeros((n, n), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
    ctx.poisson(L.POISSON_SPECTRAL, 1.0, tags.ravel(), rtol=1e-11)
    ctx.tag_value(1, 0.0)
    got = 0
    for rho in rhos:
        ctx.set_field(L.F_RHO, rho)
        ctx.solve(L.WARM_ZERO)
        got += ws.capture_iterates(iters, omega, sweeps)
    return got


def train(ws, first, count, steps, batch=8, lr=1e-3, lr_end=1e-5, chunk=50, seed=1, log=None):
    """Adam with a cosine learning-rate decay, `chunk` steps per device call.
    The output gain is set from the dataset so that targets have unit RMS."""
    gain = ws.dataset_rms(first, count)
    ws.set_gain(gain)
    losses = []
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        k = min(chunk, steps - done)
        frac = (done + 0.5 * k) / max(steps, 1)
        rate = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * frac))
        losses.extend(ws.train(first, count, batch, k, rate, seed=seed + done))
        done += k
        if log is not None:
            log(f"train {done}/{steps} loss {np.mean(losses[-k:]):.4e} "
                f"({time.perf_counter() - t0:.1f} s)")
    return np.asarray(losses), gain


def solve_table(ctx, ws, snapshots, kinds=(L.POISSON_PCG, L.POISSON_MG), rtol=1e-8,
                fcg_omega=None, fcg_sweeps=1, fcg_max_iter=3000):
    """For each solver kind (configured once, then one untimed solve), the
    iterations and wall ms of solving every snapshot (rho, phi_prev) from a
    zero start, from phi_prev and from the network's prediction (inference
    time included).  With fcg_omega set, the PCG row also holds "fcg": the
    network inside the iteration (iskra_ws_solve from a zero start,
    fcg_sweeps Jacobi sweeps of weight fcg_omega after the network).  Every
    solve must reach rtol, except the network-preconditioned one, which stops
    after fcg_max_iter iterations (its row then shows a relres above rtol).
    Returns
    {kind: {start: {"iters": [...], "ms": [...], "relres": [...]}}}."""
    n = ws.side + 1
    tags = np.zeros((n, n), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
    res = {}
    for kind in kinds:
        ctx.poisson(kind, 1.0, tags.ravel(), rtol=rtol, max_iter=100000)
        ctx.tag_value(1, 0.0)
        ctx.set_field(L.F_RHO, snapshots[0][0])
        ctx.solve(L.WARM_ZERO)
        starts = ["zero", "prev", "learned"]
        if fcg_omega is not None and kind == L.POISSON_PCG:
            starts.append("fcg")
            # untimed: first-call allocations
            ws.solve(L.WARM_ZERO, fcg_omega, fcg_sweeps, max_iter=fcg_max_iter,
                     raise_unconverged=False)
        row = {s: {"iters": [], "ms": [], "relres": []} for s in starts}
        for rho, prev in snapshots:
            ctx.set_field(L.F_RHO, rho)
            for start in starts:
                if start == "prev":
                    ctx.set_field(L.F_PHI, prev)
                ctx.sync()
                t0 = time.perf_counter()
                if start == "zero":
                    it, rr = ctx.solve(L.WARM_ZERO)
                elif start == "fcg":
                    it, rr = ws.solve(L.WARM_ZERO, fcg_omega, fcg_sweeps,
                                      max_iter=fcg_max_iter, raise_unconverged=False)
                else:
                    if start == "learned":
                        ws.predict()
                    it, rr = ctx.solve(L.WARM_PREV)
                ctx.sync()
                row[start]["iters"].append(it)
                row[start]["ms"].append(1e3 * (time.perf_counter() - t0))
                row[start]["relres"].append(rr)
        res[kind] = row
    return res


def experiment(side=512, levels=4, base=32, runs=8, steps_per_run=500, train_steps=1000,
               batch=8, lr=1e-3, held_steps=41, every=10, seed=0, log=print,
               objective=L.WS_ENERGY_NET, omega=0.8, sweeps=2, rounds=4, round_snaps=100,
               round_iters=(1, 2, 4, 8, 16, 32, 64, 128, 256), round_steps=1000, round_lr=3e-4,
               round_keep=1000, max_rounds=8, val_snaps=3, val_iter=400):
    """Capture, train, evaluate; returns a dict of the measured figures.

    objective L.WS_ENERGY_NET or L.WS_ENERGY trains the network as the
    coarse solver of iskra_ws_solve's two-level cycle (`sweeps` Jacobi
    sweeps of weight omega on either side); then rounds: the pairs of the
    solve's own iterates (`round_iters` iterations) on `round_snaps`
    training snapshots are added to the dataset and training continues
    (`round_steps` steps at `round_lr`) on the iterate pairs and the last
    `round_keep` snapshots, so the network learns the residuals it meets
    inside the iteration without them being diluted by the snapshots.  After
    `rounds` rounds, each further round (up to `max_rounds`) is run only
    while the learned solve of `val_snaps` training snapshots misses 1e-8
    within `val_iter` iterations (a check on training data; the held-out
    snapshots are never used).  The evaluation then includes that solve.
    L.WS_PHI_MSE trains a warm start only (no rounds).  The held-out
    snapshots come from a run the training never sees.  Everything runs on
    device 0 in one context, closed before returning."""
    t0 = time.perf_counter()

    def say(m):
        if log is not None:
            log(f"[{time.perf_counter() - t0:7.1f} s] {m}")

    energy = objective in (L.WS_ENERGY, L.WS_ENERGY_NET)
    rounds = rounds if energy else 0
    ctx = Context(0)
    ws = None
    try:
        ctx.grid(2, (side + 1, side + 1), (1.0, 1.0))
        ws = WarmStart(ctx, levels=levels, base=base, side=side, max_batch=batch)
        ws.set_params(he_init(levels, base, seed))
        ws.set_objective(objective, omega if energy else 0.0, sweeps if energy else 0)
        n = runs * steps_per_run
        ws.reserve(held_steps + n + (max(rounds, max_rounds) if rounds else 0) * round_snaps * len(round_iters))
        held = []
        capture_two_stream(ws, runs=1, steps=held_steps, seed=seed * 100 + 99, record=held)
        pool = []
        pool_every = max(1, (n // max(round_snaps, 1)) // 2) if rounds else 0
        t = time.perf_counter()
        for r in range(runs):
            capture_two_stream(ws, runs=1, steps=steps_per_run, seed=seed * 100 + 10 + r,
                               record=pool if rounds else None, record_every=pool_every or 1)
        ctx.sync()
        cap_s = time.perf_counter() - t
        say(f"captured {n} training and {held_steps} held-out snapshots in {cap_s:.1f} s")
        h0 = held_steps
        ws.set_gain(ws.dataset_rms(h0, n))
        e0 = ws.eval(0, held_steps)
        t = time.perf_counter()
        losses, gain = train(ws, h0, n, steps=train_steps, batch=batch, lr=lr,
                             chunk=max(1, train_steps // 10), log=say)
        ws.set_gain(gain)
        pairs = 0
        rng = np.random.default_rng(seed + 17)
        val = [pool[j][0] for j in range(max(0, len(pool) - val_snaps), len(pool))]
        tags = np.zeros((side + 1, side + 1), np.int8)
        tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1

        def validate():
            ctx.poisson(L.POISSON_PCG, 1.0, tags.ravel(), rtol=1e-8, max_iter=100000)
            ctx.tag_value(1, 0.0)
            worst = 0.0
            for rho in val:
                ctx.set_field(L.F_RHO, rho)
                _, rr = ws.solve(L.WARM_ZERO, omega, sweeps, max_iter=val_iter,
                                 raise_unconverged=False)
                worst = max(worst, rr)
            return worst

        done_rounds = 0
        for rd in range(max_rounds if rounds else 0):
            pick = rng.choice(len(pool), min(round_snaps, len(pool)), replace=False)
            pairs += capture_iterate_pairs(ws, [pool[j][0] for j in pick], round_iters, omega,
                                           sweeps)
            first = h0 + max(0, n - round_keep)
            k = ws.size() - first
            done = 0
            while done < round_steps:
                c = min(max(1, round_steps // 5), round_steps - done)
                ws.train(first, k, batch, c, round_lr, seed=seed + 7919 * (rd + 1) + done)
                done += c
            done_rounds = rd + 1
            if done_rounds >= rounds:
                worst = validate()
                say(f"round {done_rounds}: {pairs} iterate pairs, {round_steps} more steps; "
                    f"training-snapshot check: worst relres {worst:.2e} after {val_iter} iterations")
                if worst <= 1e-8:
                    break
            else:
                say(f"round {done_rounds}: {pairs} iterate pairs, {round_steps} more steps")
        ctx.sync()
        tr_s = time.perf_counter() - t
        all_steps = train_steps + done_rounds * round_steps
        e1 = ws.eval(0, held_steps)
        say(f"held-out loss {e0:.4e} -> {e1:.4e}")
        ctx.set_field(L.F_RHO, held[0][0])
        ctx.profile(True)
        for _ in range(10):
            ws.predict()
        st = ctx.stats()
        ctx.profile(False)
        inf_ms = st["kernel_ms"]["warmstart"] / max(st["kernel_calls"]["warmstart"], 1)
        names = {L.POISSON_PCG: "pcg", L.POISSON_MG: "mg"}
        objname = {L.WS_ENERGY: f"energy of the cycle ({sweeps} Jacobi sweeps each side, "
                                f"omega {omega})",
                   L.WS_ENERGY_NET: f"energy of the network's own correction inside the cycle "
                                    f"({sweeps} Jacobi sweeps each side, omega {omega})",
                   L.WS_PHI_MSE: "phi mean squared error"}[objective]
        snaps = [(rho, prev) for rho, prev, _phi in held[1::every]]
        table = solve_table(ctx, ws, snaps, fcg_omega=omega if energy else None,
                            fcg_sweeps=sweeps)
        summary = {names[k]: {s: {"iters": float(np.mean(v["iters"])),
                                  "ms": float(np.median(v["ms"])),
                                  "max_relres": float(np.max(v["relres"])),
                                  "converged": bool(np.max(v["relres"]) <= 1e-8)}
                              for s, v in row.items()} for k, row in table.items()}
        flops = forward_flops(levels, base, side)
        out = dict(side=side, levels=levels, base=base, params=ws.nparam, snapshots=n,
                   capture_s=cap_s, train_steps=all_steps, batch=batch, train_s=tr_s,
                   rounds=done_rounds, iterate_pairs=pairs,
                   train_ms_per_step=1e3 * tr_s / max(all_steps, 1),
                   objective=objname, heldout_loss_before=e0, heldout_loss_after=e1, gain=gain,
                   inference_ms=inf_ms, inference_tflops=flops / (inf_ms * 1e-3) / 1e12,
                   solves=summary, evaluated=len(snaps))
        say(json.dumps(summary))
        return out
    finally:
        if ws is not None:
            ws.close()
        ctx.close()


def forward_flops(levels, base, side):
    """Multiply-add flops (x2) of one forward pass on one sample, counting the
    padded channels the kernels compute."""
    f = 0
    shapes = layer_shapes(levels, base)
    lev = [l for l in range(levels) for _ in (0, 1)] + \
          [l for l in range(levels - 2, -1, -1) for _ in (0, 1)] + [0]
    for (cin, cout), l in zip(shapes, lev):
        f += 2 * (side >> l) ** 2 * 9 * cin * cout
    return f


def inference_only(passes, side=512, levels=4, base=32, seed=0):
    """`passes` forward passes of a He-initialised network on a random rho
    (the kernels' own timing, e.g. under a profiler); returns the mean ms
    per pass from the library's HIP events."""
    ctx = Context(0)
    ws = None
    try:
        ctx.grid(2, (side + 1, side + 1), (1.0, 1.0))
        ws = WarmStart(ctx, levels=levels, base=base, side=side, max_batch=1)
        ws.set_params(he_init(levels, base, seed))
        ws.set_gain(1.0)
        ctx.set_field(L.F_RHO, np.random.default_rng(seed).standard_normal((side + 1) ** 2))
        ws.predict()
        ctx.profile(True)
        for _ in range(passes):
            ws.predict()
        st = ctx.stats()
        ctx.profile(False)
        return st["kernel_ms"]["warmstart"] / max(st["kernel_calls"]["warmstart"], 1)
    finally:
        if ws is not None:
            ws.close()
        ctx.close()


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="learned warm start: capture, train, evaluate")
    ap.add_argument("--inference-only", type=int, default=0,
                    help="only time this many forward passes of an untrained network")
    ap.add_argument("--side", type=int, default=512)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--runs", type=int, default=40)
    ap.add_argument("--steps-per-run", type=int, default=500)
    ap.add_argument("--train-steps", type=int, default=1000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--objective", choices=("net", "cycle", "mse"), default="net")
    ap.add_argument("--omega", type=float, default=0.8)
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--rounds", type=int, default=4)
    ap.add_argument("--round-snaps", type=int, default=100)
    ap.add_argument("--round-steps", type=int, default=1000)
    ap.add_argument("--round-iters", default="1,2,4,8,16,32,64,128,256")
    ap.add_argument("--max-rounds", type=int, default=8)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    if a.inference_only:
        ms = inference_only(a.inference_only, a.side, a.levels, a.base)
        flops = forward_flops(a.levels, a.base, a.side)
        print(json.dumps({"side": a.side, "levels": a.levels, "base": a.base,
                          "forward_ms": ms, "tflops": flops / (ms * 1e-3) / 1e12}), flush=True)
        return
    res = experiment(a.side, a.levels, a.base, a.runs, a.steps_per_run, a.train_steps, a.batch,
                     a.lr, log=lambda m: print(m, flush=True),
                     objective={"net": L.WS_ENERGY_NET, "cycle": L.WS_ENERGY,
                                "mse": L.WS_PHI_MSE}[a.objective],
                     omega=a.omega, sweeps=a.sweeps, rounds=a.rounds, round_snaps=a.round_snaps,
                     round_steps=a.round_steps,
                     round_iters=tuple(int(v) for v in a.round_iters.split(",")),
                     max_rounds=a.max_rounds)
    s = json.dumps(res)
    print(s, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(s + "\n")
    print("closed", flush=True)


if __name__ == "__main__":
    main()
