"""GPU parity of the learned warm-start network (BF16 MFMA forward, input-
and weight-gradient kernels) against the CPU oracle (oracle/ws_oracle.c in
its bfloat16-rounding mode: same rounding points, FP64 sums), and the effect
of a trained network on the Poisson solver."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L_, BASE_, S_ = 3, 16, 64


def _he_params(L, base, seed):
    from iskra_amd import warmstart as W
    return W.he_init(L, base, seed)


def _net(ctx, L=L_, base=BASE_, S=S_, batch=2):
    from iskra_amd import _lib as Lb
    ctx.grid(2, (S + 1, S + 1), (1.0, 1.0))
    return Lb.WarmStart(ctx, levels=L, base=base, side=S, max_batch=batch)


def _relerr(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_ws_forward_matches_oracle(gpu_ready):
    import oracle as O
    from iskra_amd import Context
    ctx = Context(0)
    ws = _net(ctx)
    p = _he_params(L_, BASE_, 1)
    assert p.size == ws.nparam == O.ws_param_count(L_, BASE_)
    ws.set_params(p)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((2, S_, S_)).astype(np.float32)
    out = ws.forward(x)
    ref, _, _ = O.ws_run(L_, BASE_, p.astype(np.float64), x.astype(np.float64), round_bf16=True)
    # same bf16 rounding points; FP32 (GPU) vs FP64 (oracle) sums may flip a
    # rounding now and then: relative L2 error below 5e-3
    e = _relerr(out, ref)
    print("forward relerr", e)
    assert e < 5e-3
    ws.close()
    ctx.close()


def test_ws_gradients_match_oracle(gpu_ready):
    import oracle as O
    from iskra_amd import Context
    ctx = Context(0)
    ws = _net(ctx)
    p = _he_params(L_, BASE_, 2)
    ws.set_params(p)
    rng = np.random.default_rng(12)
    x = rng.standard_normal((2, S_, S_)).astype(np.float32)
    t = rng.standard_normal((2, S_, S_)).astype(np.float32)
    loss, g = ws.backward(x, t)
    _, lref, gref = O.ws_run(L_, BASE_, p.astype(np.float64), x.astype(np.float64),
                             t.astype(np.float64), round_bf16=True)
    assert loss == pytest.approx(lref, rel=1e-3)
    from iskra_amd import warmstart as W
    for k, (off, n, cin, cout) in enumerate(W.layer_slices(L_, BASE_)):
        e = _relerr(g[off:off + n], gref[off:off + n])
        print("layer", k, cin, cout, "grad relerr", e)
        assert e < 3e-2, (k, e)
    ws.close()
    ctx.close()


def _bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _presmooth(v, omega, k):
    """(I - omega/4 A)^k on the interior (S-1)^2 pixels of each plane, 0
    outside: the pair transform of the two-level objective."""
    v = np.array(v, dtype=np.float64)
    S = v.shape[-1]
    v[:, S - 1, :] = 0.0
    v[:, :, S - 1] = 0.0
    for _ in range(k):
        a = 4 * v
        a[:, 1:, :] -= v[:, :-1, :]
        a[:, :-1, :] -= v[:, 1:, :]
        a[:, :, 1:] -= v[:, :, :-1]
        a[:, :, :-1] -= v[:, :, 1:]
        v = v - omega / 4 * a
        v[:, S - 1, :] = 0.0
        v[:, :, S - 1] = 0.0
    return v


def test_ws_energy_gradients_match_oracle(gpu_ready):
    """The two-level objective's loss and parameter gradients (pair mapped by
    two pre-sweeps, network times gain, two post-sweeps, A-norm) against the
    oracle fed the same pre-smoothed pair."""
    import oracle as O
    from iskra_amd import Context, _lib as Lb
    from iskra_amd import warmstart as W
    ctx = Context(0)
    ws = _net(ctx)
    try:
        p = _he_params(L_, BASE_, 4)
        ws.set_params(p)
        ws.set_gain(1.3)
        ws.set_objective(Lb.WS_ENERGY, 0.8, 2)
        rng = np.random.default_rng(13)
        x = rng.standard_normal((2, S_, S_)).astype(np.float32)
        t = (0.05 * rng.standard_normal((2, S_, S_))).astype(np.float32)
        loss, g = ws.backward(x, t)
        xp = _presmooth(_bf16(x), 0.8, 2)
        tp = _presmooth(t, 0.8, 2)
        _, lref, gref = O.ws_run(L_, BASE_, p.astype(np.float64), xp, tp, round_bf16=True,
                                 objective=1, gain=1.3, omega=0.8, sweeps=2)
        print("energy loss", loss, lref)
        assert loss == pytest.approx(lref, rel=1e-3)
        for k, (off, n, cin, cout) in enumerate(W.layer_slices(L_, BASE_)):
            e = _relerr(g[off:off + n], gref[off:off + n])
            print("layer", k, cin, cout, "grad relerr", e)
            assert e < 3e-2, (k, e)
        # the network's own error (no post-sweeps in the loss), same pairs
        ws.set_objective(Lb.WS_ENERGY_NET, 0.8, 2)
        loss, g = ws.backward(x, t)
        _, lref, gref = O.ws_run(L_, BASE_, p.astype(np.float64), xp, tp, round_bf16=True,
                                 objective=1, gain=1.3, omega=0.8, sweeps=0)
        print("network energy loss", loss, lref)
        assert loss == pytest.approx(lref, rel=1e-3)
        # its output gradient A e is not smoothed by post-sweeps, so BF16
        # rounding flips of the gradient plane weigh more (a wrong stencil or
        # sign gives errors of order 1)
        for k, (off, n, cin, cout) in enumerate(W.layer_slices(L_, BASE_)):
            e = _relerr(g[off:off + n], gref[off:off + n])
            print("layer", k, cin, cout, "network-objective grad relerr", e)
            assert e < 6e-2, (k, e)
    finally:
        ws.close()
        ctx.close()


def _A_interior(v):
    y = 4 * v
    y[1:] -= v[:-1]
    y[:-1] -= v[1:]
    y[:, 1:] -= v[:, :-1]
    y[:, :-1] -= v[:, 1:]
    return y


def test_ws_capture_iterates_pairs_are_the_solve_iterates(gpu_ready):
    """iskra_ws_capture_iterates: each stored pair is (r_k, phi - x_k) / rms(r_k)
    of iskra_ws_solve's k-th iterate (checked against a separate solve capped
    at k iterations) and is exact: A target = input within the BF16 rounding
    of the input; the context's phi is left as it was."""
    from iskra_amd import Context, _lib as Lb
    from iskra_amd import warmstart as W
    ctx = Context(0)
    ctx.grid(2, (S_ + 1, S_ + 1), (1.0, 1.0))
    ws = Lb.WarmStart(ctx, levels=L_, base=BASE_, side=S_, max_batch=2)
    try:
        ws.set_params(W.he_init(L_, BASE_, 6))
        ws.set_gain(0.05)
        ws.reserve(8)
        n = S_ + 1
        m = S_ - 1
        rng = np.random.default_rng(21)
        rho = np.zeros((n, n))
        rho[1:-1, 1:-1] = rng.standard_normal((m, m))
        got = W.capture_iterate_pairs(ws, [rho.ravel()], iters=(1, 2, 5), omega=0.8, sweeps=2)
        assert got == 3 and ws.size() == 3
        phi_exact = ctx.field(Lb.F_PHI).copy()
        f = rho[1:-1, 1:-1]
        for idx, k in enumerate((1, 2, 5)):
            xin, tgt = ws.dataset_get(idx)
            xin, tgt = xin[:m, :m].astype(np.float64), tgt[:m, :m].astype(np.float64)
            assert np.linalg.norm(_A_interior(tgt) - xin) <= 5e-3 * np.linalg.norm(xin)
            ctx.set_field(Lb.F_RHO, rho.ravel())
            with pytest.raises(Lb.IskraError):
                ws.solve(Lb.WARM_ZERO, 0.8, 2, max_iter=k)
            x = ctx.field(Lb.F_PHI).reshape(n, n)[1:-1, 1:-1]
            r = f - _A_interior(x)
            s = np.sqrt(np.mean(r * r))
            assert np.allclose(xin, r / s, rtol=0, atol=8e-3 * np.max(np.abs(r / s))), k
            ex = phi_exact.reshape(n, n)[1:-1, 1:-1]
            assert np.allclose(tgt, (ex - x) / s, rtol=1e-5, atol=1e-6 * np.max(np.abs(tgt))), k
            ctx.set_field(Lb.F_PHI, phi_exact)
        with pytest.raises(Lb.IskraError) as ei:
            ws.capture_iterates([3, 2])
        assert ei.value.code == Lb.E_ARG
    finally:
        ws.close()
        ctx.close()


def test_ws_network_objective_with_iterate_rounds_cuts_iterations_3x(gpu_ready):
    """The training recipe of the learned solve on a 65 x 65-node grid: the
    network-error objective on two-stream snapshots, then one round of the
    solve's own iterate pairs; held-out snapshots of another run reach the
    FP64 residual 1e-8 in at least 3x fewer iterations than Jacobi-PCG."""
    from iskra_amd import Context, _lib as Lb
    from iskra_amd import warmstart as W
    ctx = Context(0)
    ctx.grid(2, (S_ + 1, S_ + 1), (1.0, 1.0))
    ws = Lb.WarmStart(ctx, levels=L_, base=BASE_, side=S_, max_batch=8)
    try:
        ws.set_params(W.he_init(L_, BASE_, 3))
        ws.set_objective(Lb.WS_ENERGY_NET, 0.8, 2)
        ws.reserve(600)
        held = []
        W.capture_two_stream(ws, runs=1, steps=40, seed=2, record=held)
        pool = []
        W.capture_two_stream(ws, runs=2, steps=150, seed=1, record=pool, record_every=5)
        _, gain = W.train(ws, 40, 300, steps=1000, batch=8, lr=2e-3)
        ws.set_gain(gain)
        W.capture_iterate_pairs(ws, [p[0] for p in pool[:40]], iters=(1, 2, 4, 8, 16))
        ws.train(40, ws.size() - 40, 8, 500, 5e-4, seed=9)
        snaps = [(rho, prev) for rho, prev, _phi in held[::8]]
        table = W.solve_table(ctx, ws, snaps, kinds=(Lb.POISSON_PCG,), fcg_omega=0.8,
                              fcg_sweeps=2)[Lb.POISSON_PCG]
        print({k: v["iters"] for k, v in table.items()}, flush=True)
        for v in table.values():
            assert max(v["relres"]) <= 1e-8
        plain, fcg = np.mean(table["zero"]["iters"]), np.mean(table["fcg"]["iters"])
        assert plain >= 3.0 * fcg, (plain, fcg)
    finally:
        ws.close()
        ctx.close()


def _dst_solve(f):
    """Exact solution of 4 x - (4 neighbours) = f on the interior nodes, 0 on
    the walls (DST-I diagonalises the operator)."""
    from scipy.fft import dstn, idstn
    m = f.shape[0]
    k = np.arange(1, m + 1)
    lam1 = 2 - 2 * np.cos(np.pi * k / (m + 1))
    return idstn(dstn(f, type=1) / (lam1[:, None] + lam1[None, :]), type=1)


def test_ws_in_iteration_cuts_pcg_iterations_3x(gpu_ready):
    """The network inside the iteration on a 65 x 65-node grid: trained on
    300 two-stream snapshots with the two-level energy objective (two Jacobi
    sweeps before and after the network), the cycle preconditions a
    flexible CG (iskra_ws_solve) that reaches the same FP64 residual 1e-8 as
    Jacobi-PCG in at least 3x fewer iterations on held-out snapshots of
    another run; each solution is checked on the host (residual recomputed,
    error against the exact DST-I solution)."""
    from iskra_amd import Context, _lib as Lb
    from iskra_amd import warmstart as W
    ctx = Context(0)
    ctx.grid(2, (S_ + 1, S_ + 1), (1.0, 1.0))
    ws = Lb.WarmStart(ctx, levels=L_, base=BASE_, side=S_, max_batch=8)
    try:
        ws.set_params(W.he_init(L_, BASE_, 3))
        ws.set_objective(Lb.WS_ENERGY, 0.8, 2)
        ws.reserve(400)
        W.capture_two_stream(ws, runs=2, steps=150, seed=1)
        held = []
        W.capture_two_stream(ws, runs=1, steps=40, seed=2, record=held)
        ws.set_gain(ws.dataset_rms(0, 300))
        before = ws.eval(300, 40)
        losses, _ = W.train(ws, 0, 300, steps=1500, batch=8, lr=2e-3)
        after = ws.eval(300, 40)
        print("held-out energy loss", before, "->", after, flush=True)
        assert np.all(np.isfinite(losses)) and after < 0.2 * before
        snaps = [(rho, prev) for rho, prev, _phi in held[::8]]
        table = W.solve_table(ctx, ws, snaps, kinds=(Lb.POISSON_PCG,),
                              fcg_omega=0.8, fcg_sweeps=2)[Lb.POISSON_PCG]
        print({k: v["iters"] for k, v in table.items()}, flush=True)
        for v in table.values():
            assert max(v["relres"]) <= 1e-8
        plain, fcg = np.mean(table["zero"]["iters"]), np.mean(table["fcg"]["iters"])
        assert plain >= 3.0 * fcg, (plain, fcg)
        n = S_ + 1
        for rho, _prev in snaps:
            ctx.set_field(Lb.F_RHO, rho)
            it, rr = ws.solve(Lb.WARM_ZERO, 0.8, 2)
            phi = ctx.field(Lb.F_PHI).reshape(n, n)
            f = rho.reshape(n, n)[1:-1, 1:-1]
            x = phi[1:-1, 1:-1]
            assert np.all(phi[0] == 0) and np.all(phi[-1] == 0)
            assert np.all(phi[:, 0] == 0) and np.all(phi[:, -1] == 0)
            r = f - (4 * x - phi[:-2, 1:-1] - phi[2:, 1:-1] - phi[1:-1, :-2] - phi[1:-1, 2:])
            assert np.linalg.norm(r) / np.linalg.norm(f) <= 1.01e-8, (it, rr)
            ex = _dst_solve(f)
            assert np.linalg.norm(x - ex) / np.linalg.norm(ex) <= 1e-6
    finally:
        ws.close()
        ctx.close()


def test_ws_solve_refuses_other_systems(gpu_ready):
    """iskra_ws_solve solves only the network's system (phi = 0 on the four
    walls): an extra Dirichlet node is refused with ISKRA_E_UNSUPPORTED."""
    from iskra_amd import Context, _lib as Lb
    ctx = Context(0)
    ctx.grid(2, (S_ + 1, S_ + 1), (1.0, 1.0))
    ws = Lb.WarmStart(ctx, levels=L_, base=BASE_, side=S_, max_batch=1)
    try:
        n = S_ + 1
        tags = np.zeros((n, n), np.int8)
        tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
        tags[n // 2, n // 2] = 2
        ctx.poisson(Lb.POISSON_PCG, 1.0, tags.ravel(), rtol=1e-8)
        ctx.tag_value(1, 0.0)
        ctx.tag_value(2, 1.0)
        with pytest.raises(Lb.IskraError) as ei:
            ws.solve(Lb.WARM_ZERO, 0.8)
        assert ei.value.code == Lb.E_UNSUPPORTED
    finally:
        ws.close()
        ctx.close()


def test_ws_training_cuts_pcg_iterations(gpu_ready):
    """Train on snapshots of a 64 x 64 two-stream run, then solve held-out
    snapshots of another run: the learned start must need fewer PCG
    iterations than a zero start, and every solve still reaches 1e-8."""
    from iskra_amd import Context, _lib as Lb
    from iskra_amd import warmstart as W
    ctx = Context(0)
    ctx.grid(2, (S_ + 1, S_ + 1), (1.0, 1.0))
    ws = Lb.WarmStart(ctx, levels=L_, base=BASE_, side=S_, max_batch=8)
    ws.set_params(W.he_init(L_, BASE_, 3))
    ws.reserve(400)
    W.capture_two_stream(ws, runs=2, steps=150, seed=1)
    assert ws.size() == 300
    held = []
    W.capture_two_stream(ws, runs=1, steps=40, seed=2, record=held)
    before = None
    ws.set_gain(ws.dataset_rms(0, 300))
    before = ws.eval(300, 40)
    losses, gain = W.train(ws, 0, 300, steps=400, batch=8, lr=2e-3)
    after = ws.eval(300, 40)
    print("held-out loss", before, "->", after, "gain", gain)
    assert np.all(np.isfinite(losses))
    assert after < 0.2 * before
    snaps = [(rho, prev) for rho, prev, _phi in held[::8]]
    table = W.solve_table(ctx, ws, snaps, kinds=(Lb.POISSON_PCG,))[Lb.POISSON_PCG]
    print({k: v["iters"] for k, v in table.items()})
    for v in table.values():
        assert max(v["relres"]) <= 1e-8
    zero, learned = table["zero"]["iters"], table["learned"]["iters"]
    assert np.mean(learned) < np.mean(zero)
    ws.close()
    ctx.close()


def test_ws_cfg5_network_against_oracle(gpu_ready):
    """The learned warm start's own architecture (4 levels, 32-256 channels,
    3x3 convolutions, BF16 activations and weights, FP32 accumulation):
    forward and every layer's weight gradient at 256^2 pixels, and the forward
    pass at 512^2 (the 513^2-node grid of the configuration), against the FP64
    oracle in its BF16-rounding mode."""
    import oracle as O
    from iskra_amd import Context, _lib as Lb
    from iskra_amd import warmstart as W
    L, base = 4, 32
    p = W.he_init(L, base, 5)
    for S, grads in ((256, True), (512, False)):
        ctx = Context(0)
        ctx.grid(2, (S + 1, S + 1), (1.0, 1.0))
        ws = Lb.WarmStart(ctx, levels=L, base=base, side=S, max_batch=1)
        try:
            ws.set_params(p)
            rng = np.random.default_rng(S)
            x = rng.standard_normal((1, S, S)).astype(np.float32)
            t = rng.standard_normal((1, S, S)).astype(np.float32)
            out = ws.forward(x)
            ref, lref, gref = O.ws_run(L, base, p.astype(np.float64), x.astype(np.float64),
                                       t.astype(np.float64) if grads else None, round_bf16=True,
                                       grads=grads)
            e = _relerr(out, ref)
            print(f"{S}^2 forward relerr {e:.2e}", flush=True)
            assert e < 5e-3
            if grads:
                loss, g = ws.backward(x, t)
                assert loss == pytest.approx(lref, rel=1e-3)
                for k, (off, n, cin, cout) in enumerate(W.layer_slices(L, base)):
                    ek = _relerr(g[off:off + n], gref[off:off + n])
                    print(f"  layer {k} {cin}->{cout} grad relerr {ek:.2e}", flush=True)
                    assert ek < 3e-2, (k, ek)
        finally:
            ws.close()
            ctx.close()
