"""Pinning of the learned warm-start oracle (oracle/ws_oracle.c), CPU only:
FP64 gradients against central finite differences, and the whole forward and
backward pass against the same network written with torch.nn.functional on
the CPU in float64 (a library routine, used as an independent check)."""
import numpy as np
import pytest

import oracle as O

WS_IN, WS_OUT = 8, 16


def _ch(base, l):
    """Channels of level l: base * 2^l, held at base * 8 from level 3 on."""
    return base << min(l, 3)


def _layers(L, base):
    """(cin, cout) of each layer in parameter order."""
    out = []
    for l in range(L):
        out.append((WS_IN if l == 0 else _ch(base, l - 1), _ch(base, l)))
        out.append((_ch(base, l), _ch(base, l)))
    for l in range(L - 2, -1, -1):
        out.append((_ch(base, l + 1), _ch(base, l)))
        out.append((_ch(base, l), _ch(base, l)))
    out.append((base, WS_OUT))
    return out


def _torch_net(L, base, params, x, t, energy=None):
    import torch
    import torch.nn.functional as F
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    ws, off = [], 0
    for cin, cout in _layers(L, base):
        n = 9 * cin * cout
        # [cout][tap][cin] -> [cout][cin][3][3]
        ws.append(p[off:off + n].view(cout, 3, 3, cin).permute(0, 3, 1, 2))
        off += n
    B, S = x.shape[0], x.shape[1]
    inp = torch.zeros((B, WS_IN, S, S), dtype=torch.float64)
    inp[:, 0] = torch.tensor(x)
    conv = lambda a, k: F.conv2d(a, ws[k], padding=1)  # noqa: E731
    E = []
    a = inp
    for l in range(L):
        if l > 0:
            a = F.avg_pool2d(E[l - 1], 2)
        ea = F.relu(conv(a, 2 * l))
        E.append(F.relu(conv(ea, 2 * l + 1)))
    d = E[L - 1]
    for l in range(L - 2, -1, -1):
        k = 2 * L + 2 * (L - 2 - l)
        u = F.interpolate(d, scale_factor=2, mode="nearest")
        da = F.relu(conv(u, k) + E[l])
        d = F.relu(conv(da, k + 1))
    o = conv(d, len(ws) - 1)[:, 0]
    tt = torch.tensor(t)
    if energy is None:
        diff = (o - tt)[:, :S - 1, :S - 1]
        loss = (diff ** 2).sum() / (B * (S - 1) ** 2)
    else:
        # the 5-point operator on the interior (S-1)^2 pixels, zero outside
        gain, omega, sweeps = energy
        k5 = torch.tensor([[0., -1., 0.], [-1., 4., -1.], [0., -1., 0.]],
                          dtype=torch.float64)[None, None]

        def A(v):
            return F.conv2d(v[:, None], k5, padding=1)[:, 0]
        xi = inp[:, 0, :S - 1, :S - 1]
        y = gain * o[:, :S - 1, :S - 1]
        z = y
        for _ in range(sweeps):
            z = z + (omega / 4) * (xi - A(z))
        tv = gain * tt[:, :S - 1, :S - 1]
        e = z - tv
        loss = (e * A(e)).sum() / (tv * A(tv)).sum()
    loss.backward()
    return o.detach().numpy(), loss.item(), p.grad.numpy()


@pytest.mark.parametrize("L,base,S", [(1, 16, 8), (2, 16, 16), (3, 16, 32), (5, 16, 32)])
def test_ws_oracle_matches_torch(L, base, S):
    pytest.importorskip("torch")
    rng = np.random.default_rng(L)
    n = O.ws_param_count(L, base)
    assert n == sum(9 * a * b for a, b in _layers(L, base))
    params = rng.standard_normal(n) * 0.15
    x = rng.standard_normal((2, S, S))
    t = rng.standard_normal((2, S, S))
    out, loss, g = O.ws_run(L, base, params, x, t, round_bf16=False)
    ot, lt, gt = _torch_net(L, base, params, x, t)
    # FP64 sums in different orders: agreement to ~1e-12 of the output scale
    assert np.allclose(out, ot, rtol=1e-12, atol=1e-12 * np.max(np.abs(ot)))
    assert loss == pytest.approx(lt, rel=1e-12)
    assert np.allclose(g, gt, rtol=1e-10, atol=1e-12 * np.max(np.abs(gt)))


def test_ws_oracle_finite_differences():
    L, base, S = 2, 16, 16
    rng = np.random.default_rng(7)
    n = O.ws_param_count(L, base)
    p = rng.standard_normal(n) * 0.1
    x = rng.standard_normal((2, S, S))
    t = rng.standard_normal((2, S, S))
    _, _, g = O.ws_run(L, base, p, x, t, round_bf16=False)
    # parameters with a non-zero gradient, spread over all layers
    cand = np.flatnonzero(np.abs(g) > 1e-6)
    for k in rng.choice(cand, 12, replace=False):
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += e
        pm[k] -= e
        _, lp, _ = O.ws_run(L, base, pp, x, t, round_bf16=False, grads=False)
        _, lm, _ = O.ws_run(L, base, pm, x, t, round_bf16=False, grads=False)
        assert g[k] == pytest.approx((lp - lm) / (2 * e), rel=1e-5, abs=1e-10)


def test_ws_oracle_bf16_mode_is_close_to_fp64_and_homogeneous():
    L, base, S = 2, 16, 16
    rng = np.random.default_rng(3)
    n = O.ws_param_count(L, base)
    p = rng.standard_normal(n) * 0.1
    x = rng.standard_normal((1, S, S))
    o64, _, _ = O.ws_run(L, base, p, x, round_bf16=False)
    o16, _, _ = O.ws_run(L, base, p, x, round_bf16=True)
    assert np.linalg.norm(o16 - o64) / np.linalg.norm(o64) < 3e-2
    # no biases and ReLU: positively homogeneous of degree 1
    o2, _, _ = O.ws_run(L, base, p, 2.0 * x, round_bf16=False)
    assert np.allclose(o2, 2.0 * o64, rtol=1e-12, atol=1e-14)


def test_ws_oracle_energy_objective_matches_torch():
    """The preconditioner objective (network times gain, one damped Jacobi
    sweep, A-norm of the error relative to the solution's) against the same
    loss written with torch on the CPU: loss and every parameter gradient."""
    pytest.importorskip("torch")
    L, base, S = 2, 16, 16
    rng = np.random.default_rng(21)
    params = rng.standard_normal(O.ws_param_count(L, base)) * 0.15
    x = rng.standard_normal((2, S, S))
    t = rng.standard_normal((2, S, S))
    for gain, omega, sweeps in ((1.7, 0.8, 1), (0.6, 0.0, 1), (1.1, 0.9, 3)):
        out, loss, g = O.ws_run(L, base, params, x, t, round_bf16=False, objective=1, gain=gain,
                                omega=omega, sweeps=sweeps)
        ot, lt, gt = _torch_net(L, base, params, x, t, energy=(gain, omega, sweeps))
        assert np.allclose(out, ot, rtol=1e-12, atol=1e-12)
        assert loss == pytest.approx(lt, rel=1e-12)
        assert np.allclose(g, gt, rtol=1e-10, atol=1e-13)


def test_ws_oracle_energy_objective_finite_differences_and_closed_form():
    """Central differences of the energy loss; and a closed form: with all
    weights zero the output is 0, and for gain t = v, x = A v with v a sine
    mode of the interior grid (A v = lam v,
    lam = 4 - 2 cos(pi k / S) - 2 cos(pi l / S)), k sweeps give
    y_k = (1 - (1 - omega lam / 4)^k) v, so the loss is
    (1 - omega lam / 4)^(2k)."""
    L, base, S = 2, 16, 16
    rng = np.random.default_rng(9)
    p = rng.standard_normal(O.ws_param_count(L, base)) * 0.1
    x = rng.standard_normal((1, S, S))
    t = rng.standard_normal((1, S, S))
    _, _, g = O.ws_run(L, base, p, x, t, round_bf16=False, objective=1, gain=1.3, omega=0.8,
                       sweeps=2)
    cand = np.flatnonzero(np.abs(g) > 1e-4)
    for k in rng.choice(cand, 10, replace=False):
        e = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[k] += e
        pm[k] -= e
        _, lp, _ = O.ws_run(L, base, pp, x, t, False, False, objective=1, gain=1.3, omega=0.8,
                            sweeps=2)
        _, lm, _ = O.ws_run(L, base, pm, x, t, False, False, objective=1, gain=1.3, omega=0.8,
                            sweeps=2)
        assert g[k] == pytest.approx((lp - lm) / (2 * e), rel=1e-5, abs=1e-10)
    m = S - 1
    zero = np.zeros(O.ws_param_count(1, 16))
    for (k, l), gain, omega, ks in (((1, 1), 1.0, 0.8, 1), ((3, 7), 2.5, 1.2, 1),
                                    ((14, 2), 0.7, 0.0, 1), ((5, 9), 1.5, 0.8, 4)):
        ii = np.arange(1, m + 1)
        v = np.outer(np.sin(np.pi * k * ii / S), np.sin(np.pi * l * ii / S))
        lam = 4 - 2 * np.cos(np.pi * k / S) - 2 * np.cos(np.pi * l / S)
        xs = np.zeros((1, S, S))
        ts = np.zeros((1, S, S))
        xs[0, :m, :m] = lam * v
        ts[0, :m, :m] = v / gain
        _, loss, _ = O.ws_run(1, 16, zero, xs, ts, round_bf16=False, grads=False, objective=1,
                              gain=gain, omega=omega, sweeps=ks)
        assert loss == pytest.approx((1 - omega * lam / 4) ** (2 * ks), rel=1e-12, abs=1e-15)
