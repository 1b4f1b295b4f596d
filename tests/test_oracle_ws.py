"""Pinning of the learned warm-start oracle (oracle/ws_oracle.c), CPU only:
FP64 gradients against central finite differences, and the whole forward and
backward pass against the same network written with torch.nn.functional on
the CPU in float64 (a library routine, used as an independent check)."""
import numpy as np
import pytest

import oracle as O

WS_IN, WS_OUT = 8, 16


def _layers(L, base):
    """(cin, cout) of each layer in parameter order."""
    out = []
    for l in range(L):
        out.append((WS_IN if l == 0 else base << (l - 1), base << l))
        out.append((base << l, base << l))
    for l in range(L - 2, -1, -1):
        out.append((base << (l + 1), base << l))
        out.append((base << l, base << l))
    out.append((base, WS_OUT))
    return out


def _torch_net(L, base, params, x, t):
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
    diff = (o - tt)[:, :S - 1, :S - 1]
    loss = (diff ** 2).sum() / (B * (S - 1) ** 2)
    loss.backward()
    return o.detach().numpy(), loss.item(), p.grad.numpy()


@pytest.mark.parametrize("L,base,S", [(1, 16, 8), (2, 16, 16), (3, 16, 32)])
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
    assert np.allclose(out, ot, rtol=1e-12, atol=1e-12)
    assert loss == pytest.approx(lt, rel=1e-12)
    assert np.allclose(g, gt, rtol=1e-10, atol=1e-13)


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
