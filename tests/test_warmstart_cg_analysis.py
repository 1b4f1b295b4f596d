"""Why a learned warm start does not cut Jacobi-PCG iterations at rtol 1e-8.

A CPU check of the mathematics behind the warm-start measurements (no GPU,
no network): on the 5-point Dirichlet Poisson problem of a 129^2 grid with a
right-hand side of particle-like noise plus one smooth mode, conjugate
gradients started from the exact solution plus a *random smooth* error of
1 % of max|phi| need about as many iterations as from zero, and more with a
5 % or 20 % error.  CG's iteration count is set by how many slow eigenmodes
the initial error holds, not by its size: the zero start's error is phi
itself, dominated by a few modes, while any approximate solution whose error
is spread over the smooth modes (as a network's is) gives CG more slow modes
to resolve.  High-frequency error, by contrast, is removed in a few
iterations.
"""
import numpy as np


def _apply(x):
    y = 4.0 * x
    y[1:, :] -= x[:-1, :]
    y[:-1, :] -= x[1:, :]
    y[:, 1:] -= x[:, :-1]
    y[:, :-1] -= x[:, 1:]
    return y


def _pcg(b, x0, tol, maxit=20000):
    """Jacobi-preconditioned CG (diagonal 4) to ||r|| <= tol ||b||."""
    x = x0.copy()
    r = b - _apply(x)
    z = r / 4.0
    p = z.copy()
    rz = float((r * z).sum())
    bn = np.linalg.norm(b)
    for k in range(maxit):
        if np.linalg.norm(r) <= tol * bn:
            return k, x
        ap = _apply(p)
        al = rz / float((p * ap).sum())
        x += al * p
        r -= al * ap
        z = r / 4.0
        rzn = float((r * z).sum())
        p = z + (rzn / rz) * p
        rz = rzn
    return maxit, x


def _smooth(rng, m, level, scale):
    e = rng.standard_normal((m, m))
    for _ in range(200):  # low-pass: repeated 4-neighbour averaging
        e = 0.25 * (np.roll(e, 1, 0) + np.roll(e, -1, 0) + np.roll(e, 1, 1) + np.roll(e, -1, 1))
    return e / np.abs(e).max() * level * scale


def test_smooth_start_error_does_not_cut_cg_iterations():
    n = 129
    m = n - 2
    rng = np.random.default_rng(0)
    X, Y = np.meshgrid(np.arange(1, n - 1) / (n - 1), np.arange(1, n - 1) / (n - 1))
    b = 0.05 * rng.standard_normal((m, m)) + np.sin(3 * np.pi * X) * np.sin(2 * np.pi * Y)
    _, phi = _pcg(b, np.zeros_like(b), 1e-13)
    assert np.linalg.norm(b - _apply(phi)) <= 1e-12 * np.linalg.norm(b)
    k0, _ = _pcg(b, np.zeros_like(b), 1e-8)
    amp = np.abs(phi).max()
    ks = {lvl: _pcg(b, phi + _smooth(rng, m, lvl, amp), 1e-8)[0] for lvl in (0.2, 0.05, 0.01)}
    kw, _ = _pcg(b, phi * (1.0 + 4e-3 * rng.standard_normal((m, m))), 1e-8)
    print("zero start", k0, "smooth error 20/5/1 %", ks, "white 0.4 %", kw)
    # within 1 % of phi in the smooth modes: no more than a 10 % cut
    assert ks[0.01] >= 0.9 * k0
    # larger smooth errors cost iterations instead of saving them
    assert ks[0.05] > k0 and ks[0.2] > ks[0.05]
    # a purely high-frequency error of the same order is cheap to remove
    assert kw < k0
