"""Reference-mode restatements of the surface tracker, the floating-electrode
Poisson rows and the RLC circuit -- TEST INFRASTRUCTURE ONLY (see the
package docstring for who may import it).

FP64, the reference's data model (absolute positions, 1-based cells, Julia's
operation order), pure Python: the cases the tests run are small.

  SurfaceTracker   ParticleInCell/src/pic/surfaces/track.jl:1-161,
                   check.jl:1-68, hit.jl:1-56, circuit_coupling.jl:45-62
  electrode_system FiniteDifferenceMethod/src/generalized_poisson.jl:22-68,
                   205-269 and problem/configuration.jl:37-76
  CircuitRLC       Circuit/src/Circuit.jl:13-66,117-136 and
                   ParticleInCell/src/pic/circuit_coupling.jl:17-43
"""
import math

import numpy as np

NONE, ABSORB, REFLECT, PASS = None, "absorbing", "reflective", "periodic"


class Electrode:
    """Fixed or floating electrode as a surface (circuit_coupling.jl:4-15)."""

    def __init__(self, fixed, area, sigma=0.0):
        self.fixed = fixed
        self.area = area
        self.sigma = sigma
        self.dq = 0.0


class SurfaceTracker:
    """SurfaceTracker{2} (track.jl:14-19): surface per directed pair of cells
    ((i, j), (i', j')), 1-based; the grid spacing dh used by check."""

    def __init__(self, nx, ny, dh, default=ABSORB):
        self.nx, self.ny, self.dh = nx, ny, dh
        self.surface = {}
        # build_default_surface! (track.jl:34-43), get! keeps existing entries
        for i in range(1, nx):
            self.surface.setdefault(((i, 1), (i, 0)), default)
            self.surface.setdefault(((i, ny - 1), (i, ny)), default)
        for j in range(1, ny):
            self.surface.setdefault(((1, j), (0, j)), default)
            self.surface.setdefault(((nx - 1, j), (nx, j)), default)

    def track_surface(self, nodes, s):
        """build_surface_lookup! (track.jl:45-60); nodes: bool (nx, ny)."""
        for i in range(1, self.nx):
            for j in range(1, self.ny):
                a = nodes[i - 1, j - 1]
                b = nodes[i, j - 1]
                c = nodes[i, j]
                d = nodes[i - 1, j]
                if a and b:
                    self.surface[((i, j), (i, j - 1))] = s
                if b and c:
                    self.surface[((i, j), (i + 1, j))] = s
                if c and d:
                    self.surface[((i, j), (i, j + 1))] = s
                if d and a:
                    self.surface[((i, j), (i - 1, j))] = s

    def contains(self, ij):
        """Base.in(::BoundaryCell, st) (track.jl:136-143)."""
        return any(a == ij or b == ij for a, b in self.surface)

    @staticmethod
    def check(pt, v, dh):
        """check(::TrackedParticle{2}, pv, dh) (check.jl:17-37)."""
        dt, p, (i, j), (hx, hy) = pt
        vx, vy = v[p, 0], v[p, 1]
        dx = dh * (1 - hx) if vx > 0 else dh * hx
        dy = dh * (1 - hy) if vy > 0 else dh * hy
        with np.errstate(divide="ignore", invalid="ignore"):
            dtx = np.float64(dx) / np.float64(abs(vx))
            dty = np.float64(dy) / np.float64(abs(vy))
        if dt < dtx and dt < dty:
            return pt
        if dtx < dty:
            return ((dt - dtx, p, (i + 1, j), (0.0, hy + vy * dtx / dh)) if vx > 0 else
                    (dt - dtx, p, (i - 1, j), (1.0, hy + vy * dtx / dh)))
        return ((dt - dty, p, (i, j + 1), (hx + vx * dty / dh, 0.0)) if vy > 0 else
                (dt - dty, p, (i, j - 1), (hx + vx * dty / dh, 1.0)))

    def step(self, x, v, np_, dt, q, wg, pushed):
        """track! before the push (track.jl:152-161), the push itself by the
        caller-supplied `pushed` (positions and velocities after
        push_in_cartesian!), then check! (check.jl:39-68).  x, v: (np, 2) and
        (np, 3) float64 arrays, modified in place.  Returns the absorbed
        particle indices (0-based, descending, as the reverse SortedSet
        removes them)."""
        tracked = []
        for p in range(np_):
            f = 1.0 + x[p] / self.dh  # particle_cell (ParticleInCell.jl:28-35)
            ij = (int(math.floor(f[0])), int(math.floor(f[1])))
            h = (f[0] - ij[0], f[1] - ij[1])
            if self.contains(ij):
                tracked.append((dt, p, ij, h))
        x[:np_], v[:np_] = pushed(x[:np_].copy(), v[:np_].copy())
        absorbed = set()
        queue = list(tracked)
        while queue:
            pt = queue.pop(0)
            pt2 = self.check(pt, v, self.dh)
            _, p, ij, _ = pt
            _, _, ij2, _ = pt2
            if ij == ij2:
                continue
            s = self.surface.get((ij, ij2))
            if s is None:
                queue.append(pt2)
            elif s == ABSORB:
                absorbed.add(p)
            elif isinstance(s, Electrode):
                if not s.fixed:
                    dq = q * wg
                    s.dq += dq
                    s.sigma += dq / s.area
                absorbed.add(p)
            elif s == REFLECT:
                dt2 = pt2[0]
                n = (ij2[0] - ij[0], ij2[1] - ij[1])
                x[p] -= v[p, :2] * dt2
                for k in range(2):
                    if n[k] != 0:
                        v[p, k] *= -1
                x[p] += v[p, :2] * dt2
                # scattered! (hit.jl:13-21)
                d2, p2, (i, j), (hx, hy) = pt2
                i2, j2, hx2, hy2 = i, j, hx, hy
                if hx == 0.0:
                    i2, hx2 = i - 1, 1.0
                if hx == 1.0:
                    i2, hx2 = i + 1, 0.0
                if hy == 0.0:
                    j2, hy2 = j - 1, 1.0
                if hy == 1.0:
                    j2, hy2 = j + 1, 0.0
                queue.append((d2, p2, (i2, j2), (hx2, hy2)))
            # any other surface: hit! does nothing
        return sorted(absorbed, reverse=True)


def electrode_area(nodes, dx, dy):
    """calculate_area (configuration.jl:41-50), dz = 1; nodes (nx, ny)."""
    nx, ny = nodes.shape
    area = 0.0
    for j in range(ny):
        for i in range(nx):
            if nodes[i, j]:
                area += dx * 1.0 if (i < nx - 1 and nodes[i + 1, j]) else 0
                area += dy * 1.0 if (j < ny - 1 and nodes[i, j + 1]) else 0
    return area


def electrode_system(nx, ny, dx, dy, fixed=(), floating=()):
    """create_poisson_solver (generalized_poisson.jl:34-68) on a 2-D grid,
    then create_electrode for each fixed (mask, phi) -- apply_dirichlet
    (:205-215) -- and floating (mask, sigma) -- add_new_dof (:217-232) and
    apply_neumann (:237-269) -- in that order.  Returns (A, b, rho_dofs,
    sigma_rows): the dense (nn + n_float) square system, b with the
    Dirichlet values and sigmas, a bool mask of the nodes whose b takes
    -rho/eps0, and the row of each sigma dof.  Node k = i + nx j."""
    nn = nx * ny
    nf = len(floating)
    N = nn + nf
    A = np.zeros((N, N))
    for j in range(ny):
        for i in range(nx):
            k = i + nx * j
            for di, dj, ok in ((1, 0, i < nx - 1), (-1, 0, i > 0), (0, 1, j < ny - 1),
                               (0, -1, j > 0)):
                if ok:
                    A[k, k] -= 1.0
                    A[k, k + di + nx * dj] += 1.0
    A[:nn, :nn] /= dx ** 2  # A ./= grid.dh[1]^2 (:65)
    b = np.zeros(N)
    rho = np.ones(nn, dtype=bool)
    for mask, phi in fixed:
        for k in np.flatnonzero(mask.T.ravel()):
            A[k, :] = 0.0
            A[k, k] = 1.0
            b[k] = phi
            rho[k] = False
    sig_rows = []
    for f, (mask, sigma) in enumerate(floating):
        s = nn + f
        A[s, s] = 1.0
        b[s] = sigma
        sig_rows.append(s)
        for i1 in range(1, nx + 1):  # findall order is column-major; rows are independent
            for j1 in range(1, ny + 1):
                if not mask[i1 - 1, j1 - 1]:
                    continue
                a_ = mask[i1 - 2, j1 - 1] if i1 > 1 else False
                b_ = mask[i1, j1 - 1] if i1 < nx else False
                c_ = mask[i1 - 1, j1 - 2] if j1 > 1 else True
                d_ = mask[i1 - 1, j1] if j1 < ny else True
                k = (i1 - 1) + nx * (j1 - 1)
                if c_ == d_ == True and a_ == b_ == False:  # noqa: E712 (Julia's chained ==)
                    ip = i1 + 1 if i1 == 1 else i1 - 1
                    A[k, :] = 0.0
                    A[k, k] -= 2.0 / dx
                    A[k, (ip - 1) + nx * (j1 - 1)] += 2.0 / dx
                    A[k, s] += 2.0
                if c_ != d_:
                    ip = i1 + 1 if i1 == 1 else i1 - 1
                    jp = j1 + 1 if c_ else j1 - 1
                    A[k, :] = 0.0
                    A[k, k] -= 4.0 / (3.0 * dx ** 2)
                    A[k, k] -= 4.0 / (3.0 * dy ** 2)
                    A[k, (ip - 1) + nx * (j1 - 1)] += 4.0 / (3.0 * dx ** 2)
                    A[k, (i1 - 1) + nx * (jp - 1)] += 4.0 / (3.0 * dy ** 2)
                    A[k, s] += (2.0 / 3.0) * (dx + dy) / (dx * dy)
    return A, b, rho, sig_rows


def electrode_potential(A, b, rho_dofs, rho, eps0):
    """calculate_electric_potential(ps, -rho) (:372-378) on that system."""
    bb = b.copy()
    nn = len(rho)
    bb[:nn][rho_dofs] = -np.asarray(rho)[rho_dofs] / eps0
    return np.linalg.solve(A, bb)[:nn]


class CircuitRLC:
    """CircuitRLC (Circuit.jl:13-22) with V(t) = v0 + v1 sin(2 pi f t + ph)
    and advance_circuit! (Circuit.jl:117-136)."""

    def __init__(self, R, L, C, v0=0.0, v1=0.0, f=0.0, ph=0.0):
        self.R, self.L, self.C = R, L, C
        self.V = lambda t: v0 + v1 * math.sin(2 * math.pi * f * t + ph)
        self.i = self.q = self.t = 0.0

    def advance(self, vext, dt):
        t, i, q = self.t, self.i, self.q
        R, L, C = self.R, self.L, self.C
        ni = (L / dt - R / 2) * i + vext - self.V(t)
        if C > 0.0:
            ni -= q / C
            ni /= (L / dt + R / 2)
            self.q = q + dt * i
        else:
            ni /= (L / dt + R / 2)
        self.i = ni
        self.t += dt


def tracker_table(st, nx, ny, electrodes=()):
    """The restated tracker as the device's edge table (codes 0 none,
    1 absorbing or fixed electrode, 2 reflective, 3 periodic, 16 + k the k-th
    floating electrode of `electrodes`) and its tracked cells."""
    ncx, ncy = nx - 1, ny - 1
    codes = np.zeros(4 * ncx * ncy, np.uint8)
    tracked = np.zeros(ncx * ncy, np.uint8)
    for ((i, j), (k, l)), s in st.surface.items():
        ci, cj = i - 1, j - 1
        d = {(-1, 0): 0, (1, 0): 1, (0, -1): 2, (0, 1): 3}[(k - i, l - j)]
        if s == ABSORB:
            code = 1
        elif s == REFLECT:
            code = 2
        elif s == PASS:
            code = 3
        else:
            code = 1 if s.fixed else 16 + electrodes.index(s)
        codes[4 * (ci + ncx * cj) + d] = code
        for a, b in ((i, j), (k, l)):
            if 1 <= a <= ncx and 1 <= b <= ncy:
                tracked[(a - 1) + ncx * (b - 1)] = 1
    return codes, tracked


def boundaries_geometry(nx=11, ny=11):
    """The surfaces of problem/07_boundaries.jl: a floating (driven) left wall,
    a floating segment on the right wall, two fixed nodes, a reflective block."""
    bcs = np.zeros((nx, ny), np.int8)
    bcs[0, 1:ny - 1] = 1
    bcs[nx - 1, 4:7] = 2
    bcs[nx - 2, 0] = 3
    bcs[nx - 2, ny - 1] = 3
    bcs[5:8, 4:7] = 4
    return bcs
