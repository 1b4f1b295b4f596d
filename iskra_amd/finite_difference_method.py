"""FiniteDifferenceMethod: the Poisson solver and E = -grad(phi)
(FiniteDifferenceMethod/src/generalized_poisson.jl).

`create_poisson_solver` records the operator (Dirichlet masks and values,
periodic rings, eps0); the solve itself runs on the GPU (libiskra_hip.so):
the tridiagonal direct solve in 1-D, multigrid or preconditioned CG in 2-D,
all to a relative residual ||b - A phi|| / ||b|| of `rtol` in FP64.  The
reference factors a dense matrix every step (generalized_poisson.jl:201-203);
the same operator is solved here, Dirichlet columns moved to the right-hand
side.
"""
import numpy as np

from . import _lib as L
from ._lib import Context, IskraError


class PoissonSolver:
    """PoissonSolver{:xy, D} (generalized_poisson.jl:11-20): the grid, eps0,
    the Dirichlet masks (one tag per distinct mask) and periodic links."""

    def __init__(self, grid, eps0, kind=None, rtol=1e-8, max_iter=0):
        self.grid = grid
        self.eps0 = float(eps0)
        self.tags = np.zeros(grid.n, dtype=np.int8)
        self.masks = []          # list of boolean masks, tag = index + 1
        self.values = []         # phi0 per tag
        self.periodic = [False, False]  # ring along x (first index), y (second index)
        self.kind = kind
        self.rtol = float(rtol)
        self.max_iter = int(max_iter)
        self.warm = L.WARM_PREV
        self.last_iters = 0
        self.last_relres = 0.0
        self._ctx = None         # context the operator is bound to
        self._own = False        # True when the solver created its context itself
        self._dirty = True       # operator changed since it was bound

    # -- binding -----------------------------------------------------------
    def _solver_kind(self):
        if self.kind is not None:
            return self.kind
        if self.grid.dim == 1:
            return L.POISSON_DIRECT
        if getattr(self.grid, "kind", "xy") == "rz":
            return L.POISSON_PCG  # the axisymmetric operator (generalized_poisson.jl:70-199)
        walls = (self.tags[0, :].all() and self.tags[-1, :].all() and self.tags[:, 0].all()
                 and self.tags[:, -1].all())
        interior = self.tags[1:-1, 1:-1].any()
        return L.POISSON_MG if walls and not interior else L.POISSON_PCG

    def _configure(self, ctx):
        """Send the operator to `ctx` (its grid must be this solver's grid)."""
        kind = self._solver_kind()
        tags = self.tags if self.masks else None
        try:
            ctx.poisson(kind, self.eps0, None if tags is None else tags.T.ravel(),
                        periodic_x=self.periodic[0], periodic_y=self.periodic[1],
                        rtol=self.rtol, max_iter=self.max_iter)
        except IskraError as err:
            if kind != L.POISSON_MG or err.code != 7:
                raise
            # multigrid needs 2^k+1 nodes per side: use preconditioned CG instead
            ctx.poisson(L.POISSON_PCG, self.eps0, None if tags is None else tags.T.ravel(),
                        periodic_x=self.periodic[0], periodic_y=self.periodic[1],
                        rtol=self.rtol, max_iter=self.max_iter)
        for t, v in enumerate(self.values, start=1):
            ctx.tag_value(t, v)
        self._ctx = ctx
        self._dirty = False

    def _context(self):
        if self._ctx is None:
            from .particle_in_cell import _grid_to
            ctx = Context(0)
            _grid_to(ctx, self.grid)
            self._own = True
            self._configure(ctx)
        elif self._dirty:
            self._configure(self._ctx)
        return self._ctx


def create_poisson_solver(grid, eps0, kind=None, rtol=1e-8, max_iter=0):
    """create_poisson_solver(grid, eps0) (generalized_poisson.jl:22-30): the
    natural-edge 3/5-point Laplacian over the grid's nodes.  kind (optional)
    picks the device solver (POISSON_DIRECT, PCG, MG, MGPCG); by default 1-D
    uses the direct solve, 2-D multigrid when all four walls are Dirichlet and
    PCG otherwise."""
    return PoissonSolver(grid, eps0, kind, rtol, max_iter)


def apply_dirichlet(ps, nodes, phi0):
    """apply_dirichlet(ps, nodes, phi0) (generalized_poisson.jl:205-215):
    phi = phi0 on the masked nodes.  Calling it again with the same mask only
    changes the value (as scripts do for an RF electrode every step)."""
    mask = np.asarray(nodes, dtype=bool).reshape(ps.grid.n)
    for t, m in enumerate(ps.masks, start=1):
        if np.array_equal(m, mask):
            ps.values[t - 1] = float(phi0)
            if ps._ctx is not None and not ps._dirty:
                ps._ctx.tag_value(t, float(phi0))
            return
    if len(ps.masks) >= 127:
        raise ValueError("apply_dirichlet: at most 127 distinct Dirichlet masks")
    ps.masks.append(mask.copy())
    ps.values.append(float(phi0))
    ps.tags[mask] = len(ps.masks)
    ps._dirty = True


def apply_periodic(ps, axis=None):
    """apply_periodic(ps[, axis]) (generalized_poisson.jl:271-324): link the
    first and last node of an axis as neighbours.  In 2-D the reference's
    axis = 1 links the second index (j) and axis = 2 the first (i); that
    labelling is kept.  1-D takes no axis.  On an axial grid the reference
    cannot apply it (generalized_poisson.jl:326-361 asserts axis != 1 and
    reads undefined names for axis 2): refused."""
    if getattr(ps.grid, "kind", "xy") == "rz":
        raise IskraError(L.E_UNSUPPORTED, "apply_periodic: not available on an axial grid")
    if ps.grid.dim == 1:
        if axis not in (None, 1):
            raise ValueError("apply_periodic: a 1-D solver has no axis argument")
        ps.periodic[0] = True
    elif axis == 1:
        ps.periodic[1] = True
    elif axis == 2:
        ps.periodic[0] = True
    else:
        raise ValueError("apply_periodic: axis must be 1 or 2")
    ps._dirty = True


def calculate_electric_potential(ps, f):
    """calculate_electric_potential(ps, f) (generalized_poisson.jl:372-378):
    solve A phi = f/eps0 at the free nodes (the loop passes f = -rho); returns
    phi with the grid's shape.  When the solver is bound to a running
    simulation and f is None, the device's own rho is used."""
    ctx = ps._context()
    if f is not None:
        rho = -np.asarray(f, dtype=np.float64).reshape(ps.grid.n)
        ctx.set_field(L.F_RHO, rho.T.ravel())
    it, rr = ctx.solve(ps.warm)
    ps.last_iters, ps.last_relres = it, rr
    return ctx.field(L.F_PHI).reshape(ps.grid.n[::-1]).T.copy()


def calculate_electric_field(ps, phi):
    """calculate_electric_field(ps, phi) (generalized_poisson.jl:380-410):
    E = -grad(phi), central differences inside, one-sided at every end.
    Returns nx x 3 (1-D) or nx x ny x 3 (2-D); E is stored in FP32 on the
    device."""
    ctx = ps._context()
    g = ps.grid
    if phi is not None:
        ctx.set_field(L.F_PHI, np.asarray(phi, dtype=np.float64).reshape(g.n).T.ravel())
    ctx.efield()
    E = np.zeros(g.n + (3,))
    E[..., 0] = ctx.field(L.F_EX).reshape(g.n[::-1]).T
    if g.dim == 2:
        E[..., 1] = ctx.field(L.F_EY).reshape(g.n[::-1]).T
    return E


def calculate_magnetic_field(ps):
    """calculate_magnetic_field (generalized_poisson.jl:412-419): B = 0."""
    return np.zeros(ps.grid.n + (3,))
