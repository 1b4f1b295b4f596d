"""RegularGrids: uniform Cartesian and axial (r, z) grids
(RegularGrids/src/RegularGrids.jl).

Grid objects are host metadata (node counts, spacing, boundary symbols,
origin); the device copy is made when a solver or a simulation binds the grid
to a context.  Node counts follow the reference: n cells give n+1 nodes
(`xs = 0:dh:L`).
"""
import numpy as np

OPEN = "open"
PERIODIC = "periodic"
OTHER = "other"


class UniformGrid:
    """CartesianGrid{D} (RegularGrids.jl:7-20): n nodes per dimension, spacing
    dh, boundary symbols bcs = ((left, right), (bottom, top)), origin, node
    coordinates."""

    def __init__(self, n, dh, bcs, origin, coords, kind="xy"):
        self.kind = kind  # "xy": CartesianGrid, "rz": AxialGrid (RegularGrids.jl:17-18)
        self.n = tuple(int(k) for k in n)
        self.dh = tuple(float(h) for h in dh)
        self.bcs = bcs
        self.origin = tuple(float(o) for o in origin)
        self.coords = coords
        self.data = {}

    @property
    def dim(self):
        return len(self.n)

    # Base.size / Base.length (RegularGrids.jl:22-23)
    def size(self):
        return self.n

    def __len__(self):
        return int(np.prod(self.n))

    def __getitem__(self, k):
        return self.data[k]

    def __setitem__(self, k, v):
        if np.size(v) != len(self):
            raise ValueError("Size mismatch!")
        self.data[k] = v

    def vol_periodic(self):
        """Flags (left, right, bottom, top) declared :periodic, as the device
        grid takes them (they switch off the edge halving of cell_volume)."""
        flags = [b == PERIODIC for pair in self.bcs for b in pair]
        return tuple(int(f) for f in (flags + [False] * 4)[:4])


def _spacing(xx):
    xx = np.asarray(xx, dtype=np.float64)
    return float(xx[1] - xx[0]) if len(xx) > 1 else 1.0


def create_uniform_grid(xx, yy=None, left=OPEN, right=OPEN, bottom=OPEN, top=OPEN):
    """create_uniform_grid(xx[, yy]; left, right, bottom, top)
    (RegularGrids.jl:55-82): node coordinates xx (and yy) as ranges."""
    xx = np.asarray(xx, dtype=np.float64)
    if yy is None:
        return UniformGrid((len(xx),), (_spacing(xx),), ((left, right),), (xx[0],), (xx.copy(),))
    yy = np.asarray(yy, dtype=np.float64)
    nx, ny = len(xx), len(yy)
    x = np.repeat(xx[:, None], ny, axis=1)
    y = np.repeat(yy[None, :], nx, axis=0)
    return UniformGrid((nx, ny), (_spacing(xx), _spacing(yy)), ((left, right), (bottom, top)),
                       (xx[0], yy[0]), (x, y))


def create_axial_grid(rr, zz, bottom=OPEN, top=OPEN):
    """create_axial_grid(rr, zz; bottom, top) (RegularGrids.jl:84-97): radial
    node coordinates rr (the axis at rr[0]) and axial zz; the radial bcs are
    (:other, :other)."""
    rr = np.asarray(rr, dtype=np.float64)
    zz = np.asarray(zz, dtype=np.float64)
    nr, nz = len(rr), len(zz)
    r = np.repeat(rr[:, None], nz, axis=1)
    z = np.repeat(zz[None, :], nr, axis=0)
    return UniformGrid((nr, nz), (_spacing(rr), _spacing(zz)), ((OTHER, OTHER), (bottom, top)),
                       (rr[0], zz[0]), (r, z), kind="rz")


def cell_volume(g):
    """cell_volume(::CartesianGrid{2}) (RegularGrids.jl:26-38): dx*dy per node,
    halved on each edge that is not :periodic (corners quartered).  The 1-D
    form (absent from the reference) is dx with the same end halving."""
    if g.dim == 1:
        (left, right), = g.bcs
        V = np.full(g.n[0], g.dh[0])
        if left != PERIODIC:
            V[0] *= 0.5
        if right != PERIODIC:
            V[-1] *= 0.5
        return V
    (left, right), (bottom, top) = g.bcs
    if g.kind == "rz":
        # cell_volume(::AxialGrid{2}) (RegularGrids.jl:40-53), 1-based i
        dr, dz = g.dh
        nr, nz = g.n
        V = np.zeros(g.n)
        for i in range(2, nr):
            V[i - 1, :] = np.pi * dz * ((i * dr - 0.5 * dr) ** 2 - (i * dr - 1.5 * dr) ** 2)
        V[0, :] = np.pi * dz * (0.5 * dr) ** 2
        V[nr - 1, :] = np.pi * dz * ((nr * dr - 1.0 * dr) ** 2 - (nr * dr - 1.5 * dr) ** 2)
        if bottom != PERIODIC:
            V[:, 0] *= 0.5
        if top != PERIODIC:
            V[:, -1] *= 0.5
        return V
    V = np.full(g.n, g.dh[0] * g.dh[1])
    if left != PERIODIC:
        V[0, :] *= 0.5
    if right != PERIODIC:
        V[-1, :] *= 0.5
    if bottom != PERIODIC:
        V[:, 0] *= 0.5
    if top != PERIODIC:
        V[:, -1] *= 0.5
    return V
