"""Problem definitions of the benchmark workloads (synthetic inputs only).

  two_stream_1d   1-D periodic two-stream instability, 64 cells, 20,000
                  electrons in two beams, immobile uniform ion background,
                  normalised units (eps0 = 1, omega_p dt = 0.1, k v0/omega_p = 0.9),
                  nearest reference script problem/ts_unstable.jl.
  slab_2d         2-D plasma slab, 256 x 256 cells, Dirichlet phi = 0 walls,
                  electrons + ions (m_i = 1836 m_e), Te = 2 eV, Ti = 0.03 eV,
                  absorbing walls.
  ccp_1d          1-D capacitively coupled RF discharge, 512 cells, 2.5 cm gap,
                  13.56 MHz 200 V electrode, argon-like gas at 1 Torr / 300 K,
                  synthetic cross sections, nearest reference script
                  problem/11_rf_discharge.jl.
"""
import numpy as np

QE = 1.6021766208e-19       # problem/units_and_constants.jl:39
ME = 9.1093837015e-31       # problem/units_and_constants.jl:40
MP = 1.6726218982e-27       # problem/units_and_constants.jl:41
EPS0 = 8.8541878128e-12     # Unitful eps0 (CODATA 2018)
KB = 1.380649e-23           # Unitful k (CODATA 2018)


def thermal_speed(T, m):
    """problem/configuration.jl:79-81 (sqrt(2 kB T / m))."""
    return np.sqrt(2 * KB * T / m)


# --------------------------------------------------------------------------
def two_stream_1d(cells=64, particles=20000, eps=1e-3, v0=1.0, kv0_wp=0.9, wp_dt=0.1):
    """Returns dict with grid, species state and parameters (normalised units)."""
    wp = 1.0
    k = kv0_wp * wp / v0
    L = 2 * np.pi / k
    dx = L / cells
    n0 = 1.0
    q, m, eps0 = -1.0, 1.0, 1.0
    w0 = n0 * L / particles
    dt = wp_dt / wp
    half = particles // 2
    xu = (np.arange(particles) + 0.5) * (L / particles)
    x = xu + (eps / k) * np.sin(k * xu)
    x = np.mod(x, L)
    v = np.zeros((particles, 3))
    v[:half, 0] = +v0
    v[half:, 0] = -v0
    # interleave the beams so both fill the domain uniformly
    order = np.empty(particles, dtype=np.int64)
    order[0::2] = np.arange(half)
    order[1::2] = np.arange(half, particles)
    v = v[order]
    return dict(nodes=cells + 1, dx=dx, L=L, x=x[:, None], v=v, q=q, m=m, w0=w0, eps0=eps0,
                dt=dt, bg=-q * n0, k=k, wp=wp)


# --------------------------------------------------------------------------
def slab_2d(cells=256, ppc=256, n0=1e16, Te=2.0, Ti=0.03, mi_over_me=1836.0):
    """2-D slab; Debye length = dx, omega_pe dt = 0.1 (reference Maxwellian
    convention sigma = sqrt(2 kB T / m))."""
    Te_K = Te * QE / KB
    Ti_K = Ti * QE / KB
    wpe = np.sqrt(n0 * QE ** 2 / (EPS0 * ME))
    lam = np.sqrt(EPS0 * Te * QE / (n0 * QE ** 2))
    dx = lam
    L = cells * dx
    n_part = ppc * cells * cells
    w0 = n0 * L * L / n_part
    dt = 0.1 / wpe
    mi = mi_over_me * ME
    return dict(nodes=cells + 1, dx=dx, L=L, n_part=n_part, w0=w0, dt=dt, eps0=EPS0,
                ve=thermal_speed(Te_K, ME), vi=thermal_speed(Ti_K, mi), me=ME, mi=mi, q=QE)


# --------------------------------------------------------------------------
def _log_peak(threshold, amp, emax=1000.0, n=120):
    """sigma = 0 at threshold, amp * ln(E/Eth)/(E/Eth) * e above it."""
    e = threshold * np.exp(np.linspace(0.0, np.log(emax / threshold), n))
    r = e / threshold
    s = amp * np.e * np.log(r) / r
    return e, s


def ccp_cross_sections():
    """Synthetic electron-argon-like and ion cross sections (m^2 vs eV)."""
    e_el = np.concatenate([[0.0], np.exp(np.linspace(np.log(0.01), np.log(1000.0), 150))])
    s_el = 2.0e-19 / np.sqrt(np.maximum(e_el, 0.04))
    ex = _log_peak(11.5, 1.0e-20)
    iz = _log_peak(15.8, 2.5e-20)
    e_ion = np.array([0.0, 1000.0])
    s_ion = np.array([5.0e-19, 5.0e-19])
    return dict(electron=[(e_el, s_el), ex, iz], energies=[0.0, 11.5, 15.8],
                ion=[(e_ion, s_ion)])


def ccp_1d(cells=512, gap=0.025, electrons=1_000_000, ions=1_000_000, Te=2.5, Ti=0.026,
           pressure_torr=1.0, Tg=300.0, volts=200.0, freq=13.56e6, n0=1e15, dt=None):
    dx = gap / cells
    ng = pressure_torr * 133.322368 / (KB * Tg)
    mi = 39.948 * 1.66053906660e-27
    if dt is None:
        dt = 1.0 / (freq * 6400)
    w0 = n0 * gap / electrons
    return dict(nodes=cells + 1, dx=dx, L=gap, ng=ng, mi=mi, me=ME, q=QE, w0=w0, dt=dt,
                ne=electrons, ni=ions, ve=thermal_speed(Te * QE / KB, ME),
                vi=thermal_speed(Ti * QE / KB, mi), volts=volts, freq=freq, Tg=Tg,
                eps0=EPS0, xs=ccp_cross_sections())
