"""Circuit: the series RLC circuit and its netlist (Circuit/src/Circuit.jl),
advanced on the GPU inside the time loop (advance!(::CircuitRLC),
ParticleInCell/src/pic/circuit_coupling.jl:35-43).

`netlist([...])` plays the part of the `@netlist` macro: elements named
V*, R*, L*, C*, EXT* with their nodes (recorded, unused, as in the reference)
and value.  The voltage source is a sine, V(t) = offset + amplitude
sin(2 pi freq t + phase) (the reference takes any Julia function of t).
"""
from . import _lib as L


class Resistor:
    def __init__(self, name, val):
        self.name, self.val = name, float(val)


class Inductor:
    def __init__(self, name, val):
        self.name, self.val = name, float(val)


class Capacitor:
    def __init__(self, name, val):
        self.name, self.val = name, float(val)


class SineVoltage:
    """V(t) = offset + amplitude sin(2 pi freq t + phase)."""

    def __init__(self, amplitude=0.0, freq=0.0, phase=0.0, offset=0.0):
        self.amplitude, self.freq, self.phase, self.offset = amplitude, freq, phase, offset


class VoltageSource:
    def __init__(self, name, val):
        if not isinstance(val, SineVoltage):
            raise L.IskraError(L.E_UNSUPPORTED, "VoltageSource: the device evaluates "
                               "V(t) = offset + amplitude sin(2 pi freq t + phase) "
                               "(pass a SineVoltage)")
        self.name, self.val = name, val


class PlasmaDevice:
    """PlasmaDevice (circuit_coupling.jl:17-24): a floating (positive) and a
    fixed (negative) electrode."""

    def __init__(self, positive, negative):
        self.positive, self.negative = positive, negative


class ExternalDevice:
    def __init__(self, name, val):
        self.name, self.val = name, val


class CircuitRLC:
    """CircuitRLC (Circuit.jl:13-22) built by rlc; i, q, t from the device."""

    def __init__(self):
        self.R = self.L = self.C = 0.0
        self.V = SineVoltage()
        self.ext = None
        self._ctx = None

    def _bind(self, ctx):
        pos = neg = -1
        if isinstance(self.ext, PlasmaDevice):
            pos, neg = self.ext.positive._id, self.ext.negative._id
        ctx.circuit(self.R, self.L, self.C, self.V.offset, self.V.amplitude, self.V.freq,
                    self.V.phase, positive=pos, negative=neg)
        self._ctx = ctx

    def state(self):
        """current i, capacitor charge q, time t and the last external voltage"""
        return self._ctx.circuit_state()


def netlist(entries):
    """@netlist (Circuit.jl:67-108): entries (name, node+, node-, value)."""
    out = []
    for name, _np, _nm, val in entries:
        if name.startswith("V"):
            out.append(VoltageSource(name, val))
        elif name.startswith("R"):
            out.append(Resistor(name, val))
        elif name.startswith("L"):
            out.append(Inductor(name, val))
        elif name.startswith("C"):
            out.append(Capacitor(name, val))
        elif name.startswith("EXT"):
            out.append(ExternalDevice(name, val))
        else:
            raise ValueError("unknown element")
    return out


def rlc(elements):
    """rlc(elements) (Circuit.jl:58-66): assign each element in order."""
    c = CircuitRLC()
    for e in elements:
        if isinstance(e, Resistor):
            c.R = e.val
        elif isinstance(e, Inductor):
            c.L = e.val
        elif isinstance(e, Capacitor):
            c.C = e.val
        elif isinstance(e, VoltageSource):
            c.V = e.val
        elif isinstance(e, ExternalDevice):
            c.ext = e.val
    return c
