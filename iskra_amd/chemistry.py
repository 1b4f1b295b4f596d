"""Chemistry: cross sections, reactions and null-collision Monte Carlo
collisions (Chemistry/src/{cross_section,reactions,mcc}.jl).

`reactions([...])` plays the part of the `@reactions` macro: each entry is
(rate, reactants, products[, type]) with species objects instead of symbols,
in the order the macro would list them.  `mcc(reactions)` builds the
MonteCarloCollisions interaction; it runs on the GPU inside the loop
(PIC.perform!, mcc.jl:231-289).
"""
import numpy as np

from . import _lib as L
from .particle_in_cell import FluidSpecies, KineticSpecies


class CrossSection:
    """CrossSection (cross_section.jl:3-17): piecewise-linear sigma(eps) over
    table nodes (eps in eV, sigma in m^2), flat outside the table."""

    def __init__(self, xs, ys=None):
        if ys is None:
            nodes = np.asarray(xs, dtype=np.float64)
            xs, ys = nodes[:, 0], nodes[:, 1]
        self.nodes = np.column_stack([np.asarray(xs, dtype=np.float64),
                                      np.asarray(ys, dtype=np.float64)])

    def maximum(self):
        return float(self.nodes[:, 1].max())

    def argmax(self):
        return float(self.nodes[np.argmax(self.nodes[:, 1]), 0])


# collision types (mcc.jl:2-16)
class ElasticIsotropic:
    code = L.COLL_ELASTIC_ISOTROPIC
    energy = 0.0


class ElasticBackward:
    code = L.COLL_ELASTIC_BACKWARD
    energy = 0.0


class InelasticBackward:
    """Backward scattering that keeps a uniform fraction of the relative
    speed (mcc.jl:144-157)."""
    code = L.COLL_INELASTIC_BACKWARD
    energy = 0.0


class Excitation:
    code = L.COLL_EXCITATION

    def __init__(self, energy):
        self.energy = float(energy)


class Ionization:
    code = L.COLL_IONIZATION

    def __init__(self, energy):
        self.energy = float(energy)


class ChemicalReaction:
    """ChemicalReaction (reactions.jl:7-12): type, rate, reactants and the
    net stoichiometry (species -> net count)."""

    def __init__(self, type_, rate, reactants, stoichiometry):
        self.type = type_
        self.rate = rate
        self.reactants = reactants
        self.stoichiometry = stoichiometry


def reactions(entries):
    """The @reactions macro (reactions.jl:3-55): entries of
    (rate, reactants, products[, type]); reactants/products are lists of
    species (repeats count)."""
    out = []
    for e in entries:
        rate, reac, prod = e[0], list(e[1]), list(e[2])
        type_ = e[3] if len(e) > 3 else None
        order = []
        for s in reac + prod:
            if not any(s is o for o in order):
                order.append(s)
        stoich = []
        reactants = []
        for s in order:
            r = sum(1 for t in reac if t is s)
            p = sum(1 for t in prod if t is s)
            if r:
                reactants.append((s, r))
            if p - r:
                stoich.append((s, p - r))
        out.append(ChemicalReaction(type_, rate, reactants, stoich))
    return out


class Collision:
    """MCC.Collision (mcc.jl:10-16)."""

    def __init__(self, type_, rate, source, target, products):
        self.type = type_
        self.rate = rate
        self.source = source
        self.target = target
        self.products = products


class MonteCarloCollisions:
    """MonteCarloCollisions (mcc.jl:18-51): the collisions of one source
    species with one fluid target, in reaction order."""

    def __init__(self, collisions):
        if not collisions:
            raise ValueError("mcc: no collisions")
        src = collisions[0].source
        tgt = collisions[0].target
        for c in collisions:
            if c.source is not src or c.target is not tgt:
                raise ValueError("mcc: all collisions must share one source and one target")
        self.collisions = collisions
        self.source = src
        self.target = tgt
        self._ctx = None
        self._id = -1
        self.events = np.zeros(len(collisions), dtype=np.int64)

    def _bind(self, ctx, grid):
        types, energy, tables = [], [], []
        ion_sid = -1
        for c in self.collisions:
            types.append(c.type.code)
            energy.append(c.type.energy)
            tables.append((c.rate.nodes[:, 0], c.rate.nodes[:, 1]))
            if c.type.code == L.COLL_IONIZATION:
                ions = [p for p in c.products if p is not self.source]
                if len(ions) != 1:
                    raise ValueError("mcc: ionisation needs exactly one ion product species")
                ion_sid = ions[0]._sid
        tgt = self.target
        self._id = ctx.mcc(self.source._sid, types, energy, tables, ion=ion_sid,
                           target_m=tgt.m, target_q=tgt.q, target_T=tgt.T)
        n = np.broadcast_to(tgt.n, grid.n)
        ctx.set_field(L.F_TARGET, np.ascontiguousarray(n.T).ravel(), index=self._id)
        self._ctx = ctx

    def counts(self, reset=False):
        """nu per process and node (mcc.jl:283-288), shape (processes, nodes),
        summed over the steps since binding (or since the last reset)."""
        return self._ctx.mcc_counts(self._id, len(self.collisions), reset)

    def _last_step(self):
        """the reference's nu: the events of the last step"""
        return self._ctx.mcc_counts_last(self._id, len(self.collisions))


def _accept(r):
    """accept(::ChemicalReaction) (mcc.jl:291-312)."""
    source = target = None
    if len(r.reactants) != 2:
        raise ValueError("Monte Carlo Collisions support only two reacting species: one fluid "
                         "and one kinetic")
    for s, _ in r.reactants:
        if isinstance(s, FluidSpecies):
            target = s
        elif isinstance(s, KineticSpecies):
            source = s
    if source is None:
        raise ValueError("Reaction without particle species")
    if target is None:
        raise ValueError("Reaction without fluid species")
    products = [s for s, c in r.stoichiometry if c > 0]
    type_ = r.type if r.type is not None else ElasticIsotropic()
    return Collision(type_, r.rate, source, target, products)


def mcc(reaction_list):
    """mcc(reactions) (mcc.jl:314-320)."""
    return MonteCarloCollisions([_accept(r) for r in reaction_list])


# ---------------------------------------------------------------------------
# direct simulation Monte Carlo (Chemistry/src/dsmc.jl)

class DSMCElasticCollision:
    """DSMC.ElasticCollision (dsmc.jl:2-6): rate is a CrossSection over the
    relative speed (m/s)."""

    def __init__(self, rate, source, target):
        self.rate = rate
        self.source = source
        self.target = target


class DSMCIonizationCollision:
    """DSMC.IonizationCollision (dsmc.jl:8-13).  The reference defines no
    perform! for it, so a run that draws one fails there; binding it here is
    refused."""

    def __init__(self, rate, source, target, products):
        self.rate = rate
        self.source = source
        self.target = target
        self.products = products


class DirectSimulationMonteCarlo:
    """DirectSimulationMonteCarlo (dsmc.jl:16-23): pairing collisions between
    two kinetic species per cell, run on the GPU inside the loop
    (PIC.perform!, dsmc.jl:81-142).  One reaction per interaction (the
    reference's candidate lists are shared between the reactions of one
    interaction and only one reaction gives defined results)."""

    def __init__(self, collisions):
        self.collisions = list(collisions)
        self._ctx = None
        self._id = -1

    def _bind(self, ctx, grid):
        if len(self.collisions) != 1:
            raise L.IskraError(L.E_UNSUPPORTED, "dsmc: one reaction per interaction (the "
                               "reference appends every reaction's particles to the same "
                               "cell lists, dsmc.jl:94-96)")
        c = self.collisions[0]
        if isinstance(c, DSMCIonizationCollision):
            raise L.IskraError(L.E_UNSUPPORTED, "dsmc: the reference has no perform! for "
                               "DSMC.IonizationCollision")
        self._id = ctx.dsmc(c.source._sid, c.target._sid, c.rate.nodes[:, 0], c.rate.nodes[:, 1])
        self._ctx = ctx

    def counts(self, reset=False):
        """nuDSMC per node (the cell's lower-left node), summed over the steps
        since binding (or since the last reset)."""
        return self._ctx.dsmc_counts(self._id, reset)

    def _last_step(self):
        """the reference's nuDSMC: the events of the last step"""
        return self._ctx.dsmc_counts_last(self._id)


def dsmc(reaction_list):
    """dsmc(reactions) (dsmc.jl:144-167): every reaction has exactly two
    kinetic reactants; with products it is an ionisation, else elastic."""
    out = []
    for r in reaction_list:
        if len(r.reactants) != 2:
            raise ValueError("Direct Simulation Monte Carlo support only two reacting, "
                             "kinetic species")
        source, _ = r.reactants[0]
        target, _ = r.reactants[1]
        products = [s for s, c in r.stoichiometry if c > 0]
        if source is None:
            raise ValueError("Reaction without source species")
        if target is None:
            raise ValueError("Reaction without target species")
        if products:
            out.append(DSMCIonizationCollision(r.rate, source, target, products))
        else:
            out.append(DSMCElasticCollision(r.rate, source, target))
    return DirectSimulationMonteCarlo(out)
