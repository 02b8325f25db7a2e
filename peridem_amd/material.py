"""Peridynamic material parameters, derived the way the reference derives them.

Host-side setup: the per-particle constants the time step consumes.
  influence functions     src/material/mparticle/influenceFn.cpp:14-122
  RNP  computeParameters  src/material/mparticle/material.h:541-620
  RNP  material props     material.h:452-489 (K used for the broken-bond Kn)
  PDState parameters      material.h:1580-1700
  MatData conversions     src/inp/decks/materialDeck.h:125-228
  Kn, Rc of a particle    src/particle/baseParticle.cpp:170-179
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

from .capi import INF_CONST, INF_GAUSSIAN, INF_LINEAR, MAT_PDELASTIC, MAT_PDSTATE, MAT_PMB, MAT_RNP


@dataclass
class MaterialDeck:
    """Fields of inp::MaterialDeck (src/inp/decks/materialDeck.h:230-301) that the
    time step depends on.  Negative elastic constants mean "not given"."""
    type: str = "RNPBond"            # RNPBond | PMBBond | PDElasticBond | PDState
    E: float = -1.0
    K: float = -1.0
    G: float = -1.0
    nu: float = -1.0
    Gc: float = -1.0
    KIc: float = -1.0
    density: float = 1.0
    horizon: float = -1.0
    horizon_mesh_ratio: float = -1.0
    influence_fn_type: int = INF_CONST
    influence_fn_params: List[float] = field(default_factory=list)
    check_sc_factor: float = 1.0
    irreversible: bool = True
    compute_from_classical: bool = True
    bond_potential_params: List[float] = field(default_factory=list)
    plane_strain: bool = False       # Is_Plane_Strain (strain / stress output only)


class InfluenceFn:
    """Const / Linear / Gaussian influence function J(r), r = |xi| / delta."""

    def __init__(self, kind: int, params, dim: int):
        self.kind = kind
        params = list(params or [])
        if kind == INF_CONST:
            self.a0 = params[0] if params else float(dim + 1)
            self.p = (self.a0, 0.0)
        elif kind == INF_LINEAR:
            if not params:
                self.a0 = {1: 6.0, 2: 12.0, 3: 20.0}[dim]
                self.a1 = -self.a0
            else:
                self.a0 = params[0]
                self.a1 = -self.a0 if len(params) < 2 else params[1]
            self.p = (self.a0, self.a1)
        elif kind == INF_GAUSSIAN:
            if not params:
                b = 0.2
                if dim == 1:
                    a = 2.0 / (b * (1.0 - math.exp(-1.0 / b)))
                elif dim == 2:
                    a = (4.0 / b) * 1.0 / (math.sqrt(math.pi * b) * math.erf(1.0 / math.sqrt(b))
                                           - 2.0 * math.exp(-1.0 / b))
                else:
                    a = (2.0 / b) * 1.0 / (b - (b + 1.0) * math.exp(-1.0 / b))
                self.alpha, self.beta = a, b
            else:
                self.alpha, self.beta = params[0], params[1]
            self.p = (self.alpha, self.beta)
        else:
            raise ValueError(f"influence function type {kind} is invalid")

    def __call__(self, r: float) -> float:
        if self.kind == INF_CONST:
            return self.a0
        if self.kind == INF_LINEAR:
            return self.a0 + self.a1 * r
        return self.alpha * math.exp(-r * r / self.beta)

    def moment(self, i: int) -> float:
        if self.kind == INF_CONST:
            return self.a0 / float(i + 1)
        if self.kind == INF_LINEAR:
            return (self.a0 / float(i + 1)) + (self.a1 / float(i + 2))
        a, b = self.alpha, self.beta
        sq1, sq2 = math.sqrt(b), math.sqrt(math.pi)
        if i == 0:
            return 0.5 * a * sq1 * sq2 * math.erf(1.0 / sq1)
        if i == 1:
            return 0.5 * a * b * (1.0 - math.exp(-1.0 / b))
        if i == 2:
            return 0.5 * a * b * sq1 * (0.5 * sq2 * math.erf(1.0 / sq1)
                                       - (1.0 / sq1) * math.exp(-1.0 / b))
        if i == 3:
            return 0.5 * a * b * b * (1.0 - (1.0 + 1.0 / b) * math.exp(-1.0 / b))
        raise ValueError("moment index must be 0..3")


def _toK(E, nu):
    return E / (3.0 * (1.0 - 2.0 * nu))


def _toE(K, nu):
    return K * (3.0 * (1.0 - 2.0 * nu))


def _toLambdaE(E, nu):
    return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def _toGE(E, nu):
    return E / (2.0 * (1.0 + nu))


def _toELambda(lam, nu):
    return lam * (1.0 + nu) * (1.0 - 2.0 * nu) / nu


@dataclass
class MaterialParams:
    material_type: int
    mat: List[float]          # RNP {C, beta, rbar, inv_factor, check_sc_factor}; PDState {K, G, s0}
    K_for_contact: float      # computeMaterialProperties(dim).d_K
    irreversible: bool


def rnp_params(deck: MaterialDeck, dim: int, horizon: float, inf: InfluenceFn) -> MaterialParams:
    """RnpMaterial ctor + computeParameters (material.h:258-311, 541-620)."""
    if dim == 1:
        invF = math.pow(horizon, 2) * 2.0
    elif dim == 2:
        invF = math.pow(horizon, 3) * math.pi
    else:
        invF = math.pow(horizon, 4) * 4.0 * math.pi / 3.0
    if deck.compute_from_classical:
        E, K, Gc, KIc = deck.E, deck.K, deck.Gc, deck.KIc
        if E < 0 and K < 0:
            raise ValueError("RNP needs E or K")
        if Gc < 0 and KIc < 0:
            raise ValueError("RNP needs Gc or KIc")
        nu = 0.25
        if E > 0.0:
            K = _toK(E, nu)
        if K > 0.0 and E < 0.0:
            E = _toE(K, nu)
        if Gc > 0.0:
            KIc = math.sqrt(Gc * E)
        if KIc > 0.0 and Gc < 0.0:
            Gc = KIc * KIc / E
        lam = _toLambdaE(E, nu)
        M = inf.moment(dim)
        if dim == 2:
            Cc = math.pi * Gc / (4.0 * M)
            beta = 4.0 * lam / (Cc * M)
        elif dim == 3:
            Cc = 2.0 * Gc / (3.0 * M)
            beta = 5.0 * lam / (Cc * M)
        else:
            raise ValueError("RNP parameters are defined for dim 2 and 3")
    else:
        Cc, beta = deck.bond_potential_params[0], deck.bond_potential_params[1]
    rbar = math.sqrt(0.5 / beta)
    # computeMaterialProperties (material.h:452-489)
    M = inf.moment(dim)
    nu = 0.25
    lam = Cc * M * beta / 4.0 if dim == 2 else Cc * M * beta / 5.0
    Ep = _toELambda(lam, nu)
    Kp = _toK(Ep, nu)
    return MaterialParams(MAT_RNP, [Cc, beta, rbar, invF, deck.check_sc_factor, 0.0], Kp,
                          deck.irreversible)


def pdstate_params(deck: MaterialDeck, dim: int, horizon: float) -> MaterialParams:
    """PdState ctor + computeParameters (material.h:1336-1380, 1580-1700)."""
    if deck.compute_from_classical:
        E, K, G, nu, Gc, KIc = deck.E, deck.K, deck.G, deck.nu, deck.Gc, deck.KIc
        gt = lambda a: (a - 0.0) > abs(a) * 1e-5  # util::isGreater(a, 0.)
        fE, fK, fG, fnu = gt(E), gt(K), gt(G), gt(nu)
        if sum([fE, fK, fG, fnu]) != 2:
            raise ValueError("PDState needs exactly two of E, K, G, nu")
        if not fnu:
            if fE and fG:
                nu = (0.5 * E / G) - 1.0
            if fE and fK:
                nu = (3.0 * K - E) / (6.0 * K)
            if fG and fK:
                nu = (3.0 * K - 2.0 * G) / (2.0 * (3.0 * K + G))
        if not fE:
            if fK:
                E = _toE(K, nu)
            if fG:
                E = 2.0 * G * (1.0 + nu)
        if not fK:
            K = _toK(E, nu)
        if not fG:
            G = _toGE(E, nu)
        if Gc > 0.0:
            KIc = math.sqrt(Gc * E)
        if KIc > 0.0 and Gc < 0.0:
            Gc = KIc * KIc / E
        d = (3.0 * G + math.pow(3.0 / 4.0, 4) * (K - 5.0 * G / 3.0)) * horizon
        s0 = math.sqrt(Gc / d)
    else:
        K, G, s0 = deck.bond_potential_params[:3]
    return MaterialParams(MAT_PDSTATE, [K, G, s0, 0.0, 0.0, 0.0], K, True)


def _bond_elastic(deck: MaterialDeck):
    """E (or K) with nu = 1/4 and lambda, as PmbMaterial / PdElastic
    computeParameters set them (material.h:925-1000, 1260-1306)."""
    E, K = deck.E, deck.K
    if E < 0 and K < 0:
        raise ValueError("bond-based material needs E or K")
    nu = 0.25
    if E > 0.0:
        K = _toK(E, nu)
    if K > 0.0 and E < 0.0:
        E = _toE(K, nu)
    return E, nu, _toLambdaE(E, nu)


def pmb_params(deck: MaterialDeck, dim: int, horizon: float) -> MaterialParams:
    """PmbMaterial ctor + computeParameters + computeMaterialProperties
    (material.h:674-716, 846-870, 925-1000)."""
    if deck.compute_from_classical:
        E, nu, lam = _bond_elastic(deck)
        Gc, KIc = deck.Gc, deck.KIc
        if Gc < 0 and KIc < 0:
            raise ValueError("PMB needs Gc or KIc")
        if KIc > 0.0 and Gc < 0.0:
            Gc = KIc * KIc / E
        if dim == 2:
            c = 24.0 * E / (math.pi * math.pow(horizon, 3.0) * (1.0 - nu))
        else:
            c = 24.0 * lam / (math.pi * math.pow(horizon, 3.0))
        s0 = math.sqrt(5.0 * math.pi * Gc / (9.0 * E * horizon))
    else:
        c, s0 = deck.bond_potential_params[0], deck.bond_potential_params[1]
    lam_p = c * (math.pi * math.pow(horizon, 3.0)) / 24.0
    Kp = _toK(_toELambda(lam_p, 0.25), 0.25)
    return MaterialParams(MAT_PMB, [c, s0, 0.0, 0.0, 0.0, 0.0], Kp, True)


def pdelastic_params(deck: MaterialDeck, dim: int, horizon: float) -> MaterialParams:
    """PdElastic ctor + computeParameters + computeMaterialProperties
    (material.h:1034-1076, 1189-1213, 1260-1306)."""
    if deck.compute_from_classical:
        E, nu, lam = _bond_elastic(deck)
        c = 24.0 * lam / (math.pi * math.pow(horizon, 3.0))
    else:
        c = deck.bond_potential_params[0]
    lam_p = c * (math.pi * math.pow(horizon, 3.0)) / 24.0
    Kp = _toK(_toELambda(lam_p, 0.25), 0.25)
    return MaterialParams(MAT_PDELASTIC, [c, 0.0, 0.0, 0.0, 0.0, 0.0], Kp, False)


def material_params(deck: MaterialDeck, dim: int, horizon: float, inf: InfluenceFn) -> MaterialParams:
    if deck.type == "RNPBond":
        return rnp_params(deck, dim, horizon, inf)
    if deck.type == "PDState":
        return pdstate_params(deck, dim, horizon)
    if deck.type == "PMBBond":
        return pmb_params(deck, dim, horizon)
    if deck.type == "PDElasticBond":
        return pdelastic_params(deck, dim, horizon)
    raise ValueError(f"material type {deck.type} is not supported (RNPBond, PMBBond, PDElasticBond, PDState)")


def elastic_constants(mp: MaterialParams, dim: int, horizon: float, inf: InfluenceFn):
    """(nu, lambda, mu) of computeMaterialProperties(dim): RNP
    (material.h:461-489), PMB (:846-870), PdElastic (:1189-1213), PdState
    (:1504-1530).  Used by the strain / stress post-processing."""
    if mp.material_type == MAT_RNP:
        Cc, beta = mp.mat[0], mp.mat[1]
        M = inf.moment(dim)
        lam = Cc * M * beta / 4.0 if dim == 2 else Cc * M * beta / 5.0
        return 0.25, lam, lam
    if mp.material_type in (MAT_PMB, MAT_PDELASTIC):
        lam = mp.mat[0] * (math.pi * math.pow(horizon, 3.0)) / 24.0
        return 0.25, lam, lam
    K, G = mp.mat[0], mp.mat[1]
    nu = (3.0 * K - 2.0 * G) / (2.0 * (3.0 * K + G))
    E = _toE(K, nu)
    return nu, _toLambdaE(E, nu), G


def particle_contact_constants(K: float, horizon: float, h: float):
    """BaseParticle: d_Rc = 0.95 h, d_Kn = 18 K / (pi delta^5) (baseParticle.cpp:170-179)."""
    Rc = 0.95 * h
    Kn = (18.0 / (math.pi * math.pow(horizon, 5))) * K
    return Kn, Rc
