"""Pins of the omexdia-type kinetics (oracle/sedcol_oracle.c oc_kinetics).

FABM's hzg_omexdia_p is not part of the reference tree, so the kinetics are a
library definition after the OMEXDIA equations of Soetaert et al. (1996), with
the parameter values of examples/standalone/omexdia_p/fabm_sed.nml:51-77.
These tests pin the oracle by what those equations fix, each at a point where
the answer is a plain number:

  * half-saturation points: a Monod term c/(c+ks) is 1/2 at c = ks, an
    inhibition term kin/(c+kin) is 1/2 at c = kin, so a rate there is half
    its saturated (or uninhibited) value;
  * saturation limits: with O2 >> every constant, nitrification, ODU
    oxidation and P adsorption run at their maximum rates, given per day in
    the namelist (unit conversion to s-1 pinned here, from the namelist
    values typed below, not from sedcol.params);
  * the production cap CprodMax/(CprodMax + sum k C): Cprod = CprodMax/2 when
    the uncapped decay equals CprodMax, Cprod -> CprodMax for unlimited
    detritus, Cprod -> k C for little;
  * the partition of Cprod into oxic mineralisation, denitrification and
    anoxic mineralisation: complete, with ratios between the shares that the
    published Monod/inhibition products fix at the half-saturation points;
  * the electron balance -dO2 + dODU - 2 dNO3 = Cprod + 0.75 N2, which ties
    the coefficients 2 (O2 per nitrified N), 1 (O2 per ODU, ODU per anoxic C)
    and 0.8 (NO3 per denitrified C) together;
  * the P exchange: equilibrium adsP/PO4 = [O2/(O2+ksO2oduox)] /
    [ODU/(ODU+PAdsODU)], equal to 2 for O2 >> ksO2oduox and ODU = PAdsODU;
  * the temperature factor of the right-hand side
    exp(-4500 (1/(T+273) - 1/288)) (fabm_sediment_driver.F90:648): 1 at
    15 degC, the Arrhenius ratio between 25 and 15 degC.

Every pin is a function of the kinetics under test, so the same suite runs
against planted mistakes (oracle.planted / a perturbed parameter set) and
must reject each of them.  The GPU kinetics are bitwise equal to the oracle's
(test_gpu_parity.py), so these pins hold for the device code too.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from sedcol import cases
from sedcol.params import OmexdiaNamelist, SedParams
import oracle as O

DAY = 86400.0

# &hzg_omexdia_p of examples/standalone/omexdia_p/fabm_sed.nml:51-77, typed
# from the namelist file (rates per day, half-saturations mmol m-3)
NML = {
    "rLabile": 0.043, "rSemilabile": 0.001, "CprodMax": 9600.0, "rnit": 200.0, "ksO2nitri": 20.0,
    "rODUox": 20.0, "ksO2oduox": 1.0, "ksO2oxic": 3.0, "ksNO3denit": 1.0, "kinO2denit": 70.0,
    "kinNO3anox": 1.0, "kinO2anox": 1.0, "PAds": 0.01, "PAdsODU": 70.0,
}
BIG = 1.0e12  # "unlimited" concentration: c/(c+ks) = 1 - O(1e-10)

LDC, SDC, LDN, SDN, DETP, O2, NO3, NH4, ODU, PO4, ADSP = range(11)


def state(**kw):
    c = np.zeros(11)
    for k, v in kw.items():
        c[globals()[k.upper()]] = v
    return c


def close(a, b, rel):
    return abs(a - b) <= rel * max(abs(a), abs(b))


# ---------------------------------------------------------------------------
# the pins: each takes kin(c, fT) -> (R, sC, sN) and returns a list of
# failures (empty = pinned)
# ---------------------------------------------------------------------------

def pin_nitrification(kin):
    bad = []
    nh4 = 10.0
    sat = NML["rnit"] / DAY * nh4  # fT = 1
    R, _, _ = kin(state(nh4=nh4, o2=BIG), 1.0)
    if not close(-R[NH4], sat, 1e-9):
        bad.append(f"saturated nitrification {-R[NH4]:.6e} != rnit NH4 {sat:.6e}")
    R, _, _ = kin(state(nh4=nh4, o2=NML["ksO2nitri"]), 1.0)
    if not close(-R[NH4], 0.5 * sat, 1e-12):
        bad.append(f"nitrification at O2 = ksO2nitri {-R[NH4]:.6e} != half {0.5 * sat:.6e}")
    if not (close(R[NO3], -R[NH4], 1e-14) and close(R[O2], 2.0 * R[NH4], 1e-14)):
        bad.append("nitrification stoichiometry: NO3 gain = NH4 loss, O2 loss = 2 NH4 loss")
    # first order in NH4
    R2, _, _ = kin(state(nh4=3.0 * nh4, o2=NML["ksO2nitri"]), 1.0)
    if not close(R2[NH4], 3.0 * R[NH4], 1e-13):
        bad.append("nitrification not first order in NH4")
    return bad


def pin_odu_oxidation(kin):
    bad = []
    odu = 50.0
    sat = NML["rODUox"] / DAY * odu
    R, _, _ = kin(state(odu=odu, o2=BIG), 1.0)
    if not close(-R[ODU], sat, 1e-9):
        bad.append(f"saturated ODU oxidation {-R[ODU]:.6e} != rODUox ODU {sat:.6e}")
    R, _, _ = kin(state(odu=odu, o2=NML["ksO2oduox"]), 1.0)
    if not close(-R[ODU], 0.5 * sat, 1e-12):
        bad.append(f"ODU oxidation at O2 = ksO2oduox {-R[ODU]:.6e} != half {0.5 * sat:.6e}")
    if not close(R[O2], R[ODU], 1e-14):
        bad.append("ODU oxidation uses one O2 per ODU")
    return bad


def pin_production_cap(kin):
    bad = []
    rL, rS, cmax = NML["rLabile"] / DAY, NML["rSemilabile"] / DAY, NML["CprodMax"] / DAY
    # little detritus: first-order decay at the namelist rates
    for name, idx, r in (("ldetC", LDC, rL), ("sdetC", SDC, rS)):
        c = np.zeros(11)
        c[idx] = 1.0e-6
        c[O2] = BIG
        R, sC, _ = kin(c, 1.0)
        if not close(-R[idx], r * 1.0e-6, 1e-10):
            bad.append(f"{name} decay {-R[idx]:.6e} != rate * C {r * 1e-6:.6e}")
    # uncapped decay equal to CprodMax: half of it
    R, sC, _ = kin(state(ldc=cmax / rL, o2=BIG), 1.0)
    if not close(sC, 0.5 * cmax, 1e-13):
        bad.append(f"Cprod at k C = CprodMax {sC:.6e} != CprodMax/2 {0.5 * cmax:.6e}")
    # the cap acts on the sum of both pools, shared in proportion to k C
    R, sC, _ = kin(state(ldc=0.25 * cmax / rL, sdc=0.75 * cmax / rS, o2=BIG), 1.0)
    if not (close(sC, 0.5 * cmax, 1e-13) and close(R[LDC] / R[SDC], 1.0 / 3.0, 1e-13)):
        bad.append("cap on the summed decay of both pools, shared in proportion")
    # unlimited detritus: CprodMax
    R, sC, _ = kin(state(ldc=1.0e6 * cmax / rL, o2=BIG), 1.0)
    if not close(sC, cmax, 2e-6):
        bad.append(f"Cprod for unlimited detritus {sC:.6e} != CprodMax {cmax:.6e}")
    # detritus N and P leave with the carbon of their pool, under the same cap
    ldc, sdc = 0.6 * cmax / rL, 0.9 * cmax / rS
    R, sC, _ = kin(state(ldc=ldc, sdc=sdc, ldn=0.15 * ldc, sdn=0.1 * sdc, detp=0.01 * ldc, o2=BIG), 1.0)
    if not (close(R[LDN], 0.15 * R[LDC], 1e-13) and close(R[SDN], 0.1 * R[SDC], 1e-13)
            and close(R[DETP], 0.01 * R[LDC], 1e-13)):
        bad.append("detritus N and P decay with their pool's carbon (labile P)")
    if not close(-(R[LDN] + R[SDN]), R[NH4], 1e-13):
        bad.append("detritus N released as NH4")
    return bad


def _shares(kin, o2, no3):
    """(OxicMin, Denit, AnoxMin, Cprod) from a state without NH4 and ODU,
    where O2 loss, NO3 loss and ODU gain are one process each."""
    R, sC, sN = kin(state(ldc=1.0e4, o2=o2, no3=no3), 1.0)
    return -R[O2], -R[NO3] / 0.8, R[ODU], sC, sN


def pin_partition(kin):
    bad = []
    ox, dn, an, cp, sn = _shares(kin, 37.0, 4.0)
    if not close(ox + dn + an, cp, 1e-13):
        bad.append("oxic + denitrification + anoxic != Cprod")
    if not close(sn, 0.8 * dn, 1e-13):
        bad.append("N2 release != 0.8 Denit")
    ox, dn, an, cp, _ = _shares(kin, 0.0, 0.0)
    if not (ox == 0.0 and dn == 0.0 and close(an, cp, 1e-15)):
        bad.append("no O2, no NO3: all anoxic")
    ox, dn, an, cp, _ = _shares(kin, BIG, 5.0)
    if not close(ox, cp, 1e-9):
        bad.append("O2 >> ks: all oxic")
    ox, dn, an, cp, _ = _shares(kin, 0.0, BIG)
    if not close(dn, cp, 1e-9):
        bad.append("NO3 >> ks, no O2: all denitrification")
    # no NO3: OxicMin/AnoxMin = [O2/(O2+ksO2oxic)] / [kinO2anox/(O2+kinO2anox)]
    for o2, want in ((NML["ksO2oxic"], 0.5 * (NML["ksO2oxic"] + NML["kinO2anox"]) / NML["kinO2anox"]),
                     (NML["kinO2anox"], NML["kinO2anox"] / (NML["kinO2anox"] + NML["ksO2oxic"]) * 2.0)):
        ox, dn, an, cp, _ = _shares(kin, o2, 0.0)
        if not close(ox / an, want, 1e-13):
            bad.append(f"OxicMin/AnoxMin at O2 = {o2}: {ox / an:.15g} != {want:.15g}")
    # no O2, NO3 at the denitrification half-saturation = the anoxic NO3
    # inhibition constant: both NO3 terms are 1/2
    ox, dn, an, cp, _ = _shares(kin, 0.0, NML["ksNO3denit"])
    want = (0.5) / (NML["kinNO3anox"] / (NML["ksNO3denit"] + NML["kinNO3anox"]))
    if not close(dn / an, want, 1e-13):
        bad.append(f"Denit/AnoxMin at NO3 = ksNO3denit, O2 = 0: {dn / an:.15g} != {want:.15g}")
    # O2 inhibitions: at NO3 = ksNO3denit the NO3 terms cancel in
    # Denit/AnoxMin, leaving [kinO2denit/(O2+kinO2denit)] / [kinO2anox/(O2+kinO2anox)];
    # each inhibition is 1/2 at O2 = its own constant
    kd, ka = NML["kinO2denit"], NML["kinO2anox"]
    for o2, want in ((kd, 0.5 / (ka / (kd + ka))), (ka, (kd / (ka + kd)) / 0.5)):
        _, dn0, an0, _, _ = _shares(kin, 0.0, NML["ksNO3denit"])
        _, dn, an, _, _ = _shares(kin, o2, NML["ksNO3denit"])
        got = (dn / an) / (dn0 / an0)
        if not close(got, want, 1e-12):
            bad.append(f"O2 inhibition ratio at O2 = {o2}: {got:.15g} != {want:.15g}")
    return bad


def pin_electron_balance(kin):
    bad = []
    rng = np.random.default_rng(17)
    for _ in range(50):
        c = rng.uniform(0.0, 1.0, 11) * np.array([5e4, 5e4, 8e3, 8e3, 500, 300, 40, 15, 200, 5, 20])
        R, sC, sN = kin(c, rng.uniform(0.5, 1.5))
        lhs = -R[O2] + R[ODU] - 2.0 * R[NO3]
        rhs = sC + 0.75 * sN
        scale = abs(R[O2]) + abs(R[ODU]) + 2 * abs(R[NO3]) + sC + 1e-300
        if abs(lhs - rhs) > 1e-14 * scale:
            bad.append(f"-dO2 + dODU - 2 dNO3 = {lhs:.6e} != Cprod + 0.75 N2 = {rhs:.6e}")
            break
    return bad


def pin_phosphorus(kin):
    bad = []
    kP = NML["PAds"] * NML["rODUox"] / DAY
    R, _, _ = kin(state(po4=2.0, o2=BIG), 1.0)
    if not close(R[ADSP], kP * 2.0, 1e-9):
        bad.append(f"saturated P adsorption {R[ADSP]:.6e} != PAds rODUox PO4 {kP * 2.0:.6e}")
    R, _, _ = kin(state(po4=2.0, o2=NML["ksO2oduox"]), 1.0)
    if not close(R[ADSP], 0.5 * kP * 2.0, 1e-12):
        bad.append("P adsorption at O2 = ksO2oduox is not half the saturated rate")
    # release: half of its ODU-saturated value at ODU = PAdsODU (no O2)
    R1, _, _ = kin(state(adsp=3.0, odu=BIG), 1.0)
    R2, _, _ = kin(state(adsp=3.0, odu=NML["PAdsODU"]), 1.0)
    if not (close(-R1[ADSP], kP * 3.0, 1e-9) and close(R2[ADSP], 0.5 * R1[ADSP], 1e-9)):
        bad.append("P release: saturated rate PAds rODUox adsP, half at ODU = PAdsODU")
    # equilibrium adsP/PO4 = 2 at O2 >> ksO2oduox, ODU = PAdsODU
    R, _, _ = kin(state(po4=1.5, adsp=3.0, o2=BIG, odu=NML["PAdsODU"]), 1.0)
    if abs(R[ADSP]) > 1e-9 * kP * 3.0:
        bad.append(f"adsP/PO4 = 2 is not the exchange equilibrium (rate {R[ADSP]:.3e})")
    if R[PO4] + R[ADSP] + R[DETP] != 0.0 and abs(R[PO4] + R[ADSP] + R[DETP]) > 1e-15 * abs(R[ADSP] + 1e-300):
        bad.append("P not conserved by the exchange")
    return bad


def _one_layer(T):
    """One 1 mm layer with no transport (closed boundaries, no bioturbation,
    no molecular diffusion), so that the right-hand side is the kinetics."""
    s = {k: np.array([v]) for k, v in {
        "phi0": 0.8, "phi_inf": 0.6, "log10_Db": -11.0, "T": T, "O2": 0.0, "NO3": 0.0, "NH4": 0.0,
        "PO4": 0.0, "Cflux": 1.0}.items()}
    d = cases.build_columns(np.array([1]), 1.0e-3, s, params=SedParams(bcup_dis=3, bcup_part=3, clamp=0))
    d.biot[:] = 0.0
    d.params.D0 = 0.0
    return d


def pin_temperature(kin, rhs_at):
    """rhs_at(T, c) -> rhs of one transport-free layer at temperature T."""
    bad = []
    c = state(nh4=10.0, o2=NML["ksO2nitri"], ldc=100.0, odu=5.0)
    R1, _, _ = kin(c, 1.0)
    r15 = rhs_at(15.0, c)
    if not np.array_equal(r15, R1):
        bad.append("f_T(15 degC) != 1")
    # first-order nitrification scales with f_T; the Arrhenius ratio 25 : 15 degC
    r25 = rhs_at(25.0, state(nh4=10.0, o2=BIG))
    r15 = rhs_at(15.0, state(nh4=10.0, o2=BIG))
    want = math.exp(4500.0 * (1.0 / 288.0 - 1.0 / 298.0))
    if not close(r25[NH4] / r15[NH4], want, 1e-14):
        bad.append(f"f_T(25)/f_T(15) = {r25[NH4] / r15[NH4]:.16g} != {want:.16g}")
    return bad


PINS = (pin_nitrification, pin_odu_oxidation, pin_production_cap, pin_partition, pin_electron_balance,
        pin_phosphorus)


def oracle_kin(p):
    return lambda c, fT: O.kinetics(p, fT, c)


def oracle_rhs_at(p):
    def rhs_at(T, c):
        d = _one_layer(T)
        pp = dict(d.params.si(), **{k: v for k, v in p.items() if k not in ("bcup_dis", "bcup_part", "D0",
                                                                              "clamp", "arith")})
        pp["arith"] = p["arith"]
        r, _ = O.Oracle(d, pp).rhs(np.asarray(c, dtype=np.float64).reshape(11, 1))
        return r[:, 0]
    return rhs_at


def all_failures(p):
    kin = oracle_kin(p)
    out = []
    for pin in PINS:
        out += [f"{pin.__name__}: {m}" for m in pin(kin)]
    out += [f"pin_temperature: {m}" for m in pin_temperature(kin, oracle_rhs_at(p))]
    return out


# ---------------------------------------------------------------------------
# tests
# ---------------------------------------------------------------------------

def test_namelist_defaults_are_the_reference_values():
    nml = OmexdiaNamelist()
    for k, v in NML.items():
        assert getattr(nml, k) == v, k


@pytest.mark.parametrize("arith", [0, 1])
def test_kinetics_pins_hold(arith):
    p = SedParams(arith=arith).si()
    assert all_failures(p) == []


# planted mistakes in the oracle's kinetics (oc_set_plant) ...
PLANTS = {
    1: "nitrification O2 term written as inhibition",
    2: "hard production cap",
    3: "one O2 per nitrified N",
    4: "one NO3 per denitrified C",
    5: "anoxic O2 inhibition with kinO2denit",
    6: "detritus N without the production cap",
    7: "P release term inverted",
    8: "temperature factor referenced to 20 degC",
    9: "oxic half-saturation kinO2anox",
    10: "ODU oxidation half-saturation ksO2oxic",
    11: "labile detritus P at the semi-labile rate",
}


@pytest.mark.parametrize("which", sorted(PLANTS))
def test_pins_reject_planted_kinetics(which):
    p = SedParams().si()
    with O.planted(which):
        fails = all_failures(p)
    assert fails, f"planted mistake {which} ({PLANTS[which]}) passes every pin"


def _swap(p, a, b):
    q = dict(p)
    q[a], q[b] = p[b], p[a]
    return q


# ... and in the parameter set handed to it
PARAM_PLANTS = {
    "ksO2nitri <-> ksO2oduox": lambda p: _swap(p, "ksO2nitri", "ksO2oduox"),
    "kinO2denit <-> kinO2anox": lambda p: _swap(p, "kinO2denit", "kinO2anox"),
    "ksO2oxic <-> kinO2anox": lambda p: _swap(p, "ksO2oxic", "kinO2anox"),
    "rLabile <-> rSemilabile": lambda p: _swap(p, "rL", "rS"),
    "rPads without rODUox": lambda p: dict(p, rPads=NML["PAds"] / DAY),
    "rnit per day": lambda p: dict(p, rnit=NML["rnit"]),
    "CprodMax per day": lambda p: dict(p, CprodMax=NML["CprodMax"]),
    "PAdsODU x 2": lambda p: dict(p, PAdsODU=2.0 * NML["PAdsODU"]),
}


@pytest.mark.parametrize("name", sorted(PARAM_PLANTS))
def test_pins_reject_planted_parameters(name):
    p = PARAM_PLANTS[name](SedParams().si())
    assert all_failures(p), f"planted parameter mistake '{name}' passes every pin"
