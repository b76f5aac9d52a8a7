"""BSIM4v7 host constants (setup defaulting, BSIM4v7temp, BSIM4v7PAeffGeo)
against the reference: the golden temperature-stage values its own test
data holds (tests/goldens/bsim4v7_nmos_setup.json, relative 1e-10) and,
through the compiled reference, every constant the load reads, bit for bit,
on several cards and sizes.  Also: cards outside the covered default-mode
subset are refused."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import neospice_amd as na
from neospice_amd._lib import NsaError, lib

HERE = os.path.dirname(os.path.abspath(__file__))
NPARAM = 110
NAMES = ("type vtm coxe coxp toxe toxp factor1 vcrit bdos ados temp_ratio PhiBS PhiBSWS PhiBSWGS "
         "PhiBD PhiBSWD PhiBSWGD MJS MJSWS MJSWGS MJD MJSWD MJSWGD czbs czbssw czbsswg czbd czbdsw "
         "czbdswg leff weff leffCV weffCV phi sqrtPhi Xdep0 vbi k1 k1ox k3 k3b dvt0 dvt1 dvt2 dvt0w "
         "dvt1w dvt2w lpe0 lpeb kt1 kt1l kt2 w0 theta0vb0 etab nfactor cdsc cdscb cdscd cit mstar "
         "voffcbn cdep0 dwg dwb prwg prwb rds0 rdswmin xj a0 ags b0 b1 keta ua ub uc ud a2 delta pvag "
         "pclm litl thetaRout pdiblb pditsd pscbe1 pscbe2 noff voffcv abulkCVfactor acde ldeb moin "
         "cgdo cgso cgbo cgdl cgsl ckappad ckappas vth0 k2ox eta0 vbsc u0temp vsattemp vtfbphi2 "
         "vfbzb").split()
assert len(NAMES) == NPARAM

CARDS = [
    ".model nmod nmos level=14 vth0=0.4 u0=0.04 toxe=2e-9",
    ".model pmod pmos level=14 vth0=-0.4 u0=0.01 toxe=2e-9",
    ".model n2 nmos level=14 vth0=0.35 u0=0.05 toxe=1.8e-9 vsat=1e5 rdsw=150 ndep=3e17 eta0=0.1 "
    "pclm=1.1 cgso=2e-10 cgdo=2e-10 cgbo=1e-11",
    ".model p2 pmos level=14 vth0=-0.42 toxe=1.7e-9 toxm=1.8e-9 dvt0=2.0 k3=60 keta=-0.03",
]


def params(ckt, dev):
    out = np.zeros(NPARAM)
    rc = lib().nsa_circuit_bsim4_params(ckt._native, dev, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    return dict(zip(NAMES, out)), out


def one_device(card, w, l):
    c = na.Circuit()
    c.model(card)
    d, g = c.node("d"), c.node("g")
    pos = " nmos " in card
    c.V("vd", d, na.GND, 0.5 if pos else -0.5)
    c.V("vg", g, na.GND, 0.6 if pos else -0.6)
    c.M("m1", d, g, na.GND, na.GND, card.split()[1], w, l)
    return c.finalize()


def test_golden_temp_stage():
    gold = json.load(open(os.path.join(HERE, "golden", "bsim4v7_nmos_setup.json")))
    card = ".model " + gold["model_card"].lower()
    c = one_device(card, gold["W"], gold["L"])
    p, _ = params(c, 2)
    e = gold["expected"]
    rt = gold["tolerance_rel"]
    # vfb follows from vtfbphi2 = 4 (type vth0 - vfb - phi) (bsim4v7_temp.cpp:1658-1666)
    vfb = p["type"] * p["vth0"] - p["phi"] - p["vtfbphi2"] / 4.0
    checks = {"pParam_BSIM4leff": p["leff"], "pParam_BSIM4weff": p["weff"],
              "pParam_BSIM4vth0": p["vth0"], "pParam_BSIM4u0temp": p["u0temp"],
              "pParam_BSIM4vfb": vfb, "pParam_BSIM4vsattemp": p["vsattemp"],
              "pParam_BSIM4k1ox": p["k1ox"], "pParam_BSIM4cdep0": p["cdep0"],
              "pParam_BSIM4phi": p["phi"], "here_BSIM4vth0": p["vth0"],
              "here_BSIM4u0temp": p["u0temp"], "here_BSIM4vfb": vfb,
              "here_BSIM4vsattemp": p["vsattemp"]}
    for k, v in checks.items():
        assert abs(v - e[k]) <= rt * abs(e[k]), (k, v, e[k])


@pytest.mark.ref
@pytest.mark.parametrize("card", CARDS)
def test_constants_match_reference(card):
    import refbind as RB
    for w, l in ((1e-6, 1e-7), (0.5e-6, 0.05e-6), (2.2e-6, 0.18e-6)):
        c = one_device(card, w, l)
        _, mine = params(c, 2)
        r = RB.RefCircuit(c, probe=False)
        r.dc()
        ref = np.zeros(NPARAM)
        RB._chk(r.L.ref_bsim4_fields(r.h, 2, ref.ctypes.data_as(C.POINTER(C.c_double))))
        bad = [(NAMES[k], mine[k], ref[k]) for k in range(NPARAM) if mine[k] != ref[k]]
        assert not bad, (card, w, l, bad[:5])


@pytest.mark.parametrize("extra", ["capmod=1", "mobmod=1", "rgatemod=1", "igcmod=1", "lvth0=1e-9",
                                   "agidl=1e-10", "alpha0=1e-6", "tnom=30", "a1=0.1"])
def test_unsupported_cards_refused(extra):
    c = na.Circuit()
    c.model(".model nb nmos level=14 vth0=0.4 " + extra)
    d = c.node("d")
    with pytest.raises(NsaError) as e:
        c.M("m1", d, d, na.GND, na.GND, "nb", 1e-6, 1e-7)
        c.finalize()
    assert e.value.code == -5  # NSA_ERR_UNSUPPORTED


def derivs(ckt, dev):
    val, dw, dl = np.zeros(NPARAM), np.zeros(NPARAM), np.zeros(NPARAM)
    P = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert lib().nsa_circuit_bsim4_param_derivs(ckt._native, dev, P(val), P(dw), P(dl)) == 0
    return val, dw, dl


@pytest.mark.parametrize("card", CARDS)
def test_size_derivatives(card):
    """d(record)/dW and d(record)/dL by dual numbers: the value part is the
    plain record bit for bit, and each derivative matches a central
    difference of the plain record (h = 1e-6 relative; Richardson-checked
    against h = 2e-6 so that a wrong term cannot hide in FD noise)."""
    for w, l in ((1e-6, 1e-7), (0.5e-6, 0.05e-6), (3.1e-6, 0.37e-6)):
        val, dw, dl = derivs(one_device(card, w, l), 2)
        _, plain = params(one_device(card, w, l), 2)
        assert np.array_equal(val, plain)
        for which, x0, d_ad in (("w", w, dw), ("l", l, dl)):
            def rec(x):
                ww, ll = (x, l) if which == "w" else (w, x)
                return params(one_device(card, ww, ll), 2)[1]
            fd = {}
            for rel in (1e-6, 2e-6):
                h = rel * x0
                fd[rel] = (rec(x0 + h) - rec(x0 - h)) / (2 * h)
            scale = np.maximum(np.abs(val) / x0, 1e-300)
            for q in range(NPARAM):
                a, f1, f2 = d_ad[q], fd[1e-6][q], fd[2e-6][q]
                tol = 1e-6 * max(abs(a), abs(f1)) + 1e-9 * scale[q] + abs(f1 - f2)
                assert abs(a - f1) <= tol, (card, w, l, which, NAMES[q], a, f1, f2)
        # fields that do not depend on the size have zero derivative
        for nm in ("type", "vtm", "coxe", "toxe", "phi", "vbi", "vth0"):
            q = NAMES.index(nm)
            assert dw[q] == 0.0 and dl[q] == 0.0, nm
        # leff = L, weff = W here (no offsets): unit derivatives
        assert dl[NAMES.index("leff")] == 1.0 and dw[NAMES.index("weff")] == 1.0


@pytest.mark.parametrize("card", CARDS)
def test_level54_alias(card):
    """ngspice's LEVEL=54 names the same BSIM4v7 model as the reference's
    LEVEL=14: the same constants, bit for bit."""
    c14 = one_device(card, 0.5e-6, 0.05e-6)
    c54 = one_device(card.replace("level=14", "level=54"), 0.5e-6, 0.05e-6)
    _, a = params(c14, 2)
    _, b = params(c54, 2)
    assert np.array_equal(a, b)
