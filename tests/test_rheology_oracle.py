"""The oracle's full-rheology restatement against the reference's own
output (goldens g14, g15, written by the reference's movemark and
ronurecalc with the rock table of its deck init.t3c), plus closed forms of
viscalc's branches that a wrong exponent, sign or clamp would fail."""
import numpy as np
import pytest

import oracle
from conftest import golden

ST = ("exx", "eyy", "ezz", "exy", "exz", "eyz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def _rocks(g):
    return [dict(zip(oracle.ROCKX_FIELDS, row)) for row in g["rocks"]]


@pytest.mark.parametrize("case", [0, 1])
def test_movemark_rheology_matches_reference(case):
    """props at the old position (full viscalc, strain-rate correction,
    strain growth and healing, grain roughness, dislocation share), then
    the motion (RK4 or one stage, outgrid 2 or 1, fluid and immobile
    markers): every output equal to the reference's to the bit."""
    g = golden("g14_movemark_rheology.npz")
    st12 = [g[k] for k in ST]
    dt = float(g["timestep"])
    eta, e, rr, rn = oracle.mk_props_x(g["x"], g["y"], g["z"], g["mtype"], g["tk"], g["eta"], g["marke"],
                                       g["markrr"], g["markrn"], _rocks(g), st12, g["pr"], g["gx"], g["gy"], g["gz"],
                                       float(g["nubeg"]), float(g["nuend"]), dt, float(g["timesum"][case]),
                                       float(g["convend"]), float(g["healrate"]), int(g["outgrid"][case]))
    x, y, z, tk, eta, t = oracle.rk4_x(g["x"], g["y"], g["z"], g["tk"], eta, g["mtype"], g["markd"], g["vx"], g["vy"],
                                       g["vz"], g["pr"], g["gx"], g["gy"], g["gz"], dt,
                                       outgrid=int(g["outgrid"][case]), markmod=int(g["markmod"][case]),
                                       vyfluid=float(g["vyfluid"]), vymelt=float(g["vymelt"]), g=tuple(g["g"]))
    for name, a in (("eta", eta), ("e", e), ("rr", rr), ("rn", rn), ("x", x), ("y", y), ("z", z), ("tk", tk),
                    ("type", t)):
        ref = g[f"{name}_out{case}"]
        bad = np.flatnonzero(~((a == ref) | (np.isnan(a) & np.isnan(ref))))
        assert bad.size == 0, (name, bad[:5], a[bad[:5]], ref[bad[:5]])
    # the case exercises what it is meant to
    moved = g[f"x_out{case}"] != g["x"]
    assert moved[(g["mtype"] >= 50) & (g["mtype"] < 100)].any() and moved[g["mtype"] >= 100].any()
    assert (g[f"rr_out{case}"] != g["markrr"]).any()
    if case == 0:
        assert (g["rn_out0"] != g["markrn"]).any() and (g["e_out0"] > 0).any()


def test_ronurecalc_rheology_matches_reference():
    """retyping, water/air mixing, viscalc from strain rates, phase
    densities: node fields equal to the reference's to the bit, marker
    viscosity, type, strain, grain roughness and dislocation share too"""
    g = golden("g15_ronurecalc_rheology.npz")
    st12 = [g[k] for k in ST]
    fld, eta, t, e, rr, rn = oracle.ronurecalc_x(
        g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"], g["marke"], g["markrr"], g["markrn"], g["gx"], g["gy"],
        g["gz"], g["pr"], st12, _rocks(g), g["init"], float(g["nubeg"]), float(g["nuend"]), float(g["timestep"]),
        float(g["timesum"]), float(g["convend"]), nukoef=float(g["nukoef"]), densimod=int(g["densimod"]),
        vdeep=float(g["vdeep"]), nudeep=float(g["nudeep"]), eroslev=float(g["eroslev"]), sedilev=float(g["sedilev"]),
        waterlev=float(g["waterlev"]), sedtype=int(g["sedtype"]), mnumy=int(g["mnumy"]), teclmin=float(g["teclmin"]),
        teclmax=float(g["teclmax"]))
    fld[6] = oracle.tkrecalc(fld[6], g["bondm"], g["bondv"], g["bondn"])
    assert np.array_equal(t, g["type_out"])
    for name, a in (("eta", eta), ("e", e), ("rr", rr), ("rn", rn)):
        assert np.array_equal(a, g[f"{name}_out"]), name
    for q in range(11):
        bad = np.flatnonzero(fld[q] != g["out"][q])
        assert bad.size == 0, (q, bad[:5], fld[q][bad[:5]], g["out"][q][bad[:5]])
    assert (g["type_out"] != g["mtype"]).any()


def test_ronurecalc_rejects_melt_types():
    g = golden("g15_ronurecalc_rheology.npz")
    mt = np.array(g["mtype"], np.int8, copy=True)
    mt[:] = 25
    with pytest.raises(ValueError):
        oracle.ronurecalc_x(g["x"], g["y"], g["z"], g["tk"], g["eta"], mt, g["marke"], g["markrr"], g["markrn"],
                            g["gx"], g["gy"], g["gz"], g["pr"], None, _rocks(g) + [{}] * 10, g["init"], 1e18, 1e25, 0.0,
                            0.0, 0.0)


def _rock(**kw):
    r = dict(n0=1.0, n1=1e40)
    r.update(kw)
    return r


def test_viscalc_closed_forms():
    """each ductile branch against its formula written independently:
    C: NU1 = nu0/ss^(mm-1), power-law stress (nu0 eps)^(1/mm); D: the same
    with exp((dh + dv P)/RT); F: Karato & Wu diffusion and dislocation
    viscosities; the brittle cap 0.5 sbrit/eps with sbrit = a + b P; the
    limits"""
    T, P, eps = 1300.0, 2.0e4, 1e-14
    # C
    r = _rock(nu0=1e20, ss=1e6, mm=3.0)
    v, *_ = oracle.viscalc(r, 5, T, P, eps, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, nubeg=1, nuend=1e40)
    nunewt = 1 / (0.5 * 1e20 / 1e6 ** 2)
    nupowl = 1 / (0.5 * (1e20 * eps) ** (1 / 3) / eps)
    assert v == pytest.approx(1 / (nunewt + nupowl), rel=1e-13)
    # D
    r = _rock(nu0=1e16, dh=2e5, dv=1.0, ss=1e6, mm=3.5)
    v, *_ = oracle.viscalc(r, 5, T, P, eps, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, nubeg=1, nuend=1e40)
    ex = np.exp((2e5 + 1.0 * P) / (8.314 * T))
    nunewt = 1 / (0.5 * 1e16 / 1e6 ** 2.5 * ex)
    nupowl = 1 / (0.5 * (1e16 * ex * eps) ** (1 / 3.5) / eps)
    assert v == pytest.approx(1 / (nunewt + nupowl), rel=1e-12)
    # F (nu0 < 0)
    r = _rock(nu0=-2e18, dh=4.3e5, dv=1.0, ss=2.4e5, mm=3.0, ll=0.4, s1=5.3e15)
    v, *_ = oracle.viscalc(r, 11, T, P, eps, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, nubeg=1, nuend=1e40)
    nunewt = 1 / (0.5 * (1e-3 / 5e-10) ** 2.5 * 8e10 / 5.3e15 * np.exp((2.4e5 + P * 0.4) / (8.314 * T)))
    nupowl = 1 / (0.5 * eps ** (1 / 3 - 1) * 8e10 / 2e18 ** (1 / 3) * np.exp((4.3e5 + P) / (8.314 * T) / 3))
    assert v == pytest.approx(1 / (nunewt + nupowl), rel=1e-12)
    # brittle cap below a stiff Newtonian rock, yield flag set on the strain
    r = _rock(nu0=1e25, a0=1e7, a1=1e7, b0=0.6, b1=0.6, e0=0, e1=1)
    v, e, *_ = oracle.viscalc(r, 5, T, P, eps, 0.0, 1.0, 0.0, -0.3, 0.0, 0.0, nubeg=1, nuend=1e40)
    sb = 1e7 + 0.6 * P * 1e5
    assert v == pytest.approx(0.5 * sb / eps, rel=1e-13) and e == pytest.approx(0.3)
    # weakening halfway between e0 and e1
    r = _rock(nu0=1e25, a0=2e7, a1=0.0, b0=0.6, b1=0.2, e0=0.0, e1=1.0)
    v, *_ = oracle.viscalc(r, 5, T, P, eps, 0.0, 1.0, 0.0, 0.5, 0.0, 0.0, nubeg=1, nuend=1e40)
    sb = 1e7 + 0.4 * P * 1e5
    assert v == pytest.approx(0.5 * sb / eps, rel=1e-13)
    # not yielding: strain made non-positive; limits
    r = _rock(nu0=1e21)
    v, e, *_ = oracle.viscalc(r, 5, T, P, eps, 0.0, 1.0, 0.0, 0.7, 0.0, 0.0, nubeg=1e22, nuend=1e25)
    assert v == 1e22 and e == pytest.approx(-0.7)


def test_viscalc_grain_size_model():
    """G (dh < 0): the roughness starts at 2e-3/pi when unset; with timesum
    <= convend it is kept and the Newtonian part is r^3 / BB / 2 with
    BB = s1 exp(-(ss + P ll)/RT) / (pi/2)^3; the dislocation part
    (eps/AA)^(1/n) / 2 / eps with AA = nu0 exp(-(-dh + P dv)/RT); past
    convend the roughness evolves and stays within [1e-6, 1e-2]"""
    T, P, eps, sig = 1500.0, 3.0e4, 1e-14, 1e7
    r = _rock(nu0=1.1e-16, dh=-5.3e5, dv=2.6, ss=3.75e5, mm=3.5, ll=0.7, s1=1.5e-15)
    v, e, rr, rn = oracle.viscalc(r, 9, T, P, eps, sig, 1.0, 1e-7, 0.0, 0.0, 0.5, nubeg=1, nuend=1e40,
                                  timestep=1e11, timesum=0.0, convend=1e13)
    r0 = np.float32(2e-3 / np.pi)
    assert rr == r0
    BB = 1.5e-15 * np.exp(-(3.75e5 + P * 0.7) / (8.314 * T)) / (np.pi / 2) ** 3
    AA = 1.1e-16 * np.exp(-(5.3e5 + P * 2.6) / (8.314 * T))
    nunewt = 1 / (float(r0) ** 3 / BB / 2)
    nupowl = 1 / ((eps / AA) ** (1 / 3.5) / 2 / eps)
    assert v == pytest.approx(1 / (nunewt + nupowl), rel=1e-12)
    assert rn == 0.5  # timesum <= 1e11: the share is not updated
    v2, _, rr2, rn2 = oracle.viscalc(r, 9, T, P, eps, sig, 1.0, 1e-7, 0.0, 1e-4, 0.5, nubeg=1, nuend=1e40,
                                     timestep=1e12, timesum=5e13, convend=1e13)
    assert 1e-6 <= rr2 <= 1e-2 and rr2 != np.float32(1e-4) and 0 < rn2 < 1
