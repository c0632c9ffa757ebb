"""The oracle's full-rheology restatement against the reference's own
output (goldens g14, g15, g17, written by the reference's movemark and
ronurecalc with the rock table of its deck init.t3c), plus closed forms of
viscalc's branches and of the melting parameterisations that a wrong
exponent, sign or clamp would fail."""
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


def test_ronurecalc_rejects_densimod_2():
    g = golden("g15_ronurecalc_rheology.npz")
    with pytest.raises(ValueError):
        oracle.ronurecalc_x(g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"], g["marke"], g["markrr"],
                            g["markrn"], g["gx"], g["gy"], g["gz"], g["pr"], None, _rocks(g), g["init"], 1e18, 1e25,
                            0.0, 0.0, 0.0, densimod=2)


def _g17(g, densimod):
    st12 = [g[k] for k in ST]
    out = oracle.ronurecalc_x(
        g["x"], g["y"], g["z"], g["tk"], g["eta"], g["mtype"], g["marke"], g["markrr"], g["markrn"], g["gx"], g["gy"],
        g["gz"], g["pr"], st12, _rocks(g), np.zeros((11, len(g["pr"])), np.float32), float(g["nubeg"]),
        float(g["nuend"]), float(g["timestep"]), float(g["timesum"]), float(g["convend"]), nukoef=float(g["nukoef"]),
        densimod=densimod, vdeep=float(g["vdeep"]), nudeep=float(g["nudeep"]), eroslev=float(g["eroslev"]),
        sedilev=float(g["sedilev"]), waterlev=float(g["waterlev"]), sedtype=int(g["sedtype"]), mnumy=int(g["mnumy"]),
        teclmin=float(g["teclmin"]), teclmax=float(g["teclmax"]), maxpmelt=float(g["maxpmelt"]), markw=g["markw"],
        markex=g["markex"])
    return out


@pytest.mark.parametrize("densimod", [0, 1])
def test_ronurecalc_melt_matches_reference(densimod):
    """partially molten rocks (meltpart0/meltpart11 in ronurecalc): node
    fields and marker viscosities equal to the reference's to the bit"""
    g = golden("g17_ronurecalc_melt.npz")
    fld, eta, t, e, rr, rn = _g17(g, densimod)
    fld[6] = oracle.tkrecalc(fld[6], g["bondm"], g["bondv"], g["bondn"])
    assert np.array_equal(t, g[f"type_out{densimod}"])
    for name, a in (("eta", eta), ("e", e), ("rr", rr), ("rn", rn)):
        assert np.array_equal(a, g[f"{name}_out{densimod}"]), name
    ref = g[f"out{densimod}"]
    for q in range(11):
        bad = np.flatnonzero(fld[q] != ref[q])
        assert bad.size == 0, (q, bad[:5], fld[q][bad[:5]], ref[q][bad[:5]])
    # the case covers what it is meant to: melt markers inside their melting
    # interval, the basaltic melt viscosity taking over
    melt = g["mtype"] > 20
    xm = np.array([oracle.meltpart11(int(tt), float(T), 3e3, float(w))[0]
                   for tt, T, w in zip(g["mtype"][melt], g["tk"][melt], g["markw"][melt])])
    assert ((xm > 0) & (xm < 1)).sum() > 50
    assert (eta != g["eta"])[np.isin(g["mtype"], (27, 28))].any()


def test_meltpart11_closed_forms():
    """melt fraction against the published parameterisations written out
    here: linear between solidus and liquidus for crustal rocks (basalt:
    Schmidt & Poli wet solidus above 48 km, Hess liquidus); dry peridotite
    (Katz et al. 2003): F = ((T - Ts)/(Tlherz - Ts))^1.5 up to cpx
    exhaustion F_cpx = 0.15/(0.5 + 0.08 P), then
    F_cpx + (1 - F_cpx) ((T - T_cpx)/(Tliq - T_cpx))^1.5; no dry melting
    above maxpmelt; latent heat 380 / 400 kJ/kg times F"""
    mpb = 20000.0  # 2 GPa, 60 km
    ykm = mpb * 3e-3
    ts = 935.4 + 0.1162 * ykm + 0.006937 * ykm * ykm
    tl = 1423.15 + 3.5 * ykm
    for T in (900.0, ts + 1.0, 0.5 * (ts + tl), tl - 1.0, tl + 50):
        xm, hl = oracle.meltpart11(27, T, mpb)
        f = min(max((T - ts) / (tl - ts), 0.0), 1.0)
        assert abs(xm - f) <= 1e-12 and abs(hl - 380000.0 * f) <= 1e-6
    P = mpb * 1e-4
    Ts = 273.15 + 1085.7 + 132.9 * P - 5.1 * P * P
    Tlz = 273.15 + 1475.0 + 80.0 * P - 3.2 * P * P
    Tlq = 273.15 + 1780.0 + 45.0 * P - 2.0 * P * P
    fcpx = 0.15 / (0.5 + 0.08 * P)
    tcpx = fcpx ** (1 / 1.5) * (Tlz - Ts) + Ts
    for T in (Ts - 5, Ts + 20, tcpx - 1, tcpx + 1, 0.5 * (tcpx + Tlq), Tlq + 10):
        if T < Ts:
            f = 0.0
        elif T >= Tlq:
            f = 1.0
        elif T <= tcpx:
            f = ((T - Ts) / (Tlz - Ts)) ** 1.5
        else:
            f = fcpx + (1 - fcpx) * ((T - tcpx) / (Tlq - tcpx)) ** 1.5
        xm, hl = oracle.meltpart11(29, T, mpb)
        assert abs(xm - f) <= 1e-9, (T, xm, f)
        assert abs(hl - 400000.0 * f) <= 1e-3
    assert oracle.meltpart11(29, Tlq - 10, mpb, maxpmelt=1.9e9) == (0.0, 0.0)
    assert oracle.meltpart11(29, Tlq - 10, mpb, mwa=0.1, maxpmelt=1.9e9)[0] > 0
    # water lowers the mantle solidus
    assert oracle.meltpart11(34, Ts - 40, mpb, mwa=0.5)[0] > 0 == oracle.meltpart11(29, Ts - 40, mpb)[0]


def test_meltpart0_mixture():
    """partially molten rock: properties mixed linearly by the melt fraction
    less the extracted melt (densimod 0), zero below 22"""
    rocks = [dict() for _ in range(40)]
    rocks[7] = dict(ro=3000.0, bro=3e-5, aro=1e-3, cp=1000.0, kt=2.0, kf=300.0, kp=1e-5, n0=1e18, n1=1e25)
    rocks[27] = dict(ro=2700.0, bro=5e-5, aro=2e-3, cp=1300.0, kt=1.0, kf=100.0, kp=2e-5, n0=1e17, n1=1e17)
    mpb, mex = 20000.0, 0.1
    ykm = mpb * 3e-3
    ts = 935.4 + 0.1162 * ykm + 0.006937 * ykm * ykm
    tl = 1423.15 + 3.5 * ykm
    T = ts + 0.6 * (tl - ts)
    e, rc = oracle.meltpart0(rocks, 27, T, mpb, mex=mex, densimod=0)
    f = 0.6 - mex
    assert rc == 1
    assert abs(e[3] - (2700.0 * f + 3000.0 * (1 - f))) <= 1e-9
    ktm = (1.0 + 100.0 / (T + 77.0)) * np.exp(2e-5 * mpb)
    kts = (2.0 + 300.0 / (T + 77.0)) * np.exp(1e-5 * mpb)
    assert abs(e[6] - (ktm * f + kts * (1 - f))) <= 1e-12
    # latent heat: dH/dT = 380 kJ/kg / (tl - ts) inside the interval
    assert abs(e[5] - (1300.0 * f + 1000.0 * (1 - f) + 380000.0 / (tl - ts))) <= 1e-6 * e[5]
    e0, rc0 = oracle.meltpart0(rocks, 21, T, mpb)
    assert rc0 == 0 and e0 == [0.0] * 7


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
