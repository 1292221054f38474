"""Pins the oracle's spectral predictor/corrector (Guess/Iterate,
movefieldshydrolinlayersplit.cpp:807-901, 1373-1478) with single-mode
closed forms and invariants rather than by restating its update formulas:

* continuity: with single half-integer modes the corrector's density obeys
  rho - Jx - (rhoP - JxP) = -hx/2 d/dy(Jy+JyP) - hx/2 d/dz(Jz+JzP), where the
  derivatives are the analytic derivatives of the basis functions
  (cos(a y) -> -a sin, sin(a y) -> a cos at y = (j+1/2) h); a wrong transform
  kind, sign or factor fails;
* the corrector's Ex solves laplacian(Ex) = div(J_perp) and psi solves
  -laplacian(psi) = rho - Jx, both exactly per mode;
* the predictor's (rho, Ex) map of a plasma wave without sources is the
  trapezoidal rotation: it conserves rho^2 + (k^2+n)^2/n Ex^2 per mode and
  turns by 2 atan(hx sqrt(n)/2), i.e. the plasma frequency sqrt(n) to O(hx^2);
* linearity in all sources and zero in -> zero out.
"""
import numpy as np
import pytest

from tertia_amd.params import small


def basis(p, kz_kind, ky_kind, kz0, ky0):
    k = np.arange(p.nz)[:, None] + 0.5
    j = np.arange(p.ny)[None, :] + 0.5
    az = np.pi * k * (kz0 + 0.5) / p.nz
    ay = np.pi * j * (ky0 + 0.5) / p.ny
    fz = np.cos(az) if kz_kind == 0 else np.sin(az)
    fy = np.cos(ay) if ky_kind == 0 else np.sin(ay)
    return fz * fy


def wavenumbers(p, ky0, kz0):
    aky = np.pi * (ky0 + 0.5) / p.ly
    akz = np.pi * (kz0 + 0.5) / p.lz
    return aky, akz, aky * aky + akz * akz


ZERO_P = ("Rho", "Ex", "Ey", "Ez", "Jx", "Jy", "Jz", "Bx", "By", "Bz")


def fresh(oracle_mod, p):
    o = oracle_mod.Oracle(p, history=())
    return o


@pytest.mark.parametrize("ky0,kz0", [(0, 0), (3, 1), (5, 7)])
def test_corrector_continuity_ex_psi(oracle_mod, ky0, kz0):
    p = small(ny=24, nz=20, nslices=3, hx=0.07)
    o = fresh(oracle_mod, p)
    aky, akz, k2 = wavenumbers(p, ky0, kz0)
    cc = basis(p, 0, 0, kz0, ky0)
    s_y = basis(p, 0, 1, kz0, ky0)     # Jy kind: cos along z, sin along y
    s_z = basis(p, 1, 0, kz0, ky0)     # Jz kind: sin along z, cos along y
    A = dict(rhoP=0.3, jxP=-0.2, jyP=0.5, jzP=-0.4, jx=0.7, jy=-0.6, jz=0.25)
    P = {"Rho": A["rhoP"] * cc, "Jx": A["jxP"] * cc, "Jy": A["jyP"] * s_y,
         "Jz": A["jzP"] * s_z}
    for f in ZERO_P:
        o.set_field(f, P.get(f, np.zeros_like(cc)), which=1)
    o.guess(False)        # loads the P spectra (RhoP, JxP, JyP, JzP)
    o.set_field("Jx", A["jx"] * cc)
    o.set_field("Jy", A["jy"] * s_y)
    o.set_field("Jz", A["jz"] * s_z)
    o.iterate(True)
    # d/dy of sin(aky y) = aky cos(aky y); d/dz of sin(akz z) = akz cos(akz z)
    div_sum = (aky * (A["jy"] + A["jyP"]) + akz * (A["jz"] + A["jzP"])) * cc
    rho_expect = A["rhoP"] * cc + (A["jx"] - A["jxP"]) * cc - p.hx / 2 * div_sum
    assert np.abs(o.get_field("Rho") - rho_expect).max() < 1e-12
    div_c = (aky * A["jy"] + akz * A["jz"]) * cc          # div J_perp of C
    ex_expect = -div_c / k2                               # laplacian(Ex) = div J
    assert np.abs(o.get_field("Ex") - ex_expect).max() < 1e-12
    psi_expect = (rho_expect - A["jx"] * cc) / k2         # -laplacian psi = rho - Jx
    assert np.abs(o.get_field("Psi") - psi_expect).max() < 1e-12


@pytest.mark.parametrize("ky0,kz0", [(0, 0), (2, 5)])
@pytest.mark.parametrize("dens", [1.0, 2.5])
def test_predictor_plasma_rotation(oracle_mod, ky0, kz0, dens):
    p = small(ny=16, nz=16, nslices=3, hx=0.05, dens=dens)
    o = fresh(oracle_mod, p)
    aky, akz, k2 = wavenumbers(p, ky0, kz0)
    cc = basis(p, 0, 0, kz0, ky0)
    rho, ex = 0.4, -0.15
    for f in ZERO_P:
        o.set_field(f, np.zeros_like(cc), which=1)
    o.set_field("Rho", rho * cc, which=1)
    o.set_field("Ex", ex * cc, which=1)
    steps = 6
    amps = [(rho, ex)]
    for _ in range(steps):
        o.guess(False)
        r1 = float((o.get_field("Rho") * cc).sum() / (cc * cc).sum())
        e1 = float((o.get_field("Ex") * cc).sum() / (cc * cc).sum())
        # no deposit: feed the prediction back as the next P layer
        for f in ZERO_P:
            o.set_field(f, np.zeros_like(cc), which=1)
        o.set_field("Rho", r1 * cc, which=1)
        o.set_field("Ex", e1 * cc, which=1)
        amps.append((r1, e1))
    w = (k2 + dens) ** 2 / dens
    Q = [r * r + w * e * e for r, e in amps]
    assert max(abs(q - Q[0]) for q in Q) < 1e-12 * Q[0]
    # rotation angle per step in the (rho, sqrt(w) Ex) plane
    ang = [np.arctan2(np.sqrt(w) * e, r) for r, e in amps]
    d = np.diff(np.unwrap(ang))
    theta = 2 * np.arctan(p.hx * np.sqrt(dens) / 2)
    assert np.abs(np.abs(d) - theta).max() < 1e-12
    assert abs(theta - p.hx * np.sqrt(dens)) < p.hx ** 3 * dens ** 1.5


def test_linearity_and_zero(oracle_mod):
    p = small(ny=16, nz=12, nslices=3)
    rng = np.random.default_rng(4)

    def run(state, J):
        o = fresh(oracle_mod, p)
        for f in ZERO_P:
            o.set_field(f, state[f], which=1)
        o.set_field("RhoBeam", state["RhoBeam"])
        o.set_field("JxBeam", state["JxBeam"])
        o.set_spectral("RhoBeam", state["SRB"])
        o.set_spectral("JxBeam", state["SJB"])
        o.guess(False)
        out = {f: o.get_field(f) for f in ("Ex", "Ey", "Ez", "Bx", "By", "Bz")}
        for f in ("Jx", "Jy", "Jz"):
            o.set_field(f, J[f])
        o.iterate(True)
        out.update({"I" + f: o.get_field(f) for f in ("Ex", "Ey", "Ez", "Bx", "By", "Bz",
                                                       "Rho", "Psi")})
        return out

    def rand_state():
        s = {f: rng.standard_normal((p.nz, p.ny)) for f in ZERO_P + ("RhoBeam", "JxBeam",
                                                                     "SRB", "SJB")}
        J = {f: rng.standard_normal((p.nz, p.ny)) for f in ("Jx", "Jy", "Jz")}
        return s, J

    s1, J1 = rand_state()
    s2, J2 = rand_state()
    a, b = 0.7, -1.3
    s3 = {k: a * s1[k] + b * s2[k] for k in s1}
    J3 = {k: a * J1[k] + b * J2[k] for k in J1}
    o1, o2, o3 = run(s1, J1), run(s2, J2), run(s3, J3)
    for k in o1:
        ref = a * o1[k] + b * o2[k]
        assert np.abs(o3[k] - ref).max() < 1e-10 * max(1.0, np.abs(ref).max()), k
    z = {k: np.zeros_like(v) for k, v in s1.items()}
    zj = {k: np.zeros_like(v) for k, v in J1.items()}
    o0 = run(z, zj)
    assert all(np.abs(v).max() == 0.0 for v in o0.values())
