"""Pins of the oracle's source terms and of its reconstruction composition
by what the mathematics fixes.

source (reference src/physics/adiabatic.hpp:396-541, area_coeff :70-90):
- the accretion profile is a Gaussian in the tilted latitude whose full
  width at half maximum is acc_width, centred on the great circle normal to
  omega_acc, and normalised so that with tilt 0 it deposits acc_rate * 4 pi
  per unit time over the mesh; the continuum limit of area_coeff is a 1-D
  integral evaluated here by quadrature;
- without friction the depletion term removes acc_rate * 4 pi at t = 0, so
  the total mass source vanishes then;
- accreted matter carries the angular momentum r x (omega_acc x r) =
  omega_acc - (omega_acc . r) r per unit mass and the energy
  e_acc + |omega_acc x r - v|^2 / 2;
- the radiative cooling law in its closed form on both branches of the
  beta(C) switch, continuous at the switch point, with both clamps and the
  negative-gravity case;
- friction vanishes for gas co-rotating with the star and otherwise
  decelerates v - v0 at alpha rho |v - v0|^2 / Sigma while removing mass at
  alpha rho |v - v0| without changing the velocity.

find_U_edges (MUSCL_base.hpp:361-515, E_fc :541-562, limiter Courant number
adiabatic.hpp:741-758):
- on smooth fields the reconstructed edge states converge to the exact
  edge-midpoint values at second order (a swapped H+/H- or beta pair, a
  wrong BM distance or a slot-4 E_fc taken at the wrong point leaves first
  order or worse);
- at a large step, where the superbee bound depends on the Courant number,
  each half-edge's state equals the pinned limiter applied to the slope
  ratio from the pinned stencil tables with nu = (c_s + v . n) dt arc / A.
"""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle_lib as O

GAM = 2 - 1 / (4 / 3)
GM = 0.217909


def _unit(x):
    return x / np.linalg.norm(x, axis=-1)[..., None]


def _solver(mesh, U=None, **params):
    U = np.tile([1.0, 0, 0, 0, 0.3], (mesh.nf, 1)) if U is None else U
    return O.Solver(mesh, U, 5, O.HLLC, GAM, params.pop("omega_ns", 0.0), O.make_params(**params))


def _pform(sig, L, p):
    return np.array([sig, L[0], L[1], L[2], p])


# ------------------------------------------------------------- accretion

@pytest.mark.parametrize("width", [0.1, 0.3])
def test_accretion_profile_fwhm_and_centre(width):
    m = O.Mesh("hex", 3)
    s = _solver(m, accretion_on=True, friction_on=True, alpha=0.0, acc_rate=1e-3, tilt_angle=0.3,
                acc_width=width)
    om = np.array([0, math.sin(0.3), math.cos(0.3)])
    u = _pform(1.0, [0, 0, 0], 0.1)
    x = np.array([1.0, 0, 0])
    peak = s.source(x, u)[0]
    for sgn in (1, -1):
        r = math.cos(width / 2) * x + sgn * math.sin(width / 2) * om
        assert s.source(r, u)[0] / peak == pytest.approx(0.5, abs=1e-3)
    # the maximum sits on the ring normal to omega_acc, not on the equator
    lat = [s.source(_unit(math.cos(a) * x + math.sin(a) * om), u)[0] for a in np.linspace(-0.05, 0.05, 11)]
    assert int(np.argmax(lat)) == 5


def test_accretion_mass_balance_at_t0():
    """tilt 0, friction off: sum A * S_Sigma = acc_rate 4 pi - acc_rate 4 pi = 0
    at t = 0 (each half to 1e-13 of its size)."""
    m = O.Mesh("quad", 5)
    g = O.Geometry(m)
    fc = _unit(g.fc)
    rng = np.random.default_rng(1)
    sig = 1 + 0.1 * rng.standard_normal(m.nf)
    U = np.zeros((m.nf, 5))
    U[:, 0] = sig
    U[:, 4] = 0.1 * sig / (GAM - 1)
    rate = 2e-3
    s = _solver(m, U, accretion_on=True, friction_on=False, acc_rate=rate, tilt_angle=0.0, acc_width=0.2)
    src = np.array([s.source(fc[f], _pform(sig[f], [0, 0, 0], 0.1 * sig[f]))[0] for f in range(m.nf)])
    total = float(np.sum(g.area * src))
    assert abs(total) < 1e-13 * rate * 4 * np.pi
    # the accretion half alone (friction on with alpha 0 drops the depletion)
    s2 = _solver(m, U, accretion_on=True, friction_on=True, alpha=0.0, acc_rate=rate, tilt_angle=0.0, acc_width=0.2)
    acc = np.array([s2.source(fc[f], _pform(sig[f], [0, 0, 0], 0.1 * sig[f]))[0] for f in range(m.nf)])
    assert float(np.sum(g.area * acc)) == pytest.approx(rate * 4 * np.pi, rel=1e-13)


def test_area_coeff_continuum_limit():
    """area_coeff -> (1/2) int_0^pi sin(t) N(t; pi/2, w/2.355) dt as the mesh is
    refined, the error falling with the face size."""
    w = 0.3
    sig = w / 2.355
    want = 0.5 * integrate.quad(lambda t: math.sin(t) * math.exp(-(t - math.pi / 2) ** 2 / (2 * sig * sig))
                                / math.sqrt(2 * math.pi * sig * sig), 0, math.pi, epsabs=1e-14)[0]
    errs = []
    for n in (4, 5, 6):
        m = O.Mesh("quad", n)
        errs.append(abs(_solver(m, accretion_on=True, acc_width=w).area_coeff() - want))
    assert errs[2] < 1e-4 and errs[2] < errs[1] < errs[0]


def _accretion_part(fcn, u, **kw):
    """The accretion terms alone: two acc_rates with friction on at alpha 0
    (no depletion, no drag) differ only by them; cooling cancels."""
    m = O.Mesh("hex", 2)
    a = _solver(m, accretion_on=True, friction_on=True, alpha=0.0, acc_rate=2e-3, **kw).source(fcn, u)
    b = _solver(m, accretion_on=True, friction_on=True, alpha=0.0, acc_rate=1e-3, **kw).source(fcn, u)
    return a - b


@pytest.mark.parametrize("tilt", [0.0, 0.3, -0.7])
def test_accreted_angular_momentum_and_energy(tilt):
    om_abs, e_acc = 0.8, 0.1
    om = om_abs * np.array([0, math.sin(tilt), math.cos(tilt)])
    rng = np.random.default_rng(7)
    for _ in range(6):
        r = _unit(rng.standard_normal(3))
        v = 0.2 * np.cross(rng.standard_normal(3), r)  # tangential
        sig, p = 1.3, 0.2
        u = _pform(sig, sig * np.cross(r, v), p)
        d = _accretion_part(r, u, tilt_angle=tilt, omega_acc_abs=om_abs, e_acc=e_acc, acc_width=2.0)
        assert d[0] > 0
        np.testing.assert_allclose(d[1:4] / d[0], om - np.dot(om, r) * r, rtol=0, atol=1e-12)
        assert d[4] / d[0] == pytest.approx(e_acc + np.sum((np.cross(om, r) - v) ** 2) / 2, rel=1e-12)


# --------------------------------------------------------------- cooling

K_M, C_SIGMA, KAPPA = 1.6e-13, 4.85e36, 3.4e6
BETA_SWITCH = 0.736194670678821
C_SWITCH = -1 - 1 / (BETA_SWITCH - 1)


def _cooling(C, sig=1.0, v=(0.0, 0.0, 0.0)):
    """res[4] of a gas at rest (or velocity v at x-hat) with acc_rate 0 and no
    drag: the cooling term only; p is chosen so that the cooling parameter
    takes the value C."""
    r = np.array([1.0, 0.0, 0.0])
    v = np.asarray(v, dtype=float)
    g_eff = GM - float(np.dot(v, v))
    p = 12 / 5 * K_M * sig * (0.75 * C_SIGMA * abs(g_eff) * sig) ** 0.25 / C
    m = O.Mesh("hex", 1)
    s = _solver(m, accretion_on=True, friction_on=True, alpha=0.0, acc_rate=0.0)
    res = s.source(r, _pform(sig, sig * np.cross(r, v), p))
    assert np.all(res[:4] == 0)
    return res[4], g_eff


@pytest.mark.parametrize("C", [0.1, 1.0, 2.0, 2.7])
def test_cooling_low_branch(C):
    got, g = _cooling(C)
    assert got == pytest.approx(-g / KAPPA / (1 + C), rel=1e-12)


@pytest.mark.parametrize("C", [3.0, 4.0, 10.0, 100.0])
def test_cooling_high_branch(C):
    got, g = _cooling(C)
    assert got == pytest.approx(-g / KAPPA * (2 / C) ** 4, rel=1e-12)


def test_cooling_continuous_at_switch_and_clamped():
    lo, g = _cooling(C_SWITCH * (1 - 1e-9))
    hi, _ = _cooling(C_SWITCH * (1 + 1e-9))
    assert hi == pytest.approx(lo, rel=1e-8)  # the branches meet only at the right switch point
    # beta clamps: 1 - beta stays within [1e-8, 1 - 1e-8]
    assert _cooling(1e3)[0] == pytest.approx(-g / KAPPA * 1e-8, rel=1e-12)
    assert _cooling(1e-10)[0] == pytest.approx(-g / KAPPA * (1 - 1e-8), rel=1e-12)


def test_cooling_with_negative_effective_gravity():
    """|v|^2 > GM makes g_eff < 0: C is NaN, beta takes the floor, and the
    term turns into heating -g_eff (1 - 1e-8) / kappa."""
    got, g = _cooling(1.0, v=(0.0, 0.6, 0.0))
    assert g < 0
    assert got == pytest.approx(-g / KAPPA * (1 - 1e-8), rel=1e-12) and got > 0


# -------------------------------------------------------------- friction

def _friction(r, v, omega_ns, alpha=1e-3, sig=1.2, p=0.15):
    m = O.Mesh("hex", 1)
    u = _pform(sig, sig * np.cross(r, v), p)
    a = _solver(m, accretion_on=True, friction_on=True, alpha=alpha, acc_rate=0.0, omega_ns=omega_ns).source(r, u)
    b = _solver(m, accretion_on=True, friction_on=True, alpha=0.0, acc_rate=0.0, omega_ns=omega_ns).source(r, u)
    return a - b, u  # cooling cancels


def test_friction_vanishes_for_corotation():
    rng = np.random.default_rng(3)
    for _ in range(5):
        r = _unit(rng.standard_normal(3))
        v0 = np.cross([0, 0, 0.1], r)
        d, _ = _friction(r, v0, 0.1)
        assert np.max(np.abs(d)) < 1e-18


def test_friction_decelerates_relative_velocity():
    gam3d = 1 / (2 - GAM)
    rng = np.random.default_rng(4)
    for _ in range(6):
        r = _unit(rng.standard_normal(3))
        v0 = np.cross([0, 0, 0.1], r)
        dv = 0.1 * np.cross(rng.standard_normal(3), r)
        v = v0 + dv
        sig, p, alpha = 1.2, 0.15, 1e-3
        d, u = _friction(r, v, 0.1, alpha, sig, p)
        g_eff = GM - np.dot(v, v)
        rho = gam3d / (2 * gam3d - 1) * g_eff * sig * sig / p
        ndv = np.linalg.norm(dv)
        assert d[0] == pytest.approx(-alpha * rho * ndv, rel=1e-12)
        # velocity change of the remaining gas: d(r x v)/dt = (dL - (L/Sigma) dSigma) / Sigma
        a = (d[1:4] - u[1:4] / sig * d[0]) / sig
        dvdt = -np.cross(r, a)
        np.testing.assert_allclose(dvdt, -alpha * rho * ndv * dv / sig, rtol=1e-10, atol=1e-18)


# ------------------------------------------------------- reconstruction

def _smooth_state(fc, gam, with_p=True):
    """Smooth Sigma, p and tangential v at unit points fc (E form) and the
    p-form values (Sigma, L, p)."""
    x, y, z = fc[:, 0], fc[:, 1], fc[:, 2]
    sig = 1 + 0.2 * x + 0.1 * y * z
    p = 0.1 + 0.05 * z * z + 0.02 * x
    v = np.cross([0.1, 0.2, 0.3], fc) + 0.05 * x[:, None] * np.cross([0, 0, 1.0], fc)
    L = sig[:, None] * np.cross(fc, v)
    E = p / (gam - 1) + sig * np.sum(v * v, axis=1) / 2
    U = np.column_stack([sig, L, E])
    return U


@pytest.mark.parametrize("kind", ["hex", "quad"])
def test_reconstruction_second_order(kind):
    """Area-weighted mean error of U_plus against the exact field at the unit
    edge midpoint (slot 4: total energy there) falls by ~4x per refinement."""
    errs = []
    levels = (3, 4, 5) if kind == "hex" else (4, 5, 6)
    for n in levels:
        m = O.Mesh(kind, n)
        g = O.Geometry(m)
        U = _smooth_state(_unit(g.fc), GAM)
        s = O.Solver(m, U, 5, O.HLLC, GAM)
        up = s.reconstruct(1e-7)
        a = m.fv
        b = np.concatenate([np.roll(m.fv[m.foff[f]:m.foff[f + 1]], -1) for f in range(m.nf)])
        ec = _unit(m.xyz[a] + m.xyz[b])
        want = _smooth_state(ec, GAM)
        owner = np.repeat(np.arange(m.nf), np.diff(m.foff))
        w = g.area[owner]
        e = np.sum(w[:, None] * np.abs(up - want), axis=0) / np.sum(w) / np.max(np.abs(want), axis=0)
        errs.append(e)
    errs = np.array(errs)
    ratios = errs[:-1] / errs[1:]
    print(f"{kind} reconstruction errors per level {errs.tolist()}, ratios {ratios.tolist()}")
    assert np.all(ratios[-1] > 3.0)


def _reference_composition(m, g, U, gam, dt):
    """U_plus from the pinned stencil tables and the pinned limiter with
    nu = (c_s + v . n) dt arc / A of the owner face (p-form U)."""
    nh = m.nh
    owner = np.repeat(np.arange(m.nf), np.diff(m.foff))
    val = np.diff(m.foff)[owner]
    a = m.fv
    b = np.concatenate([np.roll(m.fv[m.foff[f]:m.foff[f + 1]], -1) for f in range(m.nf)])
    Ui = U[owner]
    uhp = g.beta_plus[:, :1] * U[g.ff_plus[:, 0]] + g.beta_plus[:, 1:] * U[g.ff_plus[:, 1]]
    uhm = g.beta_minus[:, :1] * U[g.ff_minus[:, 0]] + g.beta_minus[:, 1:] * U[g.ff_minus[:, 1]]
    pp = (uhp - Ui) / g.h_plus[:, None]
    pm = (Ui - uhm) / g.h_minus[:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        r = pm / pp
    ec = _unit(m.xyz[a] + m.xyz[b])
    vel = np.cross(ec, Ui[:, 1:4]) / (-Ui[:, :1])
    c = np.sqrt(gam * Ui[:, 4] / Ui[:, 0])
    arc = 2 * np.arcsin(np.linalg.norm(m.xyz[a] - m.xyz[b], axis=1) / 2)
    nu = (c + np.sum(vel * g.en, axis=1)) * dt * arc / g.area[owner]
    out = np.empty((nh, 5))
    for h in range(nh):
        lim = O.limiter_adiabatic(r[h], nu[h], g.h_minus[h] / g.bm_dist[h], g.h_plus[h] / g.bm_dist[h], int(val[h]))
        out[h] = Ui[h] + pp[h] * lim * g.bm_dist[h]
    # slot 4: E at the edge midpoint from the reconstructed pressure
    v_e = np.cross(ec, out[:, 1:4]) / (-out[:, :1])
    out[:, 4] = out[:, 4] / (gam - 1) + out[:, 0] * np.sum(v_e * v_e, axis=1) / 2
    return out, nu


def test_reconstruction_courant_composition():
    m = O.Mesh("hex", 3)
    g = O.Geometry(m)
    rng = np.random.default_rng(11)
    fc = _unit(g.fc)
    U = _smooth_state(fc, GAM)
    U[:, 0] *= 1 + 0.3 * rng.random(m.nf)  # rough density: the limiters bind
    U[:, 4] *= 1 + 0.3 * rng.random(m.nf)
    s = O.Solver(m, U, 5, O.HLLC, GAM)
    dt = 0.02
    got = s.reconstruct(dt)
    W = U.copy()
    v = np.cross(fc, U[:, 1:4]) / U[:, :1]
    W[:, 4] = np.maximum(1e-16, (U[:, 4] - U[:, 0] * np.sum(v * v, axis=1) / 2) * (GAM - 1))
    want, nu = _reference_composition(m, g, W, GAM, dt)
    scale = np.max(np.abs(want), axis=0)
    assert np.max(np.abs(got - want) / scale) < 1e-12
    # the step matters here: nu at twice the step moves many states
    wrong, _ = _reference_composition(m, g, W, GAM, 2 * dt)
    moved = np.sum(np.max(np.abs(wrong - got) / scale, axis=1) > 1e-6)
    print(f"nu in [{nu.min():.3f}, {nu.max():.3f}], states moved by a doubled nu: {moved} of {m.nh}")
    assert moved > m.nh // 20
