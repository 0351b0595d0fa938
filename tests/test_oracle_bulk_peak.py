"""The bulk viscosity with zeta/s peaked at the transition temperature
(orc_set_bulk_peak; no counterpart in the reference, whose only model is
zeta = 2 eta (1/3 - c_s^2), viscous.f90:186-190).

Pin: a Bjorken grid of cells at rest with Pi = 0, temperatures across T_c
(tabulated EoS), one small first-order stage.  There the bulk pressure
grows only through the Navier-Stokes term, Pi(dt) proportional to
-zeta theta dt / tau_Pi with theta = 1 / tau, so the ratio of Pi between a
run with the peaked model and one with the reference's model is
zeta_peak / zeta_ref = zs_max exp(-(T - T_c)^2 / 2w^2) / (2 (eta/s)
(1/3 - c_s^2)) in every cell, with T and c_s^2 from the EoS.  A wrong
centre, width, temperature unit or scale fails it; the reference model is
pinned bit for bit against the reference build elsewhere."""
import numpy as np
import pytest

from libs import make_config

ZS, TC, W = 0.08, 0.155, 0.02


def _grid(oracle, peak, nx=24):
    ic, rc = make_config(nx=nx, ny=4, nz=4, coordinates=2, viscosity=1, bulk=1, eqstate=3, rec="WENO5", nrk=2,
                         bc=0, init_type=0, xlim=(-1, 1), ylim=(-1, 1), zlim=(-0.2, 0.2), tmin=0.6, cfl=0.4,
                         maxdt=1e-4, eta_over_s=0.08, tau_pi_coeff=5.0)
    oracle.setup(ic, rc)
    if peak:
        oracle.set_bulk_peak(ZS, TC, W)
    t, e, p, c = oracle.eos_table()
    # temperatures 0.10 .. 0.25 GeV across x, uniform in y and eta
    temps = np.linspace(0.10, 0.25, nx)
    ecell = np.interp(temps, t, e)
    v = np.zeros((12, 4, 4, nx))
    v[0] = 1.0
    for i in range(nx):
        err, out = oracle.eos_pressure(1.0, ecell[i])
        assert err == 0
        v[4, :, :, i] = out[0]
    return v


def test_peaked_bulk_viscosity_ratio(oracle):
    out = {}
    for peak in (False, True):
        v0 = _grid(oracle, peak)
        assert oracle.init_from_prim(v0) == 0
        assert oracle.stage(1) == 0
        out[peak] = oracle.state()[0][5, 1, 1].copy()
    pi_ref, pi_peak = out[False], out[True]
    assert np.all(pi_ref < 0) and np.all(pi_peak < 0)
    ratio = pi_peak / pi_ref
    temps, cs2 = [], []
    for i in range(v0.shape[3]):
        e, d = oracle.eos_derived(1.0, v0[4, 1, 1, i])
        temps.append(d[1])
        cs2.append(oracle.eos_sound(1.0, d[0], v0[4, 1, 1, i])[1])
    temps, cs2 = np.array(temps), np.array(cs2)
    expect = ZS * np.exp(-0.5 * ((temps - TC) / W) ** 2) / (2 * 0.08 * (1 / 3 - cs2))
    assert np.allclose(ratio, expect, rtol=1e-9, atol=0), np.abs(ratio / expect - 1).max()
    # the peak is there: largest ratio at the cell nearest T_c
    assert abs(temps[np.argmax(ratio * (1 / 3 - cs2))] - TC) < 0.01
