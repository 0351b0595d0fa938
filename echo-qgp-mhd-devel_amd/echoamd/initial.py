"""Run configurations and analytic initial states.

``make_config`` builds the two arrays (16 int32, 19 double) of struct
echo_config from the reference's param.dat keys (common.f90:874-1215);
``gubser_prim`` evaluates the ideal Gubser-flow initial state of the
reference initialiser (init.f90:2901-2961).
"""
import numpy as np

REC_CODES = {"TVD2": 1, "CENO3": 2, "WENO3": 3, "WENO5": 4, "PPM4": 5,
             "MPE3": 6, "MPE5": 7, "MPE7": 8}


def make_config(nx=32, ny=32, nz=1, coordinates=2, viscosity=0, bulk=0, mhd=0,
                divclean=0, eqstate=1, rec="WENO5", nrk=3, bc=2, algo_solver=3,
                obtained_zz=1, init_type=4, derivatives_out=0,
                xlim=(-3.2, 3.2), ylim=(-3.2, 3.2), zlim=(-0.5, 0.5), tmin=1.0,
                cfl=0.4, maxdt=1.0, maxspeed=0.995, eta_over_s=0.08,
                tau_pi_coeff=5.0, enezero=0.001, glm_alpha=10.0, sigma_el=0.0058,
                sigma_chi=0.0015, dtout=1.0e6, smooth_temp=0.0, tmax=1.0e6):
    """Integer / real configuration arrays in the field order of struct
    echo_config (EchoConfig.from_arrays turns them into the struct)."""
    ic = np.array([nx, ny, nz, coordinates, viscosity, bulk, mhd, divclean, eqstate,
                   REC_CODES[rec] if isinstance(rec, str) else rec, nrk, bc,
                   algo_solver, obtained_zz, init_type, derivatives_out],
                  dtype=np.int32)
    rc = np.array([xlim[0], xlim[1], ylim[0], ylim[1], zlim[0], zlim[1], tmin, cfl,
                   maxdt, maxspeed, eta_over_s, tau_pi_coeff, enezero, glm_alpha,
                   sigma_el, sigma_chi, dtout, smooth_temp, tmax], dtype=np.float64)
    return ic, rc



HBAR = 0.197326  # common.f90:47


def cell_centres(n, lo, hi):
    d = (hi - lo) / n
    return lo + (np.arange(1, n + 1) - 0.5) * d


def gubser_prim(nx, ny, xlim, ylim, tau, that0=2.0, q=1.0, nv=5, nz=1):
    """Ideal Gubser flow primitives (rho=1, v^x, v^y, 0, p=e/3) at proper
    time tau, following init.f90:2926-2935 with T0_idGub -> that0 (fm^-1)."""
    x = cell_centres(nx, *xlim)
    y = cell_centres(ny, *ylim)
    X, Y = np.meshgrid(x, y, indexing="xy")  # shape (ny, nx)
    r = np.sqrt(X ** 2 + Y ** 2)
    temp = (HBAR * that0 / tau * (4 * q ** 2 * tau ** 2) ** (1.0 / 3.0)
            / (1.0 + 2.0 * q ** 2 * (tau ** 2 + r ** 2) + q ** 4 * (tau ** 2 - r ** 2) ** 2) ** (1.0 / 3.0))
    e = 47.5 * np.pi ** 2 / (30.0 * HBAR ** 3) * temp ** 4
    k = np.arctanh(2 * q ** 2 * tau * r / (1.0 + (q * tau) ** 2 + (q * r) ** 2))
    vx = X / r * np.tanh(k)
    vy = Y / r * np.tanh(k)
    v = np.zeros((nv, nz, ny, nx))
    v[0] = 1.0
    v[1] = vx[None]
    v[2] = vy[None]
    v[4] = (e / 3.0)[None]
    return v
