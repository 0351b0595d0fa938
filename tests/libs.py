"""ctypes loaders shared by the tests.

- ``Oracle`` wraps oracle/liboracle.so (the C restatement; test
  infrastructure).
- ``Ref`` wraps oracle/_ref/libechoref.so (the reference Fortran modules
  built from the reference sources by oracle/build_ref.sh; absent on machines
  without the reference).
Both expose the same methods so a test can run one scenario on each.
"""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "echo-qgp-mhd-devel_amd")
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libechoref.so")
EOS_DIR = PKG  # holds eos_data/qcdIEOS0.dat
EOS_TABLE = os.path.join(PKG, "eos_data", "qcdIEOS0.dat")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

REC_CODES = {"TVD2": 1, "CENO3": 2, "WENO3": 3, "WENO5": 4, "PPM4": 5,
             "MPE3": 6, "MPE5": 7, "MPE7": 8}
REC_NGC = {1: 2, 2: 3, 3: 2, 4: 3, 5: 2, 6: 3, 7: 3, 8: 4}


def make_config(nx=32, ny=32, nz=1, coordinates=2, viscosity=0, bulk=0, mhd=0,
                divclean=0, eqstate=1, rec="WENO5", nrk=3, bc=2, algo_solver=3,
                obtained_zz=1, init_type=4, derivatives_out=0,
                xlim=(-3.2, 3.2), ylim=(-3.2, 3.2), zlim=(-0.5, 0.5), tmin=1.0,
                cfl=0.4, maxdt=1.0, maxspeed=0.995, eta_over_s=0.08,
                tau_pi_coeff=5.0, enezero=0.001, glm_alpha=10.0, sigma_el=0.0058,
                sigma_chi=0.0015, dtout=1.0e6, smooth_temp=0.0, tmax=1.0e6):
    """Integer / real configuration arrays, laid out as oracle/ref_harness.f90."""
    ic = np.array([nx, ny, nz, coordinates, viscosity, bulk, mhd, divclean, eqstate,
                   REC_CODES[rec] if isinstance(rec, str) else rec, nrk, bc,
                   algo_solver, obtained_zz, init_type, derivatives_out],
                  dtype=np.int32)
    rc = np.array([xlim[0], xlim[1], ylim[0], ylim[1], zlim[0], zlim[1], tmin, cfl,
                   maxdt, maxspeed, eta_over_s, tau_pi_coeff, enezero, glm_alpha,
                   sigma_el, sigma_chi, dtout, smooth_temp, tmax], dtype=np.float64)
    return ic, rc


class _Lib:
    prefix = ""

    def _fn(self, name, res, args):
        f = getattr(self.lib, self.prefix + name)
        f.restype = res
        f.argtypes = args
        return f

    def _bind(self):
        dbl = C.c_double
        self.c_nv = self._fn("nv", C.c_int, [])
        self.c_init = self._fn("init_from_prim", C.c_int, [_dp])
        self.c_set_state = self._fn("set_state", None, [_dp, _dp])
        self.c_get_state = self._fn("get_state", None, [_dp, _dp])
        self.c_get_deriv = self._fn("get_deriv", None, [_dp])
        self.c_get_scalars = self._fn("get_scalars", None, [_dp])
        self.c_set_time = self._fn("set_time", None, [dbl])
        self.c_rec = self._fn("rec", None, [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp])
        self.c_p2c = self._fn("prim2cons", C.c_int, [_dp, _dp])
        self.c_c2p = self._fn("cons2prim", C.c_int, [_dp, _dp])
        self.c_flux = self._fn("flux", C.c_int, [_dp, C.c_int, _dp, _dp, _dp])
        self.c_eos_p = self._fn("eos_pressure", C.c_int, [dbl, dbl, _dp])
        self.c_eos_e = self._fn("eos_energy", C.c_int, [dbl, dbl, C.POINTER(dbl)])
        self.c_eos_c = self._fn("eos_sound", C.c_int, [dbl, dbl, dbl, C.POINTER(dbl)])
        self.c_eos_d = self._fn("eos_derived", C.c_int, [dbl, dbl, _dp])
        self.c_eos_tab = self._fn("eos_table", C.c_int, [C.c_int, _dp, _dp, _dp, _dp])
        self.c_przero = self._fn("przero", dbl, [])

    # -- grid-level -------------------------------------------------------
    def setup(self, ic, rc):
        self.ic, self.rc = ic, rc
        self.nx, self.ny, self.nz = int(ic[0]), int(ic[1]), int(ic[2])
        nv = self._setup(ic, rc)
        if nv <= 0:
            raise RuntimeError(f"setup failed: {nv}")
        self.nv = nv
        return nv

    @property
    def shape(self):
        return (self.nv, self.nz, self.ny, self.nx)

    def init_from_prim(self, vin):
        vin = np.ascontiguousarray(vin, dtype=np.float64)
        assert vin.shape == self.shape
        return self.c_init(vin.ravel())

    def set_state(self, v, u):
        self.c_set_state(np.ascontiguousarray(v).ravel(), np.ascontiguousarray(u).ravel())

    def state(self):
        v = np.zeros(self.shape)
        u = np.zeros(self.shape)
        self.c_get_state(v.ravel(), u.ravel())
        return v, u

    def deriv(self):
        d = np.zeros((20, self.nz, self.ny, self.nx))
        self.c_get_deriv(d.ravel())
        return d

    def scalars(self):
        s = np.zeros(10)
        self.c_get_scalars(s)
        return dict(t=s[0], dt=s[1], timeold=s[2], timeinterval=s[3],
                    g_cov_3_old=s[4], glm_ch_old=s[5], tout=s[6], iter=int(s[7]),
                    crashed=int(s[8]), errcode=int(s[9]))

    def steps(self, n):
        return self.c_steps(int(n))

    # -- pointwise --------------------------------------------------------
    def set_time(self, t):
        self.c_set_time(float(t))

    def rec(self, code, f):
        ngc = REC_NGC[code]
        f = np.ascontiguousarray(f, dtype=np.float64)
        n = f.size - 2 * ngc
        gl = np.zeros(n + 1)
        gr = np.zeros(n + 1)
        self.c_rec(code, n, ngc, f, gl, gr)
        return gl, gr

    def prim2cons(self, v):
        v = np.array(v, dtype=np.float64)
        u = np.zeros(self.nv)
        e = self.c_p2c(v, u)
        return e, u, v

    def cons2prim(self, u, vguess):
        u = np.array(u, dtype=np.float64)
        v = np.array(vguess, dtype=np.float64)
        e = self.c_c2p(u, v)
        return e, v, u

    def flux(self, v, d):
        v = np.array(v, dtype=np.float64)
        u = np.zeros(self.nv)
        f = np.zeros(self.nv)
        vf = np.zeros(2)
        e = self.c_flux(v, d, u, f, vf)
        return e, u, f, vf

    def eos_pressure(self, rh, en):
        o = np.zeros(3)
        e = self.c_eos_p(rh, en, o)
        return e, o

    def eos_energy(self, rh, pr):
        x = C.c_double(0.0)
        e = self.c_eos_e(rh, pr, C.byref(x))
        return e, x.value

    def eos_sound(self, rh, en, pr):
        x = C.c_double(0.0)
        e = self.c_eos_c(rh, en, pr, C.byref(x))
        return e, x.value

    def eos_derived(self, rh, pr):
        o = np.zeros(3)
        e = self.c_eos_d(rh, pr, o)
        return e, o

    def eos_table(self):
        n = 4096
        t, e, p, c = (np.zeros(n) for _ in range(4))
        m = self.c_eos_tab(n, t, e, p, c)
        return t[:m], e[:m], p[:m], c[:m]


class Oracle(_Lib):
    prefix = "orc_"

    def __init__(self):
        self.lib = C.CDLL(ORACLE_SO)
        self._bind()
        f = self.lib.orc_setup
        f.restype = C.c_int
        f.argtypes = [_ip, _dp, C.c_char_p]
        self._csetup = f
        self.c_steps = self._fn("steps", C.c_int, [C.c_int])

    def _setup(self, ic, rc):
        return self._csetup(ic, rc, EOS_TABLE.encode())


class Ref(_Lib):
    prefix = "ref_"

    def __init__(self):
        self.lib = C.CDLL(REF_SO)
        self._bind()
        f = self.lib.ref_setup
        f.restype = C.c_int
        f.argtypes = [_ip, _dp, C.c_int, C.c_char_p]
        self._csetup = f
        self.c_steps = self._fn("steps", C.c_int, [C.c_int])

    def _setup(self, ic, rc):
        name = b"qcdIEOS0.dat"
        cwd = os.getcwd()
        os.chdir(EOS_DIR)  # the reference opens eos_data/<name> relative to cwd
        try:
            return self._csetup(ic, rc, len(name), name)
        finally:
            os.chdir(cwd)


def ref_available():
    return os.path.exists(REF_SO)


# ---------------------------------------------------------------------------
# initial conditions (restated from init.f90, evaluated here in numpy)
# ---------------------------------------------------------------------------
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
