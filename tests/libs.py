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
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "echo-qgp-mhd-devel_amd")
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libechoref.so")
EOS_DIR = PKG  # holds eos_data/qcdIEOS0.dat
EOS_TABLE = os.path.join(PKG, "eos_data", "qcdIEOS0.dat")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

if PKG not in sys.path:
    sys.path.insert(0, PKG)
from echoamd.initial import (HBAR, REC_CODES, cell_centres, gubser_prim,  # noqa: E402,F401
                             make_config)

REC_NGC = {1: 2, 2: 3, 3: 2, 4: 3, 5: 2, 6: 3, 7: 3, 8: 4}


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
        self.c_stage = self._fn("stage", C.c_int, [C.c_int])
        self.c_u2v = self._fn("u2v", C.c_int, [])

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

    def stage(self, irk):
        """one RK stage (evolve_main) without the step bookkeeping"""
        return self.c_stage(int(irk))

    def u2v(self):
        """evolve_u2v alone on the current v, u"""
        return self.c_u2v()

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

    # names of the repair counters (REP_* of oracle/oracle_int.h)
    REPAIRS = ("rec_pr", "rec_v", "rec_vmax", "u2v", "u2v_v", "u2v_pr", "u2v_fail", "c2p_undef")

    def repair_counts(self):
        """how often each repair branch fired since setup (dict by REPAIRS)"""
        a = (C.c_long * 8)()
        self.lib.orc_get_repair_counts(a)
        return dict(zip(self.REPAIRS, list(a)))

    def set_bulk_peak(self, zs_max, t_c=0.155, width=0.02):
        """zeta/s peaked at t_c (orc_set_bulk_peak)"""
        f = self.lib.orc_set_bulk_peak
        f.restype = None
        f.argtypes = [C.c_double] * 3
        f(float(zs_max), float(t_c), float(width))

    def set_maxdt(self, maxdt):
        """MAXDT from now on"""
        f = self.lib.orc_set_maxdt
        f.restype = None
        f.argtypes = [C.c_double]
        f(float(maxdt))

    def crash_cell(self):
        """the cell whose con2prim repair crashed the run (-1: none)"""
        f = self.lib.orc_crash_cell
        f.restype = C.c_long
        return f()

    def repair_cells(self):
        """cells (x fastest) of the first 256 con2prim repairs since setup"""
        a = (C.c_long * 256)()
        f = self.lib.orc_get_repair_cells
        f.restype = C.c_int
        m = f(a, 256)
        return list(a[:m])


def oracle_set_stop(oracle, temp_end, freeze_type=0):
    f = oracle.lib.orc_set_stop
    f.restype = None
    f.argtypes = [C.c_double, C.c_int]
    f(temp_end, freeze_type)


def oracle_get_stop(oracle):
    f = oracle.lib.orc_get_stop
    f.restype = None
    f.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_double)]
    st, mx = C.c_int(), C.c_double()
    f(C.byref(st), C.byref(mx))
    return st.value, mx.value


def oracle_glauber_prim(oracle, params, part_a, part_b, coll):
    """Primitive Glauber-MC state of the oracle's current grid (oracle_init.c)."""
    f = oracle.lib.orc_glauber_prim
    f.restype = C.c_int
    f.argtypes = [_dp, _dp, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp]
    par = np.array([params[k] for k in ("kappa", "ah", "sig_mc", "ybeam", "deta", "sigeta",
                                        "rhocenter")], dtype=np.float64)
    arrs = [np.ascontiguousarray(np.reshape(a, (-1, 2)), dtype=np.float64) for a in (part_a, part_b, coll)]
    flat = [a.ravel() if a.size else np.zeros(2) for a in arrs]
    n = oracle.nx * oracle.ny * oracle.nz
    v = np.zeros(oracle.nv * n)
    e = f(par, flat[0], arrs[0].shape[0], flat[1], arrs[1].shape[0], flat[2], arrs[2].shape[0], v)
    return e, v.reshape(oracle.nv, oracle.nz, oracle.ny, oracle.nx)


def oracle_glauber_bfield(oracle, bparams, nucleons, v):
    """Adds the oracle's initial magnetic field (oracle_init.c,
    orc_glauber_bfield) to the primitives v (nv, nz, ny, nx), in place.
    nucleons: four (xy (n, 2), type (n,)) pairs: spectators A, participants
    A, spectators B, participants B."""
    f = oracle.lib.orc_glauber_bfield
    f.restype = C.c_int
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    f.argtypes = [_dp] + [_dp, ip, C.c_int] * 4 + [_dp]
    par = np.array([bparams["gamma_col"], bparams["b_amplify"], bparams["pw"],
                    float(bparams["magfield_type"])], dtype=np.float64)
    args = []
    for xy, ty in nucleons:
        xy = np.ascontiguousarray(np.reshape(xy, (-1, 2)), dtype=np.float64)
        ty = np.ascontiguousarray(ty, dtype=np.int32)
        args += [xy.ravel() if xy.size else np.zeros(2), ty if ty.size else np.zeros(1, np.int32),
                 xy.shape[0]]
    flat = np.ascontiguousarray(v, dtype=np.float64).ravel()
    e = f(par, *args, flat)
    v[...] = flat.reshape(v.shape)
    return e


class OracleFO:
    """The oracle's freeze-out hypersurface finder (oracle_fo.c)."""

    def __init__(self, oracle):
        lib = oracle.lib
        self.o = oracle
        self.c_setup = lib.orc_fo_setup
        self.c_setup.restype = C.c_int
        self.c_setup.argtypes = [C.c_int, C.c_double, _dp]
        self.c_slice = lib.orc_fo_slice
        self.c_slice.restype = C.c_int
        self.c_slice.argtypes = [C.POINTER(C.c_long)]
        self.c_width = lib.orc_fo_width
        self.c_width.restype = C.c_int
        self.c_get = lib.orc_fo_get
        self.c_get.restype = None
        self.c_get.argtypes = [_dp]
        self.c_write = lib.orc_fo_write
        self.c_write.restype = C.c_int
        self.c_write.argtypes = [C.c_char_p, C.c_int]
        self.n = 0

    def setup(self, freeze_type, freeze_value):
        hdr = np.zeros(2)
        e = self.c_setup(int(freeze_type), float(freeze_value), hdr)
        return e, hdr

    def slice(self):
        n = C.c_long(0)
        e = self.c_slice(C.byref(n))
        self.n = n.value
        return e, n.value

    @property
    def width(self):
        return self.c_width()

    def elements(self):
        a = np.zeros((max(self.n, 1), self.width))
        self.c_get(a.ravel())
        return a[:self.n]

    def write(self, path, header):
        return self.c_write(path.encode(), int(header))


class RefFO:
    """The reference's own freeze-out path (out_alloc + get_hysufile_ready,
    out_hyp) through the harness; files go to <outdir>/outr0001/."""

    def __init__(self, ref, outdir):
        self.r = ref
        self.outdir = outdir
        self.path = os.path.join(outdir, "outr0001", "hypersurface.txt")
        os.makedirs(os.path.join(outdir, "outr0001"), exist_ok=True)
        lib = ref.lib
        self.c_setup = lib.ref_fo_setup
        self.c_setup.restype = None
        self.c_setup.argtypes = [C.c_int, C.c_double, C.c_int, C.c_char_p]
        self.c_slice = lib.ref_fo_slice
        self.c_slice.restype = C.c_int
        self.c_get = lib.ref_fo_get
        self.c_get.restype = None
        self.c_get.argtypes = [C.c_int, C.c_int, _dp]

    def _in_eos_dir(self, fn, *a):
        cwd = os.getcwd()
        os.chdir(EOS_DIR)
        try:
            return fn(*a)
        finally:
            os.chdir(cwd)

    def setup(self, freeze_type, freeze_value):
        self._in_eos_dir(self.c_setup, int(freeze_type), float(freeze_value), len(self.outdir),
                         self.outdir.encode())

    def slice(self):
        return self._in_eos_dir(self.c_slice)

    def last_count(self):
        """element count of the last slice in the file (its last "(i6)" line)"""
        lines = open(self.path).read().splitlines()
        idx = [i for i, ln in enumerate(lines) if len(ln) == 6 and ln.strip().isdigit()]
        return int(lines[idx[-1]]) if idx else 0

    def elements(self, n, width):
        a = np.zeros((max(n, 1), width))
        self.c_get(n, width, a.ravel())
        return a[:n]


def oracle_dump(oracle, path, iout, thyp=0.0):
    f = oracle.lib.orc_dump
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_int, C.c_double]
    return f(str(path).encode(), int(iout), float(thyp))


def oracle_restart(oracle, path):
    """returns (error, iout, thyp)"""
    f = oracle.lib.orc_restart
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    io, th = C.c_int(0), C.c_double(0.0)
    e = f(str(path).encode(), C.byref(io), C.byref(th))
    return e, io.value, th.value


def ref_dump(ref, outdir, iout, thyp=0.0, setup=True):
    """the reference's out_dump of its current state into <outdir>/outr0001/dump.dat;
    setup=False when ref_fo_setup has already made the output set-up"""
    os.makedirs(os.path.join(outdir, "outr0001"), exist_ok=True)
    if setup:
        f = ref.lib.ref_dump_setup
        f.restype = None
        f.argtypes = [C.c_int, C.c_char_p]
        f(len(outdir), outdir.encode())
    g = ref.lib.ref_out_dump
    g.restype = None
    g.argtypes = [C.c_int, C.c_double]
    g(int(iout), float(thyp))
    return os.path.join(outdir, "outr0001", "dump.dat")


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


class OracleMLP:
    """ctypes wrapper of the learned-con2prim oracle (oracle/oracle_mlp.c)."""

    NPARAM = 9221

    def __init__(self):
        lib = C.CDLL(ORACLE_SO)
        fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        lib.orc_f16.restype = C.c_double
        lib.orc_f16.argtypes = [C.c_double]
        lib.orc_act_h.restype = C.c_double
        lib.orc_act_h.argtypes = [C.c_float]
        lib.orc_act.restype = C.c_float
        lib.orc_act.argtypes = [C.c_float]
        lib.orc_act_d.restype = C.c_float
        lib.orc_act_d.argtypes = [C.c_float]
        lib.orc_mlp_features.restype = None
        lib.orc_mlp_features.argtypes = [dp, C.c_double, fp]
        lib.orc_mlp_targets.restype = None
        lib.orc_mlp_targets.argtypes = [dp, dp, C.c_double, dp]
        lib.orc_mlp_forward.restype = None
        lib.orc_mlp_forward.argtypes = [fp, fp, C.c_int, fp]
        lib.orc_mlp_loss.restype = C.c_double
        lib.orc_mlp_loss.argtypes = [fp, dp, dp, C.c_long, C.c_double, C.c_int, C.c_void_p]
        lib.orc_mlp_w0.restype = C.c_double
        lib.orc_mlp_w0.argtypes = [dp, C.c_double]
        lib.orc_e3p_w.restype = None
        lib.orc_e3p_w.argtypes = [C.c_double] * 5 + [dp]
        lib.orc_mlp_refine.restype = C.c_int
        lib.orc_mlp_refine.argtypes = [dp, fp, C.c_double, C.c_double, C.c_int, C.c_double, dp]
        self.lib = lib

    def f16(self, x):
        return self.lib.orc_f16(x)

    def act_h(self, z):
        return self.lib.orc_act_h(z)

    def act(self, x):
        return self.lib.orc_act(x)

    def act_d(self, x):
        return self.lib.orc_act_d(x)

    def features(self, u, g33=1.0):
        x = np.zeros(8, dtype=np.float32)
        self.lib.orc_mlp_features(np.ascontiguousarray(u, dtype=np.float64), g33, x)
        return x

    def targets(self, u, v, g33=1.0):
        t = np.zeros(5)
        self.lib.orc_mlp_targets(np.ascontiguousarray(u, dtype=np.float64),
                                 np.ascontiguousarray(v, dtype=np.float64), g33, t)
        return t

    def forward(self, P, x, rnd=1):
        y = np.zeros(5, dtype=np.float32)
        self.lib.orc_mlp_forward(np.ascontiguousarray(P, dtype=np.float32),
                                 np.ascontiguousarray(x, dtype=np.float32), rnd, y)
        return y

    def loss(self, P, u, v, rnd=1, grad=False, g33=1.0):
        """u, v: (8, n) (u de-densitised).  Returns loss or (loss, grad)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float32)
        g = np.zeros(self.NPARAM) if grad else None
        L = self.lib.orc_mlp_loss(P, u, v, u.shape[1], g33, rnd,
                                  g.ctypes.data if grad else None)
        return (L, g) if grad else L

    def w0(self, u, g33=1.0):
        return self.lib.orc_mlp_w0(np.ascontiguousarray(u, dtype=np.float64), g33)

    def e3p_w(self, W, S2, B2, SB2, U):
        """(f, df/dW, x) of the W-refinement residual."""
        o = np.zeros(3)
        self.lib.orc_e3p_w(W, S2, B2, SB2, U, o)
        return o

    def refine(self, u, y, g33=1.0, gm=1.0, iters=2, tol=1e-11):
        v = np.zeros(8)
        st = self.lib.orc_mlp_refine(np.ascontiguousarray(u, dtype=np.float64),
                                     np.ascontiguousarray(y, dtype=np.float32), g33, gm,
                                     int(iters), float(tol), v)
        return st, v


def ref_available():
    return os.path.exists(REF_SO)


