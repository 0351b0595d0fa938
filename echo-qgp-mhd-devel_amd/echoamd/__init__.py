"""Python binding of the C-ABI in include/echo_amd.h (libechoamd.so).

The library is the MI355X product path; there is no fallback.  Loading raises
``RuntimeError`` when libechoamd.so has not been built, and every call raises
``EchoError`` when the C-ABI returns an error code.

Arrays use the reference's Fortran layout: shape (nv, nz, ny, nx) in numpy
C order, i.e. a(ix, iy, iz, k) of the Fortran arrays v/u (common.f90:203-214).
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ECHO_AMD_LIB selects another build of the library (e.g. an experiment)
LIB_PATH = os.environ.get("ECHO_AMD_LIB", os.path.join(PKG_DIR, "libechoamd.so"))
EOS_DIR = os.path.join(PKG_DIR, "eos_data")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")

# the exported symbols of include/echo_amd.h
EXPORTS = (
    "echo_create", "echo_destroy", "echo_nv", "echo_init_from_prim",
    "echo_upload_state", "echo_download_state", "echo_steps", "echo_rk_stage",
    "echo_end_step", "echo_get_scalars", "echo_set_tout", "echo_cons2prim_batch",
    "echo_set_timing", "echo_get_timing", "echo_last_error",
    "echo_cons2prim_batch_dev", "echo_mlp_create", "echo_mlp_destroy", "echo_mlp_set_params",
    "echo_mlp_get_params", "echo_mlp_forward", "echo_mlp_grad", "echo_mlp_train",
    "echo_c2p_surrogate", "echo_c2p_surrogate_noalloc", "echo_set_bulk_peak", "echo_glauber_generate",
    "echo_glauber_read_event", "echo_set_tile", "echo_get_tile", "echo_tune",
    "echo_ensemble_create", "echo_ensemble_destroy", "echo_ensemble_size", "echo_ensemble_grid",
    "echo_ensemble_tune", "echo_ensemble_steps", "echo_ensemble_last_ms",
    "echo_dev_alloc", "echo_dev_free", "echo_dev_upload", "echo_dev_download",
    "echo_init_glauber", "echo_set_stop", "echo_get_stop", "echo_write_frame", "echo_wait_frames",
    "echo_fo_setup", "echo_fo_slice", "echo_fo_width", "echo_fo_elements", "echo_fo_write",
    "echo_set_dump", "echo_dump", "echo_restart", "echo_init_glauber_mhd", "echo_get_repairs",
    "echo_set_c2p_net", "echo_get_c2p_fallbacks",
)

# out_vars field flags (type out_type of the reference), in file order
OUT_FIELDS = ("density", "vx", "vy", "vz", "pressure", "energy_density", "temperature",
              "entropy_density", "bulk", "pitt", "pitx", "pity", "pitz", "pixy", "pixz", "piyz",
              "pixx", "piyy", "pizz", "v0", "bx", "by", "bz", "ex", "ey", "ez", "glm", "rc")

MLP_NPARAM = 9221

REC_CODES = {"TVD2": 1, "CENO3": 2, "WENO3": 3, "WENO5": 4, "PPM4": 5, "MPE3": 6, "MPE5": 7}


class EchoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"echo error {code}: {msg}")
        self.code = code


class EchoConfig(C.Structure):
    """Mirror of struct echo_config (16 int32 followed by 19 double)."""
    _fields_ = [(n, C.c_int32) for n in (
        "nx", "ny", "nz", "coordinates", "viscosity", "bulk", "mhd", "divclean",
        "eqstate", "recal", "nrk", "boundary_cond", "algo_solver", "obtained_zz",
        "init_type", "derivatives_out")] + [(n, C.c_double) for n in (
        "x1", "x2", "y1", "y2", "z1", "z2", "tmin", "cfl", "maxdt", "maxspeed",
        "eta_over_s", "tau_pi_coeff", "enezero", "glm_alpha", "sigma_el", "sigma_chi",
        "dtout", "smooth_temp", "tmax")]

    @classmethod
    def from_arrays(cls, ic, rc):
        """Build from the (16 int32, 19 double) arrays used by the oracle."""
        ic = np.ascontiguousarray(ic, dtype=np.int32)
        rc = np.ascontiguousarray(rc, dtype=np.float64)
        assert ic.size == 16 and rc.size == 19
        return cls.from_buffer_copy(ic.tobytes() + rc.tobytes())


class GlauberParams(C.Structure):
    """Mirror of struct echo_glauber_params."""
    _fields_ = [(n, C.c_double) for n in ("kappa", "ah", "sig_mc", "ybeam", "deta", "sigeta",
                                          "rhocenter")]


class GlauberSampler(C.Structure):
    _fields_ = [("A", C.c_int32), ("Z", C.c_int32), ("nconf", C.c_int32), ("nbcoll", C.c_int32),
                ("min_participants", C.c_int32), ("fixed_b", C.c_int32), ("radius", C.c_double),
                ("diffuse", C.c_double), ("rho0", C.c_double), ("sigma_nn", C.c_double),
                ("bimpact", C.c_double), ("bmax", C.c_double)]


class GlauberEvent(C.Structure):
    _fields_ = [("b", C.c_double), ("n_part_a", C.c_int32), ("n_part_b", C.c_int32),
                ("n_spec_a", C.c_int32), ("n_spec_b", C.c_int32), ("n_coll", C.c_int32),
                ("part_xy", C.c_void_p), ("part_type", C.c_void_p), ("spec_xy", C.c_void_p),
                ("spec_type", C.c_void_p), ("coll_xy", C.c_void_p)]


# nuclei of nuclear_data.dat (Woods-Saxon R, a in fm, roze in fm^-3)
NUCLEI = {
    "Au": dict(A=197, Z=79, radius=6.38, diffuse=0.535, rho0=0.1693),
    "Pb": dict(A=208, Z=82, radius=6.62, diffuse=0.546, rho0=0.1604),
}


def glauber_generate(directory, seed, nucleus="Au", sigma_nn=42.0, nconf=100, nbcoll=1, b=None, bmax=20.0,
                     min_participants=18):
    """Glauber-MC events (echo_glauber_generate) into directory; b: fixed
    impact parameter in fm, or None for p(b) ~ b on [0, bmax].  Returns the
    number of events written."""
    lib = load_library()
    nuc = NUCLEI[nucleus] if isinstance(nucleus, str) else nucleus
    sp = GlauberSampler(nconf=nconf, nbcoll=nbcoll, min_participants=min_participants,
                        fixed_b=0 if b is None else 1, sigma_nn=sigma_nn, bimpact=0.0 if b is None else b,
                        bmax=bmax, **nuc)
    n = C.c_int32()
    _check(lib, lib.echo_glauber_generate(C.byref(sp), C.c_uint64(seed), str(directory).encode(), C.byref(n)))
    return n.value


def glauber_read_event(directory, iev, cap=600):
    """Event iev (1-based) of directory (echo_glauber_read_event): dict of b,
    part_a, part_b, spec_a, spec_b ((n, 2) arrays, protons first), their
    types (0 proton, 1 neutron) and coll."""
    lib = load_library()
    pxy, sxy, cxy = np.zeros(2 * cap), np.zeros(2 * cap), np.zeros(2 * cap * cap // 4)
    pt, st = np.zeros(cap, dtype=np.int32), np.zeros(cap, dtype=np.int32)
    ev = GlauberEvent(part_xy=pxy.ctypes.data, part_type=pt.ctypes.data, spec_xy=sxy.ctypes.data,
                      spec_type=st.ctypes.data, coll_xy=cxy.ctypes.data)
    _check(lib, lib.echo_glauber_read_event(str(directory).encode(), int(iev), C.byref(ev), cap, cap,
                                            cap * cap // 8))
    na, nb, sa, sb, nc = ev.n_part_a, ev.n_part_b, ev.n_spec_a, ev.n_spec_b, ev.n_coll
    P = pxy[:2 * (na + nb)].reshape(-1, 2)
    S = sxy[:2 * (sa + sb)].reshape(-1, 2)
    return dict(b=ev.b, part_a=P[:na].copy(), part_b=P[na:].copy(), type_part_a=pt[:na].copy(),
                type_part_b=pt[na:na + nb].copy(), spec_a=S[:sa].copy(), spec_b=S[sa:].copy(),
                type_spec_a=st[:sa].copy(), type_spec_b=st[sa:sa + sb].copy(),
                coll=cxy[:2 * nc].reshape(-1, 2).copy())


class BFieldParams(C.Structure):
    """Mirror of struct echo_bfield_params."""
    _fields_ = [("gamma_col", C.c_double), ("b_amplify", C.c_double), ("pw", C.c_double),
                ("magfield_type", C.c_int32)]


class Nucleons(C.Structure):
    """Mirror of struct echo_nucleons."""
    _fields_ = [("xy", C.POINTER(C.c_double)), ("type", C.POINTER(C.c_int32)), ("n", C.c_int32)]


class TrainConfig(C.Structure):
    """Mirror of struct echo_mlp_train_cfg."""
    _fields_ = [("steps", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float),
                ("lr_final", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float)]


_lib = None


def load_library(path=LIB_PATH):
    """Load libechoamd.so and declare the C-ABI signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run echo-qgp-mhd-devel_amd/build.sh (or "
            "__graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(path)
    P = C.c_void_p
    sig = {
        "echo_create": (C.c_int, [C.POINTER(EchoConfig), _dp, C.c_int32, C.c_int32, C.POINTER(P)]),
        "echo_destroy": (None, [P]),
        "echo_nv": (C.c_int, [P]),
        "echo_init_from_prim": (C.c_int, [P, _dp]),
        "echo_upload_state": (C.c_int, [P, _dp, _dp]),
        "echo_download_state": (C.c_int, [P, _dp, _dp]),
        "echo_steps": (C.c_int, [P, C.c_int32]),
        "echo_rk_stage": (C.c_int, [P, C.c_int32]),
        "echo_end_step": (C.c_int, [P]),
        "echo_get_scalars": (C.c_int, [P, _dp]),
        "echo_get_repairs": (C.c_int, [P, C.POINTER(C.c_int64)]),
        "echo_set_tout": (C.c_int, [P, C.c_double]),
        "echo_cons2prim_batch": (C.c_int, [P, _dp, _dp, _ip, C.c_int64]),
        "echo_set_timing": (C.c_int, [P, C.c_int32]),
        "echo_get_timing": (C.c_int, [P, _dp]),
        "echo_last_error": (C.c_char_p, []),
        "echo_cons2prim_batch_dev": (C.c_int, [P, P, P, P, C.c_int64]),
        "echo_mlp_create": (C.c_int, [C.c_int32, C.POINTER(P)]),
        "echo_mlp_destroy": (None, [P]),
        "echo_mlp_set_params": (C.c_int, [P, _fp, C.c_int64]),
        "echo_mlp_get_params": (C.c_int, [P, _fp, C.c_int64]),
        "echo_mlp_forward": (C.c_int, [P, P, P, C.c_int64, P]),
        "echo_mlp_grad": (C.c_int, [P, P, P, P, C.c_int64, _fp, C.POINTER(C.c_double)]),
        "echo_mlp_train": (C.c_int, [P, P, P, P, C.c_int64, C.POINTER(TrainConfig), P]),
        "echo_c2p_surrogate": (C.c_int, [P, P, P, P, P, C.c_int64, C.c_int32, C.c_double,
                                         C.POINTER(C.c_int64)]),
        "echo_c2p_surrogate_noalloc": (C.c_int, [P, P, P, P, P, C.c_int64, C.c_int32, C.c_double,
                                                 C.POINTER(C.c_int64)]),
        "echo_set_bulk_peak": (C.c_int, [P, C.c_double, C.c_double, C.c_double]),
        "echo_set_c2p_net": (C.c_int, [P, P, C.c_int32, C.c_double]),
        "echo_get_c2p_fallbacks": (C.c_int, [P, C.POINTER(C.c_int64)]),
        "echo_glauber_generate": (C.c_int, [P, C.c_uint64, C.c_char_p, C.POINTER(C.c_int32)]),
        "echo_glauber_read_event": (C.c_int, [C.c_char_p, C.c_int32, P, C.c_int32, C.c_int32, C.c_int32]),
        "echo_set_tile": (C.c_int, [P, C.c_int32, C.c_int32]),
        "echo_get_tile": (C.c_int, [P, _ip]),
        "echo_tune": (C.c_int, [P, C.c_int32, _dp]),
        "echo_ensemble_create": (C.c_int, [C.POINTER(EchoConfig), C.c_int32, _dp, C.c_int32,
                                           C.c_int32, C.POINTER(P)]),
        "echo_ensemble_destroy": (None, [P]),
        "echo_ensemble_size": (C.c_int, [P]),
        "echo_ensemble_grid": (P, [P, C.c_int32]),
        "echo_ensemble_tune": (C.c_int, [P, C.c_int32]),
        "echo_ensemble_steps": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int32)]),
        "echo_ensemble_last_ms": (C.c_int, [P, C.POINTER(C.c_double)]),
        "echo_init_glauber": (C.c_int, [P, C.POINTER(GlauberParams), _dp, C.c_int32, _dp,
                                        C.c_int32, _dp, C.c_int32]),
        "echo_init_glauber_mhd": (C.c_int, [P, C.POINTER(GlauberParams), _dp, C.c_int32, _dp,
                                            C.c_int32, _dp, C.c_int32, C.POINTER(BFieldParams),
                                            C.POINTER(Nucleons)]),
        "echo_set_stop": (C.c_int, [P, C.c_double, C.c_int32]),
        "echo_write_frame": (C.c_int, [P, _ip, C.c_int32, C.c_int32, C.c_char_p]),
        "echo_wait_frames": (C.c_int, [P]),
        "echo_get_stop": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
        "echo_set_dump": (C.c_int, [P, C.c_int32]),
        "echo_dump": (C.c_int, [P, C.c_char_p, C.c_int32, C.c_double]),
        "echo_restart": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
        "echo_fo_setup": (C.c_int, [P, C.c_int32, C.c_double, _dp]),
        "echo_fo_slice": (C.c_int, [P, C.POINTER(C.c_int64)]),
        "echo_fo_width": (C.c_int, [P]),
        "echo_fo_elements": (C.c_int, [P, _dp, C.c_int64]),
        "echo_fo_write": (C.c_int, [P, C.c_char_p, C.c_int32]),
        "echo_dev_alloc": (C.c_int, [C.c_int32, C.c_int64, C.POINTER(P)]),
        "echo_dev_free": (C.c_int, [C.c_int32, P]),
        "echo_dev_upload": (C.c_int, [C.c_int32, P, P, C.c_int64]),
        "echo_dev_download": (C.c_int, [C.c_int32, P, P, C.c_int64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def read_eos_rows(name="qcdIEOS0.dat"):
    """Rows (T, e/T^4, p/T^4, c_s^2) of a reference EoS table file."""
    path = os.path.join(EOS_DIR, name)
    with open(path) as fh:
        n = int(fh.readline().split()[0])
        rows = np.loadtxt(fh, max_rows=n)
    return np.ascontiguousarray(rows, dtype=np.float64)


def _check(lib, rc):
    if rc != 0:
        raise EchoError(rc, lib.echo_last_error().decode())
    return rc


class DeviceArray:
    """A device buffer allocated through the library (echo_dev_alloc), for
    the entry points that take device pointers.  ``ptr`` is the address."""

    def __init__(self, shape, dtype=np.float64, device=0):
        self.lib = load_library()
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.device = device
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(self.lib, self.lib.echo_dev_alloc(device, self.nbytes, C.byref(p)))
        self._p = p

    @property
    def ptr(self):
        return self._p.value or 0

    @classmethod
    def from_numpy(cls, a, device=0):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _check(self.lib, self.lib.echo_dev_upload(self.device, self._p, a.ctypes.data, self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.lib, self.lib.echo_dev_download(self.device, out.ctypes.data, self._p, self.nbytes))
        return out

    def free(self):
        if getattr(self, "_p", None) is not None and self._p.value:
            self.lib.echo_dev_free(self.device, self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class EchoGrid:
    """One event's grid on one GPU (echo_grid of the C-ABI)."""

    def __init__(self, ic, rc, device=0, eos_rows=None, _handle=None):
        self.lib = load_library()
        if _handle is not None:
            # a grid owned by an ensemble
            self.cfg = EchoConfig.from_arrays(ic, rc)
            self.nx, self.ny, self.nz = self.cfg.nx, self.cfg.ny, self.cfg.nz
            self.h = C.c_void_p(_handle)
            self._owned = False
            self.nv = self.lib.echo_nv(self.h)
            return
        self._owned = True
        self.cfg = EchoConfig.from_arrays(ic, rc)
        self.nx, self.ny, self.nz = self.cfg.nx, self.cfg.ny, self.cfg.nz
        if self.cfg.eqstate == 3 and eos_rows is None:
            eos_rows = read_eos_rows()
        rows = np.zeros((1, 4)) if eos_rows is None else np.ascontiguousarray(eos_rows, dtype=np.float64)
        n_rows = 0 if eos_rows is None else rows.shape[0]
        h = C.c_void_p()
        self._check(self.lib.echo_create(C.byref(self.cfg), rows.ravel(), n_rows, device, C.byref(h)))
        self.h = h
        self.nv = self.lib.echo_nv(self.h)

    def _check(self, rc):
        if rc != 0:
            raise EchoError(rc, self.lib.echo_last_error().decode())
        return rc

    @property
    def shape(self):
        return (self.nv, self.nz, self.ny, self.nx)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.lib.echo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile(self, sweep, rows):
        self._check(self.lib.echo_set_tile(self.h, int(sweep), int(rows)))

    def tile(self):
        t = np.zeros(3, dtype=np.int32)
        self._check(self.lib.echo_get_tile(self.h, t))
        return t

    def tune(self, reps=5):
        """Autotune the flux tile shapes; returns the 3x3 per-launch times (ms)."""
        ms = np.zeros(9)
        self._check(self.lib.echo_tune(self.h, int(reps), ms))
        return ms.reshape(3, 3)

    def init_from_prim(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.shape == self.shape, (v.shape, self.shape)
        return self._check(self.lib.echo_init_from_prim(self.h, v.ravel()))

    def write_frame(self, fields, iout, outdir, precision=8):
        """Queue an output frame of the named fields (OUT_FIELDS) as
        <outdir>/out<iout>.dat; returns at once (see wait_frames)."""
        sel = np.zeros(len(OUT_FIELDS), dtype=np.int32)
        for f in fields:
            sel[OUT_FIELDS.index(f)] = 1
        self._check(self.lib.echo_write_frame(self.h, sel, int(precision), int(iout),
                                              os.fsencode(outdir)))

    def wait_frames(self):
        self._check(self.lib.echo_wait_frames(self.h))

    def set_stop(self, temp_end, freeze_type=0):
        """check_temp stop criterion (temp_end <= 0 disables it)."""
        self._check(self.lib.echo_set_stop(self.h, float(temp_end), int(freeze_type)))

    def stop_state(self):
        """(stopped, maximum of the last check)."""
        st, mx = C.c_int32(), C.c_double()
        self._check(self.lib.echo_get_stop(self.h, C.byref(st), C.byref(mx)))
        return st.value, mx.value

    # -- restart dump (out_dump / out_restart) ---------------------------
    def set_dump(self, on=True):
        """keep vold / uold on the device for echo_dump"""
        self._check(self.lib.echo_set_dump(self.h, int(bool(on))))

    def dump(self, path, iout, thyp=0.0):
        self._check(self.lib.echo_dump(self.h, os.fsencode(path), int(iout), float(thyp)))

    def restart(self, path):
        """read a dump.dat into the grid; returns (iout, thyp)"""
        io, th = C.c_int32(0), C.c_double(0.0)
        self._check(self.lib.echo_restart(self.h, os.fsencode(path), C.byref(io), C.byref(th)))
        return io.value, th.value

    # -- freeze-out hypersurface (out_hyp / find_hypersurface_*) ----------
    def fo_setup(self, freeze_type, freeze_value):
        """Freeze-out on the temperature (0) or energy density (1) at
        freeze_value; returns the header's (T_f, e_f)."""
        hdr = np.zeros(2)
        self._check(self.lib.echo_fo_setup(self.h, int(freeze_type), float(freeze_value), hdr))
        return hdr

    def fo_slice(self):
        """One out_hyp: the number of elements found against the stored
        slice (0 on the first call)."""
        n = C.c_int64(0)
        self._check(self.lib.echo_fo_slice(self.h, C.byref(n)))
        self._fo_n = n.value
        return n.value

    @property
    def fo_width(self):
        return self.lib.echo_fo_width(self.h)

    def fo_elements(self):
        """The elements of the last slice, shape (n, fo_width)."""
        n = getattr(self, "_fo_n", 0)
        a = np.zeros((max(n, 1), self.fo_width))
        self._check(self.lib.echo_fo_elements(self.h, a.ravel(), n))
        return a[:n]

    def fo_write(self, path, header=False):
        """hypersurface.txt: header=True creates it with the header line,
        otherwise the last slice is appended."""
        self._check(self.lib.echo_fo_write(self.h, os.fsencode(path), int(bool(header))))

    def init_glauber(self, params, part_a, part_b, coll):
        """Glauber-MC initial state on the device.  params: dict of
        GlauberParams fields; part_a, part_b, coll: (n, 2) arrays of (x, y)."""
        gp = GlauberParams(**params)
        arrs = [np.ascontiguousarray(np.reshape(a, (-1, 2)), dtype=np.float64) for a in
                (part_a, part_b, coll)]
        flat = [a.ravel() if a.size else np.zeros(2) for a in arrs]
        return self._check(self.lib.echo_init_glauber(self.h, C.byref(gp), flat[0], arrs[0].shape[0],
                                                      flat[1], arrs[1].shape[0], flat[2],
                                                      arrs[2].shape[0]))

    def init_glauber_mhd(self, params, part_a, part_b, coll, bparams, nucleons):
        """Glauber-MC initial state of an ideal-MHD grid with the initial
        magnetic field of the protons.  bparams: dict of BFieldParams fields;
        nucleons: four (xy (n, 2), type (n,)) pairs for the spectators of A,
        the participants of A, the spectators of B and the participants of
        B (type 0 proton, 1 neutron)."""
        gp = GlauberParams(**params)
        bp = BFieldParams(**bparams)
        arrs = [np.ascontiguousarray(np.reshape(a, (-1, 2)), dtype=np.float64) for a in
                (part_a, part_b, coll)]
        flat = [a.ravel() if a.size else np.zeros(2) for a in arrs]
        keep = []
        lists = (Nucleons * 4)()
        for l, (xy, ty) in enumerate(nucleons):
            xy = np.ascontiguousarray(np.reshape(xy, (-1, 2)), dtype=np.float64)
            ty = np.ascontiguousarray(ty, dtype=np.int32)
            assert ty.shape[0] == xy.shape[0]
            keep += [xy, ty]
            lists[l].xy = xy.ctypes.data_as(C.POINTER(C.c_double))
            lists[l].type = ty.ctypes.data_as(C.POINTER(C.c_int32))
            lists[l].n = xy.shape[0]
        return self._check(self.lib.echo_init_glauber_mhd(
            self.h, C.byref(gp), flat[0], arrs[0].shape[0], flat[1], arrs[1].shape[0], flat[2],
            arrs[2].shape[0], C.byref(bp), lists))

    def upload(self, v, u):
        v = np.ascontiguousarray(v, dtype=np.float64)
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert v.shape == self.shape and u.shape == self.shape
        self._check(self.lib.echo_upload_state(self.h, v.ravel(), u.ravel()))

    def state(self):
        v = np.zeros(self.shape)
        u = np.zeros(self.shape)
        self._check(self.lib.echo_download_state(self.h, v.ravel(), u.ravel()))
        return v, u

    def steps(self, n):
        return self._check(self.lib.echo_steps(self.h, int(n)))

    def set_bulk_peak(self, zs_max, t_c=0.155, width=0.02):
        """zeta/s = zs_max exp(-(T - t_c)^2 / (2 width^2)) (GeV) instead of the
        reference's 2 eta (1/3 - c_s^2) / s."""
        self._check(self.lib.echo_set_bulk_peak(self.h, float(zs_max), float(t_c), float(width)))

    def rk_stage(self, irk):
        return self._check(self.lib.echo_rk_stage(self.h, int(irk)))

    def end_step(self):
        return self._check(self.lib.echo_end_step(self.h))

    def scalars(self):
        s = np.zeros(10)
        self._check(self.lib.echo_get_scalars(self.h, s))
        return dict(t=s[0], dt=s[1], timeold=s[2], timeinterval=s[3], g_cov_3_old=s[4],
                    glm_ch_old=s[5], tout=s[6], iter=int(s[7]), crashed=int(s[8]),
                    errcode=int(s[9]))

    def repairs(self):
        """Cells repaired by the con2prim repair (evolve_u2v) since set-up."""
        n = C.c_int64()
        self._check(self.lib.echo_get_repairs(self.h, C.byref(n)))
        return n.value

    def set_c2p_net(self, net, iters=3, tol=1e-12):
        """Invert each stage with the learned con2prim (C2pNet; None
        detaches it).  Ideal MHD with p = e/3.  The grid keeps a reference
        to the network until it is detached."""
        self._check(self.lib.echo_set_c2p_net(self.h, None if net is None else net.h, int(iters), float(tol)))
        self._net = net

    def c2p_fallbacks(self):
        """States of the steps that the learned inversion left to the plain
        solver, since set-up."""
        n = C.c_int64()
        self._check(self.lib.echo_get_c2p_fallbacks(self.h, C.byref(n)))
        return n.value

    def set_tout(self, tout):
        self._check(self.lib.echo_set_tout(self.h, float(tout)))

    def cons2prim_batch(self, u, vguess):
        """u, vguess: (nv, n) arrays.  Returns (v, status)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.array(vguess, dtype=np.float64, order="C")
        n = u.shape[1]
        st = np.zeros(n, dtype=np.int32)
        self._check(self.lib.echo_cons2prim_batch(self.h, u.ravel(), v.ravel(), st, n))
        return v, st

    def cons2prim_batch_dev(self, d_u, d_v, d_status, n):
        """Plain solver on device pointers (ints): u, v [nv][n] FP64, status [n] int32."""
        self._check(self.lib.echo_cons2prim_batch_dev(self.h, d_u, d_v, d_status, int(n)))

    def set_timing(self, on):
        self._check(self.lib.echo_set_timing(self.h, 1 if on else 0))

    def timing_raw(self):
        t = np.zeros(4)
        self._check(self.lib.echo_get_timing(self.h, t))
        return t

    def timing(self):
        t = self.timing_raw()
        return dict(flux_x=t[0], flux_y=t[1], flux_z=t[2], update=t[3])


class C2pNet:
    """The learned con2prim network on one GPU (echo_mlp of the C-ABI).

    Device arrays are passed as integer device pointers (DeviceArray.ptr);
    every call returns after its device work has finished."""

    def __init__(self, device=0, params=None):
        self.lib = load_library()
        h = C.c_void_p()
        self._check(self.lib.echo_mlp_create(int(device), C.byref(h)))
        self.h = h
        if params is not None:
            self.set_params(params)

    def _check(self, rc):
        if rc != 0:
            raise EchoError(rc, self.lib.echo_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.echo_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.lib.echo_mlp_set_params(self.h, p, p.size))

    def params(self):
        p = np.zeros(MLP_NPARAM, dtype=np.float32)
        self._check(self.lib.echo_mlp_get_params(self.h, p, p.size))
        return p

    def forward(self, grid, d_u, n, d_y):
        self._check(self.lib.echo_mlp_forward(self.h, grid.h, d_u, int(n), d_y))

    def grad(self, grid, d_u, d_v, n):
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.echo_mlp_grad(self.h, grid.h, d_u, d_v, int(n), g, C.byref(loss)))
        return loss.value, g

    def train(self, grid, d_u, d_v, n, steps, batch, lr=3e-3, lr_final=3e-5, beta1=0.9,
              beta2=0.999, eps=1e-8):
        cfg = TrainConfig(int(steps), int(batch), lr, lr_final, beta1, beta2, eps)
        log = np.zeros(int(steps), dtype=np.float32)
        self._check(self.lib.echo_mlp_train(self.h, grid.h, d_u, d_v, int(n), C.byref(cfg),
                                            log.ctypes.data))
        return log

    def invert(self, grid, d_u, d_v, d_status, n, iters=2, tol=1e-11):
        """The learned inversion; returns the number of fallback states."""
        nf = C.c_int64()
        self._check(self.lib.echo_c2p_surrogate(grid.h, self.h, d_u, d_v, d_status, int(n),
                                                int(iters), float(tol), C.byref(nf)))
        return nf.value

    def invert_noalloc(self, grid, d_u, d_v, d_status, n, iters=2, tol=1e-11):
        """invert() through the allocation-free entry point (its scratch is a
        static device array; n <= ECHO_NOALLOC_MAX_STATES = 2^22)."""
        nf = C.c_int64()
        self._check(self.lib.echo_c2p_surrogate_noalloc(grid.h, self.h, d_u, d_v, d_status, int(n),
                                                        int(iters), float(tol), C.byref(nf)))
        return nf.value


class Ensemble:
    """Independent event grids advanced together (echo_ensemble)."""

    def __init__(self, configs, device=0, eos_rows=None):
        """configs: list of (ic, rc) arrays."""
        self.lib = load_library()
        cfgs = (EchoConfig * len(configs))(*[EchoConfig.from_arrays(ic, rc) for ic, rc in configs])
        need_tab = any(c.eqstate == 3 for c in cfgs)
        if need_tab and eos_rows is None:
            eos_rows = read_eos_rows()
        rows = np.zeros((1, 4)) if eos_rows is None else np.ascontiguousarray(eos_rows, dtype=np.float64)
        n_rows = 0 if eos_rows is None else rows.shape[0]
        h = C.c_void_p()
        _check(self.lib, self.lib.echo_ensemble_create(cfgs, len(configs), rows.ravel(), n_rows,
                                                       device, C.byref(h)))
        self.h = h
        self.grids = [EchoGrid(ic, rc, _handle=self.lib.echo_ensemble_grid(self.h, i))
                      for i, (ic, rc) in enumerate(configs)]

    def __len__(self):
        return len(self.grids)

    def tune(self, reps=3):
        _check(self.lib, self.lib.echo_ensemble_tune(self.h, int(reps)))

    def steps(self, n):
        """Advance every event by n steps; returns the number of crashed events."""
        nc = C.c_int32()
        _check(self.lib, self.lib.echo_ensemble_steps(self.h, int(n), C.byref(nc)))
        return nc.value

    def last_ms(self):
        ms = C.c_double()
        _check(self.lib, self.lib.echo_ensemble_last_ms(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            for g in self.grids:
                g.h = None
            self.lib.echo_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
