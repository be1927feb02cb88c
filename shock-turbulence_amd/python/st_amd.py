"""Python host binding of the MI355X per-stage pipeline (ctypes over libst.so).

The classes mirror the reference's subroutine interface for this path
(src/lddrk4_periodic_sf.f, src/rhsv_periodic_sf.f, src/hyper_rhs.f,
src/derivCDopt6_periodic.f, src/filters_periodic.f): DFDX/DFDY/DFDZ,
D2FDX2/..., Compact_Filter, Gaussian_Filter, RHS, Hyper_RHS, Fix_Pressure,
Compute_CFL, Compute_Tstep, LDDRK4.  Arrays are numpy float64 in Fortran
order, (My, Mz, Mx) per field and (My, Mz, Mx, 5) for the state.

There is no CPU fallback: if the HIP library is missing or a call fails,
an exception is raised.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.dirname(_HERE)
LIB_PATH = os.environ.get("ST_AMD_LIB", os.path.join(_PKG, "libst.so"))

_dp = ctypes.POINTER(ctypes.c_double)


class StError(RuntimeError):
    pass


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("period_x", ctypes.c_double), ("period_y", ctypes.c_double), ("period_z", ctypes.c_double),
                ("x_start", ctypes.c_double), ("y_start", ctypes.c_double), ("z_start", ctypes.c_double),
                ("nonperiodic_x", ctypes.c_int)]


class Deck(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_char * 24) for n in ("vis_path", "restart_path", "stats_path", "probe_path",
                                                   "source_path", "prefix")]
                + [("grid_x", ctypes.c_int), ("grid_y", ctypes.c_int)]
                + [(n, ctypes.c_double) for n in ("x_start", "y_start", "z_start", "period_x", "period_y",
                                                  "period_z", "Re", "Pr", "Sc", "gamma", "vis_exp", "bulk_ratio",
                                                  "T_ratio")]
                + [("hyper_param", ctypes.c_int), ("part_param", ctypes.c_int), ("march_mode", ctypes.c_int),
                   ("Tstep", ctypes.c_double), ("cfl_cal", ctypes.c_double), ("Nsteps", ctypes.c_int),
                   ("run_time", ctypes.c_double), ("comp_stats", ctypes.c_int)]
                + [(n, ctypes.c_int) for n in ("out_grid", "out_init", "out_incr_x", "out_incr_y", "out_incr_z",
                                               "out_restart", "out_vis", "out_probe", "out_stats", "out_source")])


class Params(ctypes.Structure):
    _fields_ = [("Re", ctypes.c_double), ("Pr", ctypes.c_double), ("gamma", ctypes.c_double),
                ("vis_exp", ctypes.c_double), ("bulk_ratio", ctypes.c_double),
                ("hyper_param", ctypes.c_int), ("cfl_cal", ctypes.c_double),
                ("d2_mode", ctypes.c_int), ("rk_scheme", ctypes.c_int)]


EXPORTS = [
    "st_ctx_create", "st_ctx_destroy", "st_last_error", "st_sync", "st_state_upload", "st_state_download",
    "st_state_dev", "st_d1", "st_d2", "st_filter", "st_filter_dir", "st_gauss", "st_gauss_dir", "st_rhs",
    "st_hyper_rhs", "st_get_hyper", "st_set_hyper", "st_get_grad", "st_fix_pressure", "st_cfl", "st_tstep",
    "st_rk_stage", "st_step", "st_march", "st_diagnostics", "st_profile", "st_profile_read", "st_pass_name",
    "st_ic_isotropic", "st_batch_create", "st_batch_slots", "st_slot_load", "st_slot_store", "st_slot_release",
    "st_batch_march", "st_batch_status", "st_autotune", "st_pass_config", "st_pass_config_set", "st_query",
    "st_restart_write", "st_restart_read", "st_deck_read", "st_sponge", "st_ic_shock", "st_compute_hyper",
    "st_compute_grad", "st_cfl_noalloc", "st_build_id", "st_group_create", "st_group_march", "st_group_sync",
    "st_group_destroy",
]

_lib = None


def lib():
    """Load libst.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise StError("HIP library not built: %s (run __graft_entry__.build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    L.st_ctx_create.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(Params), ctypes.POINTER(vp)]
    L.st_ctx_destroy.argtypes = [vp]
    L.st_ctx_destroy.restype = None
    L.st_last_error.argtypes = [vp]
    L.st_last_error.restype = ctypes.c_char_p
    L.st_build_id.argtypes = []
    L.st_build_id.restype = ctypes.c_char_p
    L.st_sync.argtypes = [vp]
    L.st_state_upload.argtypes = [vp, _dp]
    L.st_state_download.argtypes = [vp, _dp]
    L.st_state_dev.argtypes = [vp, ctypes.POINTER(_dp)]
    L.st_d1.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double]
    L.st_d2.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
    L.st_filter.argtypes = [vp, _dp, _dp, ctypes.c_int]
    L.st_filter_dir.argtypes = [vp, ctypes.c_int, _dp, _dp]
    L.st_gauss.argtypes = [vp, _dp, _dp, ctypes.c_int]
    L.st_gauss_dir.argtypes = [vp, ctypes.c_int, _dp, _dp]
    L.st_rhs.argtypes = [vp, _dp, _dp]
    L.st_hyper_rhs.argtypes = [vp, _dp]
    L.st_get_hyper.argtypes = [vp, _dp]
    L.st_compute_hyper.argtypes = [vp, _dp]
    L.st_compute_grad.argtypes = [vp, _dp, _dp]
    L.st_cfl_noalloc.argtypes = [ctypes.POINTER(Grid), ctypes.c_double, vp, vp, vp]
    L.st_set_hyper.argtypes = [vp, _dp]
    L.st_get_grad.argtypes = [vp, _dp]
    L.st_fix_pressure.argtypes = [vp, _dp]
    L.st_cfl.argtypes = [vp, _dp, _dp]
    L.st_tstep.argtypes = [vp, ctypes.c_double, ctypes.c_int, _dp]
    L.st_rk_stage.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.st_step.argtypes = [vp, ctypes.c_double, _dp, ip]
    L.st_march.argtypes = [vp, ctypes.c_int, _dp, ip, _dp]
    L.st_diagnostics.argtypes = [vp, _dp]
    L.st_profile.argtypes = [vp, ctypes.c_int]
    L.st_profile_read.argtypes = [vp, ctypes.c_int, ip, _dp, _dp]
    L.st_pass_name.argtypes = [ctypes.c_int]
    L.st_pass_name.restype = ctypes.c_char_p
    L.st_ic_isotropic.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong, _dp]
    L.st_batch_create.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(vp)]
    L.st_batch_slots.argtypes = [vp]
    L.st_slot_load.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    L.st_slot_store.argtypes = [vp, ctypes.c_int, _dp]
    L.st_slot_release.argtypes = [vp, ctypes.c_int]
    L.st_batch_march.argtypes = [vp, ctypes.c_int]
    L.st_batch_status.argtypes = [vp, ip, _dp, _dp, ip]
    L.st_autotune.argtypes = [vp, ctypes.c_int]
    L.st_query.argtypes = [vp]
    L.st_restart_write.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                                   ctypes.c_int]
    L.st_restart_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, ip]
    L.st_deck_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(Deck)]
    L.st_sponge.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
    L.st_ic_shock.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              ctypes.c_ulonglong, _dp, _dp]
    L.st_pass_config.argtypes = [vp, ctypes.c_int, ip, ip]
    L.st_pass_config_set.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.st_group_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    L.st_group_march.argtypes = [vp, ctypes.c_int]
    L.st_group_sync.argtypes = [vp]
    L.st_group_destroy.argtypes = [vp]
    L.st_group_destroy.restype = None
    _lib = L
    return L


def source_id():
    """The hash st_build_id() carries, recomputed from the source files in
    this tree (the same file list, order and framing as the library's
    Makefile)."""
    import glob
    import hashlib
    csrc = os.path.join(_PKG, "csrc")
    files = (glob.glob(os.path.join(csrc, "*.hip")) + glob.glob(os.path.join(csrc, "*.h"))
             + glob.glob(os.path.join(csrc, "*.hpp"))
             + [os.path.join(csrc, f) for f in ("st_api.cpp", "st_coef.cpp", "st_ic.cpp", "st_io.cpp")]
             + [os.path.join(os.path.dirname(_PKG), "include", "st_api.h")])
    root = os.path.dirname(_PKG)
    h = hashlib.sha256()
    for f in sorted(set(os.path.abspath(x) for x in files)):
        h.update((os.path.relpath(f, root) + "\n").encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return "src-sha256:" + h.hexdigest()[:16]


def build_id():
    """The library's st_build_id()."""
    return lib().st_build_id().decode()


def _p(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not (a.flags.f_contiguous or a.ndim == 1):
        raise StError("arrays must be float64, Fortran-ordered")
    return a.ctypes.data_as(_dp)


def _f(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def restart_write(path, Q, time, step):
    """Write_Restart format (src/out.f:493-591) for Q(Ny,Nz,Nx,5)."""
    Q = _f(Q)
    ny, nz, nx, _ = Q.shape
    r = lib().st_restart_write(str(path).encode(), nx, ny, nz, _p(Q), time, step)
    if r:
        raise StError("st_restart_write failed (%d)" % r)


def restart_read(path, nx, ny, nz):
    """Read_Restart (src/out.f:598-692): returns (Q, time, step)."""
    Q = np.zeros((ny, nz, nx, 5), dtype=np.float64, order="F")
    t = np.zeros(1)
    st = ctypes.c_int()
    r = lib().st_restart_read(str(path).encode(), nx, ny, nz, _p(Q), _p(t), ctypes.byref(st))
    if r:
        raise StError("st_restart_read failed (%d)" % r)
    return Q, float(t[0]), st.value


def deck_read(path):
    """Define_Parameters' input deck (src/define_periodic_sf.f:20-80) as a dict."""
    d = Deck()
    r = lib().st_deck_read(str(path).encode(), ctypes.byref(d))
    if r:
        raise StError("st_deck_read failed (%d)" % r)
    out = {}
    for name, _ in Deck._fields_:
        v = getattr(d, name)
        out[name] = v.decode() if isinstance(v, bytes) else v
    return out


def ic_shock(nx, ny, nz, period, mach=1.5, frac=0.25, width=None, k0=4.0, urms=0.1, gamma=1.4, seed=12345):
    """Planar shock with upstream turbulence on the non-periodic x grid;
    returns (Q, upstream state, downstream state)."""
    if width is None:
        width = 2.0 * period[0] / (nx - 1)
    Q = np.zeros((ny, nz, nx, 5), dtype=np.float64, order="F")
    per = (ctypes.c_double * 3)(*period)
    st = np.zeros(10)
    r = lib().st_ic_shock(nx, ny, nz, ctypes.cast(per, _dp), mach, frac, width, k0, urms, gamma, seed, _p(Q), _p(st))
    if r:
        raise StError("st_ic_shock failed (%d)" % r)
    return Q, st[:5].copy(), st[5:].copy()


def ic_isotropic(nx, ny, nz, k0=4.0, urms=1.0, mt=0.3, gamma=1.4, seed=12345, period=(2 * np.pi,) * 3):
    """Seeded decaying-turbulence state Q(My,Mz,Mx,5) (host, deterministic per seed)."""
    Q = np.zeros((ny, nz, nx, 5), dtype=np.float64, order="F")
    per = (ctypes.c_double * 3)(*period)
    r = lib().st_ic_isotropic(nx, ny, nz, ctypes.cast(per, _dp), k0, urms, mt, gamma, seed, _p(Q))
    if r:
        raise StError("st_ic_isotropic failed (%d)" % r)
    return Q


def re_for_taylor(re_lambda, k0, urms, mt, gamma=1.4, vis_exp=0.76):
    """Reynolds number giving Taylor-microscale Reynolds number re_lambda at t=0.

    For E(k) ~ k^4 exp(-2 (k/k0)^2) the Taylor microscale is lambda = 2/k0, and
    the kinematic viscosity is ((gamma-1) T)^vis_exp / Re = c^(2 vis_exp) / Re
    with c = 1/mt at rho = 1.
    """
    lam = 2.0 / k0
    nu = urms * lam / re_lambda
    return (1.0 / mt ** 2) ** vis_exp / nu


class Solver:
    """One device context (grid + parameters + resident state); nslots > 1
    makes a batch (see Batch)."""

    def __init__(self, nx, ny, nz, period=(2 * np.pi,) * 3, Re=100.0, Pr=0.7, gamma=1.4, vis_exp=0.76,
                 bulk_ratio=0.0, hyper_param=1, cfl_cal=0.5, d2_mode=0, rk_scheme=0, start=(0.0, 0.0, 0.0),
                 nonperiodic_x=0, nslots=1):
        self.grid = Grid(nx, ny, nz, period[0], period[1], period[2], start[0], start[1], start[2],
                         1 if nonperiodic_x else 0)
        self.params = Params(Re, Pr, gamma, vis_exp, bulk_ratio, hyper_param, cfl_cal, d2_mode, rk_scheme)
        self.shape = (ny, nz, nx)
        self.nslots = nslots
        h = ctypes.c_void_p()
        r = lib().st_batch_create(ctypes.byref(self.grid), ctypes.byref(self.params), nslots, ctypes.byref(h))
        self.h = h
        if r:
            msg = lib().st_last_error(h).decode() if h.value else "context creation failed"
            self.close()
            raise StError("st_ctx_create: %s (%d)" % (msg, r))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().st_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def _chk(self, r, what):
        if r < 0:
            raise StError("%s: %s (%d)" % (what, lib().st_last_error(self.h).decode(), r))
        return r

    def _z(self, nf=None):
        return np.zeros(self.shape if nf is None else self.shape + (nf,), dtype=np.float64, order="F")

    # state
    def upload(self, Q):
        Q = _f(Q)
        self._chk(lib().st_state_upload(self.h, _p(Q)), "st_state_upload")

    def download(self):
        Q = self._z(5)
        self._chk(lib().st_state_download(self.h, _p(Q)), "st_state_download")
        return Q

    # operators: DFDX/DFDY/DFDZ, D2FDX2..., filters
    def dfd(self, d, F, dF=None, iadd=0, cadd=1.0):
        F = _f(F)
        out = self._z() if dF is None else np.array(dF, dtype=np.float64, order="F", copy=True)
        self._chk(lib().st_d1(self.h, d, _p(F), _p(out), iadd, cadd), "st_d1")
        return out

    def d2fd2(self, d, F, iadd=0):
        F = _f(F)
        dF, d2F = self._z(), self._z()
        self._chk(lib().st_d2(self.h, d, _p(F), _p(dF), _p(d2F), iadd), "st_d2")
        return (dF, d2F) if iadd else d2F

    def filter_dir(self, d, F):
        F = _f(F)
        out = self._z()
        self._chk(lib().st_filter_dir(self.h, d, _p(F), _p(out)), "st_filter_dir")
        return out

    def gauss_dir(self, d, F):
        F = _f(F)
        out = self._z()
        self._chk(lib().st_gauss_dir(self.h, d, _p(F), _p(out)), "st_gauss_dir")
        return out

    def compact_filter(self, F):
        F = _f(F)
        nf = F.shape[3] if F.ndim == 4 else 1
        out = np.zeros_like(F, order="F")
        self._chk(lib().st_filter(self.h, _p(F), _p(out), nf), "st_filter")
        return out

    def gaussian_filter(self, F):
        F = _f(F)
        nf = F.shape[3] if F.ndim == 4 else 1
        out = np.zeros_like(F, order="F")
        self._chk(lib().st_gauss(self.h, _p(F), _p(out), nf), "st_gauss")
        return out

    # physics
    def rhs(self, Q=None):
        out = self._z(5)
        Qf = None if Q is None else _f(Q)
        self._chk(lib().st_rhs(self.h, _p(Qf), _p(out)), "st_rhs")
        return out

    def hyper_rhs(self, Qrhs):
        out = np.array(Qrhs, dtype=np.float64, order="F", copy=True)
        self._chk(lib().st_hyper_rhs(self.h, _p(out)), "st_hyper_rhs")
        return out

    def full_rhs(self, Q=None):
        return self.hyper_rhs(self.rhs(Q))

    def grad(self):
        out = self._z(12)
        self._chk(lib().st_get_grad(self.h, _p(out)), "st_get_grad")
        return out

    def hyper(self):
        out = self._z(3)
        self._chk(lib().st_get_hyper(self.h, _p(out)), "st_get_hyper")
        return out

    def compute_hyper(self, Q=None):
        """Compute_hyper(Q) (src/main.f:100): the coefficients from the state
        (Q, or the resident one) and the current /GRAD/; returns them."""
        Qf = None if Q is None else _f(Q)
        self._chk(lib().st_compute_hyper(self.h, _p(Qf)), "st_compute_hyper")
        return self.hyper()

    def compute_grad(self):
        """Compute_Grad (src/grad.f:13-115) on the resident state: (MaxDil, Enstrophy)."""
        m, e = np.zeros(1), np.zeros(1)
        self._chk(lib().st_compute_grad(self.h, _p(m), _p(e)), "st_compute_grad")
        return m[0], e[0]

    def set_hyper(self, H):
        H = _f(H)
        self._chk(lib().st_set_hyper(self.h, _p(H)), "st_set_hyper")

    def fix_pressure(self, Q):
        out = np.array(Q, dtype=np.float64, order="F", copy=True)
        self._chk(lib().st_fix_pressure(self.h, _p(out)), "st_fix_pressure")
        return out

    def cfl(self):
        a, b = np.zeros(1), np.zeros(1)
        self._chk(lib().st_cfl(self.h, _p(a), _p(b)), "st_cfl")
        return a[0], b[0]

    def tstep(self, cfl_resv, step):
        dt = np.zeros(1)
        self._chk(lib().st_tstep(self.h, cfl_resv, step, _p(dt)), "st_tstep")
        return dt[0]

    def rk_stage(self, alpha, beta, dt, reset_qw=False):
        self._chk(lib().st_rk_stage(self.h, alpha, beta, dt, 1 if reset_qw else 0), "st_rk_stage")

    def lddrk4(self, dt, time=0.0, step=0):
        t = np.array([time])
        s = ctypes.c_int(step)
        self._chk(lib().st_step(self.h, dt, _p(t), ctypes.byref(s)), "st_step")
        return t[0], s.value

    def march(self, nsteps, time=0.0, step=0):
        t = np.array([time])
        s = ctypes.c_int(step)
        dts = np.zeros(max(nsteps, 1))
        self._chk(lib().st_march(self.h, nsteps, _p(t), ctypes.byref(s), _p(dts)), "st_march")
        return t[0], s.value, dts[:nsteps]

    def diagnostics(self):
        out = np.zeros(7)
        self._chk(lib().st_diagnostics(self.h, _p(out)), "st_diagnostics")
        return out

    def sync(self):
        self._chk(lib().st_sync(self.h), "st_sync")

    def sponge(self, amp, power, frac, ref_left, ref_right):
        """Sponge_Forcing in every stage (amp 0: off)."""
        L = np.ascontiguousarray(ref_left, dtype=np.float64)
        R = np.ascontiguousarray(ref_right, dtype=np.float64)
        self._chk(lib().st_sponge(self.h, amp, power, frac, _p(L), _p(R)), "st_sponge")

    def autotune(self, reps=2):
        """Per-pass lane layout and tile width by timing; returns {pass: (narrow, nt)}."""
        self._chk(lib().st_autotune(self.h, reps), "st_autotune")
        return self.pass_configs()

    def pass_configs(self):
        """{"kind.direction": (narrow, threads per workgroup)} of every pass."""
        out = {}
        for k in range(72):
            name = lib().st_pass_name(k).decode()
            if not name or name.startswith("."):
                continue
            n, t = ctypes.c_int(), ctypes.c_int()
            lib().st_pass_config(self.h, k, ctypes.byref(n), ctypes.byref(t))
            out[name] = (n.value, t.value)
        return out

    def set_pass_configs(self, cfg):
        """Apply {"kind.direction": (narrow, nt)} (as pass_configs returns)."""
        idx = {}
        for k in range(72):
            name = lib().st_pass_name(k).decode()
            if name and not name.startswith("."):
                idx[name] = k
        for name, (n, t) in cfg.items():
            self._chk(lib().st_pass_config_set(self.h, idx[name], int(n), int(t)), "st_pass_config_set")

    # profiling
    def profile(self, on=True):
        self._chk(lib().st_profile(self.h, 1 if on else 0), "st_profile")

    def profile_read(self):
        n = 72
        cnt = (ctypes.c_int * n)()
        ms = np.zeros(n)
        fr = np.zeros(n)
        k = self._chk(lib().st_profile_read(self.h, n, cnt, _p(ms), _p(fr)), "st_profile_read")
        out = {}
        for i in range(k):
            if cnt[i]:
                out[lib().st_pass_name(i).decode()] = {"count": cnt[i], "ms": ms[i], "fields": fr[i]}
        return out


class Batch(Solver):
    """A batch of same-size instances advanced together (st_batch_*)."""

    def __init__(self, nx, ny, nz, nslots, **kw):
        super().__init__(nx, ny, nz, nslots=nslots, **kw)

    def load(self, slot, Q, Re, nsteps, time=0.0, step=0):
        Q = _f(Q)
        self._chk(lib().st_slot_load(self.h, slot, _p(Q), Re, time, step, nsteps), "st_slot_load")

    def store(self, slot):
        Q = self._z(5)
        self._chk(lib().st_slot_store(self.h, slot, _p(Q)), "st_slot_store")
        return Q

    def release(self, slot):
        self._chk(lib().st_slot_release(self.h, slot), "st_slot_release")

    def march(self, nsteps):
        """Enqueue nsteps batch steps (asynchronous; status() or sync() waits)."""
        self._chk(lib().st_batch_march(self.h, nsteps), "st_batch_march")

    def sync(self):
        self._chk(lib().st_sync(self.h), "st_sync")

    def idle(self):
        """True when the enqueued steps have finished (does not wait)."""
        return self._chk(lib().st_query(self.h), "st_query") == 1

    def status(self):
        n = self.nslots
        step = (ctypes.c_int * n)()
        act = (ctypes.c_int * n)()
        t = np.zeros(n)
        dt = np.zeros(n)
        self._chk(lib().st_batch_status(self.h, step, _p(t), _p(dt), act), "st_batch_status")
        return np.array(step[:]), t, dt, np.array(act[:])


# the grid sizes a Group takes (cubic, one batch each)
GROUP_SIZES = (48, 64, 96, 128)


def narrow_configs(solver):
    """Every pass in its M/2-point layout on 256-thread tiles: the layouts a
    Group uses, so a Batch with these configs gives the same bits."""
    return {k: (1, 256) for k in solver.pass_configs()}


class Group:
    """Batches of different sizes (one per size in GROUP_SIZES) stepped
    together, one launch per pass over all of them (st_group_*).  The
    batches stay usable for load/store/release/status; close the group
    before them."""

    def __init__(self, batches):
        self.batches = list(batches)
        arr = (ctypes.c_void_p * len(self.batches))(*[b.h.value for b in self.batches])
        h = ctypes.c_void_p()
        r = lib().st_group_create(arr, len(self.batches), ctypes.byref(h))
        if r:
            raise StError("st_group_create: %s (%d)" % (lib().st_last_error(None).decode(), r))
        self.h = h

    def _chk(self, r, what):
        if r < 0:
            raise StError("%s: %s (%d)" % (what, lib().st_last_error(None).decode(), r))
        return r

    def march(self, nsteps):
        """Enqueue nsteps group steps (asynchronous; sync() or a member's status() waits)."""
        self._chk(lib().st_group_march(self.h, nsteps), "st_group_march")

    def sync(self):
        self._chk(lib().st_group_sync(self.h), "st_group_sync")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().st_group_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()
