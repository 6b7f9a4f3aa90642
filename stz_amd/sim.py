"""Python mirror of the reference's ``shear_sim`` class (sb/main/shear_sim.hh:20-151)
over the gfx950 C ABI.

Method names, argument meaning and failure behaviour follow the reference:
``step_forward_quasistatic(dt)``, ``advection_step(dt)``,
``projection_step(dt)``, ``init_fields(case, T0, dT)``, ``avg_iters()``,
``bdry_vel()``, ``bdry_pos()``, public ``time`` and ``chi_inf``.  Where the
reference calls exit(1) on a failed solve (qs_multi.hh:52), this raises
``StzError``.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FIELD_INDEX, FIELD_NAMES, NFIELDS, check


class ShearSim:
    """Quasi-static simple-shear STZ simulation resident on one GPU."""

    def __init__(self, params, precision="fp64"):
        lib = _lib.load()
        self.params = dict(params)
        self._p = _lib.make_params(params)
        self.m, self.n = self._p.m, self._p.n
        prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
        self.precision = precision
        h = C.c_void_p()
        check(lib.stz_sim_create(C.byref(self._p), prec, C.byref(h)))
        self._h = h
        self._lib = lib

    @classmethod
    def _borrow(cls, owner, handle, params, precision):
        """a member of an ensemble: the ensemble owns the handle"""
        self = cls.__new__(cls)
        self.params = dict(params)
        self._p = _lib.make_params(params)
        self.m, self.n = self._p.m, self._p.n
        self.precision = precision
        self._h = handle
        self._lib = _lib.load()
        self._owner = owner
        return self

    # ---- lifetime
    def close(self):
        if getattr(self, "_owner", None) is not None:
            self._h = None
            self._owner = None
            return
        if getattr(self, "_h", None):
            self._lib.stz_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- fields (public fm, vel, src)
    @property
    def plane_shape(self):
        return (self.n + 2, self.m)

    def _fid(self, k):
        return FIELD_INDEX[k] if isinstance(k, str) else int(k)

    def set_field(self, k, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.plane_shape)
        check(self._lib.stz_sim_upload_field(self._h, self._fid(k), a))

    def get_field(self, k):
        a = np.empty(self.plane_shape, dtype=np.float64)
        check(self._lib.stz_sim_download_field(self._h, self._fid(k), a))
        return a

    def set_fields(self, planes):
        for k in range(NFIELDS):
            self.set_field(k, planes[k])

    def get_fields(self):
        return np.stack([self.get_field(k) for k in range(NFIELDS)])

    def set_vel(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.n + 1, self.m, 2)
        check(self._lib.stz_sim_upload_vel(self._h, a))

    def get_vel(self):
        a = np.empty((self.n + 1, self.m, 2), dtype=np.float64)
        check(self._lib.stz_sim_download_vel(self._h, a))
        return a

    def get_src(self):
        a = np.empty((self.n + 1, self.m, 2), dtype=np.float64)
        check(self._lib.stz_sim_download_src(self._h, a))
        return a

    def field_device_ptr(self, k):
        p = C.c_void_p()
        check(self._lib.stz_sim_field_device_ptr(self._h, self._fid(k), C.byref(p)))
        return p.value

    def vel_device_ptr(self):
        p = C.c_void_p()
        check(self._lib.stz_sim_vel_device_ptr(self._h, C.byref(p)))
        return p.value

    # ---- public scalars
    @property
    def time(self):
        t = C.c_double()
        check(self._lib.stz_sim_get_time(self._h, C.byref(t)))
        return t.value

    @time.setter
    def time(self, t):
        check(self._lib.stz_sim_set_time(self._h, float(t)))

    def set_chi_inf(self, c):
        check(self._lib.stz_sim_set_chi_inf(self._h, float(c)))

    def bdry_vel(self):
        """shear_sim.hh:120"""
        t, u = self.time, self.params["u_bdry"]
        return t * u if t < 1 else u

    def bdry_pos(self):
        """shear_sim.hh:124"""
        t, u = self.time, self.params["u_bdry"]
        return u * t * t * 0.5 if t < 1 else u * (t - 0.5)

    # ---- stepping
    def init_fields(self, chi_case, tem_base=600.0, tem_delta=200.0):
        check(self._lib.stz_sim_init_fields(self._h, int(chi_case), float(tem_base), float(tem_delta)))

    def step_forward_quasistatic(self, dt, nsteps=1):
        check(self._lib.stz_sim_step_quasistatic(self._h, float(dt), int(nsteps)))

    def advection_step(self, dt):
        check(self._lib.stz_sim_advection_step(self._h, float(dt)))

    def advection_launch(self, dt, count=1):
        """asynchronous: `count` advection steps on the simulation's stream"""
        check(self._lib.stz_sim_advection_launch(self._h, float(dt), int(count)))

    def projection_step(self, dt):
        check(self._lib.stz_sim_projection_step(self._h, float(dt)))

    def set_boundaries(self):
        check(self._lib.stz_sim_set_boundaries(self._h))

    def synchronize(self):
        check(self._lib.stz_sim_synchronize(self._h))

    # ---- multigrid
    def mg_config(self):
        g = _lib.MGConfig()
        check(self._lib.stz_sim_get_mg_config(self._h, C.byref(g)))
        return g

    def set_mg_config(self, **kw):
        g = self.mg_config()
        for k, v in kw.items():
            setattr(g, k, v)
        check(self._lib.stz_sim_set_mg_config(self._h, C.byref(g)))

    @property
    def predictor_limit(self):
        """tgmg_predict::lim: the next solve's predicted V-cycles x predict_mult"""
        v = C.c_int32()
        check(self._lib.stz_sim_get_predictor(self._h, C.byref(v)))
        return v.value

    @predictor_limit.setter
    def predictor_limit(self, lim):
        check(self._lib.stz_sim_set_predictor(self._h, int(lim)))

    def last_solve(self):
        info = _lib.SolveInfo()
        check(self._lib.stz_sim_last_solve(self._h, C.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def avg_iters(self):
        a = C.c_double()
        check(self._lib.stz_sim_avg_iters(self._h, C.byref(a)))
        return a.value

    def mg_hierarchy(self):
        dims = np.zeros(64, dtype=np.int32)
        nl = C.c_int()
        check(self._lib.stz_sim_mg_hierarchy(self._h, 32, C.byref(nl), dims))
        return [(int(dims[2 * k]), int(dims[2 * k + 1])) for k in range(nl.value)]

    def mg_setup(self, viscdt):
        check(self._lib.stz_mg_setup(self._h, float(viscdt)))

    def mg_solve(self):
        info = _lib.SolveInfo()
        code = self._lib.stz_mg_solve(self._h, C.byref(info))
        check(code)
        return {k: getattr(info, k) for k, _ in info._fields_}

    def mg_v_cycle(self):
        check(self._lib.stz_mg_v_cycle(self._h))

    def mg_l2_error(self):
        e = C.c_double()
        check(self._lib.stz_mg_l2_error(self._h, C.byref(e)))
        return e.value

    def mg_smooth(self, level, sweeps):
        """asynchronous smoothing sweeps on one level (tgmg gauss_seidel)"""
        check(self._lib.stz_mg_smooth(self._h, int(level), int(sweeps)))

    def mg_fmg_pre(self):
        """tgmg::fmg_pre on the current src/vel"""
        check(self._lib.stz_mg_fmg_pre(self._h))

    def mg_clear_z(self):
        """tgmg_base::clear_z: vel = 0"""
        check(self._lib.stz_mg_clear_z(self._h))

    def mg_vcycle_launch(self, level, mode, count=1):
        """asynchronous: one of the V-cycle's smoothing launches on `level`,
        `count` times (mode 1 pre-smoothing with restriction, mode 2
        post-smoothing with prolongation and, mixed top level, vel += e)"""
        check(self._lib.stz_mg_vcycle_launch(self._h, int(level), int(mode), int(count)))

    def set_stream(self, stream_handle):
        """run on a HIP stream given as an integer handle (e.g. a torch
        stream's cuda_stream); None returns to the library's own stream"""
        check(self._lib.stz_sim_set_stream(self._h, C.c_void_p(stream_handle) if stream_handle else None))

    def mg_set_src(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.n + 1, self.m, 2)
        check(self._lib.stz_mg_upload_src(self._h, a))

    # ---- direct step, diagnostics, initial conditions
    def step_forward(self, dt, nsteps=1):
        """shear_sim::step_forward: direct (dynamic) explicit step"""
        check(self._lib.stz_sim_step_forward(self._h, float(dt), int(nsteps)))

    def tractions(self):
        """(bx, by, tx, ty) of shear_sim::tractions"""
        out = np.zeros(4)
        check(self._lib.stz_sim_tractions(self._h, out))
        return tuple(out)

    def qs_measure(self, staggered=False):
        """(Q, max_qs) of shear_sim::qs_measure; the integrand is left in
        get_field("cu")"""
        q, mq = C.c_double(), C.c_double()
        check(self._lib.stz_sim_qs_measure(self._h, 1 if staggered else 0, C.byref(q), C.byref(mq)))
        return q.value, mq.value

    def compute_strain(self, corrected=False):
        """shear_sim::compute_strain into get_field("cu").  corrected=True
        fixes the reference formula's two slips: a centred y difference on the
        interior rows (where the reference gets 0) and the periodic shift's
        sign in the seam column"""
        check(self._lib.stz_sim_compute_strain(self._h, 1 if corrected else 0))

    def l2_comparison(self, other):
        out = np.zeros(4)
        check(self._lib.stz_sim_l2_comparison(self._h, other._h, out))
        return out

    def initialize_random(self, l, tem_avg, tem_stdev, seed=None):
        """shear_sim::initialize_random; with seed, the C library's rand() is
        seeded first (srand), as the reference's drivers do"""
        if seed is not None:
            C.CDLL(None).srand(C.c_uint(int(seed)))
        check(self._lib.stz_sim_initialize_random(self._h, float(l), float(tem_avg), float(tem_stdev)))

    def initialize_chi_bicubic(self, arr, lo=0.0, sca=1.0):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        nn, mm = arr.shape
        check(self._lib.stz_sim_initialize_chi_bicubic(self._h, int(mm), int(nn), arr, float(lo), float(sca)))

    def output(self, directory, prefix, mode, sn):
        """shear_sim::output: one field as a Gnuplot binary matrix"""
        from . import frames
        cu = self.get_field("cu") if mode == 11 else None
        return frames.write_field(directory, prefix, self.params, self.get_fields(), mode, sn,
                                  self.params["stz"]["TZ"], cu)

    def write_files(self, directory, k, flags=None):
        """shear_sim::write_files: the reference's default field selection"""
        from . import frames
        flags = frames.DEFAULT_FLAGS if flags is None else flags
        cu = self.get_field("cu") if flags & (1 << 11) else None
        return frames.write_files(directory, self.params, self.get_fields(), k, self.params["stz"]["TZ"], flags, cu)

    # ---- decisions
    def record_decisions(self, on=True):
        check(self._lib.stz_sim_record_decisions(self._h, 1 if on else 0))

    def replay_decisions(self, records):
        """take the decisions of the next len(records) advection steps from
        records (each an (n, m) uint32 array from get_decisions) instead of
        evaluating them; an empty list ends the replay"""
        recs = [np.ascontiguousarray(r, dtype=np.uint32) for r in records]
        for r in recs:
            if r.shape != (self.n, self.m):
                raise ValueError("decision record of shape %s, want %s" % (r.shape, (self.n, self.m)))
        a = np.stack(recs) if recs else np.zeros((1, self.n, self.m), np.uint32)
        check(self._lib.stz_sim_replay_decisions(self._h, len(recs), a))

    def get_decisions(self):
        a = np.empty((self.n, self.m), dtype=np.uint32)
        check(self._lib.stz_sim_download_decisions(self._h, a))
        return a


__all__ = ["ShearSim", "FIELD_NAMES"]
