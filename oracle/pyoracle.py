"""ctypes access to the CPU oracles -- TEST INFRASTRUCTURE ONLY.

Two libraries, both built by oracle/Makefile:
  * ``liboracle.so``      -- this repository's FP64 restatement of the
    reference step (oracle/stz_oracle.c);
  * ``_ref/libstzref.so`` -- the reference's own C++ sources compiled with
    oracle/ref_harness.cc (only where the reference tree was available at
    build time; it is optional everywhere).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module, and only as the oracle.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libstzref.so")

NFIELDS = 10
FIELD_NAMES = ("u", "v", "p", "q", "s", "tau", "chi", "X", "Y", "ddev")

_MODEL_DOUBLES = ("TZ", "c0", "tau0", "kappa", "Delta", "Omegaeps0", "chi_inf",
                  "theta", "rho", "t_icpt", "t_grad")
_SIM_DOUBLES = ("ax", "bx", "ay", "by", "mu", "K", "viscosity", "chi_diff",
                "chi_len", "t_scale", "tmult", "adapt_fac", "u_bdry")


class ModelParams(C.Structure):
    _fields_ = [("kind", C.c_int32), ("t_model", C.c_int32)] + \
        [(k, C.c_double) for k in _MODEL_DOUBLES]


class SimParams(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32)] + \
        [(k, C.c_double) for k in _SIM_DOUBLES] + [("stz", ModelParams)]


def make_params(d):
    """Fill SimParams from a flat dict (keys of SimParams plus 'stz' dict)."""
    p = SimParams()
    p.m, p.n = int(d["m"]), int(d["n"])
    for k in _SIM_DOUBLES:
        setattr(p, k, float(d[k]))
    mp = d["stz"]
    p.stz.kind = int(mp.get("kind", 0))
    p.stz.t_model = int(mp.get("t_model", -1))
    for k in _MODEL_DOUBLES:
        setattr(p.stz, k, float(mp[k]))
    return p


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _bind(lib, prefix):
    vp, d, i = C.c_void_p, C.c_double, C.c_int
    sig = {
        "create": (vp, [C.POINTER(SimParams)]),
        "destroy": (None, [vp]),
        "set_field": (None, [vp, i, _dp]),
        "get_field": (None, [vp, i, _dp]),
        "set_vel": (None, [vp, _dp]),
        "get_vel": (None, [vp, _dp]),
        "set_src": (None, [vp, _dp]),
        "get_src": (None, [vp, _dp]),
        "set_time": (None, [vp, d]),
        "get_time": (d, [vp]),
        "set_chi_inf": (None, [vp, d]),
        "set_tol_mult": (None, [vp, d]),
        "tolerance": (d, [vp]),
        "init_fields": (None, [vp, i, d, d]),
        "initialize_random": (None, [vp, d, d, d, C.c_uint]),
        "advection_step": (None, [vp, d]),
        "mg_setup": (None, [vp, d]),
        "mg_v_cycle": (None, [vp]),
        "mg_l2_error": (d, [vp]),
        "mg_levels": (i, [vp, _ip, i]),
        "mg_stencil": (i, [vp, i, _dp]),
        "avg_iters": (d, [vp]),
        "tractions": (None, [vp, _dp]),
        "qs_measure": (None, [vp, i, C.POINTER(d), C.POINTER(d)]),
        "get_cu": (None, [vp, _dp]),
        "step_forward": (None, [vp, d]),
        "compute_strain": (None, [vp, i]),
        "l2_comparison": (None, [vp, vp, _dp]),
        "initialize_chi_bicubic": (None, [vp, i, i, _dp, d, d]),
    }
    out = {}
    for name, (res, args) in sig.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
        out[name] = fn
    return out


class _Sim:
    """Common driver over either oracle library."""

    def __init__(self, params):
        self.params = params
        self._p = make_params(params)
        self.m, self.n = self._p.m, self._p.n
        self._h = self._f["create"](C.byref(self._p))
        if not self._h:
            raise RuntimeError("oracle create failed")

    def close(self):
        if self._h:
            self._f["destroy"](self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields
    @property
    def plane_shape(self):
        return (self.n + 2, self.m)

    def set_field(self, k, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.plane_shape)
        self._f["set_field"](self._h, int(k), a)

    def get_field(self, k):
        a = np.empty(self.plane_shape, dtype=np.float64)
        self._f["get_field"](self._h, int(k), a)
        return a

    def set_fields(self, planes):
        for k in range(NFIELDS):
            self.set_field(k, planes[k])

    def get_fields(self):
        return np.stack([self.get_field(k) for k in range(NFIELDS)])

    def set_vel(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.n + 1, self.m, 2)
        self._f["set_vel"](self._h, a)

    def get_vel(self):
        a = np.empty((self.n + 1, self.m, 2), dtype=np.float64)
        self._f["get_vel"](self._h, a)
        return a

    def set_src(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.n + 1, self.m, 2)
        self._f["set_src"](self._h, a)

    def get_src(self):
        a = np.empty((self.n + 1, self.m, 2), dtype=np.float64)
        self._f["get_src"](self._h, a)
        return a

    time = property(lambda self: self._f["get_time"](self._h),
                    lambda self, t: self._f["set_time"](self._h, float(t)))

    def set_chi_inf(self, c):
        self._f["set_chi_inf"](self._h, float(c))

    def set_tol_mult(self, mult):
        self._f["set_tol_mult"](self._h, float(mult))

    def tolerance(self):
        return self._f["tolerance"](self._h)

    def init_fields(self, case, tem_base=600.0, tem_delta=200.0):
        self._f["init_fields"](self._h, int(case), float(tem_base), float(tem_delta))

    def initialize_random(self, l, tem_avg, tem_stdev, seed):
        self._f["initialize_random"](self._h, float(l), float(tem_avg), float(tem_stdev), int(seed))

    def advection_step(self, dt):
        self._f["advection_step"](self._h, float(dt))

    def mg_setup(self, viscdt):
        self._f["mg_setup"](self._h, float(viscdt))

    def mg_v_cycle(self):
        self._f["mg_v_cycle"](self._h)

    def mg_l2_error(self):
        return self._f["mg_l2_error"](self._h)

    def mg_levels(self):
        dims = np.zeros(2 * 32, dtype=np.int32)
        nl = self._f["mg_levels"](self._h, dims, 32)
        return [(int(dims[2 * k]), int(dims[2 * k + 1])) for k in range(nl)]

    def mg_stencil(self, level):
        m, n = self.mg_levels()[level]
        out = np.empty((n, m, 10, 4), dtype=np.float64)
        if self._f["mg_stencil"](self._h, int(level), out) != 0:
            raise ValueError("bad level")
        return out

    def avg_iters(self):
        return self._f["avg_iters"](self._h)

    def tractions(self):
        o = np.zeros(4)
        self._f["tractions"](self._h, o)
        return o

    def get_cu(self):
        a = np.empty(self.plane_shape, dtype=np.float64)
        self._f["get_cu"](self._h, a)
        return a

    def step_forward(self, dt):
        self._f["step_forward"](self._h, float(dt))

    def compute_strain(self, corrected=False):
        self._f["compute_strain"](self._h, 1 if corrected else 0)

    def l2_comparison(self, other):
        out = np.zeros(4)
        self._f["l2_comparison"](self._h, other._h, out)
        return out

    def initialize_chi_bicubic(self, arr, lo=0.0, sca=1.0):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        nn, mm = arr.shape
        self._f["initialize_chi_bicubic"](self._h, int(mm), int(nn), arr, float(lo), float(sca))

    def qs_measure(self, staggered=False):
        q, mq = C.c_double(), C.c_double()
        self._f["qs_measure"](self._h, 1 if staggered else 0, C.byref(q), C.byref(mq))
        return q.value, mq.value


class Oracle(_Sim):
    """The FP64 restatement (oracle/stz_oracle.c)."""
    _lib = None

    def __init__(self, params):
        if Oracle._lib is None:
            lib = C.CDLL(ORACLE_SO)
            Oracle._lib = lib
            f = _bind(lib, "orc_")
            for name, res, args in (
                ("step_quasistatic", C.c_int, [C.c_void_p, C.c_double, C.c_int]),
                ("projection_step", C.c_int, [C.c_void_p, C.c_double]),
                ("set_boundaries", None, [C.c_void_p]),
                ("mg_solve", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
                ("set_decisions", None, [C.c_void_p, C.c_int]),
                ("get_decisions", None, [C.c_void_p, _up]),
                ("last_cycles", C.c_int, [C.c_void_p]),
                ("eno2", C.c_double, [C.c_double] * 4),
                ("dplastic", C.c_double, [C.POINTER(ModelParams), C.c_double, C.c_double,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]),
                ("plastic_term", C.c_double, [C.c_void_p, C.c_double, C.POINTER(C.c_double),
                                              C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_int)]),
            ):
                fn = getattr(lib, "orc_" + name)
                fn.restype = res
                fn.argtypes = args
                f[name] = fn
            Oracle._fns = f
        self._f = Oracle._fns
        super().__init__(params)

    def step(self, dt, nsteps=1):
        return self._f["step_quasistatic"](self._h, float(dt), int(nsteps))

    def projection_step(self, dt):
        return self._f["projection_step"](self._h, float(dt))

    def set_boundaries(self):
        self._f["set_boundaries"](self._h)

    def mg_solve(self):
        c, r = C.c_int(), C.c_double()
        st = self._f["mg_solve"](self._h, C.byref(c), C.byref(r))
        return st, c.value, r.value

    def record_decisions(self, on=True):
        self._f["set_decisions"](self._h, 1 if on else 0)

    def get_decisions(self):
        a = np.empty((self.n, self.m), dtype=np.uint32)
        self._f["get_decisions"](self._h, a)
        return a

    def last_cycles(self):
        return self._f["last_cycles"](self._h)

    @staticmethod
    def eno2(p0, p1, p2, p3):
        lib = C.CDLL(ORACLE_SO)
        lib.orc_eno2.restype = C.c_double
        lib.orc_eno2.argtypes = [C.c_double] * 4
        return lib.orc_eno2(float(p0), float(p1), float(p2), float(p3))

    def dplastic(self, sbar, chi):
        d1, d2 = C.c_double(), C.c_double()
        r = self._f["dplastic"](C.byref(self._p.stz), float(sbar), float(chi), C.byref(d1), C.byref(d2))
        return r, d1.value, d2.value

    def plastic_term(self, sbar, chi, dt):
        c, dd, t = C.c_double(chi), C.c_double(), C.c_int()
        r = self._f["plastic_term"](self._h, float(sbar), C.byref(c), C.byref(dd), float(dt), C.byref(t))
        return r, c.value, dd.value, t.value


def reference_available():
    return os.path.exists(REF_SO)


class Reference(_Sim):
    """The reference's own compiled code (oracle/_ref/libstzref.so)."""
    _fns = None

    def __init__(self, params):
        if Reference._fns is None:
            lib = C.CDLL(REF_SO)
            f = _bind(lib, "ref_")
            for name, res, args in (
                ("output", None, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int]),
                ("step", None, [C.c_void_p, C.c_double, C.c_int]),
                ("projection_step", None, [C.c_void_p, C.c_double]),
                ("mg_solve", C.c_int, [C.c_void_p]),
                ("mg_lim", C.c_int, [C.c_void_p]),
                ("set_threads", None, [C.c_int]),
            ):
                fn = getattr(lib, "ref_" + name)
                fn.restype = res
                fn.argtypes = args
                f[name] = fn
            Reference._fns = f
        self._f = Reference._fns
        super().__init__(params)

    def step(self, dt, nsteps=1):
        self._f["step"](self._h, float(dt), int(nsteps))
        return 0

    def projection_step(self, dt):
        self._f["projection_step"](self._h, float(dt))
        return 0

    def mg_solve(self):
        return self._f["mg_solve"](self._h)

    def mg_lim(self):
        return self._f["mg_lim"](self._h)

    def output(self, directory, prefix, mode, sn):
        """shear_sim::output: Gnuplot binary matrix of one field"""
        self._f["output"](self._h, directory.encode(), prefix.encode(), int(mode), int(sn))

    @staticmethod
    def set_threads(n):
        Reference._fns["set_threads"](int(n))


def reference_dplastic(model, sbar, chi):
    """stz_dynamics::Dplastic of the reference's own law classes
    (stz_model.hh:28) for a model dict: (Dpl, dchi1, dchi2)"""
    lib = C.CDLL(REF_SO)
    fn = lib.ref_dplastic
    fn.restype = C.c_double
    fn.argtypes = [C.POINTER(ModelParams), C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    mp = make_params({"m": 4, "n": 4, "stz": model, **{k: 0.0 for k in _SIM_DOUBLES}}).stz
    d1, d2 = C.c_double(), C.c_double()
    r = fn(C.byref(mp), float(sbar), float(chi), C.byref(d1), C.byref(d2))
    return r, d1.value, d2.value


# ---------------------------------------------------------------------------
# free-boundary reference (bd_sim) through oracle/ref_bd_harness.cc
# ---------------------------------------------------------------------------
_BD_DOUBLES = ("ax", "bx", "ay", "by", "mu", "K", "viscosity", "chi_len",
               "filter_stress", "t_scale", "tmult", "adapt_fac")


class BdParams(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32)] + \
        [(k, C.c_double) for k in _BD_DOUBLES] + [("stz", ModelParams)]


def make_bd_params(d):
    p = BdParams()
    p.m, p.n = int(d["m"]), int(d["n"])
    for k in _BD_DOUBLES:
        setattr(p, k, float(d[k]))
    mp = d["stz"]
    p.stz.kind = int(mp.get("kind", 0))
    p.stz.t_model = int(mp.get("t_model", -1))
    for k in _MODEL_DOUBLES:
        setattr(p.stz, k, float(mp[k]))
    return p


class RefBd:
    """The reference's bd_sim compiled from its own sources
    (oracle/_ref/libstzref.so); same method names as stz_amd.BdSim."""
    _f = None
    CELL = ("p", "q", "s", "tau", "chi", "ddev")
    NODE = ("u", "v", "X", "Y")

    def __init__(self, params, outdir):
        if RefBd._f is None:
            lib = C.CDLL(REF_SO)
            vp, d, i = C.c_void_p, C.c_double, C.c_int
            sig = {
                "create": (vp, [C.POINTER(BdParams), C.c_char_p]),
                "destroy": (None, [vp]),
                "init_bar": (None, [vp, i]),
                "set_phi_build": (None, [vp, _dp]),
                "get_phi": (None, [vp, _dp]),
                "get_status": (None, [vp, _ip, _ip]),
                "set_cells": (None, [vp, i, _dp]),
                "get_cells": (None, [vp, i, _dp]),
                "set_nodes": (None, [vp, i, _dp]),
                "get_nodes": (None, [vp, i, _dp]),
                "get_vel": (None, [vp, _dp]),
                "get_src": (None, [vp, _dp]),
                "set_time": (None, [vp, d]),
                "set_force": (None, [vp, _dp]),
                "use_builtin_force": (None, [vp, i]),
                "advection_step": (None, [vp, d]),
                "projection_step": (None, [vp, d]),
                "post_process": (None, [vp, i]),
                "step": (None, [vp, d, i]),
                "operator": (None, [vp, d, _ip, _dp, _dp]),
                "interface_velocity": (None, [vp, _dp]),
                "set_bc_smooth": (None, [vp, d]),
                "remarch": (None, [vp]),
            }
            f = {}
            for name, (res, args) in sig.items():
                fn = getattr(lib, "ref_bd_" + name)
                fn.restype = res
                fn.argtypes = args
                f[name] = fn
            RefBd._f = f
        self.params = params
        self._p = make_bd_params(params)
        self.m, self.n = self._p.m, self._p.n
        # bd_sim keeps the directory name pointer (bd_sim.hh:83)
        self._outdir = str(outdir).encode()
        self._h = self._f["create"](C.byref(self._p), self._outdir)

    def close(self):
        if self._h:
            self._f["destroy"](self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def cell_shape(self):
        return (self.n, self.m)

    @property
    def node_shape(self):
        return (self.n + 1, self.m + 1)

    def set_level_set(self, phi):
        self._f["set_phi_build"](self._h, np.ascontiguousarray(phi, dtype=np.float64).ravel())

    def init_bar(self, ca):
        self._f["init_bar"](self._h, int(ca))

    def get_cell(self, k):
        a = np.zeros(self.m * self.n)
        if k == "phi":
            self._f["get_phi"](self._h, a)
        else:
            self._f["get_cells"](self._h, self.CELL.index(k), a)
        return a.reshape(self.cell_shape)

    def set_cell(self, k, v):
        a = np.ascontiguousarray(np.broadcast_to(v, self.cell_shape), dtype=np.float64).ravel()
        self._f["set_cells"](self._h, self.CELL.index(k), a)

    def get_node(self, k):
        a = np.zeros((self.m + 1) * (self.n + 1))
        self._f["get_nodes"](self._h, self.NODE.index(k), a)
        return a.reshape(self.node_shape)

    def set_node(self, k, v):
        a = np.ascontiguousarray(np.broadcast_to(v, self.node_shape), dtype=np.float64).ravel()
        self._f["set_nodes"](self._h, self.NODE.index(k), a)

    def status(self):
        c = np.zeros(self.m * self.n, np.int32)
        cc = np.zeros((self.m + 1) * (self.n + 1), np.int32)
        self._f["get_status"](self._h, c, cc)
        return c.reshape(self.cell_shape), cc.reshape(self.node_shape)

    def get_vel(self):
        a = np.zeros(2 * (self.m + 1) * (self.n + 1))
        self._f["get_vel"](self._h, a)
        return a.reshape(self.node_shape + (2,))

    def get_src(self):
        a = np.zeros(2 * (self.m + 1) * (self.n + 1))
        self._f["get_src"](self._h, a)
        return a.reshape(self.node_shape + (2,))

    def set_time(self, t):
        self._f["set_time"](self._h, float(t))

    def set_body_force(self, F):
        self._f["set_force"](self._h, np.ascontiguousarray(F, dtype=np.float64).ravel())

    def advection_step(self, dt):
        self._f["advection_step"](self._h, float(dt))

    def projection_step(self, dt):
        self._f["projection_step"](self._h, float(dt))

    def post_process(self, all_fields=True):
        self._f["post_process"](self._h, 1 if all_fields else 0)

    def step_forward_quasistatic(self, dt, nsteps=1):
        self._f["step"](self._h, float(dt), int(nsteps))

    def operator(self, dt):
        nn = (self.m + 1) * (self.n + 1)
        d = np.zeros(nn, np.int32)
        mats = np.zeros(36 * nn)
        b = np.zeros(2 * nn)
        self._f["operator"](self._h, float(dt), d, mats, b)
        return d.reshape(self.node_shape), mats.reshape(self.node_shape + (9, 2, 2)), \
            b.reshape(self.node_shape + (2,))

    def interface_velocity(self):
        a = np.zeros(self.m * self.n)
        self._f["interface_velocity"](self._h, a)
        return a.reshape(self.cell_shape)

    def remarch(self):
        self._f["remarch"](self._h)


def bd_remarch(phi, status, dx, dy, delta):
    """the round-parallel band rebuild of bd_oracle.c on copies of phi and
    status; returns (phi, status, rounds)"""
    lib = C.CDLL(ORACLE_SO)
    lib.orc_bd_remarch.restype = C.c_int
    lib.orc_bd_remarch.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _ip]
    p = np.ascontiguousarray(phi, dtype=np.float64).copy()
    s = np.ascontiguousarray(status, dtype=np.int32).copy()
    n, m = p.shape
    r = lib.orc_bd_remarch(m, n, float(dx), float(dy), float(delta), p.ravel(), s.ravel())
    return p, s, r
