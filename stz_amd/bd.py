"""Python mirror of the reference's free-boundary class ``bd_sim``
(sb/main/bd_sim.hh:27-190) over the gfx950 C ABI (stz_bd_*).

Method names and argument meaning follow the reference:
``init_bar(ca)``, ``step_forward_quasistatic(dt)``, ``advection_step(dt)``,
``projection_step(dt)``, ``post_process(all_fields)``, public ``time``.
Grids: cell planes are (n, m) arrays, node planes (n+1, m+1); phi < 0 inside
the body.  Failures raise ``StzError`` where the reference calls exit(1).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check

CELL_NAMES = ("p", "q", "s", "tau", "chi", "ddev", "phi")
NODE_NAMES = ("u", "v", "X", "Y")


class BdSim:
    """Free-boundary quasi-static STZ simulation resident on one GPU."""

    def __init__(self, params):
        lib = _lib.load()
        self.params = dict(params)
        self._p = _lib.make_bd_params(params)
        self.m, self.n = self._p.m, self._p.n
        h = C.c_void_p()
        check(lib.stz_bd_create(C.byref(self._p), C.byref(h)))
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.stz_bd_destroy(self._h)
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

    # ---- geometry
    @property
    def cell_shape(self):
        return (self.n, self.m)

    @property
    def node_shape(self):
        return (self.n + 1, self.m + 1)

    def cell_centres(self):
        p = self._p
        dx, dy = (p.bx - p.ax) / p.m, (p.by - p.ay) / p.n
        x = p.ax + (np.arange(p.m) + 0.5) * dx
        y = p.ay + (np.arange(p.n) + 0.5) * dy
        return np.meshgrid(x, y)

    def node_coords(self):
        p = self._p
        dx, dy = (p.bx - p.ax) / p.m, (p.by - p.ay) / p.n
        return np.meshgrid(p.ax + np.arange(p.m + 1) * dx, p.ay + np.arange(p.n + 1) * dy)

    # ---- level set
    def set_level_set(self, phi):
        """levelset::build_band on a signed function of the cells."""
        a = np.ascontiguousarray(phi, dtype=np.float64).reshape(self.cell_shape)
        check(self._lib.stz_bd_set_level_set(self._h, a.ravel()))

    def upload_level_set(self, phi, status):
        a = np.ascontiguousarray(phi, dtype=np.float64).ravel()
        s = np.ascontiguousarray(status, dtype=np.int32).ravel()
        check(self._lib.stz_bd_upload_level_set(self._h, a, s))

    def init_bar(self, ca):
        check(self._lib.stz_bd_init_bar(self._h, int(ca)))

    def status(self):
        c = np.zeros(self.m * self.n, np.int32)
        cc = np.zeros((self.m + 1) * (self.n + 1), np.int32)
        check(self._lib.stz_bd_download_status(self._h, c, cc))
        return c.reshape(self.cell_shape), cc.reshape(self.node_shape)

    # ---- fields
    def get_cell(self, k):
        a = np.zeros(self.m * self.n)
        check(self._lib.stz_bd_download_cells(self._h, CELL_NAMES.index(k), a))
        return a.reshape(self.cell_shape)

    def set_cell(self, k, v):
        a = np.ascontiguousarray(np.broadcast_to(v, self.cell_shape), dtype=np.float64)
        check(self._lib.stz_bd_upload_cells(self._h, CELL_NAMES.index(k), a.ravel()))

    def get_node(self, k):
        a = np.zeros((self.m + 1) * (self.n + 1))
        check(self._lib.stz_bd_download_nodes(self._h, NODE_NAMES.index(k), a))
        return a.reshape(self.node_shape)

    def set_node(self, k, v):
        a = np.ascontiguousarray(np.broadcast_to(v, self.node_shape), dtype=np.float64)
        check(self._lib.stz_bd_upload_nodes(self._h, NODE_NAMES.index(k), a.ravel()))

    def get_vel(self):
        a = np.zeros(2 * (self.m + 1) * (self.n + 1))
        check(self._lib.stz_bd_download_vel(self._h, a))
        return a.reshape(self.node_shape + (2,))

    def get_src(self):
        a = np.zeros(2 * (self.m + 1) * (self.n + 1))
        check(self._lib.stz_bd_download_src(self._h, a))
        return a.reshape(self.node_shape + (2,))

    @property
    def time(self):
        t = C.c_double()
        check(self._lib.stz_bd_get_time(self._h, C.byref(t)))
        return t.value

    @time.setter
    def time(self, t):
        check(self._lib.stz_bd_set_time(self._h, float(t)))

    # ---- loading
    def set_body_force(self, f):
        if f is None:
            check(self._lib.stz_bd_set_body_force(self._h, None))
            return
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(self.node_shape + (2,))
        self._force = a
        check(self._lib.stz_bd_set_body_force(self._h, a.ctypes.data_as(C.c_void_p)))

    def set_edge_velocity(self, u_left, u_right):
        check(self._lib.stz_bd_set_edge_velocity(self._h, float(u_left), float(u_right)))

    # ---- stepping (bd_sim.cc:240-259)
    def step_forward_quasistatic(self, dt, nsteps=1):
        check(self._lib.stz_bd_step_quasistatic(self._h, float(dt), int(nsteps)))

    def advection_step(self, dt):
        check(self._lib.stz_bd_advection_step(self._h, float(dt)))

    def projection_step(self, dt):
        check(self._lib.stz_bd_projection_step(self._h, float(dt)))

    def post_process(self, all_fields=True):
        check(self._lib.stz_bd_post_process(self._h, 1 if all_fields else 0))

    def last_solve(self):
        info = _lib.SolveInfo()
        check(self._lib.stz_bd_last_solve(self._h, C.byref(info)))
        return {"status": info.status, "iterations": info.total_cycles // 2,
                "vcycles": info.total_cycles, "residual": info.residual,
                "tolerance": info.tolerance}

    def set_solver(self, nu1=1, nu2=1, bottom_sweeps=20, max_iter=500, tol=1e-20):
        check(self._lib.stz_bd_set_solver(self._h, int(nu1), int(nu2), int(bottom_sweeps), int(max_iter),
                                          float(tol)))

    def set_precision(self, precision):
        """"fp64" (default) or "mixed": the projection's V-cycle
        preconditioner in FP32 inside FP64 BiCGStab"""
        check(self._lib.stz_bd_set_precision(self._h, {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]))

    def step_stats(self):
        a = np.zeros(4, np.int32)
        check(self._lib.stz_bd_step_stats(self._h, a))
        return dict(zip(("march_rounds", "cell_waves", "node_waves", "bicgstab_iterations"), a.tolist()))

    def synchronize(self):
        check(self._lib.stz_bd_synchronize(self._h))

    def set_stream(self, stream):
        check(self._lib.stz_bd_set_stream(self._h, C.c_void_p(stream)))

    # ---- diagnostics of single pieces
    def remarch(self):
        """the band rebuilt around the unmoved interface"""
        check(self._lib.stz_bd_remarch(self._h))

    def interface_velocity(self):
        a = np.zeros(self.m * self.n)
        check(self._lib.stz_bd_interface_velocity(self._h, a))
        return a.reshape(self.cell_shape)

    def operator(self, dt):
        nn = (self.m + 1) * (self.n + 1)
        d = np.zeros(nn, np.int32)
        mats = np.zeros(36 * nn)
        b = np.zeros(2 * nn)
        check(self._lib.stz_bd_operator(self._h, float(dt), d, mats, b))
        return d.reshape(self.node_shape), mats.reshape(self.node_shape + (9, 2, 2)), \
            b.reshape(self.node_shape + (2,))
