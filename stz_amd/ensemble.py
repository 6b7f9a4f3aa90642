"""Ensembles of independent quasi-static simulations stepped together.

No counterpart in the reference, which runs one ``shear_sim`` per process.
``Ensemble(params_list, precision)`` owns one simulation per parameter set
(grid sizes may differ); ``ens[k]`` is member k as a ``ShearSim`` for every
per-simulation call (fields, time, initial conditions, diagnostics);
``step_forward_quasistatic(dt, nsteps)`` advances all members with one
batched launch per kernel of the step, the members' adaptive V-cycle counts
decided on the device, and each step replayed from a HIP graph with one host
synchronisation (``use_graphs(False)`` launches directly).  Each member ends
exactly where
``ShearSim.step_forward_quasistatic`` would take it.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .sim import ShearSim


class Ensemble:
    def __init__(self, params_list, precision="mixed"):
        lib = _lib.load()
        params_list = [dict(p) for p in params_list]
        n = len(params_list)
        arr = (_lib.SimParams * n)(*[_lib.make_params(p) for p in params_list])
        prec = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED}[precision]
        h = C.c_void_p()
        check(lib.stz_ensemble_create(arr, n, prec, C.byref(h)))
        self._h, self._lib, self.precision = h, lib, precision
        self.members = []
        for k, p in enumerate(params_list):
            mh = C.c_void_p()
            check(lib.stz_ensemble_member(h, k, C.byref(mh)))
            self.members.append(ShearSim._borrow(self, mh, p, precision))

    def __len__(self):
        return len(self.members)

    def __getitem__(self, k):
        return self.members[k]

    def __iter__(self):
        return iter(self.members)

    def close(self):
        if getattr(self, "_h", None):
            for m in self.members:
                m.close()
            self._lib.stz_ensemble_destroy(self._h)
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

    def set_mg_config(self, **kw):
        """the same multigrid settings on every member"""
        g = self.members[0].mg_config()
        for k, v in kw.items():
            setattr(g, k, v)
        check(self._lib.stz_ensemble_set_mg_config(self._h, C.byref(g)))

    def use_graphs(self, on=True):
        """replay each step from a HIP graph (default) or launch directly"""
        check(self._lib.stz_ensemble_use_graphs(self._h, 1 if on else 0))

    def step_forward_quasistatic(self, dt, nsteps=1):
        check(self._lib.stz_ensemble_step_quasistatic(self._h, float(dt), int(nsteps)))

    def stats(self):
        """{steps, slots, extra_slots, captures} since creation"""
        out = np.zeros(4, dtype=np.int64)
        check(self._lib.stz_ensemble_stats(self._h, out))
        return dict(zip(("steps", "slots", "extra_slots", "captures"), (int(v) for v in out)))

    def tune_sweeps(self, dt, probe_steps=6, candidates=((1, 1), (2, 2), (3, 3), (4, 4)), reps=1):
        """The sweep-count tuner: time `probe_steps` batched steps of dt from
        the members' current state (fields, velocity, time and predictor)
        with each V(nu1, nu2) of `candidates` after one untimed step (the
        tables, the tile-shape tuner and the graph captures of the new
        setting), keep the fastest and restore the state.  Sweep counts
        change how many V-cycles the solves take to reach their threshold,
        never the threshold.  Returns {"choice": (nu1, nu2), "ms_per_step":
        {candidate: ms}}."""
        import time
        saved = [(m.get_fields(), m.get_vel(), m.time, m.predictor_limit) for m in self.members]

        def restore(nu1, nu2):
            self.set_mg_config(nu1=nu1, nu2=nu2)
            for m, (f, v, t, lim) in zip(self.members, saved):
                m.set_fields(f)
                m.set_vel(v)
                m.time = t
                m.predictor_limit = lim

        times = {}
        for cand in candidates:
            best = None
            for _ in range(reps):
                restore(*cand)
                self.step_forward_quasistatic(dt, 1)     # tables, tile tuner, graph captures
                self.members[0].synchronize()
                t0 = time.perf_counter()
                self.step_forward_quasistatic(dt, probe_steps)
                self.members[0].synchronize()
                el = (time.perf_counter() - t0) / probe_steps
                best = el if best is None else min(best, el)
            times[tuple(cand)] = best * 1e3
        choice = min(times, key=times.get)
        restore(*choice)
        return {"choice": choice, "ms_per_step": times}

    def tile_choice(self):
        """{launches kept at 32x16 tiles, at 64x34} by the tuner"""
        out = np.zeros(2, dtype=np.int32)
        check(self._lib.stz_ensemble_tile_choice(self._h, out))
        return {"32x16": int(out[0]), "64x34": int(out[1])}
