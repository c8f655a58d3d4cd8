"""Python host API of the HIP step (ctypes over include/oercel_mi.h).

Mirrors the reference engine's step interface (cel/engine/oercel.c):
  Engine(params)            sys_alloc (:714-759) for one system, or many replicas
  Engine.start()            the fresh start of main (:2187-2212): soup, velocities,
                            forces at step 0, vh = v - dt/2 f
  Engine.step(n)            n iterations of the main loop body (:2300-2362)
  Engine.sort/forces/chem   sort_cells (:828), pair_forces + bonded_forces
                            (:1198, :1317), the chemistry check
  Engine.stats()            kinetic_T / count_types (:2053-2071) + energies
  Engine.get_state/set_state  checkpoint-style state in gid order (:1886-2010)

Errors raise abi.OcError with the library's message (the reference prints and
exits).  There is no CPU fallback: without the built HIP library this module
raises at construction.
"""
import ctypes as C

import numpy as np

from . import abi
from .params import Params, ParamError


def empty_state(N):
    return {"x": np.zeros(N), "y": np.zeros(N), "vx": np.zeros(N), "vy": np.zeros(N), "fx": np.zeros(N),
            "fy": np.zeros(N), "type": np.zeros(N, np.uint8), "mpos": np.full(N, -1, np.int8),
            "nxt": np.full(N, -1, np.int32), "bond": np.full(N, -1, np.int32)}


def _contig(s):
    out = {}
    for k, dt in (("x", np.float64), ("y", np.float64), ("vx", np.float64), ("vy", np.float64), ("fx", np.float64),
                  ("fy", np.float64), ("type", np.uint8), ("mpos", np.int8), ("nxt", np.int32),
                  ("bond", np.int32)):
        if k in s and s[k] is not None:
            out[k] = np.ascontiguousarray(s[k], dtype=dt)
    return out


def soup_host(params, N=None):
    """The reference soup (init_soup + set_velocities) of one replica, built on
    the host by the library (no GPU needed).  Velocities are full v."""
    lib = abi.load_library()
    p = params.to_c() if isinstance(params, Params) else params
    if N is None:
        N = params.N
    s = empty_state(N)
    st = abi.make_state(s)
    abi.check(lib, lib.oc_soup_host(C.byref(p), N, C.byref(st)))
    return s


class Engine:
    """One engine handle over one or more replicas (a single system is an
    ensemble of one)."""

    def __init__(self, params, N=None, pair_block=0, tile_beads=0, pair_lists=False, pair_variant=None):
        plist = params if isinstance(params, (list, tuple)) else [params]
        self.params = [p if isinstance(p, Params) else p for p in plist]
        for p in self.params:
            if isinstance(p, Params):
                p.check_engine()
        if N is None:
            N = [p.N for p in self.params]
        elif np.isscalar(N):
            N = [int(N)]
        self.N = [int(n) for n in N]
        self.nrep = len(self.params)
        self.lib = abi.load_library()
        self._cp = (abi.oc_params * self.nrep)(*[p.to_c() if isinstance(p, Params) else p for p in self.params])
        nn = (C.c_int32 * self.nrep)(*self.N)
        h = C.c_void_p()
        abi.check(self.lib, self.lib.oc_create(self._cp, nn, self.nrep, C.byref(h)))
        self.h = h
        if pair_block or tile_beads:
            self.set_tuning(pair_block, tile_beads)
        if pair_variant is None:
            pair_variant = 1 if pair_lists else None
        if pair_variant is not None:
            self._chk(self.lib.oc_set_pair_variant(self.h, int(pair_variant)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.oc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        return abi.check(self.lib, rc)

    def set_tuning(self, pair_block=0, tile_beads=0):
        self._chk(self.lib.oc_set_tuning(self.h, int(pair_block), int(tile_beads)))

    @property
    def stream(self):
        return self.lib.oc_stream(self.h)

    # ------------------------------------------------------------------ state
    def start(self, states=None):
        """Fresh start of every replica.  states: list of gid-order dicts with
        full velocities v (default: the reference soup of each replica)."""
        if states is None:
            states = [soup_host(p, n) for p, n in zip(self.params, self.N)]
        if isinstance(states, dict):
            states = [states]
        cat = {k: np.concatenate([np.asarray(s[k]) for s in states]) for k in ("x", "y", "vx", "vy", "type", "mpos",
                                                                                 "nxt")}
        cat["bond"] = np.concatenate([np.asarray(s.get("bond", np.full(len(s["x"]), -1, np.int32))) for s in states])
        cat = _contig(cat)
        st = abi.make_state(cat)
        self._chk(self.lib.oc_start(self.h, C.byref(st)))
        return self

    def set_state(self, s, step=0, r=0, slot_gid=None):
        """Restart-style load of replica r: x, y, vh, f as given (f optional);
        the predicted velocity is set to vh.  slot_gid (gids in slot order, as
        a checkpoint or cells() holds them) restores the slot order too, so the
        continuation is bitwise the uninterrupted run's."""
        c = _contig(s)
        st = abi.make_state(c)
        if slot_gid is None:
            self._chk(self.lib.oc_set_state(self.h, int(r), C.byref(st), int(step)))
        else:
            o = np.ascontiguousarray(slot_gid, dtype=np.uint32)
            if o.shape != (self.N[r],):
                raise ValueError(f"slot_gid must hold {self.N[r]} gids")
            self._chk(self.lib.oc_set_state_slots(self.h, int(r), C.byref(st), int(step),
                                                  o.ctypes.data_as(C.POINTER(C.c_uint32))))

    def get_state(self, r=0):
        s = empty_state(self.N[r])
        st = abi.make_state(s)
        self._chk(self.lib.oc_get_state(self.h, int(r), C.byref(st)))
        return s

    # ------------------------------------------------------------------ stepping
    def step(self, n=1, graph=True):
        self._chk(self.lib.oc_step(self.h, int(n), 1 if graph else 0))

    def sort(self):
        self._chk(self.lib.oc_sort(self.h))

    def forces(self, step, bonded=True):
        self._chk(self.lib.oc_forces(self.h, int(step), 1 if bonded else 0))

    def chem(self, step):
        self._chk(self.lib.oc_chem(self.h, int(step)))

    def forces_theta(self, step, lo, hi, theta, bonded=True):
        """Forces with the caller's pair noise: theta[q] of the gid pair
        (lo[q] < hi[q]), entries in increasing (lo, hi) order (oc_forces_theta)."""
        lo = np.ascontiguousarray(lo, dtype=np.uint32)
        hi = np.ascontiguousarray(hi, dtype=np.uint32)
        th = np.ascontiguousarray(theta, dtype=np.float32)
        P = C.POINTER
        self._chk(self.lib.oc_forces_theta(self.h, int(step), 1 if bonded else 0, lo.ctypes.data_as(P(C.c_uint32)),
                                           hi.ctypes.data_as(P(C.c_uint32)), th.ctypes.data_as(P(C.c_float)),
                                           len(th)))

    def chem_uniforms(self, step, u_conv, u_break, lo, hi, u_form):
        """One precursor-chemistry check with the caller's draws: per-gid
        conversion and break uniforms, per-pair formation uniforms
        (oc_chem_uniforms)."""
        uc = np.ascontiguousarray(u_conv, dtype=np.float64)
        ub = np.ascontiguousarray(u_break, dtype=np.float64)
        lo = np.ascontiguousarray(lo, dtype=np.uint32)
        hi = np.ascontiguousarray(hi, dtype=np.uint32)
        uf = np.ascontiguousarray(u_form, dtype=np.float64)
        P = C.POINTER
        dp = lambda a: a.ctypes.data_as(P(C.c_double))
        self._chk(self.lib.oc_chem_uniforms(self.h, int(step), dp(uc), dp(ub), lo.ctypes.data_as(P(C.c_uint32)),
                                            hi.ctypes.data_as(P(C.c_uint32)), dp(uf), len(uf)))

    # ------------------------------------------------------------------ readouts
    def cells(self, r=0):
        n = C.c_int32()
        self._chk(self.lib.oc_get_cells(self.h, int(r), None, None, C.byref(n)))
        cs = np.zeros(n.value + 1, np.int32)
        g = np.zeros(self.N[r], np.uint32)
        self._chk(self.lib.oc_get_cells(self.h, int(r), cs.ctypes.data_as(C.POINTER(C.c_int32)),
                                        g.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(n)))
        return cs, g

    def events(self, r=0):
        """Pending chemistry events of replica r (structured array: step, kind,
        gid_a, gid_b), sorted by (step, kind, gid_a, gid_b)."""
        n = C.c_int64()
        self._chk(self.lib.oc_get_events(self.h, int(r), None, 0, C.byref(n)))
        ev = np.zeros(n.value, dtype=abi.event_dtype())
        if n.value:
            self._chk(self.lib.oc_get_events(self.h, int(r), ev.ctypes.data_as(C.POINTER(abi.oc_event)), n.value,
                                             C.byref(n)))
        return ev

    def stats(self, r=0):
        o = abi.oc_stats()
        self._chk(self.lib.oc_get_stats(self.h, int(r), C.byref(o)))
        d = {k: getattr(o, k) for k, _ in abi.oc_stats._fields_ if k != "count"}
        d["count"] = list(o.count)[:abi.NTYPES]
        return d

    # ------------------------------------------------------------------ window gradient
    THETA_NAMES = ("a_W_W", "a_W_H", "a_W_T", "a_H_H", "a_H_T", "a_T_T", "gamma")
    THETA_IJ = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))

    def set_pair_params(self, a, gamma, r=0):
        """Replace replica r's repulsion matrix (7x7 or 8x8, symmetric) and gamma."""
        a64 = np.zeros((8, 8))
        a = np.asarray(a, dtype=np.float64)
        a64[:a.shape[0], :a.shape[1]] = a
        a64 = np.ascontiguousarray(a64)
        self._chk(self.lib.oc_set_pair_params(self.h, int(r), a64.ctypes.data_as(C.POINTER(C.c_double)),
                                              float(gamma)))

    def set_theta(self, theta, r=0):
        """theta = (a_WW, a_WH, a_WT, a_HH, a_HT, a_TT, gamma) of replica r; the
        other a_ij keep their values."""
        a = np.array(self.params[r].a, dtype=np.float64)
        for q, (i, j) in enumerate(self.THETA_IJ):
            a[i][j] = a[j][i] = float(theta[q])
        self.set_pair_params(a, float(theta[6]), r)
        p = self.params[r]
        for q, (i, j) in enumerate(self.THETA_IJ):
            p.a[i][j] = p.a[j][i] = float(theta[q])
        p.gamma = float(theta[6])

    def theta(self, r=0):
        p = self.params[r]
        return np.array([p.a[i][j] for i, j in self.THETA_IJ] + [p.gamma])

    def window_forward(self, nsteps, spec, ckpt_every=20):
        """Run an nsteps window (checkpoints every ckpt_every steps); returns the
        window-mean smooth RDF per replica, shape (nrep, nbins)."""
        g = np.zeros((self.nrep, spec.nbins))
        self._chk(self.lib.oc_window_forward(self.h, int(nsteps), int(ckpt_every), C.byref(spec),
                                             g.ctypes.data_as(C.POINTER(C.c_double))))
        return g

    def window_backward(self, dL_dgbar):
        """dL/dtheta (nrep, 7) of the last forward window for dL/dgbar (nrep, nbins)."""
        dl = np.ascontiguousarray(dL_dgbar, dtype=np.float64)
        g = np.zeros((self.nrep, 7))
        self._chk(self.lib.oc_window_backward(self.h, dl.ctypes.data_as(C.POINTER(C.c_double)),
                                              g.ctypes.data_as(C.POINTER(C.c_double))))
        return g

    def window_rewind(self):
        """Back to the start of the last forward window (state and steps)."""
        self._chk(self.lib.oc_window_rewind(self.h))

    def window_grad(self, nsteps, spec, target, ckpt_every=20):
        """L2 RDF loss per replica, its gradient (nrep, 7) and the mean RDF."""
        t = np.ascontiguousarray(target, dtype=np.float64)
        stride = spec.nbins if t.ndim == 2 else 0
        loss = np.zeros(self.nrep)
        g = np.zeros((self.nrep, 7))
        gb = np.zeros((self.nrep, spec.nbins))
        P = C.POINTER(C.c_double)
        self._chk(self.lib.oc_window_grad(self.h, int(nsteps), int(ckpt_every), C.byref(spec), t.ctypes.data_as(P),
                                          stride, loss.ctypes.data_as(P), g.ctypes.data_as(P), gb.ctypes.data_as(P)))
        return loss, g, gb

    def max_bond(self, r=0, set_to=None):
        """Running max bond length of replica r (since the last stats() call);
        set_to replaces it afterwards (a checkpoint's value on restart)."""
        g = C.c_double()
        sv = None if set_to is None else C.byref(C.c_double(float(set_to)))
        self._chk(self.lib.oc_max_bond(self.h, int(r), C.byref(g), sv))
        return g.value

    def step_count(self, r=0):
        s = C.c_int64()
        self._chk(self.lib.oc_get_step(self.h, int(r), C.byref(s)))
        return s.value
