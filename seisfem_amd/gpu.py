"""Python mirror of the GPU time-loop C ABI (include/seisfem_gpu.h).

``Operator`` holds M, S1..S6 on the device, ``Medium`` the per-node
coefficients, ``Run`` a batch of shots being stepped.  ``simulate`` is the
one-call equivalent of the reference's masslump shot loop
(time_evolution/elastic_wave.c:446-818) for a problem from host.build_problem.
"""
import ctypes as C

import numpy as np

from ._native import NativeError, ShotDesc, check, lib  # noqa: F401  (NativeError re-exported)

FP64, FP32, FP32_KAHAN, FP32_KAHAN16 = 0, 1, 2, 3
PRECISION = {"fp64": FP64, "fp32": FP32, "fp32_kahan": FP32_KAHAN, "fp32_kahan16": FP32_KAHAN16}


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Operator:
    def __init__(self, node_num, csr):
        """csr: 7 tuples (p, j, x) in the order M, S1..S6."""
        L = lib()
        self._keep = [(np.ascontiguousarray(p, np.int32), np.ascontiguousarray(j, np.int32),
                       np.ascontiguousarray(x, np.float64)) for p, j, x in csr]
        P = (C.POINTER(C.c_int) * 7)(*[_pi(t[0]) for t in self._keep])
        J = (C.POINTER(C.c_int) * 7)(*[_pi(t[1]) for t in self._keep])
        X = (C.POINTER(C.c_double) * 7)(*[_pd(t[2]) for t in self._keep])
        h = C.c_void_p()
        check(L.sfg_operator_create(C.byref(h), node_num, P, J, X), "sfg_operator_create")
        self.h = h
        self.node_num = node_num
        self._keep = None

    @property
    def nnz(self):
        return lib().sfg_operator_nnz(self.h)

    def close(self):
        if self.h:
            lib().sfg_operator_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Medium:
    def __init__(self, op, c, mass_lump, dirichlet, mpml):
        L = lib()
        keep = [np.ascontiguousarray(a, np.float64) for a in c] + \
               [np.ascontiguousarray(a, np.float64) for a in mpml]
        cc = (C.POINTER(C.c_double) * 4)(*[_pd(a) for a in keep[:4]])
        pp = (C.POINTER(C.c_double) * 6)(*[_pd(a) for a in keep[4:]])
        m = np.ascontiguousarray(mass_lump, np.float64)
        f = np.ascontiguousarray(dirichlet, np.int32)
        h = C.c_void_p()
        check(L.sfg_medium_create(C.byref(h), op.h, cc, _pd(m), _pi(f), pp), "sfg_medium_create")
        self.h = h
        self.op = op

    def set_row_scale(self, scale):
        """rho_model / rho_operator per node (per-shot media only)."""
        s = np.ascontiguousarray(scale, np.float64)
        check(lib().sfg_medium_set_row_scale(self.h, _pd(s)), "sfg_medium_set_row_scale")

    def close(self):
        if self.h:
            lib().sfg_medium_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Run:
    def __init__(self, op, md, src_nodes, max_steps, rec_nodes, dt, f0, t0, amplitude=1.0e10,
                 angle_deg=90.0, precision=FP64, nsteps=None, block=0, shots_per_lane_group=0,
                 guard_every=0, use_graph=True, xcd_remap=True, sources=None, kernel_variant=0, row_order=None,
                 sell_sigma=0):
        """sources: optional per-shot list of (nodes, weights[k][2]) distributed
        sources (u and w weights on each node) replacing the point forces.
        row_order: optional permutation of the nodes giving the order in which
        rows are stepped (host.locality_order); results do not depend on it."""
        L = lib()
        self.op, self.md = op, md
        self.src = np.ascontiguousarray(np.atleast_1d(src_nodes), np.int32)
        self.rec = np.ascontiguousarray(rec_nodes, np.int32)
        self.nshots = self.src.size
        self.max_steps = int(max_steps)
        self.steps = None if nsteps is None else np.ascontiguousarray(nsteps, np.int32)
        d = ShotDesc()
        d.precision = precision
        d.nshots = self.nshots
        d.src_node = _pi(self.src)
        d.nsteps = _pi(self.steps) if self.steps is not None else C.POINTER(C.c_int)()
        d.max_steps = self.max_steps
        d.rec_num = self.rec.size
        d.rec_node = _pi(self.rec) if self.rec.size else C.POINTER(C.c_int)()
        d.dt, d.f0, d.t0, d.amplitude, d.angle_deg = dt, f0, t0, amplitude, angle_deg
        d.block, d.shots_per_lane_group = block, shots_per_lane_group
        d.guard_every, d.use_graph = guard_every, int(bool(use_graph))
        # True: contiguous ranges per XCD; False: round-robin; an int Q > 1: stripes of Q row blocks
        d.xcd_remap = xcd_remap if (type(xcd_remap) is int and xcd_remap > 1) else (0 if xcd_remap else -1)
        d.kernel_variant = int(kernel_variant)
        d.sell_sigma = int(sell_sigma)
        if row_order is not None:
            self._order = np.ascontiguousarray(row_order, np.int32)
            d.row_order = _pi(self._order)
        if sources is not None:
            if len(sources) != self.nshots:
                raise ValueError("need one source description per shot")
            self._scount = np.array([len(nd) for nd, _ in sources], np.int32)
            self._snodes = np.ascontiguousarray(np.concatenate([np.asarray(nd, np.int32) for nd, _ in sources]))
            self._swt = np.ascontiguousarray(np.concatenate([np.asarray(wt, np.float64).reshape(-1, 2)
                                                             for _, wt in sources]))
            d.src_count, d.src_nodes, d.src_weight = _pi(self._scount), _pi(self._snodes), _pd(self._swt)
        h = C.c_void_p()
        if isinstance(md, (list, tuple)):
            # one medium per shot (many models on one mesh)
            if len(md) != self.nshots:
                raise ValueError("need one Medium per shot")
            arr = (C.c_void_p * self.nshots)(*[m.h.value for m in md])
            check(L.sfg_run_create_media(C.byref(h), op.h, arr, C.byref(d)), "sfg_run_create_media")
        else:
            check(L.sfg_run_create(C.byref(h), op.h, md.h, C.byref(d)), "sfg_run_create")
        self.h = h
        self.precision = precision

    def advance(self, nsteps):
        return check(lib().sfg_run_advance(self.h, int(nsteps)), "sfg_run_advance")

    def sync(self):
        return check(lib().sfg_run_sync(self.h), "sfg_run_sync")

    def reset(self):
        """Back to the state at creation (zero fields, seismograms, step 0)."""
        return check(lib().sfg_run_reset(self.h), "sfg_run_reset")

    def seismograms(self):
        """(u, w) arrays [nshots, max_steps, rec_num]."""
        shape = (self.nshots, self.max_steps, self.rec.size)
        u = np.zeros(shape, np.float64)
        w = np.zeros(shape, np.float64)
        check(lib().sfg_run_seismograms(self.h, _pd(u), _pd(w)), "sfg_run_seismograms")
        return u, w

    def snapshot(self, shot=0):
        n = self.op.node_num
        U = np.zeros(n, np.float64)
        W = np.zeros(n, np.float64)
        check(lib().sfg_run_snapshot(self.h, shot, _pd(U), _pd(W)), "sfg_run_snapshot")
        return U, W

    def energy(self, max_checks=100000):
        buf = np.zeros((max_checks, self.nshots, 2), np.float64)
        k = check(lib().sfg_run_energy(self.h, _pd(buf), max_checks), "sfg_run_energy")
        return buf[:k]

    @property
    def last_ms_per_step(self):
        return lib().sfg_run_last_ms_per_step(self.h)

    def step_bytes(self, active=None):
        return lib().sfg_run_step_bytes(self.h, self.nshots if active is None else int(active))

    def close(self):
        if self.h:
            lib().sfg_run_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Multi:
    """Several runs (own meshes, media, shots, step counts) stepped by one
    kernel launch per step (sfg_multi_*)."""

    def __init__(self, runs):
        self.runs = list(runs)
        arr = (C.c_void_p * len(self.runs))(*[r.h.value for r in self.runs])
        h = C.c_void_p()
        check(lib().sfg_multi_create(C.byref(h), arr, len(self.runs)), "sfg_multi_create")
        self.h = h

    def advance(self, nsteps):
        return check(lib().sfg_multi_advance(self.h, int(nsteps)), "sfg_multi_advance")

    @property
    def last_ms_per_step(self):
        return lib().sfg_multi_last_ms_per_step(self.h)

    def set_persistent(self, blocks_per_cu):
        """blocks_per_cu > 0: one cooperative launch per guard interval with that
        many resident blocks per CU pulling each step's tiles from a queue; 0: the
        graph-replayed per-step launches (default)."""
        check(lib().sfg_multi_set_persistent(self.h, int(blocks_per_cu)), "sfg_multi_set_persistent")

    def step_bytes(self, step):
        return lib().sfg_multi_step_bytes(self.h, int(step))

    def close(self):
        if self.h:
            lib().sfg_multi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stream_copy(nbytes=1 << 30, reps=20):
    """(GB/s, ms per copy) of a plain device copy of `nbytes` (read + write counted)."""
    gbs = C.c_double(0.0)
    ms = C.c_double(0.0)
    check(lib().sfg_stream_copy(int(nbytes), int(reps), C.byref(gbs), C.byref(ms)), "sfg_stream_copy")
    return gbs.value, ms.value


def estimate_dt(op, md, max_iter=2000, tol=1e-9):
    """(lambda_max, dt_max, iterations) of M_L^{-1} K by FP64 power iteration on the GPU."""
    lam = C.c_double(0.0)
    dtm = C.c_double(0.0)
    it = check(lib().sfg_estimate_dt(op.h, md.h, int(max_iter), float(tol), C.byref(lam), C.byref(dtm)),
               "sfg_estimate_dt")
    return lam.value, dtm.value, it


def simulate(prob, src_nodes, nsteps_exec, rec_nodes, dt, f0, t0, precision=FP64, amplitude=1.0e10,
             angle_deg=90.0, snap_every=0, **kw):
    """Runs the loop for problem ``prob`` and returns dict(seis_u, seis_w, snaps).

    nsteps_exec is the executed step count (the reference's ``step`` - 2).
    With snap_every > 0 the snapshot rows of elastic_wave.c:797-806 are
    gathered for shot 0: the initial zero row, then every step whose
    reference index it satisfies (it+1) % snap_every == 0.
    """
    op = Operator(prob["node_num"], prob["csr"])
    md = Medium(op, prob["c"], prob["mass_lump"], prob["dirichlet"], prob["mpml"])
    run = Run(op, md, src_nodes, nsteps_exec, rec_nodes, dt, f0, t0, amplitude=amplitude,
              angle_deg=angle_deg, precision=precision, **kw)
    snaps_u, snaps_w = [], []
    if snap_every > 0:
        n = prob["node_num"]
        snaps_u.append(np.zeros(n))
        snaps_w.append(np.zeros(n))
        done = 0
        # executed step t has reference index it = t + 2; snapshot after it when (it+1) % k == 0
        marks = [t for t in range(nsteps_exec) if (t + 3) % snap_every == 0]
        for t in marks:
            run.advance(t + 1 - done)
            done = t + 1
            U, W = run.snapshot(0)
            snaps_u.append(U)
            snaps_w.append(W)
        run.advance(nsteps_exec - done)
    else:
        run.advance(nsteps_exec)
    u, w = run.seismograms()
    out = dict(seis_u=u, seis_w=w, snaps_u=np.array(snaps_u), snaps_w=np.array(snaps_w),
               ms_per_step=run.last_ms_per_step)
    run.close()
    md.close()
    op.close()
    return out
