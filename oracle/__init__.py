"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes access to
  * liboracle.so       -- the C restatement of the reference's masslump loop,
                          M-PML profile, SpMV and Ricker (sfem_oracle.c);
  * _ref/libsfem_ref.so -- the reference's own setup/SpMV/M-PML/Ricker sources
                          compiled unchanged from /root/reference (see Makefile).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this package, and only as the checker.  The product never imports it.
"""
import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build (the sanitized one of tests/sanitize.sh)
ORACLE_LIB = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libsfem_ref.so")

_i, _d = C.c_int, C.c_double
_pi, _pd = C.POINTER(C.c_int), C.POINTER(C.c_double)


def _ip(a):
    return a.ctypes.data_as(_pi)


def _dp(a):
    return a.ctypes.data_as(_pd)


_olib = None


def olib():
    global _olib
    if _olib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError("oracle/liboracle.so missing: run __graft_entry__.build()")
        L = C.CDLL(ORACLE_LIB)
        L.oracle_csr_matvec.argtypes = [_i, _pi, _pi, _pd, _pd, _pd]
        L.oracle_csr_matvec.restype = None
        L.oracle_seismic_source.argtypes = [_d, _d, _d, _d]
        L.oracle_seismic_source.restype = _d
        L.oracle_abc_mpml.argtypes = [_i, _i, _i, _pi, _pd, _pd, _i, _i, _d, _d, _d, _d, _d, _d,
                                      _i, _i, _i, _i, _pd, _pd, _pd, _pd, _pd, _pd]
        L.oracle_abc_mpml.restype = None
        L.oracle_elastic_masslump.argtypes = [
            _i, C.POINTER(_pi), C.POINTER(_pi), C.POINTER(_pd), C.POINTER(_pd), _pd, _pi,
            C.POINTER(_pd), _i, _d, _d, _d, _d, _d, _i, _i, _pi, _pd, _pd, _i, _pd, _pd, _pi, _pd,
            _pd, _pi]
        L.oracle_elastic_masslump.restype = _i
        L.oracle_elastic_masslump_timed.argtypes = L.oracle_elastic_masslump.argtypes + [C.c_char_p, _pd]
        L.oracle_elastic_masslump_timed.restype = _i
        L.oracle_set_shared_energy.argtypes = [_i]
        L.oracle_set_shared_energy.restype = None
        _olib = L
    return _olib


def ref_available():
    return os.path.exists(REF_LIB)


_rlib = None


def rlib():
    """The reference's own compiled sources (only where oracle/_ref was built)."""
    global _rlib
    if _rlib is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libsfem_ref.so not built (needs /root/reference)")
        L = C.CDLL(REF_LIB)
        P2 = C.POINTER(_pd)
        L.csr_matvec.argtypes = [_i, _pi, _pi, _pd, _pd, _pd]
        L.csr_matvec.restype = None
        L.seismic_source.argtypes = [_d, _d, _d, _d]
        L.seismic_source.restype = _d
        L.abc_mpml.argtypes = [_i, _i, _i, _pi, P2, _i, _i, _d, _d, _d, _d, _d, _d, _i, _i, _i, _i,
                               _pd, _pd, _pd, _pd, _pd, _pd]
        L.abc_mpml.restype = None
        L.mass_sparse_all.argtypes = [C.c_char_p, _i, _i, _i, _pi, P2, _pd, _pi, _pi, _pd, _i]
        L.mass_sparse_all.restype = None
        L.stif_sparse_all.argtypes = [C.c_char_p, _i, _i, _i, _pi, P2, _pi, _pi, _pd, _i]
        L.stif_sparse_all.restype = None
        L.mesh_element.argtypes = [C.c_char_p, _i, _i, _pi]
        L.mesh_element.restype = _i
        L.mesh_xy.argtypes = [C.c_char_p, _i, _i, _i, _d, _d, _d, _d, P2]
        L.mesh_xy.restype = _i
        L.node_location.argtypes = [_i, _d, P2, _d, _d]
        L.node_location.restype = _i
        L.set_rec_and_src.argtypes = [_i, _d, _d, _d, _d, _pd, _pd, _i, _d, _d, _d, _d, _pd, _pd]
        L.set_rec_and_src.restype = None
        f = L._QMcoo2csr_libPcoo2csr_canonical
        f.argtypes = [_pi, _pi, _pi, _pi, _pi, _pd, _pi, _pi, _pd]
        f.restype = None
        _rlib = L
    return _rlib


def _big_stack(fn):
    """The reference's Fortran COO->CSR keeps nnz-sized automatic arrays on the
    stack; run it on a thread with a large stack."""
    out = {}

    def body():
        out["r"] = fn()

    old = threading.stack_size()
    threading.stack_size(1 << 30)
    try:
        th = threading.Thread(target=body)
        th.start()
        th.join()
    finally:
        threading.stack_size(old)
    return out["r"]


TYPE_STR = {1: b"T3", 2: b"T6", 4: b"Q4", 5: b"Q9"}


def ref_mesh(type_code, nelemx, nelemy, xmin, xmax, ymin, ymax, node_num, element_num, order):
    L = rlib()
    en = np.zeros(element_num * order, np.int32)
    L.mesh_element(TYPE_STR[type_code], nelemx, nelemy, _ip(en))
    x = np.zeros(node_num)
    y = np.zeros(node_num)
    xy = (_pd * 2)(_dp(x), _dp(y))
    L.mesh_xy(TYPE_STR[type_code], nelemx, nelemy, node_num, xmin, xmax, ymin, ymax, xy)
    return en, x, y


def ref_coo2csr(node_num, ai, aj, ax):
    """Reference Fortran canonicalisation (copies its inputs: it mutates them)."""
    L = rlib()
    ai = np.array(ai, np.int32)
    aj = np.array(aj, np.int32)
    ax = np.array(ax, np.float64)
    nnz = C.c_int(ai.size)
    psz = C.c_int(node_num + 1)
    csz = C.c_int(0)
    bp = np.zeros(node_num + 1, np.int32)
    bj = np.zeros(max(ai.size, 1), np.int32)
    bx = np.zeros(max(ai.size, 1), np.float64)

    def go():
        L._QMcoo2csr_libPcoo2csr_canonical(C.byref(nnz), C.byref(psz), C.byref(csz), _ip(ai), _ip(aj),
                                          _dp(ax), _ip(bp), _ip(bj), _dp(bx))
        return csz.value

    k = _big_stack(go)
    return bp, bj[:k].copy(), bx[:k].copy()


def ref_assemble(type_code, element_node, x, y, rho):
    """Reference COO for M (lumpflag 0), the lumped mass, and S1..S6."""
    L = rlib()
    order = {1: 3, 2: 6, 4: 4, 5: 9}[type_code]
    ne = element_node.size // order
    n = x.size
    nnz = ne * order * order
    xy = (_pd * 2)(_dp(x), _dp(y))
    en = np.ascontiguousarray(element_node, np.int32)
    rho = np.ascontiguousarray(rho, np.float64)
    out = {}
    mi = np.zeros(nnz, np.int32)
    mj = np.zeros(nnz, np.int32)
    mx = np.zeros(max(nnz, n), np.float64)
    L.mass_sparse_all(TYPE_STR[type_code], n, ne, order, _ip(en), xy, _dp(rho), _ip(mi), _ip(mj), _dp(mx), 1)
    out["lump"] = mx[:n].copy()
    mx = np.zeros(nnz, np.float64)
    L.mass_sparse_all(TYPE_STR[type_code], n, ne, order, _ip(en), xy, _dp(rho), _ip(mi), _ip(mj), _dp(mx), 0)
    out[0] = (mi.copy(), mj.copy(), mx.copy())
    for k in range(1, 7):
        si = np.zeros(nnz, np.int32)
        sj = np.zeros(nnz, np.int32)
        sx = np.zeros(nnz, np.float64)
        L.stif_sparse_all(TYPE_STR[type_code], n, ne, order, _ip(en), xy, _ip(si), _ip(sj), _dp(sx), k)
        out[k] = (si, sj, sx)
    return out


def ref_abc_mpml(element_node, order, x, y, pml_nx, pml_ny, h, xmin, xmax, ymin, ymax, vp_max, use):
    L = rlib()
    n = x.size
    outs = [np.full(n, np.nan) for _ in range(6)]
    xy = (_pd * 2)(_dp(x), _dp(y))
    en = np.ascontiguousarray(element_node, np.int32)
    L.abc_mpml(n, en.size // order, order, _ip(en), xy, pml_nx, pml_ny, h, xmin, xmax, ymin, ymax, vp_max,
               *use, *[_dp(a) for a in outs])
    return outs


def oracle_abc_mpml(element_node, order, x, y, pml_nx, pml_ny, h, xmin, xmax, ymin, ymax, vp_max, use):
    L = olib()
    n = x.size
    outs = [np.full(n, np.nan) for _ in range(6)]
    en = np.ascontiguousarray(element_node, np.int32)
    L.oracle_abc_mpml(n, en.size // order, order, _ip(en), _dp(x), _dp(y), pml_nx, pml_ny, h, xmin, xmax,
                      ymin, ymax, vp_max, *use, *[_dp(a) for a in outs])
    return outs


def oracle_csr_matvec(p, j, x_, v):
    y = np.zeros(p.size - 1)
    olib().oracle_csr_matvec(p.size, _ip(np.ascontiguousarray(p, np.int32)), _ip(np.ascontiguousarray(j, np.int32)),
                             _dp(np.ascontiguousarray(x_, np.float64)), _dp(np.ascontiguousarray(v, np.float64)),
                             _dp(y))
    return y


def set_shared_energy(on):
    """Timing only: the loop's energy sums accumulated the reference's way (one
    shared pair of sums written by every thread, elastic_wave.c:776-784) instead
    of by a reduction.  The seismograms are unchanged; the sums are not exact."""
    olib().oracle_set_shared_energy(1 if on else 0)


def oracle_masslump(prob, src_node, step, rec_nodes, dt, f0, t0, amplitude=1.0e10, angle_deg=90.0,
                    snap_every=0, want_state=False, io_dir=None):
    """One shot of the reference loop (step = the reference's ``step``; it
    executes step-2 steps).  Returns dict(seis_u, seis_w, snaps_u, snaps_w,
    energy, state, status, loop_seconds).  With io_dir the loop also writes
    the reference's %f seismogram and wavefield text files there;
    loop_seconds is omp_get_wtime() around the step loop alone."""
    L = olib()
    n = prob["node_num"]
    csr = [(np.ascontiguousarray(p, np.int32), np.ascontiguousarray(j, np.int32),
            np.ascontiguousarray(x, np.float64)) for p, j, x in prob["csr"]]
    P = (_pi * 7)(*[_ip(t[0]) for t in csr])
    J = (_pi * 7)(*[_ip(t[1]) for t in csr])
    X = (_pd * 7)(*[_dp(t[2]) for t in csr])
    cs = [np.ascontiguousarray(a, np.float64) for a in prob["c"]]
    Cc = (_pd * 4)(*[_dp(a) for a in cs])
    pm = [np.ascontiguousarray(a, np.float64) for a in prob["mpml"]]
    Pm = (_pd * 6)(*[_dp(a) for a in pm])
    mass = np.ascontiguousarray(prob["mass_lump"], np.float64)
    flag = np.ascontiguousarray(prob["dirichlet"], np.int32)
    rec = np.ascontiguousarray(rec_nodes, np.int32)
    ns = max(step - 2, 0)
    su = np.zeros((ns, rec.size))
    sw = np.zeros((ns, rec.size))
    nsnap_max = (step // snap_every + 2) if snap_every > 0 else 0
    snu = np.zeros((max(nsnap_max, 1), n))
    snw = np.zeros((max(nsnap_max, 1), n))
    nsnap = C.c_int(0)
    energy = np.zeros((step, 2))
    state = np.zeros((26, n)) if want_state else None
    zm = C.c_int(0)
    loop_s = C.c_double(0.0)
    rc = L.oracle_elastic_masslump_timed(
        n, P, J, X, Cc, _dp(mass), _ip(flag), Pm, step, dt, f0, t0, amplitude, angle_deg, int(src_node),
        rec.size, _ip(rec) if rec.size else _pi(), _dp(su), _dp(sw), int(snap_every),
        _dp(snu) if snap_every > 0 else _pd(), _dp(snw) if snap_every > 0 else _pd(), C.byref(nsnap),
        _dp(energy), _dp(state) if want_state else _pd(), C.byref(zm),
        None if io_dir is None else os.fsencode(io_dir), C.byref(loop_s))
    k = nsnap.value
    return dict(seis_u=su, seis_w=sw, snaps_u=snu[:k], snaps_w=snw[:k], energy=energy, state=state,
                status=rc, zero_mass=zm.value, loop_seconds=loop_s.value)
