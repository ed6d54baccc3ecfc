"""Host-side setup (mesh, element matrices, canonical CSR, stations, flags).

Thin numpy wrappers over the sfh_* C functions, which rebuild the arrays the
reference's elastic_wave() consumes (time_evolution/elastic_wave.c:136-313)
bit for bit.  ``build_problem`` strings them together for one mesh + model.
"""
import ctypes as C

import numpy as np

from ._native import check, lib

T3, T6, Q4, Q9 = 1, 2, 4, 5
TYPE_NAMES = {T3: "T3", T6: "T6", Q4: "Q4", Q9: "Q9"}
MATRIX_NAMES = ("M", "S1", "S2", "S3", "S4", "S5", "S6")


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def element_order(type_code):
    return lib().sfh_element_order(type_code)


def mesh_structured(type_code, nelemx, nelemy, xmin, xmax, ymin, ymax):
    """Structured mesh: (element_node [E*order] 1-based int32, x, y)."""
    L = lib()
    n = L.sfh_mesh_node_num(type_code, nelemx, nelemy)
    ne = L.sfh_mesh_element_num(type_code, nelemx, nelemy)
    o = L.sfh_element_order(type_code)
    if n <= 0 or ne <= 0 or o <= 0:
        raise ValueError(f"bad mesh request {type_code} {nelemx}x{nelemy}")
    en = np.zeros(ne * o, np.int32)
    x = np.zeros(n, np.float64)
    y = np.zeros(n, np.float64)
    check(L.sfh_mesh_element(type_code, nelemx, nelemy, _pi(en)), "sfh_mesh_element")
    check(L.sfh_mesh_xy(type_code, nelemx, nelemy, xmin, xmax, ymin, ymax, _pd(x), _pd(y)), "sfh_mesh_xy")
    return en, x, y


def assemble_coo(type_code, element_node, x, y, rho, which=range(7)):
    """COO triplets of M, S1..S6: (ci, cj, {m: values})."""
    L = lib()
    o = element_order(type_code)
    ne = element_node.size // o
    n = x.size
    nnz = ne * o * o
    ci = np.empty(nnz, np.int32)
    cj = np.empty(nnz, np.int32)
    vals = {m: np.empty(nnz, np.float64) for m in which}
    arr = (C.POINTER(C.c_double) * 7)()
    for m in range(7):
        arr[m] = _pd(vals[m]) if m in vals else C.POINTER(C.c_double)()
    en = np.ascontiguousarray(element_node, np.int32)
    rho = np.ascontiguousarray(rho, np.float64)
    check(L.sfh_assemble_coo(type_code, n, ne, _pi(en), _pd(x), _pd(y), _pd(rho), _pi(ci), _pi(cj), arr),
          "sfh_assemble_coo")
    return ci, cj, vals


def lumped_mass(type_code, element_node, x, y, rho):
    L = lib()
    o = element_order(type_code)
    ne = element_node.size // o
    m = np.empty(x.size, np.float64)
    en = np.ascontiguousarray(element_node, np.int32)
    check(L.sfh_lumped_mass(type_code, x.size, ne, _pi(en), _pd(x), _pd(y),
                            _pd(np.ascontiguousarray(rho, np.float64)), _pd(m)), "sfh_lumped_mass")
    return m


def coo2csr(node_num, ci, cj, cx):
    """Canonical CSR (p, j, x) of one COO matrix (coo2csr_lib.f90 semantics)."""
    L = lib()
    nnz = ci.size
    bp = np.empty(node_num + 1, np.int32)
    bj = np.empty(max(nnz, 1), np.int32)
    bx = np.empty(max(nnz, 1), np.float64)
    size = C.c_int(0)
    check(L.sfh_coo2csr(nnz, node_num, _pi(np.ascontiguousarray(ci, np.int32)),
                        _pi(np.ascontiguousarray(cj, np.int32)), _pd(np.ascontiguousarray(cx, np.float64)),
                        _pi(bp), _pi(bj), _pd(bx), C.byref(size)), "sfh_coo2csr")
    k = size.value
    return bp, bj[:k].copy(), bx[:k].copy()


def node_location(edge_size, x, y, px, py):
    return lib().sfh_node_location(x.size, edge_size, _pd(x), _pd(y), px, py)


def station_line(num, x_first, y_first, x_last, y_last):
    xs = np.empty(num, np.float64)
    ys = np.empty(num, np.float64)
    lib().sfh_station_line(num, x_first, y_first, x_last, y_last, _pd(xs), _pd(ys))
    return xs, ys


def locality_order(x, y):
    """Node ids along a Hilbert curve over the coordinates (sfh_locality_order)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    order = np.empty(x.size, np.int32)
    rc = lib().sfh_locality_order(x.size, _pd(x), _pd(y), _pi(order))
    if rc != 0:
        raise ValueError(f"sfh_locality_order failed with code {rc}")
    return order


def mesh_line_length(x, y):
    """Nodes per grid line of a structured mesh numbered line by line (the
    reference's generators, mesh/mesh_xy_*.c): the first node at which the
    coordinate that grows along a line falls back.  0 if there is none."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    for c in (x, y):
        drop = np.flatnonzero(np.diff(c) < 0)
        if drop.size:
            return int(drop[0] + 1)
    return 0


def patch_order(x, y, pw=16, ph=16):
    """Rows in patches of pw nodes x ph grid lines of a structured mesh,
    patches line by line, each patch's rows line by line: a block of pw*ph
    rows then covers a compact piece of the mesh whose halo is small (the
    LDS-staged kernel variant), while every pw rows stay consecutive nodes.
    Node order when the mesh has no line structure."""
    n = len(x)
    line = mesh_line_length(x, y)
    ids = np.arange(n, dtype=np.int64)
    if line <= 0:
        return ids.astype(np.int32)
    ix, iy = ids % line, ids // line
    return np.lexsort((ix, iy, ix // pw, iy // ph)).astype(np.int32)


def dirichlet_flags(x, y, edge_size, xmin, xmax, ymin, ymax, free_surface):
    f = np.empty(x.size, np.int32)
    lib().sfh_dirichlet_flags(x.size, _pd(x), _pd(y), edge_size, xmin, xmax, ymin, ymax,
                              int(free_surface), _pi(f))
    return f


def abc_mpml(element_node, element_order_, x, y, pml_nx, pml_ny, h, xmin, xmax, ymin, ymax,
             vp_max, use=(1, 1, 1, 1)):
    """GPU M-PML profiles (drop-in abc_mpml) -> 6 arrays dx, dy, dxx, dyy, dxx_pyx, dyy_pxy."""
    L = lib()
    n = x.size
    out = [np.zeros(n, np.float64) for _ in range(6)]
    xy = (C.POINTER(C.c_double) * 2)(_pd(x), _pd(y))
    en = np.ascontiguousarray(element_node, np.int32)
    L.abc_mpml(n, en.size // element_order_, element_order_, _pi(en), xy, pml_nx, pml_ny, h,
               xmin, xmax, ymin, ymax, vp_max, *use, *[_pd(a) for a in out])
    from ._native import last_error
    err = last_error()
    if err:
        raise RuntimeError(f"abc_mpml: {err}")
    return out


def elastic_constants(rho, vp, vs):
    """Isotropic c11, c13, c33, c44 from rho, vp, vs, rounded as
    model_elastic_parameter/model_elastic_parameter.c:37-45 forms them."""
    mu = vs * vs * rho
    lam = vp * vp * rho - 2 * mu
    return [lam + 2 * mu, lam, lam + 2 * mu, mu]


def build_problem(type_code, element_node, x, y, rho, c, vp, edge_size, pml_nx, pml_ny,
                  free_surface, bbox=None, mpml=None):
    """Everything elastic_wave() assembles before its step loop
    (time_evolution/elastic_wave.c:136-313) for one mesh and model.

    c: list of 4 node arrays (c11, c13, c33, c44).  If ``mpml`` is None the
    M-PML profiles are computed by the GPU drop-in abc_mpml.
    """
    n = x.size
    if bbox is None:
        bbox = (float(x.min()), float(x.max()), float(y.min()), float(y.max()))
    xmin, xmax, ymin, ymax = bbox
    flags = dirichlet_flags(x, y, edge_size, xmin, xmax, ymin, ymax, free_surface)
    vp_max = float(np.max(vp))
    use = (1, 1 - int(free_surface), 1, 1)  # elastic_wave.c:107-110, 157
    if mpml is None:
        mpml = abc_mpml(element_node, element_order(type_code), x, y, pml_nx, pml_ny, edge_size,
                        xmin, xmax, ymin, ymax, vp_max, use)
    mass_lump = lumped_mass(type_code, element_node, x, y, rho)
    ci, cj, vals = assemble_coo(type_code, element_node, x, y, rho)
    csr = [coo2csr(n, ci, cj, vals[m]) for m in range(7)]
    return dict(type_code=type_code, node_num=n, element_node=element_node, x=x, y=y, rho=rho,
                c=[np.ascontiguousarray(a, np.float64) for a in c], vp=vp, mass_lump=mass_lump,
                dirichlet=flags, mpml=mpml, csr=csr, bbox=bbox, edge_size=edge_size,
                use_mpml=use, vp_max=vp_max)
