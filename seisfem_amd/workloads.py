"""Synthetic workloads (meshes, models, stations) for the benchmark configurations.

Coordinates follow the reference's convention for a free surface: the free
surface is the x = xmax side (time_evolution/elastic_wave.c:146-157), so x is
height above the bottom and y is the horizontal coordinate
(EXAMPLE/example6_surface_wave/README).  "Depth" below means xmax - x.

Every builder returns a dict with the assembled ``problem`` (host.build_problem)
plus source/receiver nodes, dt, f0, t0 and executed step counts.  The M-PML
profiles come from ``mpml_fn`` (default: the GPU drop-in abc_mpml).
"""
import numpy as np

from . import host


def _gpu_mpml(*args, **kw):
    return host.abc_mpml(*args, **kw)


def _stations_on_surface(x, y, edge, xs_ref, ys):
    return np.array([host.node_location(edge, x, y, xs_ref, yy) for yy in ys], np.int32)


def isotropic_model(rho, vp, vs):
    return rho, vp, vs, host.elastic_constants(rho, vp, vs)


def layered_model(depth, rng, nlayers, dmax, vmin, vmax, thomsen=False):
    """Seeded layered model: interfaces sorted U[0, dmax], vp rising linearly
    vmin->vmax over the layers, vs = vp/1.73, rho = 310 vp^0.25 (Gardner).
    With thomsen=True each layer gets eps ~ U[0,0.25], delta ~ U[-0.1,0.15]
    mapped to VTI stiffnesses (c33 = rho vp^2, c44 = rho vs^2,
    c11 = c33 (1+2 eps), c13 = sqrt((c33-c44)(c33(1+2 delta)-c44)) - c44)."""
    cuts = np.sort(rng.uniform(0.0, dmax, nlayers - 1))
    layer = np.searchsorted(cuts, depth, side="right")
    vps = np.linspace(vmin, vmax, nlayers)
    vp = vps[layer]
    vs = vp / 1.73
    rho = 310.0 * vp ** 0.25
    if not thomsen:
        return rho, vp, vs, host.elastic_constants(rho, vp, vs), cuts
    eps = rng.uniform(0.0, 0.25, nlayers)[layer]
    dlt = rng.uniform(-0.1, 0.15, nlayers)[layer]
    c33 = rho * vp * vp
    c44 = rho * vs * vs
    c11 = c33 * (1.0 + 2.0 * eps)
    c13 = np.sqrt((c33 - c44) * (c33 * (1.0 + 2.0 * dlt) - c44)) - c44
    return rho, vp, vs, [c11, c13, c33, c44], cuts


def structured_case(type_code, nex, ney, edge, model_fn, pml, dt, f0, t0, steps_exec,
                    src_depth, src_y, rec_num, rec_y0, rec_y1, mpml_fn=None):
    """Structured mesh of nex x ney elements (x = height, y = horizontal)."""
    mpml_fn = mpml_fn or _gpu_mpml
    xmax, ymax = edge * nex, edge * ney
    en, x, y = host.mesh_structured(type_code, nex, ney, 0.0, xmax, 0.0, ymax)
    rho, vp, vs, c = model_fn(xmax - x, y)[:4]
    bbox = (0.0, xmax, 0.0, ymax)
    use = (1, 0, 1, 1)
    mp = mpml_fn(en, host.element_order(type_code), x, y, pml, pml, edge, 0.0, xmax, 0.0, ymax,
                 float(np.max(vp)), use)
    prob = host.build_problem(type_code, en, x, y, rho, c, vp, edge, pml, pml, 1, bbox=bbox, mpml=mp)
    src_y = np.atleast_1d(src_y)
    src = np.array([host.node_location(edge, x, y, xmax - src_depth, sy) for sy in src_y], np.int32)
    ry = np.linspace(rec_y0, rec_y1, rec_num)
    rec = _stations_on_surface(x, y, edge, xmax, ry)
    return dict(problem=prob, src=src, rec=rec, dt=dt, f0=f0, t0=t0, steps_exec=steps_exec,
                xmax=xmax, ymax=ymax)


def config4b(mpml_fn=None, nmodels=16, seed=41, dt=5.0e-4, step_range=(1500, 3000), sizes=None):
    """16 independent Q4 models (dx = 10 m) with sizes drawn from
    {100x100, 200x150, 300x200, 500x300} elements, each with its own seeded
    4-layer isotropic model, dt 0.5 ms, 10 Hz Ricker, executed steps drawn
    U[1500, 3000]; one source at 10 m depth mid-surface, 101 receivers."""
    rng = np.random.default_rng(seed)
    choices = sizes or [(100, 100), (200, 150), (300, 200), (500, 300)]
    out = []
    for m in range(nmodels):
        nex, ney = choices[int(rng.integers(0, len(choices)))]
        H = 10.0 * nex
        mseed = int(rng.integers(0, 2 ** 31))

        def model(depth, y, H=H, mseed=mseed):
            return layered_model(depth, np.random.default_rng(mseed), 4, H, 1500.0, 4500.0, thomsen=False)

        case = structured_case(host.Q4, nex, ney, 10.0, model, 20, dt, 10.0, 0.12, int(step_range[1]), 10.0,
                               5.0 * ney, 101, 0.0, 10.0 * ney, mpml_fn)
        case["nsteps"] = np.array([int(rng.integers(step_range[0], step_range[1] + 1))], np.int32)
        case["size"] = (nex, ney)
        out.append(case)
    return out


def _shape_grad_ref(type_code, r, s):
    """dphi/dr, dphi/ds of the reference element at (r, s) (same node order as
    shape/shape_{t3,t6,q4,q9}.c)."""
    if type_code == host.T3:
        return np.array([-1.0, 1.0, 0.0]), np.array([-1.0, 0.0, 1.0])
    if type_code == host.T6:
        dr = np.array([-3 + 4 * r + 4 * s, -1 + 4 * r, 0.0, 4 - 8 * r - 4 * s, 4 * s, -4 * s])
        ds = np.array([-3 + 4 * r + 4 * s, 0.0, -1 + 4 * s, -4 * r, 4 * r, 4 - 4 * r - 8 * s])
        return dr, ds
    if type_code == host.Q4:
        return np.array([-1 + s, 1 - s, s, -s]), np.array([-1 + r, -r, r, 1 - r])
    r1, rh, s1, sh = r - 1, r - 0.5, s - 1, s - 0.5
    a, b, c = 2 * r - 1.5, 2 * r - 0.5, 2 * r - 1.0
    p, q, u = 2 * s - 1.5, 2 * s - 0.5, 2 * s - 1.0
    dr = np.array([4 * a * s1 * sh, 4 * b * s1 * sh, 4 * b * s * sh, 4 * a * s * sh, -8 * c * s1 * sh,
                   -8 * b * s * s1, -8 * c * s * sh, -8 * a * s * s1, 16 * c * s * s1])
    ds = np.array([4 * r1 * rh * p, 4 * r * rh * p, 4 * r * rh * q, 4 * r1 * rh * q, -8 * r * r1 * p,
                   -8 * r * rh * u, -8 * r * r1 * q, -8 * r1 * rh * u, 16 * r * r1 * u])
    return dr, ds


def explosive_source(type_code, element_node, x, y, xs, ys, moment=1.0):
    """Equivalent nodal forces of an isotropic moment tensor M0*I at (xs, ys):
    f_i = M0 * grad(phi_i)(xs, ys) on the nodes of the element holding the point
    (the weak form of -div(M delta)).  The gradient jumps across element edges,
    so a point on an edge or a vertex takes the mean over the elements sharing
    it (mirror-symmetric meshes then give mirror-symmetric forces).  Returns
    (nodes, weights [k][2]) with the u (x) and w (y) weights.  Not in the
    reference, whose only source is a point force
    (time_evolution/elastic_wave.c:585, 651)."""
    o = host.element_order(type_code)
    en = np.asarray(element_node, np.int64).reshape(-1, o) - 1
    quad = type_code in (host.Q4, host.Q9)
    nv = 4 if quad else 3
    cx, cy = x[en[:, :nv]], y[en[:, :nv]]
    cand = np.flatnonzero((cx.min(1) <= xs + 1e-9) & (cx.max(1) >= xs - 1e-9) &
                          (cy.min(1) <= ys + 1e-9) & (cy.max(1) >= ys - 1e-9))
    acc = {}
    hits = 0
    for e in cand:
        nd = en[e]
        vx, vy = x[nd[:nv]], y[nd[:nv]]
        if not quad:
            J = np.array([[vx[1] - vx[0], vx[2] - vx[0]], [vy[1] - vy[0], vy[2] - vy[0]]])
            r, s = np.linalg.solve(J, [xs - vx[0], ys - vy[0]])
            if r < -1e-12 or s < -1e-12 or r + s > 1 + 1e-12:
                continue
        else:
            r, s = 0.5, 0.5
            for _ in range(30):  # Newton on the bilinear corner map
                N = np.array([(1 - r) * (1 - s), r * (1 - s), r * s, (1 - r) * s])
                Jm = np.array([[-(1 - s), 1 - s, s, -s], [-(1 - r), -r, r, 1 - r]])
                J = np.array([[Jm[0] @ vx, Jm[1] @ vx], [Jm[0] @ vy, Jm[1] @ vy]])
                res = np.array([N @ vx - xs, N @ vy - ys])
                d = np.linalg.solve(J, res)
                r, s = r - d[0], s - d[1]
                if np.max(np.abs(d)) < 1e-14:
                    break
            if r < -1e-9 or s < -1e-9 or r > 1 + 1e-9 or s > 1 + 1e-9:
                continue
            Jm = np.array([[-(1 - s), 1 - s, s, -s], [-(1 - r), -r, r, 1 - r]])
            J = np.array([[Jm[0] @ vx, Jm[1] @ vx], [Jm[0] @ vy, Jm[1] @ vy]])
        dr, ds = _shape_grad_ref(type_code, r, s)
        inv = np.linalg.inv(J)  # [dr/dx dr/dy; ds/dx ds/dy]
        gx = dr * inv[0, 0] + ds * inv[1, 0]
        gy = dr * inv[0, 1] + ds * inv[1, 1]
        for k, node in enumerate(nd):
            a = acc.setdefault(int(node), np.zeros(2))
            a += (gx[k], gy[k])
        hits += 1
    if not hits:
        raise ValueError("source point outside the mesh")
    nodes = np.array(sorted(acc), np.int32)
    return nodes, moment * np.array([acc[int(i)] for i in nodes]) / hits


def topography_t6(nex, ney, edge, hills, rng=None):
    """T6 mesh of nex x ney quads split on the SE-NW diagonal (the reference's
    T6 numbering) whose top (x = xmax side) follows a sum of Gaussian hills
    h(y) = sum a_k exp(-(y - c_k)^2 / (2 w_k^2)).  Vertices are stretched in x
    per column, mid-side nodes are placed at edge midpoints so every element
    stays straight-sided.  Returns element_node (1-based), x, y, surface(y)."""
    H, L = edge * nex, edge * ney
    en, x, y = host.mesh_structured(host.T6, nex, ney, 0.0, H, 0.0, L)
    line = 2 * nex + 1

    def top(yy):
        return H + sum(a * np.exp(-(yy - c) ** 2 / (2.0 * w * w)) for a, c, w in hills)

    X = x.reshape(2 * ney + 1, line).copy()
    Y = y.reshape(2 * ney + 1, line)
    for j in range(0, 2 * ney + 1, 2):             # vertex rows of the lattice
        X[j, 0::2] = X[j, 0::2] * top(Y[j, 0]) / H
    for j in range(0, 2 * ney + 1, 2):             # mid-side nodes along x at vertex rows
        X[j, 1::2] = 0.5 * (X[j, 0:-1:2] + X[j, 2::2])
    for j in range(1, 2 * ney, 2):                 # rows between vertex rows
        X[j, 0::2] = 0.5 * (X[j - 1, 0::2] + X[j + 1, 0::2])
        # odd-odd nodes: midpoint of the SE-NW diagonal of their quad
        X[j, 1::2] = 0.5 * (X[j - 1, 2::2] + X[j + 1, 0:-1:2])
    return en, np.ascontiguousarray(X.reshape(-1)), y, top


def config3(mpml_fn=None, steps_exec=5000, seed=3, nex=200, ney=400, edge=10.0):
    """T6 mesh of a 4 km (horizontal) x 2 km (depth) domain from a 400x200
    quad grid split on the diagonal, free-surface topography of 3 seeded
    Gaussian hills (height 20-120 m, width 200-600 m); vp 2500 -> 4000 m/s
    linear in depth below the local surface with a 5 % Gaussian-correlated
    (80 m) perturbation on vp and vs, Gardner density; M-PML 25; dt 0.2 ms;
    12 Hz Ricker explosive source 50 m below the surface centre; 161
    receivers on the surface."""
    rng = np.random.default_rng(seed)
    L = edge * ney
    hills = [(rng.uniform(20, 120), rng.uniform(0.2 * L, 0.8 * L), rng.uniform(200, 600)) for _ in range(3)]
    en, x, y, top = topography_t6(nex, ney, edge, hills)
    depth = top(y) - x
    vp0 = 2500.0 + 1500.0 * np.clip(depth / (edge * nex), 0.0, 1.0)
    # Gaussian-correlated field on a 10 m grid, sampled at the nodes
    gy, gx = np.arange(0.0, L + edge, edge), np.arange(0.0, x.max() + edge, edge)
    noise = rng.standard_normal((gy.size, gx.size))
    k = np.exp(-0.5 * (np.arange(-24, 25) * edge / 80.0) ** 2)
    k /= k.sum()
    f = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 0, noise)
    f = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 1, f)
    f *= 0.05 / f.std()
    pert = f[np.clip(np.rint(y / edge).astype(int), 0, gy.size - 1), np.clip(np.rint(x / edge).astype(int), 0,
                                                                           gx.size - 1)]
    vp = vp0 * (1.0 + pert)
    vs = vp0 / 1.73 * (1.0 + pert)
    rho = 310.0 * vp ** 0.25
    c = host.elastic_constants(rho, vp, vs)
    xmax = float(x.max())
    bbox = (0.0, xmax, 0.0, L)
    pml = 25
    mp = (mpml_fn or _gpu_mpml)(en, 6, x, y, pml, pml, edge, 0.0, xmax, 0.0, L, float(vp.max()), (1, 0, 1, 1))
    prob = host.build_problem(host.T6, en, x, y, rho, c, vp, edge, pml, pml, 1, bbox=bbox, mpml=mp)
    ys = L / 2
    nodes, wts = explosive_source(host.T6, en, x, y, top(ys) - 50.0, ys)
    ry = np.linspace(0.0, L, 161)
    rec = np.array([host.node_location(edge, x, y, top(v), v) for v in ry], np.int32)
    return dict(problem=prob, src=np.array([nodes[0]], np.int32), sources=[(nodes, wts)], rec=rec, dt=2.0e-4,
                f0=12.0, t0=0.1, steps_exec=steps_exec, xmax=xmax, ymax=L, top=top)


def small_q4(nelem=16, steps=60, edge=10.0, mpml_fn=None, nshots=1, type_code=host.Q4):
    """Small homogeneous case for smoke and parity tests."""
    pml = max(2, nelem // 5)

    def model(depth, y):
        n = depth.size
        return isotropic_model(np.full(n, 2000.0), np.full(n, 3000.0), np.full(n, 1732.0))

    L = edge * nelem
    src_y = np.linspace(0.3 * L, 0.7 * L, nshots)
    # quadratic elements halve the node spacing and triangles stiffen the
    # operator: keep dt well inside each element's stability limit
    dt = {host.Q4: 5.0e-4, host.T3: 2.5e-4, host.Q9: 2.0e-4, host.T6: 1.5e-4}[type_code]
    return structured_case(type_code, nelem, nelem, edge, model, pml, dt, 25.0, 0.04, steps,
                           2 * edge, src_y, min(11, nelem + 1), 0.0, L, mpml_fn)


def config1(mpml_fn=None, steps_exec=1998):
    """Q4 200x200 (dx = 10 m), homogeneous vp 3000 / vs 1732 / rho 2000, M-PML 20,
    free surface on top, dt 0.5 ms, 10 Hz Ricker point force at 20 m depth under
    the surface centre (t0 = 0.12 s), 101 surface receivers every 20 m."""
    def model(depth, y):
        n = depth.size
        return isotropic_model(np.full(n, 2000.0), np.full(n, 3000.0), np.full(n, 1732.0))

    return structured_case(host.Q4, 200, 200, 10.0, model, 20, 5.0e-4, 10.0, 0.12, steps_exec,
                           20.0, 1000.0, 101, 0.0, 2000.0, mpml_fn)


def config2(mpml_fn=None, steps_exec=4000, seed=2, nelem=400):
    """Q9 400x400 on 4 km (801x801 nodes), seeded 6-layer VTI model, M-PML 30,
    dt 0.25 ms, 15 Hz Ricker at the surface centre, 201 surface receivers."""
    rng = np.random.default_rng(seed)
    L = 10.0 * nelem

    def model(depth, y):
        return layered_model(depth, rng, 6, L, 1500.0, 4500.0, thomsen=True)

    return structured_case(host.Q9, nelem, nelem, 10.0, model, 30 * nelem // 400, 2.5e-4, 15.0, 0.08,
                           steps_exec, 10.0, L / 2, 201, 0.0, L, mpml_fn)


def config4a(mpml_fn=None, nshots=128, seed=4, nelem=400, dt=5.0e-4, step_range=(1500, 3000)):
    """128 shots sharing one Q4 400x400 mesh (dx = 10 m) with the layered
    isotropic model of config 2 (eps = delta = 0); sources every 30 m along
    the surface, executed steps drawn U[1500, 3000] per shot."""
    rng = np.random.default_rng(seed)
    L = 10.0 * nelem

    def model(depth, y):
        return layered_model(depth, np.random.default_rng(2), 6, L, 1500.0, 4500.0, thomsen=False)

    y0 = L / 2 - 30.0 * (nshots - 1) / 2
    src_y = y0 + 30.0 * np.arange(nshots)
    case = structured_case(host.Q4, nelem, nelem, 10.0, model, 30 * nelem // 400, dt, 10.0, 0.12,
                           int(step_range[1]), 10.0, src_y, 201, 0.0, L, mpml_fn)
    case["nsteps"] = rng.integers(step_range[0], step_range[1] + 1, nshots).astype(np.int32)
    case["steps_exec"] = int(case["nsteps"].max())
    return case
