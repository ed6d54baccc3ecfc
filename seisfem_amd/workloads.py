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


def small_q4(nelem=16, steps=60, edge=10.0, mpml_fn=None, nshots=1, type_code=host.Q4):
    """Small homogeneous case for smoke and parity tests."""
    pml = max(2, nelem // 5)

    def model(depth, y):
        n = depth.size
        return isotropic_model(np.full(n, 2000.0), np.full(n, 3000.0), np.full(n, 1732.0))

    L = edge * nelem
    src_y = np.linspace(0.3 * L, 0.7 * L, nshots)
    return structured_case(type_code, nelem, nelem, edge, model, pml, 5.0e-4, 25.0, 0.04, steps,
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
