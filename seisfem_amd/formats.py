"""The reference's on-disk formats.

  par.txt                     fixed-format run parameters (seisfem/seisfem.c:121-148)
  mesh/element_and_node_num.txt, mesh/element_node.txt (1-based), mesh/node_xy.txt
                              exterior mesh (seisfem.c:157-214)
  mesh/velocity_and_density.txt
                              rho vp vs c11 c13 c33 c44 per node (elastic_wave.c:198-209)
  outputfile/seismogram_{u,w}_shot_N.txt, wavefield_{u,w}_shot_N.txt
                              "%f " text rows (elastic_wave.c:503-507, 787-806)

Readers accept plain or .gz files.
"""
import gzip
import io
import os
import re

import numpy as np

PAR_KEYS = [
    ("use_exterior_mesh", int), ("type_code", int), ("nelemx", int), ("nelemy", int),
    ("edge_size", float), ("f0", float), ("t0", float), ("src_num", int),
    ("src_x_first", float), ("src_y_first", float), ("src_x_last", float), ("src_y_last", float),
    ("rec_num", int), ("rec_x_first", float), ("rec_y_first", float), ("rec_x_last", float),
    ("rec_y_last", float), ("dt", float), ("step", int), ("solver_code", int), ("free_surface", int),
]


def _open(path):
    if os.path.exists(path):
        return gzip.open(path, "rt") if path.endswith(".gz") else open(path)
    if os.path.exists(path + ".gz"):
        return gzip.open(path + ".gz", "rt")
    raise FileNotFoundError(path)


def read_par(path):
    """par.txt -> dict.  The keys are read in the reference's fixed order;
    missing keys raise (the reference's fscanf would leave them at 0)."""
    with _open(path) as f:
        text = f.read()
    out = {}
    for key, typ in PAR_KEYS:
        m = re.search(r"^\s*" + key + r"\s*=\s*([-+0-9.eE]+)", text, re.M)
        if not m:
            raise ValueError(f"{path}: missing '{key}'")
        out[key] = typ(float(m.group(1))) if typ is int else typ(m.group(1))
    return out


def read_exterior_mesh(mesh_dir, element_order):
    """-> element_node (1-based int32, element-major), x, y, bbox.

    The bounding box starts at 0 in every direction, as seisfem.c:199-214 does."""
    with _open(os.path.join(mesh_dir, "element_and_node_num.txt")) as f:
        text = f.read()
    ne = int(re.search(r"element_num\s*=\s*(\d+)", text).group(1))
    nn = int(re.search(r"node_num\s*=\s*(\d+)", text).group(1))
    with _open(os.path.join(mesh_dir, "element_node.txt")) as f:
        en = np.loadtxt(f, dtype=np.int64, max_rows=ne).astype(np.int32).reshape(-1)
    if en.size != ne * element_order:
        raise ValueError("element_node.txt does not match element_num * order")
    with _open(os.path.join(mesh_dir, "node_xy.txt")) as f:
        xy = np.loadtxt(f, dtype=np.float64, max_rows=nn)
    x = np.ascontiguousarray(xy[:, 0])
    y = np.ascontiguousarray(xy[:, 1])
    bbox = (min(0.0, float(x.min())), max(0.0, float(x.max())), min(0.0, float(y.min())),
            max(0.0, float(y.max())))
    return en, x, y, bbox


def read_velocity(path, node_num):
    """velocity_and_density.txt -> rho, vp, vs, [c11, c13, c33, c44]."""
    with _open(path) as f:
        a = np.loadtxt(f, dtype=np.float64, max_rows=node_num)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy(), [a[:, k].copy() for k in range(3, 7)]


SEIS_SEP = "   "   # seismogram rows: "%f   " per value (elastic_wave.c:791-792)
SNAP_SEP = "\t"    # wavefield rows: "%f\t" per value (elastic_wave.c:503-504, 801-802)


def format_rows(rows, sep=SEIS_SEP):
    """Text exactly as the reference writes it: each value "%f" + sep, then a newline."""
    buf = io.StringIO()
    fmt = "%f" + sep
    for r in np.atleast_2d(rows):
        buf.write("".join(fmt % v for v in r))
        buf.write("\n")
    return buf.getvalue()


def read_rows(path):
    """Rows of a "%f " text file as float64 (values only; formatting is compared as text)."""
    with _open(path) as f:
        return np.array([[float(t) for t in line.split()] for line in f if line.strip()])


def read_text(path):
    with _open(path) as f:
        return f.read()
