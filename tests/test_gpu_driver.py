"""The drop-in driver (par.txt -> elastic_wave() -> output files) on the GPU:
example 6's run directory produces the shipped files byte for byte."""
import gzip
import os
import shutil
import tempfile

import pytest

import seisfem_amd

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "ex6")


def _unpack(dst):
    for sub in ("", "mesh"):
        os.makedirs(os.path.join(dst, sub), exist_ok=True)
        for name in os.listdir(os.path.join(GOLD, sub) if sub else GOLD):
            src = os.path.join(GOLD, sub, name)
            if name.endswith(".gz") and os.path.isfile(src):
                with gzip.open(src, "rb") as fi, open(os.path.join(dst, sub, name[:-3]), "wb") as fo:
                    shutil.copyfileobj(fi, fo)
    os.makedirs(os.path.join(dst, "outputfile"), exist_ok=True)


def test_example6_driver_files_identical():
    print("[gpu-case] seisfem_run on example 6 (T3 exterior mesh, 1198 steps, FP64)", flush=True)
    L = seisfem_amd.lib()
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        _unpack(d)
        try:
            L.sfh_driver_options(0, 1)
            rc = L.seisfem_run(d.encode())
        finally:
            os.chdir(cwd)
        assert rc == 0, seisfem_amd.last_error()
        for name in ("seismogram_u_shot_1.txt", "seismogram_w_shot_1.txt", "wavefield_u_shot_1.txt",
                     "wavefield_w_shot_1.txt", "receiver_station.txt", "source_location.txt"):
            with gzip.open(os.path.join(GOLD, "outputfile", name + ".gz"), "rb") as f:
                want = f.read()
            with open(os.path.join(d, "outputfile", name), "rb") as f:
                got = f.read()
            assert got == want, name
        assert os.path.getsize(os.path.join(d, "outputfile", "seismogram_uw_shot_1.f64")) == 2 * 1198 * 81 * 8
