"""The C-ABI library loads and exports every function include/*.h declares
(no GPU calls)."""
import ctypes
import os
import re

import seisfem_amd
from seisfem_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("seisfem_gpu.h", "seisfem_host.h", "seisfem_surrogate.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_native.LIB_PATH)
    names = declared_functions()
    assert {"csr_matvec", "abc_mpml", "seismic_source", "sfg_run_create", "sfh_coo2csr",
            "__coo2csr_lib_MOD_coo2csr_canonical", "sfg_run_snapshot_noalloc"} <= names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    """Every declared function has a typed ctypes binding in seisfem_amd._native."""
    L = seisfem_amd.lib()
    untyped = [n for n in sorted(declared_functions()) if getattr(L, n).argtypes is None]
    assert not untyped, untyped


def test_host_seismic_source_matches_formula():
    L = seisfem_amd.lib()
    import math
    f0, t0, t = 10.0, 0.1, 0.137
    a = (math.pi * f0 * (t - t0)) ** 2
    assert abs(L.seismic_source(f0, t0, 1e10, t) - 1e10 * (1 - 2 * a) * math.exp(-a)) < 1e-3


def test_stream_copy_rejects_bad_sizes_before_touching_the_device():
    import pytest
    from seisfem_amd import NativeError, gpu
    for nbytes, reps in ((15, 1), (0, 1), (1024, 0)):
        with pytest.raises(NativeError) as e:
            gpu.stream_copy(nbytes, reps)
        assert e.value.code == -1


def test_bench_refuses_more_than_one_gpu_without_touching_it():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_headers_compile_as_c99_and_link(tmp_path):
    """A C caller (the reference's seisfem.c is C) includes the three headers
    with -std=c99 -Wall -Werror, links the library and calls a host-only entry
    point (no device is touched)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    assert cc, "no C compiler"
    src = tmp_path / "caller.c"
    src.write_text(
        '#include "seisfem_gpu.h"\n#include "seisfem_host.h"\n#include "seisfem_surrogate.h"\n'
        "#include <stdio.h>\n"
        "int main(void) {\n"
        "    sfg_shot_desc d = {0};\n"
        "    d.precision = SFG_FP32_KAHAN16;\n"
        '    printf("%.6f %d\\n", seismic_source(10.0, 0.1, 1.0, 0.1), d.precision);\n'
        "    return 0;\n}\n")
    exe = tmp_path / "caller"
    libdir = os.path.join(ROOT, "seisfem_amd")
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe), "-L", libdir, "-lseisfem_amd", "-Wl,-rpath," + libdir], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert float(out[0]) == 1.0 and int(out[1]) == 3


def test_shot_descriptor_binding_matches_the_c_layout(tmp_path):
    """The ctypes mirror of sfg_shot_desc has the C struct's size and every
    field at the C offset (a drift would hand the library wrong launch fields)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    assert cc, "no C compiler"
    names = [f[0] for f in _native.ShotDesc._fields_]
    body = "".join(f'    printf("{n} %zu\\n", offsetof(sfg_shot_desc, {n}));\n' for n in names)
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "seisfem_gpu.h"\nint main(void) {\n'
                   '    printf("sizeof %zu\\n", sizeof(sfg_shot_desc));\n' + body + "    return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    assert int(got["sizeof"]) == ctypes.sizeof(_native.ShotDesc)
    for n in names:
        assert int(got[n]) == getattr(_native.ShotDesc, n).offset, n
