"""The drop-in boundary without a GPU: the C-ABI library loads and exports
every entry point that include/o3d.h declares; host-side configuration and
checkpoint formats round-trip."""
from __future__ import annotations

import ctypes
import math
import os
import re

import numpy as np
import pytest

from osinco3d_amd import capi
from osinco3d_amd.config import Config, mixing_layer_config, tgv_config
from osinco3d_amd.fields_io import read_fields, write_fields

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(o3d_\w+)\s*\(", text)) - {"o3d_params", "o3d_step_info"})


def test_library_exports_every_declared_function():
    lib = capi.load()
    declared = header_functions(os.path.join(ROOT, "include", "o3d.h"))
    assert declared, "no declarations parsed"
    missing = [f for f in declared if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(capi.EXPORTS) == declared


def test_nn_header_functions_exported():
    lib = capi.load()
    p = os.path.join(ROOT, "include", "o3d_nn.h")
    if not os.path.exists(p):
        pytest.skip("no o3d_nn.h")
    missing = [f for f in header_functions(p) if not hasattr(lib, f)]
    assert not missing, missing


def test_version_string():
    assert capi.load().o3d_version().decode().startswith("osinco3d-amd")


def test_runtime_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    rc = capi.load().o3d_init(0)
    assert rc < 0 and "device" in capi.last_error().lower()


def test_struct_layouts_match_header(tmp_path):
    """Every field of o3d_params and o3d_step_info at the same offset and the
    same total size in the ctypes mirror as in the C compiler's layout of
    include/o3d.h (a small C program prints sizeof and offsetof)."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = {"o3d_params": capi.O3DParams, "o3d_step_info": capi.O3DStepInfo}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "o3d.h"', "int main(void) {"]
    for sname, cls in structs.items():
        lines.append(f'  printf("{sname} size %zu\\n", sizeof({sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{sname} {fname} %zu\\n", offsetof({sname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        sname, key, val = line.split()
        got[(sname, key)] = int(val)
    for sname, cls in structs.items():
        assert got[(sname, "size")] == ctypes.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert got[(sname, fname)] == getattr(cls, fname).offset, (sname, fname)


def test_namelist_round_trip(tmp_path):
    cfg = mixing_layer_config()
    p = tmp_path / "parameters.o3d"
    cfg.write(str(p))
    back = Config.read(str(p))
    assert back == cfg
    text = p.read_text()
    # Groups in the order the reference reads them.
    order = [m.group(1) for m in re.finditer(r"^&(\w+)", text, re.M)]
    assert order == ["FlowParam", "Domain", "BoundaryConditions", "AdvanceTime", "PoissonEq",
                     "Scalar", "VisuParameters", "InitInflow", "LES"]


def test_reads_reference_style_parameter_file():
    text = """&FlowParam
typesim = 2  !> comment
re = 1600.
/End
&Domain
nx = 185
xlx = 3.141592653589793
/End
&PoissonEq
eps = 1.e-4
/End
"""
    cfg = Config.from_namelist(text)
    assert cfg.nx == 185 and cfg.re == 1600.0 and cfg.eps == 1e-4 and cfg.typesim == 2


def test_grid_constants():
    cfg = tgv_config(64, 100, 2)
    dx = cfg.spacing()[0]
    assert math.isclose(dx, 2 * math.pi / 64, rel_tol=1e-15)       # exactly periodic
    assert math.isclose(cfg.dt, 0.25 * dx, rel_tol=1e-15)
    assert cfg.poisson_kind() == "0000"
    assert mixing_layer_config().poisson_kind() == "0011"
    c = Config(nbcx1=1, nbcxn=1, nbcy1=0, nbcyn=0)
    assert c.poisson_kind() == ""                                  # no solver bound
    with pytest.raises(ValueError):
        Config(nbcx1=0, nbcxn=1).bc()


def test_viscous_dt_override():
    cfg = Config(nx=16, ny=16, nz=16, xlx=1.0, yly=1.0, zlz=1.0, re=0.01, dt=0.1)
    dmin = 1.0 / 15
    assert math.isclose(cfg.time_step_overrides(), 0.1 * dmin * dmin * 0.01)


def test_fields_bin_round_trip(tmp_path):
    cfg = Config(nx=7, ny=6, nz=5, xlx=1.0, yly=2.0, zlz=3.0, y0=-1.0)
    rng = np.random.default_rng(0)
    f = [rng.standard_normal((5, 6, 7)) for _ in range(5)]
    p = tmp_path / "fields.bin"
    write_fields(str(p), cfg, *f, time=1.25)
    assert p.stat().st_size == 8 + 12 + 8 * (7 + 6 + 5) + 5 * 8 * 210
    t, d = read_fields(str(p), cfg)
    assert t == 1.25
    for a, k in zip(f, ("ux", "uy", "uz", "pp", "phi")):
        assert np.array_equal(a, d[k])
    assert np.allclose(d["y"], -1.0 + np.arange(6) * 0.4)
    with pytest.raises(ValueError):
        read_fields(str(p), Config(nx=8, ny=6, nz=5))


def test_bench_runs_on_one_gpu_only():
    """bench.py's contract on the command line: it has no multi-GPU mode, so
    --gpus above 1 (or a launcher's WORLD_SIZE above 1) ends with a message
    and exit status 2 before touching a device; --help lists the flags the
    harness passes."""
    import subprocess
    import sys
    bench = os.path.join(ROOT, "bench.py")
    r = subprocess.run([sys.executable, bench, "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr, (r.returncode, r.stderr)
    env = dict(os.environ, WORLD_SIZE="4")
    r = subprocess.run([sys.executable, bench], capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode == 2 and "WORLD_SIZE 4" in r.stderr, (r.returncode, r.stderr)
    r = subprocess.run([sys.executable, bench, "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--workload", "--forms"):
        assert flag in r.stdout, flag
