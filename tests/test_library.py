"""The C-ABI library loads on the CPU and exports every entry point declared in
include/dcgrid_amd.h; defaults restate SimParams::defaultParams. No compute without a GPU."""
import ctypes as C
import os
import re

import pytest

import dcgrid_amd
from dcgrid_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "dcgrid_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(dcg_[a-z0-9_]+)\s*\(", hdr)))


def test_exports_every_declared_symbol():
    lib = dcgrid_amd.library()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_abi_version_and_defaults():
    lib = dcgrid_amd.library()
    assert lib.dcg_abi_version() == 2
    p = dcgrid_amd.default_params()
    # src/data/sim_params.cpp:6-17; rdx = 1/dx (src/main.cpp:17)
    assert (p.gx, p.gy, p.gz) == (128, 128, 128)
    assert p.dx == C.c_float(10000.0 / 128).value and p.dt == 3.0
    assert p.rdx == C.c_float(1.0 / C.c_float(10000.0 / 128).value).value
    assert p.velocity_emission_rate == 150.0 and p.emission_radius == 750.0
    assert p.density_emission_rate == C.c_float(0.002).value
    assert p.enable_additional_solids == 0
    assert (p.project_top_pairs, p.project_level_pairs, p.project_local_pairs) == (5, 5, 10)
    assert p.refine_threshold == C.c_float(1e-4).value
    q = scenes.reference_params(128)
    for f in ("dt", "gx", "dx", "rdx", "velocity_emission_rate", "density_emission_rate", "emission_radius"):
        assert getattr(p, f) == getattr(q, f)


def test_struct_layout_matches_header(tmp_path):
    """The ctypes mirror has the C compiler's layout of every public struct field."""
    from dcgrid_amd._abi import Info, Params, RenderParams, Tables
    structs = {"dcg_params": Params, "dcg_info": Info, "dcg_tables": Tables, "dcg_render_params": RenderParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dcgrid_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    import subprocess
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs.items():
        assert got[(cname, "sizeof")] == C.sizeof(py)
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_null_handle_is_an_error_not_a_crash():
    lib = dcgrid_amd.library()
    lib.dcg_step.argtypes = [C.c_void_p, C.c_int32]
    lib.dcg_step.restype = C.c_int
    assert lib.dcg_step(None, 1) == -1  # DCG_E_ARG


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_fails_loudly():
    with pytest.raises(dcgrid_amd.DCGError) as e:
        dcgrid_amd.Simulation("uniform", (16, 16, 16), scenes.config(0, scale=4).params)
    assert e.value.code == -3  # DCG_E_HIP


@pytest.mark.parametrize("size,bw,levels,budget,code", [
    ((96, 96, 96), 8, 4, 4000, -4),    # not a multiple of the 64-cell top block: DCG_E_UNSUPPORTED
    ((96, 64, 64), 4, 0, 4000, -4),    # the reference's level rule gives 5 levels; 96 is not a multiple of 64
    ((64, 64, 64), 8, 3, 7, -2),       # fewer blocks than the dense top level: DCG_E_BUDGET
    ((64, 64, 64), 8, 3, 8, -2),       # the top level takes the whole budget, level 0 gets none
])
def test_planner_deviations_are_reported(size, bw, levels, budget, code):
    """Grids this build does not plan (documented deviations from fluid_simulation_dcgrid.cu:29-53,
    which sizes partial coarse blocks with iDivUp) fail at creation with a status code, before
    any device call (so also without a GPU), and the oracle's planner agrees."""
    import oracle
    p = scenes.reference_params(size[0], size[1], size[2])
    p.block_width = bw
    p.max_levels = levels
    with pytest.raises(dcgrid_amd.DCGError) as e:
        dcgrid_amd.Simulation("dcgrid", size, p, budget)
    assert e.value.code == code
    with pytest.raises(dcgrid_amd.DCGError) as e2:
        oracle.plan(size, budget, p)
    assert e2.value.code == code
