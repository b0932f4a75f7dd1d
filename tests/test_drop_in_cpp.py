"""The C++ drop-in header (include/dcgrid_amd_fluid_simulation.hpp) compiles with g++ against the
C ABI, links libdcgrid_amd.so and drives the step the way src/simulation.cpp:93-115 does.
Without a GPU the constructor throws dcgrid_amd::Error with DCG_E_HIP; on the GPU the debugStats
value equals that of the same step sequence through the Python binding."""
import os
import subprocess

import numpy as np
import pytest

import dcgrid_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "dcgrid_amd")


def build_program(tmp_path):
    dcgrid_amd.library()  # raises when the library has not been built
    exe = tmp_path / "drop_in_main"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "drop_in_main.cpp"), "-o", str(exe),
                    "-L", LIBDIR, "-ldcgrid_amd", "-Wl,-rpath," + LIBDIR], check=True)
    return str(exe)


def test_header_compiles_and_links(tmp_path):
    exe = build_program(tmp_path)
    assert os.path.exists(exe)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_constructor_throws_without_gpu(tmp_path):
    exe = build_program(tmp_path)
    r = subprocess.run([exe, "dcgrid", "64", "64", "64", "256", "1", "4", "3"], capture_output=True, text=True)
    assert r.returncode == 3
    assert r.stdout.startswith("error -3 FluidSimulationDCGrid:")


@pytest.mark.gpu
@pytest.mark.parametrize("kind,max_blocks,bw,levels", [("dcgrid", 400, 4, 3), ("uniform", 0, 4, 3)])
def test_cpp_drop_in_matches_python_binding(tmp_path, kind, max_blocks, bw, levels):
    exe = build_program(tmp_path)
    n, steps = 64, 3
    r = subprocess.run([exe, kind, str(n), str(n), str(n), str(max_blocks), str(steps), str(bw), str(levels)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.strip().split("\n")
    cells = int(lines[0].split()[1])
    label, value = lines[1].split(":")
    assert label == ("Residual" if kind == "dcgrid" else "Smoke")

    p = dcgrid_amd.default_params()
    p.gx = p.gy = p.gz = n
    # the program computes dx and rdx in float arithmetic
    p.dx = float(np.float32(10000.0) / np.float32(n))
    p.rdx = float(np.float32(1.0) / np.float32(p.dx))
    p.block_width, p.max_levels = bw, levels
    sim = dcgrid_amd.Simulation(kind, (n, n, n), p, max_blocks if kind == "dcgrid" else None)
    sim.init()
    for _ in range(steps):
        sim.advect_velocity()
        sim.adapt_topology()
        sim.project()
        sim.advect_density()
    assert sim.info().num_cells == cells
    # printf("%g") keeps six significant digits
    expect = sim.debug_stats()
    assert float(value) == pytest.approx(expect, rel=1e-5, abs=1e-30)
