"""dcg_render_host_noalloc (the render staged in the library's file-scope device array, no device
allocation) gives, through the C ABI, exactly the image dcg_render_host gives: two calls in a row,
the second with a smaller image, on an adaptive and a uniform grid; an image beyond the
1024 x 1024-pixel array is rejected with DCG_E_ARG and leaves the handle usable."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dcgrid_amd  # noqa: E402
from dcgrid_amd import scenes  # noqa: E402
from dcgrid_amd._abi import DCGError  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(kind):
    w = scenes.config(1, scale=4) if kind == "dcgrid" else scenes.config(0, scale=2)
    s = dcgrid_amd.Simulation(w.kind, w.size, w.params, w.max_num_blocks if kind == "dcgrid" else None)
    s.seed_scene()
    s.step(3)
    return s


@pytest.mark.parametrize("kind", ["dcgrid", "uniform"])
def test_noalloc_render_equals_render_host(kind):
    sim = make(kind)
    say(f"[case] noalloc render {kind}")
    r = sim.default_render_params()
    r.width, r.height = 320, 240
    a = sim.render(r)
    b = sim.render_noalloc(r)
    assert a.shape == b.shape == (240, 320, 4)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.isfinite(b).all() and b.max() > 0
    # a second, smaller call: the array keeps the last call's contents, the image is rewritten
    r2 = sim.default_render_params()
    r2.width, r2.height = 96, 64
    a2 = sim.render(r2)
    b2 = sim.render_noalloc(r2)
    assert np.array_equal(a2.view(np.uint32), b2.view(np.uint32))
    # beyond the array: an error, not a fault, and the handle still renders
    big = sim.default_render_params()
    big.width, big.height = 2048, 1024
    with pytest.raises(DCGError) as ex:
        sim.render_noalloc(big)
    assert ex.value.code == -1
    assert np.array_equal(sim.render_noalloc(r2).view(np.uint32), a2.view(np.uint32))
    sim.close()
