"""HIP-graph replay of dcg_step: pairs of steps replayed from one capture give the same bits as
the eager launches, replay after replay, with host work (exports) and synchronisation between
them; the capture is redone when a single-stage call swaps field buffers or the parameters
change."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import dcgrid_amd  # noqa: E402
from dcgrid_amd import scenes  # noqa: E402
from parity_util import TABLE_KEYS, compare_tables  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(name):
    if name == "cfg1_64":
        w = scenes.config(1, scale=4)
        sim = dcgrid_amd.Simulation(w.kind, w.size, w.params, w.max_num_blocks)
        sim.seed_scene()
        return sim
    if name == "ref64_bw4":
        p = scenes.reference_params(64)
        return dcgrid_amd.Simulation("dcgrid", (64, 64, 64), p, 1500)
    w = scenes.config(0, scale=2)
    sim = dcgrid_amd.Simulation(w.kind, w.size, w.params)
    sim.seed_scene()
    return sim


def state(sim):
    out = {f: sim.export_field(f) for f in ("density", "velocity", "pressure", "temperature")}
    if sim.kind == "dcgrid":
        t = sim.export_tables()
        inf = sim.info()
        m = np.repeat(t["block_levels"] != 0xFF, inf.block_width ** 3)
        out = {k: (v[m] if v.ndim == 1 else v[m]) for k, v in out.items()}
        out["tables"] = t
    out["cell_steps"] = sim.cell_steps()
    return out


def same(a, b):
    for k in a:
        if k == "tables":
            assert compare_tables(a[k], b[k], TABLE_KEYS) == []
        elif k == "cell_steps":
            assert a[k] == b[k]
        else:
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


@pytest.mark.parametrize("name", ["cfg1_64", "ref64_bw4", "uniform32"])
def test_graph_replay_matches_eager(name):
    g, e = make(name), make(name)
    g.set_graph(True)
    for rep in range(4):
        say(f"[case] {name} replay {rep}")
        g.step(2)  # one replay of the captured pair (captured on the first call)
        e.step(2)
        g.synchronize()
        same(state(g), state(e))
    # an odd count: pairs replayed, the last step eager
    g.step(3)
    e.step(3)
    same(state(g), state(e))
    # single-stage calls swap buffers: the next step() must capture again
    g.advect_velocity()
    e.advect_velocity()
    g.step(4)
    e.step(4)
    same(state(g), state(e))
    # parameter change: recapture
    p = g.params.copy()
    p.dt = p.dt * 0.5
    g.set_params(p)
    e.set_params(p)
    g.step(2)
    e.step(2)
    same(state(g), state(e))
    g.close()
    e.close()


def test_graph_after_imported_fluidity():
    """An imported fluidity field on the uniform grid: the first step after the import rebuilds
    the fluidity pyramid on the host's order (host state a captured pass would not replay), so
    that step runs eagerly and the capture follows; replay after replay equals the eager run."""
    g, e = make("uniform32"), make("uniform32")
    n = g.field_size("fluidity")
    rng = np.random.default_rng(5)
    flu = (0.5 + 0.5 * rng.random(n)).astype(np.float32)
    for s in (g, e):
        s.import_field("fluidity", flu)
    g.set_graph(True)
    for rep in range(3):
        say(f"[case] imported fluidity replay {rep}")
        g.step(2)
        e.step(2)
        g.synchronize()
        same(state(g), state(e))
        assert np.array_equal(g.export_field("fluidity").view(np.uint32), e.export_field("fluidity").view(np.uint32))
    g.close()
    e.close()
