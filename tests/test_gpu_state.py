"""Versioned state (dcg_save_state / dcg_load_state, the DCGSTATE format) on the GPU.

Bar: a run resumed from a saved state continues bit for bit as the run it was saved from
(tables, hash map and fields after further steps, including moves and refinements); the file
read with numpy (dcgrid_amd.state) holds exactly what the C ABI exports; loading replaces every
part of a different prior state, restores the parameters, and rejects states of another grid,
another version, a bad magic or a truncated buffer without changing the simulation."""
import struct
import sys

import numpy as np
import pytest

import dcgrid_amd
from dcgrid_amd import scenes, state
from dcgrid_amd._abi import STORAGE_BF16, STORAGE_FP16, hash_map
from parity_util import TABLE_KEYS, compare_tables

pytestmark = pytest.mark.gpu

DC_FIELDS = ("density", "velocity", "temperature", "pressure", "t_pressure", "divergence", "fluidity", "vorticity")
U_FIELDS = ("density", "velocity", "temperature", "pressure", "t_pressure", "divergence", "fluidity")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def workload(name):
    if name == "ref64_bw4":
        p = scenes.reference_params(64)
        return "dcgrid", (64, 64, 64), p, 1500, False
    if name in ("cfg1_64", "cfg1_64_fp16", "cfg1_64_bf16"):
        w = scenes.config(1, scale=4)
        p = w.params.copy()
        if name.endswith("fp16"):
            p.storage = STORAGE_FP16
        if name.endswith("bf16"):
            p.storage = STORAGE_BF16
        return w.kind, w.size, p, w.max_num_blocks, True
    if name == "cfg4_128":
        w = scenes.config(4, scale=8)
        return w.kind, w.size, w.params, w.max_num_blocks, True
    if name == "cfg4_128_sparse":
        w = scenes.config(4, scale=8)
        w.params.score_mode = scenes.SCORE_GEOMETRIC
        w.params.refine_threshold = 1e-4
        return w.kind, w.size, w.params, 16 + 128 + 256, True
    if name == "cfg0_32":
        w = scenes.config(0, scale=2)
        return w.kind, w.size, w.params, None, True
    raise KeyError(name)


def make(name, seed_steps=0):
    kind, size, p, b, seed = workload(name)
    s = dcgrid_amd.Simulation(kind, size, p, b)
    if seed:
        s.seed_scene()
    if seed_steps:
        s.step(seed_steps)
    return s


def fields_of(s):
    return DC_FIELDS + (("vapor",) if s.params.moist else ()) if s.kind == "dcgrid" else U_FIELDS


def mask_of(s):
    if s.kind != "dcgrid":
        return None
    inf = s.info()
    return np.repeat(s.export_tables()["block_levels"] != 0xFF, inf.block_width ** 3)


def same_field(a, b, m):
    if m is not None:
        if a.size == 3 * m.size:
            a, b = a.reshape(-1, 3)[m], b.reshape(-1, 3)[m]
        else:
            a, b = a[m], b[m]
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def assert_same_state(a, b, where):
    if a.kind == "dcgrid":
        ta, tb = a.export_tables(), b.export_tables()
        assert compare_tables(ta, tb, TABLE_KEYS) == [], where
        ia, ib = a.info(), b.info()
        assert hash_map(ta, ia) == hash_map(tb, ib), where
        assert list(ia.block_loads) == list(ib.block_loads) and list(ia.move_limit) == list(ib.move_limit), where
    m = mask_of(a)
    for f in fields_of(a):
        assert same_field(a.export_field(f), b.export_field(f), m), (where, f)


@pytest.mark.parametrize("name", ["ref64_bw4", "cfg1_64", "cfg1_64_fp16", "cfg1_64_bf16", "cfg4_128",
                                  "cfg4_128_sparse", "cfg0_32"])
def test_resumed_run_continues_bitwise(name, tmp_path):
    say(f"[case] {name}")
    a = make(name, seed_steps=3)
    p = a.params.copy()
    p.dt = p.dt * 0.75  # parameters are part of the state
    a.set_params(p)
    a.step(1)
    path = tmp_path / "a.dcgstate"
    n = a.save_state(str(path))
    assert path.stat().st_size == n == a.state_size()
    # a different prior state (other steps, other dt) is replaced entirely
    b = make(name, seed_steps=1)
    b.load_state(str(path))
    assert abs(b.params.dt - p.dt) == 0.0
    assert_same_state(a, b, (name, "after load"))
    for k in range(4):
        a.step(1)
        b.step(1)
        assert_same_state(a, b, (name, "step", k))
    # saving again gives the same bytes for the same state
    pa, pb = tmp_path / "a2", tmp_path / "b2"
    a.save_state(str(pa))
    b.save_state(str(pb))
    assert pa.read_bytes() == pb.read_bytes()
    a.close()
    b.close()


@pytest.mark.parametrize("name", ["cfg1_64", "cfg1_64_fp16", "cfg0_32"])
def test_state_file_holds_the_exported_state(name, tmp_path):
    s = make(name, seed_steps=3)
    path = tmp_path / "s.dcgstate"
    s.save_state(str(path))
    st = state.read(str(path))
    inf = s.info()
    assert st.kind == inf.kind and st.size == tuple(s.size) and st.num_cells == inf.num_cells
    assert st.storage == s.params.storage and st.levels == inf.levels
    assert bytes(st.params) == bytes(s.params)
    for f in fields_of(s):
        got = st.field(f).ravel()
        ref = s.export_field(f).ravel()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f
    if s.kind == "dcgrid":
        t = s.export_tables()
        nf32 = np.uint32(0xFFFFFFFF)
        assert np.array_equal(st.table("positions").ravel(), t["block_positions"].ravel())
        assert np.array_equal(st.table("levels"), t["block_levels"])
        assert np.array_equal(st.table("flags"), t["block_flags"].astype(np.uint32))
        par = st.table("parents")
        assert np.array_equal(np.where(par == nf32, np.uint64(2 ** 64 - 1), par.astype(np.uint64)),
                              t["parent_indices"])
        chi = st.table("children")
        assert np.array_equal(np.where(chi == nf32, np.uint64(2 ** 64 - 1), chi.astype(np.uint64)),
                              t["child_indices"])
        assert list(st.table("control")[:inf.levels]) == list(inf.block_loads)[:inf.levels]
        assert st.table("neighbours").shape == (inf.max_num_blocks, 27, 4)
    # the numpy writer reproduces the library's bytes
    out = tmp_path / "rewritten"
    state.write(str(out), st)
    assert out.read_bytes() == path.read_bytes()
    s.close()


def test_load_rejects_foreign_or_damaged_states(tmp_path):
    a = make("cfg1_64", seed_steps=2)
    path = tmp_path / "a.dcgstate"
    a.save_state(str(path))
    raw = path.read_bytes()
    b = make("ref64_bw4", seed_steps=1)  # another grid
    before = b.export_field("density")
    with pytest.raises(dcgrid_amd.DCGError) as e:
        b.load_state(str(path))
    assert e.value.code == -1  # DCG_E_ARG
    assert np.array_equal(b.export_field("density").view(np.uint32), before.view(np.uint32))
    c = make("cfg1_64", seed_steps=1)
    before = c.export_field("velocity")
    cases = {
        "magic": b"XCGSTATE" + raw[8:],
        "version": raw[:8] + struct.pack("<I", 99) + raw[12:],
        "truncated": raw[:len(raw) - 4096],
        "header": raw[:40],
    }
    for what, blob in cases.items():
        p = tmp_path / what
        p.write_bytes(blob)
        with pytest.raises(dcgrid_amd.DCGError) as e:
            c.load_state(str(p))
        assert e.value.code == (-4 if what == "version" else -1), what  # DCG_E_UNSUPPORTED / DCG_E_ARG
        assert np.array_equal(c.export_field("velocity").view(np.uint32), before.view(np.uint32)), what
    a.close()
    b.close()
    c.close()


@pytest.mark.parametrize("name", ["cfg4_128_sparse", "cfg1_64", "ref64_bw4"])
def test_oracle_continues_a_gpu_written_state(name, tmp_path):
    """The oracle leg of the format: a state the GPU wrote is read with numpy alone, its block
    tables rebuilt in the reference's form (dcgrid_amd.state.reference_tables, equal to what the
    C ABI exports), loaded into the CPU oracle with its fields, and the oracle then steps bit for
    bit with the GPU resumed from the same file (tables, hash map and fields)."""
    import oracle
    kind, size, p, b, seed = workload(name)
    say(f"[case] oracle resumes {name}")
    g = make(name, seed_steps=3)
    path = tmp_path / "s"
    g.save_state(str(path))
    st = state.read(str(path))
    t = state.reference_tables(st)
    assert compare_tables(t, g.export_tables(), TABLE_KEYS + ("hash_key", "hash_val")) == []
    c = oracle.OracleSim(kind, size, st.params, b, precision=32)
    loads, limits = state.control_words(st)
    c.import_tables(t, loads, limits)
    for f in fields_of(g):
        c.import_field(f, st.field(f).reshape(-1))
    r = dcgrid_amd.Simulation(kind, size, p, b)
    r.load_state(str(path))
    g.close()
    for k in range(3):
        r.step(1)
        c.step(1)
        tr, tc = r.export_tables(), c.export_tables()
        assert compare_tables(tr, tc, TABLE_KEYS) == [], (name, k)
        assert hash_map(tr, r.info()) == hash_map(tc, c.info()), (name, k)
        m = mask_of(r)
        for f in ("density", "velocity", "pressure", "temperature", "fluidity"):
            assert same_field(r.export_field(f), np.asarray(c.export_field(f), np.float32), m), (name, k, f)
    r.close()
    c.close()
