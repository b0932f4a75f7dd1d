"""dcg_state_check on the CPU: a state buffer's block tables are checked against the planner's
sizes of the grid its header describes before dcg_load_state copies anything to the device. A
consistent state of a small all-dense grid passes; each kind of damage that would send a
kernel outside an array (loads beyond capacity, a dense level not full, levels not matching the
loads, positions outside the level, parent / child / neighbour / hash links to unallocated or
out-of-range blocks) and a truncated or foreign buffer is rejected with DCG_E_ARG (an unknown
version with DCG_E_UNSUPPORTED)."""
import numpy as np
import pytest

import dcgrid_amd
from dcgrid_amd import scenes, state
from dcgrid_amd._abi import DCG_OK

DCG_E_ARG, DCG_E_UNSUPPORTED = -1, -4
NONE = 0xFFFFFFFF
OUT = 31


def dense_state():
    """State of a 16^3 grid of 8^3 blocks with two dense levels (8 blocks + 1): the tables
    k_activate_dense writes (blocks at the dense index, parents, children, same-level neighbour
    descriptors and the clamp form outside the domain), empty hash, zero fields."""
    p = scenes.reference_params(16)
    p.block_width = 8
    bw, B = 8, 9
    rx = [2, 1]
    offs = [0, 8]
    pos = np.zeros((B, 3), np.int32)
    lev = np.array([0] * 8 + [1], np.uint8)
    par = np.full(B, NONE, np.uint32)
    chi = np.full((B, 8), NONE, np.uint32)
    nd = np.zeros((B, 27, 4), np.int32)
    for l in range(2):
        r = rx[l]
        for k in range(r ** 3):
            b = offs[l] + k
            bx, by, bz = k // (r * r), (k // r) % r, k % r
            pos[b] = (bw * bx, bw * by, bw * bz)
            for d in range(27):
                dx, dy, dz = d // 9 - 1, (d // 3) % 3 - 1, d % 3 - 1
                nx, ny, nz = bx + dx, by + dy, bz + dz
                if 0 <= nx < r and 0 <= ny < r and 0 <= nz < r:
                    nb = offs[l] + (nx * r + ny) * r + nz
                    nd[b, d] = (nb, -bw * dx, -bw * dy, -bw * dz)
                else:
                    nd[b, d] = (np.int32(np.uint32(OUT << 27).view(np.int32)), 0, 0, 0)
            if l == 0:
                sb = (bx << 2) | (by << 1) | bz
                par[b] = 8 * 8 + sb
                chi[8, sb] = b
    ctl = np.zeros(98, np.uint32)
    ctl[0:2] = (8, 1)
    H = 4 * B
    st = state.State(state.VERSION, 1, 0, (16, 16, 16), bw, 2, 0, B, B * 512, p)
    st.sections["control"] = ctl.view(np.uint8)
    st.sections["positions"] = pos.view(np.uint8).ravel()
    st.sections["levels"] = lev
    st.sections["flags"] = np.zeros(B, np.uint32).view(np.uint8)
    st.sections["parents"] = par.view(np.uint8)
    st.sections["children"] = chi.view(np.uint8).ravel()
    st.sections["neighbours"] = nd.view(np.uint8).ravel()
    st.sections["hash_keys"] = np.full(H, NONE, np.uint32).view(np.uint8)
    st.sections["hash_values"] = np.full(H, NONE, np.uint32).view(np.uint8)
    st.sections["density"] = np.zeros(B * 512, np.float32).view(np.uint8)
    for k in st.sections:
        st.elem[k] = 1 if k == "levels" else 4
    return st


def tables(st):
    """Writable copies of the table sections."""
    return {k: np.array(v, copy=True) for k, v in st.sections.items()}


def check(tmp_path, st, edit=None):
    st2 = state.State(st.version, st.kind, st.storage, st.size, st.block_width, st.levels, st.flags,
                      st.max_num_blocks, st.num_cells, st.params)
    sec = tables(st)
    if edit:
        edit(sec)
    st2.sections = sec
    st2.elem = dict(st.elem)
    path = tmp_path / "st"
    state.write(str(path), st2)
    return dcgrid_amd.state_check(path.read_bytes())


@pytest.fixture(scope="module")
def lib():
    dcgrid_amd.library()


def test_consistent_state_passes(lib, tmp_path):
    st, msg = check(tmp_path, dense_state())
    assert st == DCG_OK, msg


def u32(sec, name):
    return sec[name].view(np.uint32)


DAMAGE = {
    "load beyond capacity": lambda s: u32(s, "control").__setitem__(0, 9),
    "dense level not full": lambda s: u32(s, "control").__setitem__(1, 0),
    "load of a level beyond the grid": lambda s: u32(s, "control").__setitem__(5, 3),
    "level not matching the loads": lambda s: s["levels"].__setitem__(3, 1),
    "free level inside the prefix": lambda s: s["levels"].__setitem__(8, 0xFF),
    "position outside the level": lambda s: s["positions"].view(np.int32).__setitem__(4, 64),
    "position off the block lattice": lambda s: s["positions"].view(np.int32).__setitem__(3, 3),
    "parent out of range": lambda s: u32(s, "parents").__setitem__(2, 8 * 1000),
    "parent at the wrong level": lambda s: u32(s, "parents").__setitem__(2, 8 * 3),
    "top level with a parent": lambda s: u32(s, "parents").__setitem__(8, 8 * 8),
    "child out of range": lambda s: u32(s, "children").__setitem__(8 * 8 + 1, 500),
    "finest level with a child": lambda s: u32(s, "children").__setitem__(0, 1),
    "neighbour out of range": lambda s: s["neighbours"].view(np.int32).__setitem__(4 * (27 * 0 + 22), 1 << 20),
    "neighbour level shift beyond the top": lambda s: s["neighbours"].view(np.int32).__setitem__(
        4 * (27 * 0 + 22), (3 << 27) | 1),
    "hash value out of range": lambda s: u32(s, "hash_values").__setitem__(5, 10 ** 6),
}


@pytest.mark.parametrize("name", sorted(DAMAGE))
def test_damaged_tables_rejected(lib, tmp_path, name):
    st, msg = check(tmp_path, dense_state(), DAMAGE[name])
    assert st == DCG_E_ARG, (name, msg)
    assert msg.startswith("state rejected") or "section" in msg, msg


def test_truncated_and_foreign_buffers(lib, tmp_path):
    path = tmp_path / "ok"
    state.write(str(path), dense_state())
    raw = path.read_bytes()
    assert dcgrid_amd.state_check(raw)[0] == DCG_OK
    assert dcgrid_amd.state_check(raw[:40])[0] == DCG_E_ARG
    assert dcgrid_amd.state_check(raw[:len(raw) // 2])[0] == DCG_E_ARG
    bad = bytearray(raw)
    bad[0:8] = b"NOTSTATE"
    assert dcgrid_amd.state_check(bytes(bad))[0] == DCG_E_ARG
    ver = bytearray(raw)
    ver[8] = 9
    assert dcgrid_amd.state_check(bytes(ver))[0] == DCG_E_UNSUPPORTED


def test_header_describing_an_impossible_grid(lib, tmp_path):
    st = dense_state()
    st.max_num_blocks = 0  # fewer blocks than the top level needs
    path = tmp_path / "s"
    state.write(str(path), st)
    code, msg = dcgrid_amd.state_check(path.read_bytes())
    assert code == DCG_E_ARG and "planner" in msg


def test_reference_tables_match_the_oracle_activation(lib, tmp_path):
    """dcgrid_amd.state.reference_tables rebuilds the reference-form tables from a state's
    neighbour descriptors: on the all-dense grid they equal the CPU oracle's own tables after
    its activation (k_dcgrid_activate_level semantics), apron entry by apron entry."""
    import oracle
    oracle.library()
    path = tmp_path / "d"
    state.write(str(path), dense_state())
    t = state.reference_tables(state.read(str(path)))
    p = scenes.reference_params(16)
    p.block_width = 8
    c = oracle.OracleSim("dcgrid", (16, 16, 16), p, 9, precision=32)
    tc = c.export_tables()
    for k in ("block_positions", "block_levels", "parent_indices", "child_indices", "cell_indices"):
        assert np.array_equal(t[k], tc[k]), k
    c.close()
