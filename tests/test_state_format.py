"""The DCGSTATE layout as documented in include/dcgrid_amd.h, through the numpy reader and writer
(dcgrid_amd.state), without a GPU: header fields at their byte offsets, section entries, 256-byte
aligned section data, storage decoding, and the errors of a damaged buffer. The library's own
files are checked against the same reader in the GPU tests."""
import ctypes as C
import re
import struct
import os

import numpy as np
import pytest

from dcgrid_amd import scenes, state
from dcgrid_amd._abi import Params, STORAGE_BF16, STORAGE_FP16

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def synthetic(storage=0):
    p = scenes.config(1, scale=4).params.copy()
    p.storage = storage
    rng = np.random.default_rng(3)
    B, C_ = 10, 10 * 512
    st = state.State(state.VERSION, 1, storage, (64, 64, 64), 8, 3, 1, B, C_, p)
    st.sections["control"] = rng.integers(0, 99, 97, dtype=np.uint32).view(np.uint8)
    st.sections["positions"] = rng.integers(0, 64, 3 * B, dtype=np.int32).view(np.uint8)
    st.sections["levels"] = rng.integers(0, 3, B, dtype=np.uint8)
    st.sections["neighbours"] = rng.integers(-5, 5, B * 27 * 4, dtype=np.int32).view(np.uint8)
    elem = 4 if storage == 0 else 2
    x = rng.standard_normal(C_).astype(np.float32)
    v = rng.standard_normal(3 * C_).astype(np.float32)
    if storage == 0:
        st.sections["density"], st.sections["velocity"] = x.view(np.uint8), v.view(np.uint8)
    elif storage == STORAGE_FP16:
        st.sections["density"] = x.astype(np.float16).view(np.uint8)
        st.sections["velocity"] = v.astype(np.float16).view(np.uint8)
    else:
        st.sections["density"] = (x.view(np.uint32) >> 16).astype(np.uint16).view(np.uint8)
        st.sections["velocity"] = (v.view(np.uint32) >> 16).astype(np.uint16).view(np.uint8)
    for k in st.sections:
        st.elem[k] = {"levels": 1}.get(k, elem if k in ("density", "velocity") else 4)
    return st, x, v


def test_documented_offsets(tmp_path):
    st, x, v = synthetic()
    path = tmp_path / "s"
    total = state.write(str(path), st)
    raw = path.read_bytes()
    assert len(raw) == total and total % 256 == 0
    assert raw[:8] == b"DCGSTATE"
    version, header_bytes = struct.unpack_from("<II", raw, 8)
    kind, storage, sx, sy, sz, bw, levels = struct.unpack_from("<7i", raw, 16)
    flags, pbytes, count = struct.unpack_from("<III", raw, 44)
    maxb, ncells = struct.unpack_from("<QQ", raw, 56)
    assert (version, kind, storage, (sx, sy, sz), bw, levels, flags) == (1, 1, 0, (64, 64, 64), 8, 3, 1)
    assert pbytes == C.sizeof(Params) and count == 6 and (maxb, ncells) == (10, 5120)
    assert header_bytes == 72 + pbytes + 24 * count
    assert Params.from_buffer_copy(raw[72:72 + pbytes]).seed == st.params.seed
    ids = []
    for i in range(count):
        sid, elem, off, nb = struct.unpack_from("<IIQQ", raw, 72 + pbytes + 24 * i)
        assert off % 256 == 0 and off >= header_bytes
        ids.append(sid)
        if sid == 16:  # DCG_STATE_FIELD + DCG_FIELD_DENSITY
            assert np.array_equal(np.frombuffer(raw[off:off + nb], np.float32), x)
    assert ids == [1, 2, 3, 7, 16, 17]
    # the section ids and the version are the header's
    hdr = open(os.path.join(ROOT, "include", "dcgrid_amd.h")).read()
    assert int(re.search(r"#define DCG_STATE_VERSION (\d+)", hdr).group(1)) == state.VERSION
    for name, sid in state.TABLE_IDS.items():
        cname = "DCG_STATE_" + name.upper()
        assert re.search(rf"{cname} = {sid}\b", hdr), cname
    assert re.search(rf"DCG_STATE_FIELD = {state.FIELD_BASE}\b", hdr)


@pytest.mark.parametrize("storage", [0, STORAGE_FP16, STORAGE_BF16])
def test_round_trip_and_decoding(tmp_path, storage):
    st, x, v = synthetic(storage)
    path = tmp_path / "s"
    state.write(str(path), st)
    r = state.read(str(path))
    assert r.size == (64, 64, 64) and r.flags == 1 and r.storage == storage
    assert bytes(r.params) == bytes(st.params)
    for k in st.sections:
        assert np.array_equal(np.asarray(r.sections[k]), np.asarray(st.sections[k])), k
    assert r.table("positions").shape == (10, 3) and r.table("neighbours").shape == (10, 27, 4)
    d, u = r.field("density"), r.field("velocity")
    assert u.shape == (5120, 3)
    if storage == 0:
        assert np.array_equal(d, x) and np.array_equal(u.ravel(), v)
    elif storage == STORAGE_FP16:
        assert np.array_equal(d, x.astype(np.float16).astype(np.float32))
    else:  # truncated to bfloat16 here; decoding is exact widening
        assert np.array_equal(d.view(np.uint32), x.view(np.uint32) & 0xFFFF0000)
    # rewriting what was read gives the same bytes
    again = tmp_path / "t"
    state.write(str(again), r)
    assert again.read_bytes() == path.read_bytes()


def test_damaged_buffers_are_rejected(tmp_path):
    st, _, _ = synthetic()
    path = tmp_path / "s"
    state.write(str(path), st)
    raw = path.read_bytes()
    with pytest.raises(ValueError, match="not a DCGSTATE"):
        state.read(b"XCGSTATE" + raw[8:])
    with pytest.raises(ValueError, match="version"):
        state.read(raw[:8] + struct.pack("<I", 2) + raw[12:])
    with pytest.raises(ValueError, match="truncated"):
        state.read(raw[:60])
    with pytest.raises(ValueError, match="truncated"):
        state.read(raw[:-300])
