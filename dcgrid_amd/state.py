"""The DCGSTATE format: a versioned snapshot of a simulation (include/dcgrid_amd.h,
DCG_STATE_VERSION 1), written and restored by the library (dcg_save_state / dcg_load_state,
Simulation.save_state / load_state) and read here with numpy alone, without the GPU.

Layout (little endian): a 72-byte header, the dcg_params structure, one 24-byte entry per
section (id, element bytes, offset, bytes), then the section data at 256-byte aligned offsets.
DCGrid states hold the block tables (positions, levels, flags, parent and child links, the 27
neighbour descriptors per block, the hash slots) and the device control words; both grids hold
their exported fields in the storage format (FP32, or FP16 / BF16 for DCGrid 16-bit storage).
The reference keeps this state in device memory only (src/dcgrid/dcgrid.h:5-43,
src/uniformgrid/uniformgrid.h:31-46); the format is this library's own.
"""
import ctypes as C
import struct
from dataclasses import dataclass, field

import numpy as np

from ._abi import FIELDS, Params, STORAGE_BF16, STORAGE_FP16, STORAGE_FP32

MAGIC = b"DCGSTATE"
VERSION = 1
HEADER = struct.Struct("<8sII7iIIIQQ")  # 72 bytes
ENTRY = struct.Struct("<IIQQ")           # 24 bytes
ALIGN = 256
FIELD_BASE = 16
TABLES = {1: "control", 2: "positions", 3: "levels", 4: "flags", 5: "parents", 6: "children",
          7: "neighbours", 8: "hash_keys", 9: "hash_values"}
TABLE_IDS = {v: k for k, v in TABLES.items()}
FIELD_NAMES = {v: k for k, v in FIELDS.items()}
_TABLE_DTYPES = {"levels": np.uint8, "positions": np.int32, "neighbours": np.int32}

assert HEADER.size == 72 and ENTRY.size == 24


def section_name(sid):
    if sid in TABLES:
        return TABLES[sid]
    return FIELD_NAMES.get(sid - FIELD_BASE, f"section{sid}")


def section_id(name):
    if name in TABLE_IDS:
        return TABLE_IDS[name]
    return FIELD_BASE + FIELDS[name]


@dataclass
class State:
    version: int
    kind: int
    storage: int
    size: tuple
    block_width: int
    levels: int
    flags: int
    max_num_blocks: int
    num_cells: int
    params: Params
    sections: dict = field(default_factory=dict)  # name -> raw numpy view (uint8/…)
    elem: dict = field(default_factory=dict)      # name -> element bytes

    def table(self, name):
        """A table section as its integer type (positions: (B, 3), neighbours: (B, 27, 4))."""
        a = self.sections[name].view(_TABLE_DTYPES.get(name, np.uint32))
        if name == "positions":
            return a.reshape(-1, 3)
        if name == "neighbours":
            return a.reshape(-1, 27, 4)
        return a

    def field(self, name):
        """A field widened to float32 (velocity and vorticity as (n, 3))."""
        raw = self.sections[name]
        if self.storage == STORAGE_FP32:
            v = raw.view(np.float32)
        elif self.storage == STORAGE_FP16:
            v = raw.view(np.float16).astype(np.float32)
        elif self.storage == STORAGE_BF16:
            v = (raw.view(np.uint16).astype(np.uint32) << 16).view(np.float32)
        else:
            raise ValueError(f"unknown storage {self.storage}")
        return v.reshape(-1, 3) if name in ("velocity", "vorticity") else v


def read(path_or_buffer):
    """Parse a state file (memory-mapped) or bytes-like object; raises ValueError on a bad
    magic, an unknown version or a truncated buffer."""
    if isinstance(path_or_buffer, (bytes, bytearray, memoryview, np.ndarray)):
        buf = np.frombuffer(path_or_buffer, dtype=np.uint8)
    else:
        buf = np.memmap(path_or_buffer, dtype=np.uint8, mode="r")
    if buf.size < HEADER.size:
        raise ValueError("state truncated (header)")
    (magic, version, header_bytes, kind, storage, sx, sy, sz, bw, levels, flags, pbytes, count, maxb,
     ncells) = HEADER.unpack_from(buf, 0)
    if magic != MAGIC:
        raise ValueError("not a DCGSTATE buffer")
    if version != VERSION:
        raise ValueError(f"state version {version} (this reader knows {VERSION})")
    if pbytes != C.sizeof(Params) or header_bytes != HEADER.size + pbytes + count * ENTRY.size:
        raise ValueError("state header does not match the dcg_params layout of this package")
    if buf.size < header_bytes:
        raise ValueError("state truncated (section table)")
    params = Params.from_buffer_copy(bytes(buf[HEADER.size:HEADER.size + pbytes]))
    st = State(version, kind, storage, (sx, sy, sz), bw, levels, flags, maxb, ncells, params)
    for i in range(count):
        sid, elem, off, nbytes = ENTRY.unpack_from(buf, HEADER.size + pbytes + i * ENTRY.size)
        if off % ALIGN or off + nbytes > buf.size:
            raise ValueError(f"state section {i} misplaced or truncated")
        name = section_name(sid)
        st.sections[name] = buf[off:off + nbytes]
        st.elem[name] = elem
    return st


def write(path, st):
    """Write State `st` (sections in the given order) in the library's layout."""
    names = list(st.sections)
    pbytes = C.sizeof(Params)
    header_bytes = HEADER.size + pbytes + len(names) * ENTRY.size
    off = -(-header_bytes // ALIGN) * ALIGN
    entries = []
    for n in names:
        nb = st.sections[n].nbytes
        entries.append((section_id(n), st.elem[n], off, nb))
        off = -(-(off + nb) // ALIGN) * ALIGN
    out = np.zeros(off, dtype=np.uint8)
    HEADER.pack_into(out, 0, MAGIC, VERSION, header_bytes, st.kind, st.storage, *st.size, st.block_width,
                     st.levels, st.flags, pbytes, len(names), st.max_num_blocks, st.num_cells)
    out[HEADER.size:HEADER.size + pbytes] = np.frombuffer(bytes(st.params), dtype=np.uint8)
    for i, (sid, elem, o, nb) in enumerate(entries):
        ENTRY.pack_into(out, HEADER.size + pbytes + i * ENTRY.size, sid, elem, o, nb)
        out[o:o + nb] = np.ascontiguousarray(st.sections[names[i]]).view(np.uint8).ravel()
    out.tofile(path)
    return off


def _morton(x, y, z, bb):
    """Local cell index of block coordinates (x, y, z): bit b of x at 3b + 2, of y at 3b + 1, of
    z at 3b (src/dcgrid/dcgrid_utils.cuh:50-58,69-81), for int64 arrays (negative values keep
    their two's-complement low bits, as the device's int arithmetic does)."""
    r = np.zeros(np.broadcast(x, y, z).shape, dtype=np.int64)
    for b in range(bb):
        r |= (((x >> b) & 1) << (3 * b + 2)) | (((y >> b) & 1) << (3 * b + 1)) | (((z >> b) & 1) << (3 * b))
    return r


def reference_tables(st):
    """The block tables of a DCGrid state in the reference's form (what dcg_export_tables
    returns, src/dcgrid/dcgrid.h:5-73): uint64 links with 2^64 - 1 for none, uint8 flags, and the
    (BW+2)^3 apron cell indices of every block rebuilt from its 27 neighbour descriptors (an
    entry of direction d is the descriptor's block at the offset shifted to its level, or, outside
    the domain, the block's own clamped cell; free blocks have no border entries)."""
    bw = st.block_width
    bb = {4: 2, 8: 3}[bw]
    bv, A = bw ** 3, bw + 2
    lev = st.table("levels")
    B = lev.size
    none = np.uint64(0xFFFFFFFFFFFFFFFF)

    def u64(a):
        out = a.astype(np.uint64)
        out[a == 0xFFFFFFFF] = none
        return out

    nd = st.table("neighbours").astype(np.int64)  # (B, 27, 4)
    i, j, k = (a.ravel() for a in np.meshgrid(np.arange(A), np.arange(A), np.arange(A), indexing="ij"))
    dx = np.where(i == 0, -1, np.where(i == bw + 1, 1, 0))
    dy = np.where(j == 0, -1, np.where(j == bw + 1, 1, 0))
    dz = np.where(k == 0, -1, np.where(k == bw + 1, 1, 0))
    interior = (dx == 0) & (dy == 0) & (dz == 0)
    d = nd[:, (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1), :]  # (B, A^3, 4)
    x0 = d[..., 0] & 0xFFFFFFFF
    sh, nb = x0 >> 27, x0 & 0x07FFFFFF
    b = np.arange(B, dtype=np.int64)[:, None]
    ci, cj, ck = np.clip(i, 1, bw), np.clip(j, 1, bw), np.clip(k, 1, bw)
    spread = _morton(ci & (bw - 1), cj & (bw - 1), ck & (bw - 1), bb)
    clamped = _morton(ci - 1, cj - 1, ck - 1, bb)
    out_cell = b * bv + np.where(d[..., 1] != 0, spread[None, :], clamped[None, :])
    s = np.where(sh == 31, 0, sh)
    in_cell = nb * bv + _morton((i - 1 - d[..., 1]) >> s, (j - 1 - d[..., 2]) >> s, (k - 1 - d[..., 3]) >> s, bb)
    cell = np.where(sh == 31, out_cell, in_cell)
    cell = np.where(interior[None, :], b * bv + _morton(i - 1, j - 1, k - 1, bb)[None, :], cell).astype(np.uint64)
    cell[(lev == 0xFF)[:, None] & ~interior[None, :]] = none
    return {
        "block_positions": st.table("positions").astype(np.int32).reshape(B, 3),
        "block_levels": lev.astype(np.uint8),
        "block_flags": st.table("flags").astype(np.uint8),
        "parent_indices": u64(st.table("parents")),
        "child_indices": u64(st.table("children")),
        "cell_indices": cell.reshape(B, A ** 3),
        "hash_key": st.table("hash_keys").astype(np.uint32),
        "hash_val": u64(st.table("hash_values")),
    }


def control_words(st):
    """(block loads, move limits) per level from a DCGrid state's control words."""
    ctl = st.table("control")
    return [int(v) for v in ctl[0:st.levels]], [int(v) for v in ctl[16:16 + st.levels]]
