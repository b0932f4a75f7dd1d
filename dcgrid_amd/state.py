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
