"""Brute-force geometry of a DCGrid for the oracle pins: where every cell of the field arrays
lies, and which cell the reference's apron semantics make the neighbour of a lattice position
(the cell of the finest allocated block at the same or a coarser level containing it), found by
painting the blocks into dense volumes — independent of the oracle's tables and lookups."""
import numpy as np

import oracle

NF = np.uint64(0xFFFFFFFFFFFFFFFF)


def local_table(bw):
    """local index of (x, y, z) within a block (orc_local_index), as a dense bw^3 array."""
    lib = oracle.library()
    t = np.zeros((bw, bw, bw), np.int64)
    for x in range(bw):
        for y in range(bw):
            for z in range(bw):
                t[x, y, z] = lib.orc_local_index(x, y, z, bw)
    return t


class Geometry:
    """Cells, levels and neighbour owners of a DCGrid state (tables from export_tables)."""

    def __init__(self, sim):
        self.inf = inf = sim.info()
        self.t = t = sim.export_tables()
        self.bw = bw = int(inf.block_width)
        self.bv = bv = bw ** 3
        self.L = int(inf.levels)
        self.size = np.array([int(v) for v in inf.size[:3]])
        self.lt = local_table(bw)
        locs = np.zeros((bv, 3), np.int64)
        for x in range(bw):
            for y in range(bw):
                for z in range(bw):
                    locs[self.lt[x, y, z]] = (x, y, z)
        B = int(inf.max_num_blocks)
        self.B = B
        blev = t["block_levels"]
        self.cell_lev = np.repeat(blev.astype(np.int64), bv)
        self.alloc = np.repeat(blev != 0xFF, bv)
        self.pos = np.repeat(t["block_positions"].astype(np.int64), bv, axis=0) + np.tile(locs, (B, 1))
        scale = np.where(self.alloc, 1 << np.where(self.alloc, self.cell_lev, 0), 0)
        self.scale = scale
        self.centre = (self.pos + 0.5) * scale[:, None]  # cell centre in finest-cell units
        sub = np.arange(B * bv) // (bv // 8)
        self.leaf = self.alloc & (t["child_indices"][sub] == NF)
        # owners[l][q]: finest allocated block at level >= l containing level-l position q
        self.owners = []
        for start in range(self.L):
            shape = tuple(int(v) >> start for v in self.size)
            vol = -np.ones(shape, np.int64)
            for level in range(self.L - 1, start - 1, -1):
                sh = level - start
                for b in np.nonzero(blev == level)[0]:
                    x, y, z = (int(v) << sh for v in t["block_positions"][b])
                    e = bw << sh
                    vol[x:x + e, y:y + e, z:z + e] = b
            self.owners.append(vol)

    def owner_cells(self, q, level):
        """Cell indices (and validity) of the owners of level-`level` positions q (N,3)."""
        q = np.asarray(q, np.int64)
        lim = self.size >> level
        inside = np.all((q >= 0) & (q < lim), axis=1)
        qq = np.where(inside[:, None], q, 0)
        nb = self.owners[level][qq[:, 0], qq[:, 1], qq[:, 2]]
        ok = inside & (nb >= 0)
        nbs = np.where(ok, nb, 0)
        blev = self.t["block_levels"][nbs].astype(np.int64)
        sh = blev - level
        r = (qq >> sh[:, None]) - self.t["block_positions"][nbs].astype(np.int64)
        r = np.clip(r, 0, self.bw - 1)
        cell = nbs * self.bv + self.lt[r[:, 0], r[:, 1], r[:, 2]]
        return cell, ok, blev

    def face_owners(self, cells):
        """For each cell, the owner cells of its six face-neighbour positions on its own level,
        in the order x-, x+, y-, y+, z-, z+; `inside` is False where a position is outside."""
        out, oks, same = [], [], []
        lev = self.cell_lev[cells]
        for axis in range(3):
            for sgn in (-1, 1):
                q = self.pos[cells].copy()
                q[:, axis] += sgn
                cell = np.zeros(len(cells), np.int64)
                ok = np.zeros(len(cells), bool)
                olev = np.zeros(len(cells), np.int64)
                for level in range(self.L):
                    m = lev == level
                    if m.any():
                        c, o, bl = self.owner_cells(q[m], level)
                        cell[m], ok[m], olev[m] = c, o, bl
                out.append(cell)
                oks.append(ok)
                same.append(olev == lev)
        return np.stack(out, 1), np.all(np.stack(oks, 1), axis=1), np.all(np.stack(same, 1), axis=1)
