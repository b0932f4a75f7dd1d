"""CPU oracle (TEST INFRASTRUCTURE ONLY).

A C++ restatement of the reference's simulation step (src/dcgrid/*.cu, src/uniformgrid/*.cu,
src/utils/sim_utils.cu) built into oracle/liboracle.so by oracle/Makefile. Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may load it, as the checker — never as
the thing measured or shipped.

Parity status of each restated function is documented in oracle_dcgrid.cpp / oracle_uniform.cpp
and pinned by tests/test_oracle_*.py.
"""
import ctypes as C
import os

from dcgrid_amd._abi import Handle, Info, Params, RenderParams, bind

_HERE = os.path.dirname(os.path.abspath(__file__))
# DCG_ORACLE_LIB: another build of the oracle (e.g. one with a deliberately planted mistake, to
# show that a pin catches it)
LIB_PATH = os.environ.get("DCG_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None
_fns = None


def library():
    global _lib, _fns
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} missing (make -C oracle)")
        lib = C.CDLL(LIB_PATH)
        _fns = bind(lib, "orc_", extra_create_arg=True)
        lib.orc_set_threads.argtypes = [C.c_int32]
        lib.orc_set_threads.restype = None
        lib.orc_get_threads.argtypes = []
        lib.orc_get_threads.restype = C.c_int32
        lib.orc_export_field_f64.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_uint64]
        lib.orc_export_field_f64.restype = C.c_int
        lib.orc_import_field_f64.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_uint64]
        lib.orc_import_field_f64.restype = C.c_int
        lib.orc_tape_forward.argtypes = [C.c_void_p, C.c_int32]
        lib.orc_tape_forward.restype = C.c_int
        lib.orc_tape_backward.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 6
        lib.orc_tape_backward.restype = C.c_int
        lib.orc_lookup.argtypes = [C.c_void_p] + [C.c_int32] * 4
        lib.orc_lookup.restype = C.c_int64
        lib.orc_render.argtypes = [C.c_void_p, C.POINTER(RenderParams), C.c_void_p]
        lib.orc_render.restype = C.c_int
        lib.orc_scores.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_scores.restype = C.c_int
        lib.orc_import_tables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_import_tables.restype = C.c_int
        lib.orc_plan.argtypes = [C.POINTER(C.c_int32), C.c_uint64, C.POINTER(Params), C.POINTER(Info)]
        lib.orc_plan.restype = C.c_int
        for name, args, res in [
            ("orc_bit_spread3", [C.c_uint32], C.c_uint32),
            ("orc_grid_hash", [C.c_int32] * 4, C.c_uint32),
            ("orc_local_index", [C.c_int32] * 4, C.c_uint32),
            ("orc_mipmap_cells", [C.c_uint64] * 3, C.c_uint64),
            ("orc_mipmap_idx", [C.c_int32] * 3 + [C.c_uint64] * 4, C.c_uint64),
            ("orc_value_noise", [C.c_float] * 3 + [C.c_uint32], C.c_float),
            ("orc_hash01", [C.c_uint32] + [C.c_int32] * 4, C.c_float),
            ("orc_exp_det", [C.c_float], C.c_float),
            ("orc_qsat", [C.POINTER(Params), C.c_float], C.c_float),
            ("orc_terrain_height", [C.POINTER(Params), C.c_float, C.c_float], C.c_float),
            ("orc_cell_fluidity", [C.POINTER(Params)] + [C.c_int32] * 4, C.c_float),
            ("orc_saturation_adjust", [C.POINTER(Params)] + [C.POINTER(C.c_float)] * 3, None),
            ("orc_round_storage", [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32], None),
        ]:
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


def set_threads(n):
    library().orc_set_threads(int(n))


def plan(size, max_num_blocks, params):
    """The DCGrid level/budget planner alone (no allocation); returns an Info."""
    lib = library()
    inf = Info()
    st = lib.orc_plan((C.c_int32 * 3)(*size), int(max_num_blocks), C.byref(params), C.byref(inf))
    if st != 0:
        from dcgrid_amd._abi import DCGError
        raise DCGError(st, "plan", "")
    return inf


def get_threads():
    return library().orc_get_threads()


class OracleSim(Handle):
    """Same surface as dcgrid_amd.Simulation, computed on the CPU in float (32) or double (64)."""

    def __init__(self, kind, size, params, max_num_blocks=None, precision=64):
        library()
        self._fns = _fns
        self._extra = (C.c_int32(precision),)
        self.precision = precision
        super().__init__(kind, size, params, max_num_blocks)

    def render(self, r):
        """The reference's ray marcher on the CPU: (height, width, 4) float32 RGBA."""
        import numpy as np
        out = np.empty((r.height, r.width, 4), dtype=np.float32)
        self._check(_lib.orc_render(self._h, C.byref(r), C.c_void_p(out.ctypes.data)), "render")
        return out

    def lookup(self, x, y, z, level=0):
        """getBlockIndexDeep on the oracle's hash tables (-1 when nothing contains the cell)."""
        return int(_lib.orc_lookup(self._h, x, y, z, level))

    def scores(self):
        """(subblock scores [8B + 1], block scores [B]) the next adapt_topology starts from."""
        import numpy as np
        B = int(self.info().max_num_blocks)
        sub = np.empty(8 * B + 1, np.float32)
        blk = np.empty(B, np.float32)
        self._check(_lib.orc_scores(self._h, C.c_void_p(sub.ctypes.data), C.c_void_p(blk.ctypes.data)), "scores")
        return sub, blk

    def import_tables(self, tables, loads, move_limits):
        """Replace the block tables (the dict export_tables returns), loads and move limits."""
        import numpy as np
        from dcgrid_amd._abi import Tables, _ptr
        arr = {
            "block_positions": np.ascontiguousarray(tables["block_positions"], np.int32),
            "block_levels": np.ascontiguousarray(tables["block_levels"], np.uint8),
            "block_flags": np.ascontiguousarray(tables["block_flags"], np.uint8),
            "parent_indices": np.ascontiguousarray(tables["parent_indices"], np.uint64),
            "child_indices": np.ascontiguousarray(tables["child_indices"], np.uint64),
            "cell_indices": np.ascontiguousarray(tables["cell_indices"], np.uint64),
            "hash_key": np.ascontiguousarray(tables["hash_key"], np.uint32),
            "hash_val": np.ascontiguousarray(tables["hash_val"], np.uint64),
        }
        t = Tables(_ptr(arr["block_positions"], C.c_int32), _ptr(arr["block_levels"], C.c_uint8),
                   _ptr(arr["block_flags"], C.c_uint8), _ptr(arr["parent_indices"], C.c_uint64),
                   _ptr(arr["child_indices"], C.c_uint64), _ptr(arr["cell_indices"], C.c_uint64),
                   _ptr(arr["hash_key"], C.c_uint32), _ptr(arr["hash_val"], C.c_uint64))
        ld = np.zeros(16, np.uint64)
        ml = np.zeros(16, np.uint64)
        ld[:len(loads)] = loads
        ml[:len(move_limits)] = move_limits
        st = _lib.orc_import_tables(self._h, C.byref(t), C.c_void_p(ld.ctypes.data), C.c_void_p(ml.ctypes.data))
        self._check(st, "import_tables")

    def export_field_f64(self, field):
        import numpy as np
        from dcgrid_amd._abi import FIELDS, FIELD_VELOCITY, FIELD_VORTICITY
        f = FIELDS.get(field, field)
        n = self.field_size(f)
        out = np.empty(n, dtype=np.float64)
        st = _lib.orc_export_field_f64(self._h, f, out.ctypes.data_as(C.POINTER(C.c_double)), n)
        self._check(st, "export_field_f64")
        if f in (FIELD_VELOCITY, FIELD_VORTICITY):
            out = out.reshape(-1, 3)
        return out

    def import_field_f64(self, field, values):
        import numpy as np
        from dcgrid_amd._abi import FIELDS
        f = FIELDS.get(field, field)
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        st = _lib.orc_import_field_f64(self._h, f, a.ctypes.data_as(C.POINTER(C.c_double)), a.size)
        self._check(st, "import_field_f64")

    def tape_forward(self, steps):
        """Record `steps` steps from the current state (differentiable rollout; DCGrid: on a frozen
        block topology)."""
        self._check(_lib.orc_tape_forward(self._h, int(steps)), "tape_forward")

    def tape_backward(self, g_density, g_velocity, g_temperature=None):
        """Gradient of <g, final state> with respect to the recorded initial state:
        returns (d_density, d_velocity[N,3], d_temperature or None)."""
        import numpy as np
        P = C.POINTER(C.c_double)
        gd = np.ascontiguousarray(g_density, dtype=np.float64).reshape(-1)
        gv = np.ascontiguousarray(g_velocity, dtype=np.float64).reshape(-1)
        n = gd.size
        od = np.empty(n)
        ov = np.empty(3 * n)
        if g_temperature is not None:
            gt = np.ascontiguousarray(g_temperature, dtype=np.float64).reshape(-1)
            ot = np.empty(n)
            gtp, otp = gt.ctypes.data_as(P), ot.ctypes.data_as(P)
        else:
            ot = None
            gtp = otp = None
        st = _lib.orc_tape_backward(self._h, gd.ctypes.data_as(P), gv.ctypes.data_as(P), gtp,
                                    od.ctypes.data_as(P), ov.ctypes.data_as(P), otp)
        self._check(st, "tape_backward")
        return od, ov.reshape(-1, 3), ot
