"""ctypes mirror of include/dcgrid_amd.h (structs, enums) and a generic handle wrapper.

The wrapper's method names follow the reference's step interface `FluidSimulation`
(src/fluid_simulation.h:4-26): init, reset, adapt_topology, advect_velocity, project,
project_local, advect_density, debug_stats.
"""
import ctypes as C

import numpy as np

DCG_OK = 0
DCG_E_ARG = -1
DCG_E_BUDGET = -2
DCG_E_HIP = -3
DCG_E_UNSUPPORTED = -4
DCG_E_HASH_FULL = -5
DCG_E_NOMEM = -6
ERROR_NAMES = {
    DCG_E_ARG: "DCG_E_ARG",
    DCG_E_BUDGET: "DCG_E_BUDGET",
    DCG_E_HIP: "DCG_E_HIP",
    DCG_E_UNSUPPORTED: "DCG_E_UNSUPPORTED",
    DCG_E_HASH_FULL: "DCG_E_HASH_FULL",
    DCG_E_NOMEM: "DCG_E_NOMEM",
}

DCG_MAX_LEVELS = 16
KIND_UNIFORM, KIND_DCGRID = 0, 1

FIELD_DENSITY = 0
FIELD_VELOCITY = 1
FIELD_FLUIDITY = 2
FIELD_PRESSURE = 3
FIELD_T_PRESSURE = 4
FIELD_DIVERGENCE = 5
FIELD_TEMPERATURE = 6
FIELD_VORTICITY = 7
FIELD_VAPOR = 8
FIELDS = {
    "density": FIELD_DENSITY,
    "velocity": FIELD_VELOCITY,
    "fluidity": FIELD_FLUIDITY,
    "pressure": FIELD_PRESSURE,
    "t_pressure": FIELD_T_PRESSURE,
    "divergence": FIELD_DIVERGENCE,
    "temperature": FIELD_TEMPERATURE,
    "vorticity": FIELD_VORTICITY,
    "vapor": FIELD_VAPOR,
}

PROJECT_MULTILEVEL, PROJECT_LOCAL, PROJECT_SOLVE = 0, 1, 2
SOLVE_MIXED, SOLVE_FP32, SOLVE_FP64 = 0, 1, 2
SCORE_GEOMETRIC, SCORE_FEATURES = 0, 1
SOURCE_NONE, SOURCE_SPHERE_BUOYANT = 0, 1
SOLID_SPHERE, SOLID_TERRAIN = 0, 1
STORAGE_FP32, STORAGE_FP16, STORAGE_BF16 = 0, 1, 2

STAGE_DIVERGENCE = 0
STAGE_RESTRICT = 1
STAGE_JACOBI = 2
STAGE_JACOBI_INV = 3
STAGE_PROLONGATE = 4
STAGE_APPLY_PRESSURE = 5
STAGE_VORTICITY = 6

K_NAMES = ["advect_velocity", "advect_density", "divergence", "jacobi", "prolongate",
           "apply_pressure", "restrict", "topology", "other", "solve_smooth", "solve"]


class Params(C.Structure):
    """dcg_params: the simulation fields of SimParams (src/data/sim_params.h:14-30) + knobs."""

    _fields_ = [
        ("dt", C.c_float),
        ("gx", C.c_int32),
        ("gy", C.c_int32),
        ("gz", C.c_int32),
        ("dx", C.c_float),
        ("rdx", C.c_float),
        ("velocity_emission_rate", C.c_float),
        ("density_emission_rate", C.c_float),
        ("emission_radius", C.c_float),
        ("enable_additional_solids", C.c_int32),
        ("block_width", C.c_int32),
        ("max_levels", C.c_int32),
        ("project_mode", C.c_int32),
        ("project_top_pairs", C.c_int32),
        ("project_level_pairs", C.c_int32),
        ("project_local_pairs", C.c_int32),
        ("score_mode", C.c_int32),
        ("refine_threshold", C.c_float),
        ("grad_threshold", C.c_float),
        ("vort_threshold", C.c_float),
        ("source_mode", C.c_int32),
        ("src_center", C.c_float * 3),
        ("src_radius", C.c_float),
        ("src_density", C.c_float),
        ("src_temperature", C.c_float),
        ("src_noise", C.c_float),
        ("buoy_alpha", C.c_float),
        ("buoy_beta", C.c_float),
        ("t_ambient", C.c_float),
        ("init_velocity_noise", C.c_float),
        ("noise_period", C.c_float),
        ("seed", C.c_uint32),
        ("solve_precision", C.c_int32),
        ("solve_tol", C.c_float),
        ("solve_max_cycles", C.c_int32),
        ("solve_smooth", C.c_int32),
        ("solid_mode", C.c_int32),
        ("terrain_height", C.c_float),
        ("terrain_period", C.c_float),
        ("moist", C.c_int32),
        ("lapse_rate", C.c_float),
        ("qsat_scale", C.c_float),
        ("humidity", C.c_float),
        ("moist_layer", C.c_float),
        ("cond_rate", C.c_float),
        ("latent_heat", C.c_float),
        ("wind", C.c_float * 3),
        ("adiabatic_rate", C.c_float),
        ("storage", C.c_int32),
        ("solve_warm", C.c_int32),
        ("reserved", C.c_int32 * 1),
    ]

    def copy(self):
        p = Params()
        C.pointer(p)[0] = self
        return p

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if name == "reserved":
                continue
            d[name] = list(v) if not isinstance(v, (int, float)) else v
        return d


class Info(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("size", C.c_int32 * 3),
        ("block_width", C.c_int32),
        ("levels", C.c_int32),
        ("sparse_levels", C.c_int32),
        ("apron_width", C.c_int32),
        ("max_num_blocks", C.c_uint64),
        ("num_cells", C.c_uint64),
        ("active_cells", C.c_uint64),
        ("hash_table_size", C.c_uint64),
        ("max_blocks_level", C.c_uint64 * DCG_MAX_LEVELS),
        ("full_blocks_level", C.c_uint64 * DCG_MAX_LEVELS),
        ("level_offsets", C.c_uint64 * DCG_MAX_LEVELS),
        ("block_loads", C.c_uint64 * DCG_MAX_LEVELS),
        ("move_limit", C.c_uint64 * DCG_MAX_LEVELS),
        ("hash_size_level", C.c_uint64 * DCG_MAX_LEVELS),
        ("hash_offset_level", C.c_uint64 * DCG_MAX_LEVELS),
    ]

    def levels_list(self, name):
        return [int(v) for v in getattr(self, name)[: self.levels]]


class SolveInfo(C.Structure):
    _fields_ = [
        ("levels", C.c_int32),
        ("cycles", C.c_int32 * DCG_MAX_LEVELS),
        ("rel_residual", C.c_double * DCG_MAX_LEVELS),
        ("initial_residual", C.c_double * DCG_MAX_LEVELS),
        ("singular", C.c_int32 * DCG_MAX_LEVELS),
        ("converged", C.c_int32),
    ]

    def as_dict(self):
        n = self.levels
        return {"cycles": list(self.cycles[:n]), "rel_residual": list(self.rel_residual[:n]),
                "initial_residual": list(self.initial_residual[:n]), "singular": list(self.singular[:n]),
                "converged": bool(self.converged)}


class Tables(C.Structure):
    _fields_ = [
        ("block_positions", C.POINTER(C.c_int32)),
        ("block_levels", C.POINTER(C.c_uint8)),
        ("block_flags", C.POINTER(C.c_uint8)),
        ("parent_indices", C.POINTER(C.c_uint64)),
        ("child_indices", C.POINTER(C.c_uint64)),
        ("cell_indices", C.POINTER(C.c_uint64)),
        ("hash_key", C.POINTER(C.c_uint32)),
        ("hash_val", C.POINTER(C.c_uint64)),
    ]


RENDER_FLUID, RENDER_DENSITY, RENDER_VELOCITY, RENDER_FLUIDITY, RENDER_RESOLUTION = range(5)


class RenderParams(C.Structure):
    """dcg_render_params: the render() arguments and SimParams render fields
    (src/fluid_simulation.h:19, src/data/sim_params.h:32-44)."""
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("eye_origin", C.c_float * 3), ("eye_dir", C.c_float * 3), ("fov", C.c_float),
        ("sun_dir", C.c_float * 3),
        ("channel", C.c_int32), ("solids", C.c_int32), ("shadows", C.c_int32), ("precise", C.c_int32),
        ("aa_samples", C.c_float), ("ambient", C.c_float),
        ("background_color", C.c_float * 3), ("floor_color", C.c_float * 3), ("smoke_color", C.c_float * 3),
        ("scene_color", C.c_float * 3),
    ]

    def copy(self):
        r = RenderParams()
        C.pointer(r)[0] = self
        return r


class DCGError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where} failed: {ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def bind(lib, prefix, extra_create_arg=False):
    """Declare the argtypes of the step API exported under `prefix` (dcg_ or orc_)."""
    P = C.POINTER
    sim = C.c_void_p
    ext = [C.c_int32] if extra_create_arg else []
    sig = {
        "create_uniform": ([P(C.c_int32), P(Params)] + ext + [P(sim)], C.c_int),
        "create_dcgrid": ([P(C.c_int32), C.c_uint64, P(Params)] + ext + [P(sim)], C.c_int),
        "destroy": ([sim], None),
        "set_params": ([sim, P(Params)], C.c_int),
        "init": ([sim], C.c_int),
        "reset": ([sim], C.c_int),
        "adapt_topology": ([sim], C.c_int),
        "advect_velocity": ([sim], C.c_int),
        "project": ([sim], C.c_int),
        "project_local": ([sim], C.c_int),
        "project_solve": ([sim], C.c_int),
        "solve_report": ([sim, P(SolveInfo)], C.c_int),
        "advect_density": ([sim], C.c_int),
        "step": ([sim, C.c_int32], C.c_int),
        "debug_stats": ([sim, P(C.c_double)], C.c_int),
        "seed_scene": ([sim], C.c_int),
        "info_get": ([sim, P(Info)], C.c_int),
        "field_size": ([sim, C.c_int32, P(C.c_uint64)], C.c_int),
        "export_field": ([sim, C.c_int32, P(C.c_float), C.c_uint64], C.c_int),
        "import_field": ([sim, C.c_int32, P(C.c_float), C.c_uint64], C.c_int),
        "export_tables": ([sim, P(Tables)], C.c_int),
        "run_stage": ([sim, C.c_int32, C.c_int32], C.c_int),
        "last_error": ([sim], C.c_char_p),
    }
    fns = {}
    for name, (args, res) in sig.items():
        f = getattr(lib, prefix + name)
        f.argtypes = args
        f.restype = res
        fns[name] = f
    return fns


class Handle:
    """Owns one simulation handle of a library bound with `bind`.

    Subclasses set `_fns` (the bound functions) and `_extra` (extra create arguments).
    """

    _fns = None
    _extra = ()

    def __init__(self, kind, size, params, max_num_blocks=None):
        self._h = C.c_void_p()
        self.params = params.copy()
        sz = (C.c_int32 * 3)(*size)
        if kind == "uniform":
            st = self._fns["create_uniform"](sz, C.byref(self.params), *self._extra, C.byref(self._h))
        elif kind == "dcgrid":
            st = self._fns["create_dcgrid"](sz, int(max_num_blocks), C.byref(self.params), *self._extra,
                                             C.byref(self._h))
        else:
            raise ValueError(kind)
        if st != DCG_OK:
            msg = self._fns["last_error"](None)
            raise DCGError(st, "create_" + kind, msg.decode() if msg else "")
        self.kind = kind
        self.size = tuple(size)

    def _check(self, st, where):
        if st != DCG_OK:
            msg = self._fns["last_error"](self._h)
            raise DCGError(st, where, msg.decode() if msg else "")

    def close(self):
        if self._h:
            self._fns["destroy"](self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- FluidSimulation methods (src/fluid_simulation.h:9-22)
    def set_params(self, p):
        self._check(self._fns["set_params"](self._h, C.byref(p)), "set_params")
        self.params = p.copy()

    def init(self):
        self._check(self._fns["init"](self._h), "init")

    def reset(self):
        self._check(self._fns["reset"](self._h), "reset")

    def adapt_topology(self):
        self._check(self._fns["adapt_topology"](self._h), "adapt_topology")

    def advect_velocity(self):
        self._check(self._fns["advect_velocity"](self._h), "advect_velocity")

    def project(self):
        self._check(self._fns["project"](self._h), "project")

    def project_local(self):
        self._check(self._fns["project_local"](self._h), "project_local")

    def project_solve(self):
        self._check(self._fns["project_solve"](self._h), "project_solve")

    def solve_report(self):
        out = SolveInfo()
        self._check(self._fns["solve_report"](self._h, C.byref(out)), "solve_report")
        return out.as_dict()

    def advect_density(self):
        self._check(self._fns["advect_density"](self._h), "advect_density")

    def step(self, n=1):
        self._check(self._fns["step"](self._h, int(n)), "step")

    def debug_stats(self):
        out = C.c_double()
        self._check(self._fns["debug_stats"](self._h, C.byref(out)), "debug_stats")
        return out.value

    def seed_scene(self):
        self._check(self._fns["seed_scene"](self._h), "seed_scene")

    def run_stage(self, stage, arg=0):
        """One stage of the step (STAGE_*); arg = level, or field id for STAGE_RESTRICT."""
        self._check(self._fns["run_stage"](self._h, int(stage), int(FIELDS.get(arg, arg))), "run_stage")

    def info(self):
        inf = Info()
        self._check(self._fns["info_get"](self._h, C.byref(inf)), "info_get")
        return inf

    def field_size(self, field):
        n = C.c_uint64()
        self._check(self._fns["field_size"](self._h, FIELDS.get(field, field), C.byref(n)), "field_size")
        return n.value

    def export_field(self, field):
        f = FIELDS.get(field, field)
        n = self.field_size(f)
        out = np.empty(n, dtype=np.float32)
        self._check(self._fns["export_field"](self._h, f, _ptr(out, C.c_float), n), "export_field")
        if f in (FIELD_VELOCITY, FIELD_VORTICITY):
            out = out.reshape(-1, 3)
        return out

    def import_field(self, field, arr):
        f = FIELDS.get(field, field)
        n = self.field_size(f)
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(-1)
        if a.size != n:
            raise ValueError(f"field {field}: expected {n} floats, got {a.size}")
        self._check(self._fns["import_field"](self._h, f, _ptr(a, C.c_float), n), "import_field")

    def export_tables(self):
        inf = self.info()
        B = int(inf.max_num_blocks)
        A3 = int(inf.apron_width) ** 3
        H = int(inf.hash_table_size)
        t = {
            "block_positions": np.empty(3 * B, np.int32),
            "block_levels": np.empty(B, np.uint8),
            "block_flags": np.empty(B, np.uint8),
            "parent_indices": np.empty(B, np.uint64),
            "child_indices": np.empty(8 * B, np.uint64),
            "cell_indices": np.empty(B * A3, np.uint64),
            "hash_key": np.empty(H, np.uint32),
            "hash_val": np.empty(H, np.uint64),
        }
        ct = Tables(
            _ptr(t["block_positions"], C.c_int32),
            _ptr(t["block_levels"], C.c_uint8),
            _ptr(t["block_flags"], C.c_uint8),
            _ptr(t["parent_indices"], C.c_uint64),
            _ptr(t["child_indices"], C.c_uint64),
            _ptr(t["cell_indices"], C.c_uint64),
            _ptr(t["hash_key"], C.c_uint32),
            _ptr(t["hash_val"], C.c_uint64),
        )
        self._check(self._fns["export_tables"](self._h, C.byref(ct)), "export_tables")
        t["block_positions"] = t["block_positions"].reshape(B, 3)
        t["cell_indices"] = t["cell_indices"].reshape(B, A3)
        return t


def hash_map(tables, info):
    """The hash tables as a set of (level, key, block) triples (slot placement ignored)."""
    out = set()
    keys, vals = tables["hash_key"], tables["hash_val"]
    for level in range(info.levels):
        o = int(info.hash_offset_level[level])
        n = int(info.hash_size_level[level])
        k = keys[o:o + n]
        v = vals[o:o + n]
        used = k != 0xFFFFFFFF
        for kk, vv in zip(k[used].tolist(), v[used].tolist()):
            out.add((level, kk, vv))
    return out
