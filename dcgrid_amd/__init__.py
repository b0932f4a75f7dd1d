"""dcgrid_amd — MI355X (gfx950) DCGrid / uniform-grid simulation step.

Host-side mirror of the reference's step interface (src/fluid_simulation.h:4-26) over the C ABI
of libdcgrid_amd.so (include/dcgrid_amd.h). The HIP library is required: importing works
without it, but creating a simulation raises if it has not been built.
"""
import ctypes as C
import os

from ._abi import (  # noqa: F401
    DCG_OK, DCGError, FIELDS, Info, Params, PROJECT_LOCAL, PROJECT_MULTILEVEL, SCORE_FEATURES,
    SCORE_GEOMETRIC, SOURCE_NONE, SOURCE_SPHERE_BUOYANT, K_NAMES, Handle, bind, hash_map,
    RENDER_FLUID, RENDER_DENSITY, RENDER_VELOCITY, RENDER_FLUIDITY, RENDER_RESOLUTION, RenderParams,
)
from . import scenes  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdcgrid_amd.so")

_lib = None
_fns = None


def library():
    """Load libdcgrid_amd.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _lib, _fns
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        try:
            # One HIP runtime per process. torch ships its own libamdhip64 under the same soname
            # (libamdhip64.so.7): loaded first, it also satisfies this library's dependency, so
            # torch tensors and this library share one runtime. Loaded after this library,
            # torch would map a second runtime, which then fails to initialise the device.
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        fns = bind(lib, "dcg_")
        lib.dcg_default_params.argtypes = [C.POINTER(Params)]
        lib.dcg_default_params.restype = None
        lib.dcg_abi_version.argtypes = []
        lib.dcg_abi_version.restype = C.c_int32
        lib.dcg_synchronize.argtypes = [C.c_void_p]
        lib.dcg_synchronize.restype = C.c_int
        lib.dcg_set_timing.argtypes = [C.c_void_p, C.c_int32]
        lib.dcg_set_timing.restype = C.c_int
        lib.dcg_set_graph.argtypes = [C.c_void_p, C.c_int32]
        lib.dcg_set_graph.restype = C.c_int
        lib.dcg_timing_get.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        lib.dcg_timing_get.restype = C.c_int
        lib.dcg_timing_bytes.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]
        lib.dcg_timing_bytes.restype = C.c_int
        lib.dcg_get_params.argtypes = [C.c_void_p, C.POINTER(Params)]
        lib.dcg_get_params.restype = C.c_int
        lib.dcg_cell_steps.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        lib.dcg_cell_steps.restype = C.c_int
        for name in ("dcg_export_field_device", "dcg_import_field_device"):
            f = getattr(lib, name)
            f.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64]
            f.restype = C.c_int
        lib.dcg_tape_forward.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        lib.dcg_tape_forward.restype = C.c_int
        lib.dcg_tape_backward.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        lib.dcg_tape_backward.restype = C.c_int
        lib.dcg_default_render_params.argtypes = [C.c_void_p, C.POINTER(RenderParams)]
        lib.dcg_default_render_params.restype = None
        for name in ("dcg_render", "dcg_render_host", "dcg_render_host_noalloc"):
            f = getattr(lib, name)
            f.argtypes = [C.c_void_p, C.POINTER(RenderParams), C.c_void_p]
            f.restype = C.c_int
        lib.dcg_state_check.argtypes = [C.c_void_p, C.c_uint64]
        lib.dcg_state_check.restype = C.c_int
        lib.dcg_state_size.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        lib.dcg_state_size.restype = C.c_int
        for name in ("dcg_save_state", "dcg_load_state"):
            f = getattr(lib, name)
            f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
            f.restype = C.c_int
        _lib, _fns = lib, fns
    return _lib


def state_check(buf):
    """Check a DCGSTATE buffer (bytes-like or numpy uint8) without a simulation or a device:
    returns (status, message); status 0 = the tables are consistent with the grid the header
    describes (dcg_state_check)."""
    import numpy as np
    library()
    a = np.ascontiguousarray(np.frombuffer(buf, dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf)
    st = _lib.dcg_state_check(C.c_void_p(a.ctypes.data), a.size)
    msg = _lib.dcg_last_error(None)
    return st, (msg.decode() if (st and msg) else "")


def default_params():
    """Reference defaults (src/data/sim_params.cpp:4-17, rdx = 1/dx as in src/main.cpp:17)."""
    library()
    p = Params()
    _lib.dcg_default_params(C.byref(p))
    return p


class Simulation(Handle):
    """One simulation on the GPU (FluidSimulationDCGrid or FluidSimulationUniform)."""

    def __init__(self, kind, size, params, max_num_blocks=None):
        library()
        self._fns = _fns
        super().__init__(kind, size, params, max_num_blocks)

    def synchronize(self):
        self._check(_lib.dcg_synchronize(self._h), "synchronize")

    def cell_steps(self):
        """Active cells summed over the steps run by step() since creation or reset."""
        v = C.c_uint64()
        self._check(_lib.dcg_cell_steps(self._h, C.byref(v)), "cell_steps")
        return v.value

    def set_timing(self, kernels):
        """Time every launch of the named kernel classes (see K_NAMES) with HIP events."""
        mask = 0
        for k in kernels:
            mask |= 1 << K_NAMES.index(k)
        self._check(_lib.dcg_set_timing(self._h, mask), "set_timing")

    def set_graph(self, enable=True):
        """Replay step() in pairs of steps through a captured HIP graph (off by default)."""
        self._check(_lib.dcg_set_graph(self._h, 1 if enable else 0), "set_graph")

    def timing(self):
        """{kernel name: (total ms, launches)} accumulated since timing was enabled."""
        out = {}
        for k, name in enumerate(K_NAMES):
            ms = C.c_double()
            n = C.c_uint64()
            self._check(_lib.dcg_timing_get(self._h, k, C.byref(ms), C.byref(n)), "timing_get")
            out[name] = (ms.value, n.value)
        return out

    def timing_bytes(self):
        """{kernel name: algorithmic bytes of the timed launches} for the classes the library
        counts itself (the solve's smoothing), else 0."""
        out = {}
        for k, name in enumerate(K_NAMES):
            v = C.c_uint64()
            self._check(_lib.dcg_timing_bytes(self._h, k, C.byref(v)), "timing_bytes")
            out[name] = v.value
        return out

    # -- rendering (FluidSimulation::render)
    def default_render_params(self):
        """The reference's render defaults and a camera in front of this grid's domain."""
        r = RenderParams()
        _lib.dcg_default_render_params(self._h, C.byref(r))
        return r

    def render(self, r=None):
        """Ray-marched image as a (height, width, 4) float32 RGBA array."""
        import numpy as np
        r = r or self.default_render_params()
        out = np.empty((r.height, r.width, 4), dtype=np.float32)
        self._check(_lib.dcg_render_host(self._h, C.byref(r), C.c_void_p(out.ctypes.data)), "render")
        return out

    def render_noalloc(self, r=None):
        """render() through dcg_render_host_noalloc (no device allocation; at most 1024 x 1024)."""
        import numpy as np
        r = r or self.default_render_params()
        out = np.empty((r.height, r.width, 4), dtype=np.float32)
        self._check(_lib.dcg_render_host_noalloc(self._h, C.byref(r), C.c_void_p(out.ctypes.data)), "render_noalloc")
        return out

    def render_device(self, r, ptr):
        """Render into caller-owned device memory at `ptr` (4 * width * height floats)."""
        self._check(_lib.dcg_render(self._h, C.byref(r), C.c_void_p(ptr)), "render")

    # -- versioned state (the DCGSTATE format, dcgrid_amd.state)
    def state_size(self):
        n = C.c_uint64()
        self._check(_lib.dcg_state_size(self._h, C.byref(n)), "state_size")
        return n.value

    def save_state(self, path):
        """Write the simulation's state to `path` (memory-mapped, no second host copy)."""
        import numpy as np
        n = self.state_size()
        mm = np.memmap(path, dtype=np.uint8, mode="w+", shape=(n,))
        try:
            self._check(_lib.dcg_save_state(self._h, C.c_void_p(mm.ctypes.data), n), "save_state")
            mm.flush()
        finally:
            del mm
        return n

    def load_state(self, path):
        """Restore a state written by save_state from a grid created with the same shape."""
        import numpy as np
        mm = np.memmap(path, dtype=np.uint8, mode="r")
        try:
            self._check(_lib.dcg_load_state(self._h, C.c_void_p(mm.ctypes.data), mm.size), "load_state")
        finally:
            del mm
        self.params = self.get_params()

    def get_params(self):
        p = Params()
        self._check(_lib.dcg_get_params(self._h, C.byref(p)), "get_params")
        return p

    # -- device-pointer access and the differentiable rollout (uniform grid; DCGrid on a frozen topology)
    def export_field_device(self, field, ptr, nfloats):
        """Copy a field into caller-owned device memory at `ptr` (an int address)."""
        f = FIELDS.get(field, field)
        self._check(_lib.dcg_export_field_device(self._h, f, C.c_void_p(ptr), int(nfloats)), "export_field_device")

    def import_field_device(self, field, ptr, nfloats):
        f = FIELDS.get(field, field)
        self._check(_lib.dcg_import_field_device(self._h, f, C.c_void_p(ptr), int(nfloats)), "import_field_device")

    def tape_forward(self, steps, checkpoint_every):
        """Run `steps` steps from the current state, checkpointing every `checkpoint_every`."""
        self._check(_lib.dcg_tape_forward(self._h, int(steps), int(checkpoint_every)), "tape_forward")

    def tape_backward(self, g_density, g_velocity, g_temperature, d_density, d_velocity, d_temperature):
        """Device addresses (ints; temperature ones may be None): adjoint of the final state in,
        gradient with respect to the rollout's initial state out."""
        vp = lambda p: C.c_void_p(p) if p else None  # noqa: E731
        st = _lib.dcg_tape_backward(self._h, vp(g_density), vp(g_velocity), vp(g_temperature), vp(d_density),
                                    vp(d_velocity), vp(d_temperature))
        self._check(st, "tape_backward")
