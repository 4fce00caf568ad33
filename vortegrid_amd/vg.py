"""ctypes binding of libvortegrid (include/vortegrid.h).

This is the Python host side of the drop-in boundary: it loads the in-tree
HIP library and exposes its C-ABI with numpy-friendly helpers.  There is no
CPU fallback: if the library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvortegrid.so")

VG_OK, VG_EINVAL, VG_EMPTY, VG_EDEVICE, VG_ENOMEM, VG_ESTATE = range(6)
RSQRT_REFERENCE, RSQRT_ACCURATE = 0, 1
TECH_DIRECT, TECH_TREE, TECH_POISSON = 0, 1, 2

# Exported entry points declared in include/vortegrid.h.
EXPORTS = (
    "vg_last_error vg_version vg_device_count vg_malloc vg_free vg_memcpy_htod vg_memcpy_dtoh "
    "vg_memset vg_stream_create vg_stream_destroy vg_stream_synchronize vg_profile_enable "
    "vg_profile_reset vg_profile_select vg_profile_get vg_profile_dump vg_define_shape vg_decimate "
    "vg_nested_layers vg_finalize_bounding_box vg_auto_mollify vg_find_bounding_box "
    "vg_direct_velocity vg_direct_vecpot vg_direct_velocity_grid vg_direct_velocity_f64 "
    "vg_direct_velocity_noalloc "
    "vg_tree_create vg_tree_destroy vg_tree_build vg_tree_depth vg_tree_layer "
    "vg_tree_layer_moments vg_tree_velocity_grid_mac vg_tree_mac_work "
    "vg_tree_velocity vg_tree_velocity_grid vg_tree_interaction_count vg_tree_vecpot "
    "vg_tree_vecpot_boundary "
    "vg_tree_splat_vorticity vg_tree_populate_velocity_grid vg_downsample_inject "
    "vg_upsample_interior vg_sor_solve vg_curl vg_residual_norm vg_mg_create vg_mg_destroy "
    "vg_mg_cascade vg_mg_vcycle_solve vg_mg_vcycle_compensation vg_mg_layer vg_sim_multigrid "
    "vg_assign_velocity_from_field "
    "vg_evolve vg_sort_particles_by_cell "
    "vg_advect_tracers vg_sim_config_default vg_sim_create vg_sim_destroy vg_sim_set_vortons "
    "vg_sim_find_bounding_box vg_sim_find_bounding_boxes vg_sim_update_bounding_box "
    "vg_sim_update vg_sim_velocity_grid "
    "vg_sim_vector_potential vg_sim_negative_vorticity vg_sim_grid_template vg_sim_bounds "
    "vg_sim_tree vg_sim_spreading vg_sim_poisson_residual "
    "vg_mlp_create vg_mlp_destroy vg_mlp_set_params vg_mlp_get_params vg_mlp_forward "
    "vg_mlp_gradients vg_mlp_train vg_tree_mlp_samples vg_tree_velocity_grid_mlp vg_sim_set_mlp "
    "vg_batch_config_default vg_batch_create vg_batch_destroy vg_batch_buffers vg_batch_step "
    "vg_batch_capture vg_batch_launch vg_batch_grid vg_batch_mollify vg_batch_set_launch "
    "vg_batch_get_launch vg_batch_autotune vg_batch_solve vg_batch_set_method "
    "vg_batch_sort_tracers vg_batch_clamped vg_jacobian "
    "vg_stretch_tilt "
    "vg_stretch_tilt_from_velocity vg_sim_stretch_tilt vg_tree_partition vg_tree_pse_diffuse "
    "vg_sim_pse_diffuse"
).split()

MLP_INPUTS, MLP_HIDDEN, MLP_OUTPUTS, MLP_PARAMS = 16, 64, 3, 9603


class VGError(RuntimeError):
    pass


class GridDesc(C.Structure):
    """vg_grid_desc == UniformGridGeometry (uniformGrid.h:144-860)."""

    _fields_ = [
        ("min", C.c_float * 3),
        ("extent", C.c_float * 3),
        ("cell", C.c_float * 3),
        ("cpe", C.c_float * 3),
        ("np", C.c_uint32 * 3),
    ]

    @property
    def capacity(self) -> int:
        return int(self.np[0]) * int(self.np[1]) * int(self.np[2])

    @property
    def shape(self):
        return tuple(int(v) for v in self.np)

    def as_dict(self):
        d = {k: [float(v) for v in getattr(self, k)] for k in ("min", "extent", "cell", "cpe")}
        d["np"] = [int(v) for v in self.np]
        return d

    @classmethod
    def from_fields(cls, other) -> "GridDesc":
        g = cls()
        for k in ("min", "extent", "cell", "cpe", "np"):
            for a in range(3):
                getattr(g, k)[a] = getattr(other, k)[a]
        return g


class Vortons(C.Structure):
    _fields_ = [
        ("pos", C.c_void_p),
        ("angvel", C.c_void_p),
        ("size", C.c_void_p),
        ("stride", C.c_size_t),
        ("n", C.c_size_t),
        ("uniform_size", C.c_float),
    ]


class SimConfig(C.Structure):
    _fields_ = [
        ("technique", C.c_int),
        ("rsqrt_policy", C.c_int),
        ("auto_mollify", C.c_int),
        ("mg_steps", C.c_uint32),
        ("boundary_copy_odd", C.c_int),
        ("poisson_solver", C.c_int),
        ("mg_tol", C.c_double),
        ("mg_max_cycles", C.c_uint32),
        ("tree_near_k", C.c_int),
        ("tree_order", C.c_int),
        ("tree_expand_from", C.c_int),
    ]


_lib = None


def lib():
    """Load libvortegrid.so (built by vortegrid_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VGError(f"{LIB_PATH} is missing: build it with `python -m vortegrid_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.vg_last_error.restype = C.c_char_p
        L.vg_version.restype = C.c_char_p
        for name in EXPORTS:
            fn = getattr(L, name)
            if name not in ("vg_last_error", "vg_version"):
                fn.restype = C.c_int
        _lib = L
    return _lib


def check(status: int, what: str = ""):
    if status != VG_OK:
        msg = lib().vg_last_error().decode(errors="replace")
        raise VGError(f"{what or 'vg call'} failed with status {status}: {msg}")


def call(name: str, *args):
    check(getattr(lib(), name)(*args), name)


def device_count() -> int:
    n = C.c_int(0)
    st = lib().vg_device_count(C.byref(n))
    return n.value if st == VG_OK else 0


def require_gpu():
    if device_count() < 1:
        raise VGError("no HIP device visible: the vortegrid path needs an MI355X (gfx950)")


# ---------------------------------------------------------------- device memory
class DeviceArray:
    """Owning device buffer with numpy upload/download."""

    def __init__(self, nbytes: int, dtype=np.float32, shape=None):
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)
        self.ptr = C.c_void_p()
        if self.nbytes:
            call("vg_malloc", C.byref(self.ptr), C.c_size_t(self.nbytes))

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes, a.dtype, a.shape)
        if a.nbytes:
            call("vg_memcpy_htod", d.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes))
        return d

    @classmethod
    def zeros(cls, shape, dtype=np.float32) -> "DeviceArray":
        shape = tuple(np.atleast_1d(shape).tolist())
        nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        d = cls(nbytes, dtype, shape)
        if nbytes:
            call("vg_memset", d.ptr, 0, C.c_size_t(nbytes))
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        call("vg_memcpy_htod", self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            call("vg_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), self.ptr,
                 C.c_size_t(self.nbytes))
        return out

    def offset(self, nfloats: int) -> C.c_void_p:
        return C.c_void_p(self.ptr.value + nfloats * 4)

    def free(self):
        if self.ptr and self.ptr.value:
            lib().vg_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def vortons_view(pos: DeviceArray, angvel: DeviceArray, n: int, uniform_size: float = 0.0,
                 size: DeviceArray | None = None, stride: int = 3) -> Vortons:
    v = Vortons()
    v.pos = pos.ptr.value if n else None
    v.angvel = angvel.ptr.value if n else None
    v.size = size.ptr.value if size is not None else None
    v.stride = stride
    v.n = n
    v.uniform_size = uniform_size
    return v


class Stream:
    def __init__(self):
        self.handle = C.c_void_p()
        call("vg_stream_create", C.byref(self.handle))

    def sync(self):
        call("vg_stream_synchronize", self.handle)

    def __del__(self):
        try:
            if self.handle:
                lib().vg_stream_destroy(self.handle)
        except Exception:
            pass


def synchronize(stream=None):
    call("vg_stream_synchronize", stream.handle if stream is not None else None)


# ---------------------------------------------------------------- geometry
def define_shape(n: int, vmin, vmax, pow2: bool = True) -> GridDesc:
    g = GridDesc()
    a = (C.c_float * 3)(*[float(x) for x in vmin])
    b = (C.c_float * 3)(*[float(x) for x in vmax])
    call("vg_define_shape", C.byref(g), C.c_size_t(n), a, b, int(pow2))
    return g


def nested_layers(g: GridDesc):
    layers = (GridDesc * 16)()
    dec = (C.c_uint32 * 48)()
    depth = C.c_uint32()
    call("vg_nested_layers", C.byref(g), layers, dec, 16, C.byref(depth))
    return [layers[i] for i in range(depth.value)], [tuple(dec[3 * i:3 * i + 3]) for i in range(depth.value)]


def finalize_bounding_box(mn, mx, vorton_radius: float, n: int) -> GridDesc:
    a = (C.c_float * 3)(*[float(x) for x in mn])
    b = (C.c_float * 3)(*[float(x) for x in mx])
    g = GridDesc()
    call("vg_finalize_bounding_box", a, b, C.c_float(vorton_radius), C.c_size_t(n), C.byref(g))
    return g


def auto_mollify(g: GridDesc, vorton_radius: float):
    s, c = C.c_float(), C.c_float()
    call("vg_auto_mollify", C.byref(g), C.c_float(vorton_radius), C.byref(s), C.byref(c))
    return s.value, c.value


def find_bounding_box(pos: DeviceArray, n: int, stride: int = 3, mn=None, mx=None, stream=None):
    big = np.finfo(np.float32).max
    a = (C.c_float * 3)(*(mn if mn is not None else [big] * 3))
    b = (C.c_float * 3)(*(mx if mx is not None else [-big] * 3))
    call("vg_find_bounding_box", pos.ptr, C.c_size_t(stride), C.c_size_t(n), a, b,
         stream.handle if stream else None)
    return np.array(a[:], np.float32), np.array(b[:], np.float32)


# ---------------------------------------------------------------- operators
def _s(stream):
    return stream.handle if stream is not None else None


def direct_velocity(v: Vortons, q: DeviceArray, nq: int, out: DeviceArray, spread=1.0, circ=1.0,
                    policy=RSQRT_REFERENCE, q_stride=3, stream=None):
    call("vg_direct_velocity", C.byref(v), q.ptr, C.c_size_t(q_stride), C.c_size_t(nq),
         C.c_float(spread), C.c_float(circ), policy, out.ptr, _s(stream))


def direct_velocity_noalloc(v: Vortons, q: DeviceArray, nq: int, out: DeviceArray, spread=1.0,
                            circ=1.0, policy=RSQRT_REFERENCE, q_stride=3, stream=None):
    """direct_velocity with the library's fixed device scratch (no allocation)."""
    call("vg_direct_velocity_noalloc", C.byref(v), q.ptr, C.c_size_t(q_stride), C.c_size_t(nq),
         C.c_float(spread), C.c_float(circ), policy, out.ptr, _s(stream))


def direct_vecpot(v: Vortons, q: DeviceArray, nq: int, out: DeviceArray, spread=1.0, circ=1.0,
                  policy=RSQRT_REFERENCE, q_stride=3, stream=None):
    call("vg_direct_vecpot", C.byref(v), q.ptr, C.c_size_t(q_stride), C.c_size_t(nq),
         C.c_float(spread), C.c_float(circ), policy, out.ptr, _s(stream))


def direct_velocity_grid(v: Vortons, g: GridDesc, out: DeviceArray, spread=1.0, circ=1.0,
                         policy=RSQRT_REFERENCE, stream=None):
    call("vg_direct_velocity_grid", C.byref(v), C.byref(g), C.c_float(spread), C.c_float(circ),
         policy, out.ptr, _s(stream))


def direct_velocity_f64(v: Vortons, q: DeviceArray, nq: int, out: DeviceArray, spread=1.0,
                        circ=1.0, q_stride=3, stream=None):
    call("vg_direct_velocity_f64", C.byref(v), q.ptr, C.c_size_t(q_stride), C.c_size_t(nq),
         C.c_float(spread), C.c_float(circ), out.ptr, _s(stream))


class Tree:
    """vg_tree: the NestedGrid influence tree on the device."""

    def __init__(self):
        self.h = C.c_void_p()
        call("vg_tree_create", C.byref(self.h))

    def build(self, g: GridDesc, v: Vortons, stream=None):
        call("vg_tree_build", self.h, C.byref(g), C.byref(v), _s(stream))
        self._n = int(v.n)

    @property
    def depth(self) -> int:
        d = C.c_uint32()
        call("vg_tree_depth", self.h, C.byref(d))
        return d.value

    def layer(self, l: int):
        g = GridDesc()
        call("vg_tree_layer", self.h, l, C.byref(g), None, None, None)
        cap = g.capacity
        pos = np.zeros(cap * 3, np.float32)
        om = np.zeros(cap * 3, np.float32)
        size = np.zeros(cap, np.float32)
        call("vg_tree_layer", self.h, l, C.byref(g), pos.ctypes.data_as(C.c_void_p),
             om.ctypes.data_as(C.c_void_p), size.ctypes.data_as(C.c_void_p))
        return g, pos.reshape(-1, 3), om.reshape(-1, 3), size

    def velocity(self, q: DeviceArray, nq: int, out: DeviceArray, spread=1.0, circ=1.0,
                 policy=RSQRT_REFERENCE, q_stride=3, stream=None):
        call("vg_tree_velocity", self.h, q.ptr, C.c_size_t(q_stride), C.c_size_t(nq),
             C.c_float(spread), C.c_float(circ), policy, out.ptr, _s(stream))

    def moments(self, l: int) -> np.ndarray:
        g = GridDesc()
        call("vg_tree_layer", self.h, l, C.byref(g), None, None, None)
        out = np.zeros(g.capacity * 9, np.float32)
        call("vg_tree_layer_moments", self.h, l, out.ctypes.data_as(C.c_void_p))
        return out.reshape(-1, 3, 3)

    def velocity_grid_mac(self, g: GridDesc, out: DeviceArray, near_k=2, order=1, spread=1.0,
                          circ=1.0, policy=RSQRT_ACCURATE, expand_from=-1, stream=None):
        call("vg_tree_velocity_grid_mac", self.h, C.byref(g), int(near_k), int(order),
             int(expand_from), C.c_float(spread), C.c_float(circ), policy, out.ptr, _s(stream))

    def partition(self):
        """PartitionVortons: (cell_start, cell_count, vorton_index) as numpy."""
        st, ct, ix = C.c_void_p(), C.c_void_p(), C.c_void_p()
        cells = C.c_size_t()
        call("vg_tree_partition", self.h, C.byref(st), C.byref(ct), C.byref(ix), C.byref(cells))
        out = []
        for ptr, count in ((st, cells.value), (ct, cells.value), (ix, self._n)):
            a = np.empty(count, np.uint32)
            if count:
                call("vg_memcpy_dtoh", a.ctypes.data_as(C.c_void_p), ptr, C.c_size_t(a.nbytes))
            out.append(a)
        return tuple(out)

    def pse_diffuse(self, angvel, viscosity, dt, ang_stride=3, birth=None, birth_stride=18,
                    order=0, stream=None):
        """DiffuseAndDissipateVorticityPSE on the vortons this tree was built
        from (order 0: the reference's in-place order; 1: simultaneous)."""
        call("vg_tree_pse_diffuse", self.h, _p(angvel), C.c_size_t(ang_stride),
             _p(birth) if birth is not None else None, C.c_size_t(birth_stride),
             C.c_float(viscosity), C.c_float(dt), int(order), _s(stream))

    def mac_work(self, g: GridDesc, near_k=2, expand_from=-1, stream=None):
        """(cluster interactions at points, direct pairs, clusters expanded)."""
        c, p, e = C.c_uint64(), C.c_uint64(), C.c_uint64()
        call("vg_tree_mac_work", self.h, C.byref(g), int(near_k), int(expand_from), C.byref(c),
             C.byref(p), C.byref(e), _s(stream))
        return c.value, p.value, e.value

    def mlp_samples(self, g: GridDesc, n: int, near_k=1, layer_mask=0xFFFF, spread=1.0,
                    circ=1.0, seed=1, stream=None):
        """(x [n,16], basis [n,3,3], target [n,3], meta [n,2] uint32) device
        arrays of training pairs and the number of pairs left empty."""
        x = DeviceArray.zeros((max(n, 1), MLP_INPUTS))
        b = DeviceArray.zeros((max(n, 1), 3, 3))
        t = DeviceArray.zeros((max(n, 1), MLP_OUTPUTS))
        meta = DeviceArray.zeros((max(n, 1), 2), np.uint32)
        failed = C.c_uint32()
        call("vg_tree_mlp_samples", self.h, C.byref(g), int(near_k), C.c_uint32(layer_mask),
             C.c_float(spread), C.c_float(circ), C.c_uint32(n), C.c_uint64(seed), x.ptr, b.ptr,
             t.ptr, meta.ptr, C.byref(failed), _s(stream))
        return x, b, t, meta, failed.value

    def velocity_grid_mlp(self, g: GridDesc, mlp: "Mlp", out: DeviceArray, near_k=1,
                          layer_mask=0xFFFF, spread=1.0, circ=1.0, policy=RSQRT_ACCURATE,
                          stream=None):
        call("vg_tree_velocity_grid_mlp", self.h, C.byref(g), mlp.h, int(near_k),
             C.c_uint32(layer_mask), C.c_float(spread), C.c_float(circ), policy, out.ptr,
             _s(stream))

    def interaction_count(self, g: GridDesc, stream=None) -> int:
        c = C.c_uint64()
        call("vg_tree_interaction_count", self.h, C.byref(g), C.byref(c), _s(stream))
        return c.value

    def vecpot(self, q: DeviceArray, nq: int, out: DeviceArray, spread=1.0, circ=1.0,
               policy=RSQRT_REFERENCE, q_stride=3, stream=None):
        call("vg_tree_vecpot", self.h, q.ptr, C.c_size_t(q_stride), C.c_size_t(nq),
             C.c_float(spread), C.c_float(circ), policy, out.ptr, _s(stream))

    def velocity_grid(self, g: GridDesc, out: DeviceArray, spread=1.0, circ=1.0,
                      policy=RSQRT_REFERENCE, stream=None):
        call("vg_tree_velocity_grid", self.h, C.byref(g), C.c_float(spread), C.c_float(circ),
             policy, out.ptr, _s(stream))

    def vecpot_boundary(self, g: GridDesc, out: DeviceArray, spread=1.0, circ=1.0,
                        policy=RSQRT_REFERENCE, copy_odd=True, stream=None):
        call("vg_tree_vecpot_boundary", self.h, C.byref(g), C.c_float(spread), C.c_float(circ),
             policy, int(copy_odd), out.ptr, _s(stream))

    def splat_vorticity(self, vorton_radius: float, scale: float, out: DeviceArray, stream=None):
        call("vg_tree_splat_vorticity", self.h, C.c_float(vorton_radius), C.c_float(scale),
             out.ptr, _s(stream))

    def populate_velocity_grid(self, vel: DeviceArray, out: DeviceArray, vel_stride=3,
                               stream=None):
        call("vg_tree_populate_velocity_grid", self.h, vel.ptr, C.c_size_t(vel_stride), out.ptr,
             _s(stream))

    def __del__(self):
        try:
            if self.h:
                lib().vg_tree_destroy(self.h)
        except Exception:
            pass


def downsample_inject(lo: GridDesc, lo_field: DeviceArray, hi: GridDesc, hi_field: DeviceArray,
                      stream=None):
    call("vg_downsample_inject", C.byref(lo), lo_field.ptr, C.byref(hi), hi_field.ptr, _s(stream))


def upsample_interior(hi: GridDesc, hi_field: DeviceArray, lo: GridDesc, lo_field: DeviceArray,
                      stream=None):
    call("vg_upsample_interior", C.byref(hi), hi_field.ptr, C.byref(lo), lo_field.ptr, _s(stream))


def sor_solve(g: GridDesc, soln: DeviceArray, lap: DeviceArray, steps: int, stream=None):
    call("vg_sor_solve", C.byref(g), soln.ptr, lap.ptr, C.c_uint32(steps), _s(stream))


def curl(g: GridDesc, a: DeviceArray, out: DeviceArray, stream=None):
    call("vg_curl", C.byref(g), a.ptr, out.ptr, _s(stream))


def residual_norm(g: GridDesc, u: DeviceArray, f: DeviceArray, stream=None) -> float:
    r = C.c_double()
    call("vg_residual_norm", C.byref(g), u.ptr, f.ptr, C.byref(r), _s(stream))
    return r.value


class Multigrid:
    def __init__(self):
        self.h = C.c_void_p()
        call("vg_mg_create", C.byref(self.h))

    def cascade(self, g: GridDesc, a0: DeviceArray, negw: DeviceArray, steps=16, stream=None):
        call("vg_mg_cascade", self.h, C.byref(g), a0.ptr, negw.ptr, C.c_uint32(steps), _s(stream))

    def vcycle_solve(self, g: GridDesc, a0: DeviceArray, negw: DeviceArray, tol=1e-6,
                     max_cycles=30, stream=None):
        rel = C.c_double()
        cyc = C.c_uint32()
        call("vg_mg_vcycle_solve", self.h, C.byref(g), a0.ptr, negw.ptr, C.c_double(tol),
             C.c_uint32(max_cycles), C.byref(rel), C.byref(cyc), _s(stream))
        return rel.value, cyc.value

    def compensation(self) -> np.ndarray:
        """The FP32 compensation term lo of the last V-cycle solution a0 + lo."""
        p, n = C.c_void_p(), C.c_size_t()
        call("vg_mg_vcycle_compensation", self.h, C.byref(p), C.byref(n))
        return fetch(p, n.value).reshape(-1, 3)

    def __del__(self):
        try:
            if self.h:
                lib().vg_mg_destroy(self.h)
        except Exception:
            pass


class BatchConfig(C.Structure):
    _fields_ = [("cells", C.c_uint32), ("near_k", C.c_int), ("order", C.c_int),
                ("rsqrt_policy", C.c_int), ("expand_from", C.c_int), ("max_cells", C.c_uint32)]


class Batch:
    """Many small effects advanced together (vg_batch)."""

    def __init__(self, vorton_counts, tracer_counts, radii, cells=32, near_k=2, order=1,
                 policy=1, expand_from=-1, max_cells=64):
        """cells=0: each effect's grid from DefineShape every frame (at most
        max_cells per axis)."""
        cfg = BatchConfig()
        call("vg_batch_config_default", C.byref(cfg))
        cfg.cells, cfg.near_k, cfg.order, cfg.rsqrt_policy = cells, near_k, order, policy
        cfg.expand_from, cfg.max_cells = expand_from, max_cells
        vc = np.ascontiguousarray(vorton_counts, np.uint32)
        tc = np.ascontiguousarray(tracer_counts, np.uint32)
        rr = np.ascontiguousarray(radii, np.float32)
        self.effects = len(vc)
        self.vorton_counts, self.tracer_counts = vc.copy(), tc.copy()
        self.h = C.c_void_p()
        call("vg_batch_create", C.byref(self.h), C.c_uint32(len(vc)), vc.ctypes.data_as(C.c_void_p),
             tc.ctypes.data_as(C.c_void_p), rr.ctypes.data_as(C.c_void_p), C.byref(cfg))
        ptrs = [C.c_void_p() for _ in range(4)]
        nv, nt = C.c_size_t(), C.c_size_t()
        call("vg_batch_buffers", self.h, *[C.byref(p) for p in ptrs], C.byref(nv), C.byref(nt))
        self.nv, self.nt = nv.value, nt.value
        self._ptr = {k: p.value for k, p in zip(("vpos", "vom", "vvel", "tpos"), ptrs)}

    def upload(self, vpos, vom, tpos):
        for key, a, n in (("vpos", vpos, self.nv), ("vom", vom, self.nv), ("tpos", tpos, self.nt)):
            a = np.ascontiguousarray(a, np.float32).reshape(-1, 3)
            assert len(a) == n
            if n:
                call("vg_memcpy_htod", C.c_void_p(self._ptr[key]), a.ctypes.data_as(C.c_void_p),
                     C.c_size_t(a.nbytes))

    def download(self, key):
        n = self.nt if key == "tpos" else self.nv
        out = np.zeros((n, 3), np.float32)
        if n:
            synchronize()
            call("vg_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(self._ptr[key]),
                 C.c_size_t(out.nbytes))
        return out

    def step(self, dt, stream=None):
        call("vg_batch_step", self.h, C.c_float(dt), _s(stream))

    def solve(self, stream=None):
        """Velocity grids of the current particles without moving them."""
        call("vg_batch_solve", self.h, _s(stream))

    def clamped(self):
        """(effect, frame) grid definitions limited by max_cells so far."""
        n = C.c_uint32()
        call("vg_batch_clamped", self.h, C.byref(n))
        return n.value

    def sort_tracers(self, stream=None):
        """Reorder each effect's tracers into its current grid's cell order."""
        call("vg_batch_sort_tracers", self.h, _s(stream))

    def set_method(self, near_k, order, expand_from):
        call("vg_batch_set_method", self.h, int(near_k), int(order), int(expand_from))

    def capture(self, dt, stream):
        call("vg_batch_capture", self.h, C.c_float(dt), _s(stream))

    def launch(self, stream=None):
        call("vg_batch_launch", self.h, _s(stream))

    def grid(self, effect):
        g = GridDesc()
        f = C.c_void_p()
        call("vg_batch_grid", self.h, C.c_uint32(effect), C.byref(g), C.byref(f))
        return g, fetch(f, 3 * g.capacity).reshape(-1, 3)

    def grid_ptr(self, effect):
        """The effect's grid geometry and its device field pointer."""
        g = GridDesc()
        f = C.c_void_p()
        call("vg_batch_grid", self.h, C.c_uint32(effect), C.byref(g), C.byref(f))
        return g, f

    def desc(self, effect) -> GridDesc:
        """The effect's current grid geometry (no field copy)."""
        g = GridDesc()
        call("vg_batch_grid", self.h, C.c_uint32(effect), C.byref(g), None)
        return g

    def mollify(self, effect):
        s, c = C.c_float(), C.c_float()
        call("vg_batch_mollify", self.h, C.c_uint32(effect), C.byref(s), C.byref(c))
        return s.value, c.value

    def set_launch(self, walk_waves, tracer_block, tracer_items):
        call("vg_batch_set_launch", self.h, int(walk_waves), int(tracer_block), int(tracer_items))

    def get_launch(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        call("vg_batch_get_launch", self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def autotune(self, stream=None) -> str:
        buf = C.create_string_buffer(8192)
        call("vg_batch_autotune", self.h, _s(stream), buf, C.c_size_t(len(buf)))
        return buf.value.decode()

    def __del__(self):
        try:
            if self.h:
                lib().vg_batch_destroy(self.h)
        except Exception:
            pass


class Mlp:
    """Learned far-field correction network (vg_mlp)."""

    def __init__(self, seed: int = 1):
        self.h = C.c_void_p()
        call("vg_mlp_create", C.byref(self.h), C.c_uint64(seed))

    def set_params(self, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == MLP_PARAMS
        call("vg_mlp_set_params", self.h, p.ctypes.data_as(C.c_void_p))

    def params(self) -> np.ndarray:
        p = np.zeros(MLP_PARAMS, np.float32)
        call("vg_mlp_get_params", self.h, p.ctypes.data_as(C.c_void_p))
        return p

    def forward(self, x: DeviceArray, n: int, y: DeviceArray, stream=None):
        call("vg_mlp_forward", self.h, x.ptr, C.c_uint32(n), y.ptr, _s(stream))

    def gradients(self, x: DeviceArray, t: DeviceArray, n: int, basis=None, stream=None):
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        call("vg_mlp_gradients", self.h, x.ptr, basis.ptr if basis is not None else None, t.ptr,
             C.c_uint32(n), g.ctypes.data_as(C.c_void_p), C.byref(loss), _s(stream))
        return loss.value, g

    def train(self, x: DeviceArray, t: DeviceArray, n: int, batch=4096, steps=1000, lr=1e-3,
              basis=None, stream=None) -> float:
        loss = C.c_float()
        call("vg_mlp_train", self.h, x.ptr, basis.ptr if basis is not None else None, t.ptr,
             C.c_uint32(n), C.c_uint32(batch), C.c_uint32(steps), C.c_float(lr), C.byref(loss),
             _s(stream))
        return loss.value

    def __del__(self):
        try:
            if self.h:
                lib().vg_mlp_destroy(self.h)
        except Exception:
            pass


def assign_velocity_from_field(g: GridDesc, field: DeviceArray, pos, vel, n: int, gain=1.0,
                               stride=3, stream=None):
    call("vg_assign_velocity_from_field", C.byref(g), field.ptr, pos.ptr if hasattr(pos, "ptr") else pos,
         vel.ptr if hasattr(vel, "ptr") else vel, C.c_size_t(stride), C.c_size_t(n),
         C.c_float(gain), _s(stream))


def evolve(pos, vel, n: int, dt: float, stride=3, stream=None):
    call("vg_evolve", pos.ptr if hasattr(pos, "ptr") else pos, vel.ptr if hasattr(vel, "ptr") else vel,
         C.c_size_t(stride), C.c_size_t(n), C.c_float(dt), _s(stream))


def sort_particles_by_cell(g: GridDesc, records, n: int, stride=3, perm=None, stream=None):
    """Reorder n particle records (stride floats, position first) in cell order."""
    call("vg_sort_particles_by_cell", C.byref(g), _p(records), C.c_size_t(stride), C.c_size_t(n),
         _p(perm), _s(stream))


def advect_tracers(g: GridDesc, field, pos, vel, n: int, dt: float, gain=1.0, stride=3,
                   stream=None):
    fp = field.ptr if hasattr(field, "ptr") else field
    vp = None if vel is None else (vel.ptr if hasattr(vel, "ptr") else vel)
    call("vg_advect_tracers", C.byref(g), fp, pos.ptr if hasattr(pos, "ptr") else pos, vp,
         C.c_size_t(stride), C.c_size_t(n), C.c_float(dt), C.c_float(gain), _s(stream))


def _p(x):
    return x.ptr if hasattr(x, "ptr") else x


def jacobian(g: GridDesc, vel, jac, stream=None):
    """ComputeJacobian: jac (capacity, 3, 3) with [o, a, b] = d v_b / d a."""
    call("vg_jacobian", C.byref(g), _p(vel), _p(jac), _s(stream))


def stretch_tilt(g: GridDesc, jac, pos, angvel, n: int, dt: float, pos_stride=3, ang_stride=3,
                 order=None, stream=None):
    """StretchAndTiltVortons from a Jacobian grid (angvel updated in place);
    order: optional device uint32 permutation giving the processing order."""
    call("vg_stretch_tilt", C.byref(g), _p(jac), _p(pos), C.c_size_t(pos_stride), _p(angvel),
         C.c_size_t(ang_stride), C.c_size_t(n), C.c_float(dt),
         _p(order) if order is not None else None, _s(stream))


def stretch_tilt_from_velocity(g: GridDesc, vel, pos, angvel, n: int, dt: float, pos_stride=3,
                               ang_stride=3, order=None, stream=None):
    """StretchAndTiltVortons straight from the velocity grid (no Jacobian grid)."""
    call("vg_stretch_tilt_from_velocity", C.byref(g), _p(vel), _p(pos), C.c_size_t(pos_stride),
         _p(angvel), C.c_size_t(ang_stride), C.c_size_t(n), C.c_float(dt),
         _p(order) if order is not None else None, _s(stream))


def cell_order(g: GridDesc, pos: np.ndarray) -> np.ndarray:
    """The influence tree's vorton order (by leaf cell, index order within a
    cell) computed on the host, as a uint32 permutation for `order`."""
    idx = [np.clip(np.floor((pos[:, a].astype(np.float32) - np.float32(g.min[a]))
                            * np.float32(g.cpe[a])), 0, g.np[a] - 2).astype(np.int64)
           for a in range(3)]
    key = idx[0] + g.np[0] * (idx[1] + g.np[1] * idx[2])
    return np.argsort(key, kind="stable").astype(np.uint32)


# ---------------------------------------------------------------- simulation
def sim_config(**kw) -> SimConfig:
    c = SimConfig()
    call("vg_sim_config_default", C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Sim:
    """vg_sim == VortonSim (vortonSim.h:137-623) with device-resident state."""

    def __init__(self, cfg: SimConfig | None = None, stream=None):
        self.h = C.c_void_p()
        self.stream = stream
        call("vg_sim_create", C.byref(self.h), C.byref(cfg) if cfg is not None else None,
             _s(stream))
        self._v = None

    def set_mlp(self, mlp, layer_mask=0xFFFF):
        """Learned correction for TREE with tree_near_k >= 0 (None = off)."""
        self._mlp = mlp
        call("vg_sim_set_mlp", self.h, mlp.h if mlp is not None else None, C.c_uint32(layer_mask))

    def set_vortons(self, v: Vortons):
        self._v = v
        call("vg_sim_set_vortons", self.h, C.byref(v))

    def find_bounding_box(self):
        call("vg_sim_find_bounding_box", self.h)

    def find_bounding_boxes(self, pos, n: int, stride=3):
        """Vorton box (kept) and the box of n other particles (returned), one wait."""
        mn = np.full(3, np.finfo(np.float32).max, np.float32)
        mx = -mn
        call("vg_sim_find_bounding_boxes", self.h, _p(pos), C.c_size_t(stride), C.c_size_t(n),
             mn.ctypes.data_as(C.c_void_p), mx.ctypes.data_as(C.c_void_p))
        return mn, mx

    def update_bounding_box(self, mn, mx, final=True):
        a = (C.c_float * 3)(*[float(x) for x in mn])
        b = (C.c_float * 3)(*[float(x) for x in mx])
        call("vg_sim_update_bounding_box", self.h, a, b, int(final))

    def update(self, dt=1.0 / 30.0, frame=0, vel_out=None, vel_stride=3):
        call("vg_sim_update", self.h, C.c_float(dt), C.c_uint32(frame),
             vel_out.ptr if vel_out is not None else None, C.c_size_t(vel_stride))

    def velocity_grid(self):
        g = GridDesc()
        p = C.c_void_p()
        call("vg_sim_velocity_grid", self.h, C.byref(g), C.byref(p))
        return g, p

    def stretch_tilt(self, dt, angvel, ang_stride=3):
        """StretchAndTiltVortons on the last update's velocity grid."""
        call("vg_sim_stretch_tilt", self.h, C.c_float(dt), _p(angvel), C.c_size_t(ang_stride))

    def pse_diffuse(self, viscosity, dt, angvel, ang_stride=3, birth=None, birth_stride=18,
                    order=0):
        """DiffuseAndDissipateVorticityPSE on the last update's partition."""
        call("vg_sim_pse_diffuse", self.h, C.c_float(viscosity), C.c_float(dt), _p(angvel),
             C.c_size_t(ang_stride), _p(birth) if birth is not None else None,
             C.c_size_t(birth_stride), int(order))

    def vector_potential(self):
        g = GridDesc()
        p = C.c_void_p()
        call("vg_sim_vector_potential", self.h, C.byref(g), C.byref(p))
        return g, p

    def negative_vorticity(self):
        g = GridDesc()
        p = C.c_void_p()
        call("vg_sim_negative_vorticity", self.h, C.byref(g), C.byref(p))
        return g, p

    def grid_template(self) -> GridDesc:
        g = GridDesc()
        call("vg_sim_grid_template", self.h, C.byref(g))
        return g

    def bounds(self):
        a = (C.c_float * 3)()
        b = (C.c_float * 3)()
        call("vg_sim_bounds", self.h, a, b)
        return np.array(a[:], np.float32), np.array(b[:], np.float32)

    def spreading(self):
        s, c = C.c_float(), C.c_float()
        call("vg_sim_spreading", self.h, C.byref(s), C.byref(c))
        return s.value, c.value

    def poisson_residual(self) -> float:
        r = C.c_double()
        call("vg_sim_poisson_residual", self.h, C.byref(r))
        return r.value

    def __del__(self):
        try:
            if self.h:
                lib().vg_sim_destroy(self.h)
        except Exception:
            pass


def fetch(ptr: C.c_void_p, count: int, dtype=np.float32) -> np.ndarray:
    out = np.empty(count, dtype)
    if count:
        call("vg_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), ptr, C.c_size_t(out.nbytes))
    return out


# ---------------------------------------------------------------- profiling
def profile_enable(on=True):
    call("vg_profile_enable", int(on))


def profile_select(prefixes=None):
    """Time only kernels whose names start with one of the prefixes (None: all)."""
    call("vg_profile_select", ",".join(prefixes).encode() if prefixes else None)


def profile_reset():
    call("vg_profile_reset")


def profile_get(kernel: str):
    ms = C.c_double()
    cnt = C.c_uint64()
    call("vg_profile_get", kernel.encode(), C.byref(ms), C.byref(cnt))
    return ms.value, cnt.value


def profile_dump() -> dict:
    buf = C.create_string_buffer(1 << 16)
    call("vg_profile_dump", buf, C.c_size_t(len(buf)))
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        out[name] = (int(cnt), float(ms))
    return out
