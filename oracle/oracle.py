"""ctypes front end of the CPU oracle (oracle/vg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class Grid(C.Structure):
    """Mirror of vgo_grid (UniformGridGeometry, uniformGrid.h:144-860)."""

    _fields_ = [
        ("min", C.c_float * 3),
        ("extent", C.c_float * 3),
        ("cell", C.c_float * 3),
        ("cpe", C.c_float * 3),
        ("np", C.c_uint32 * 3),
    ]

    @property
    def shape(self):
        return tuple(int(v) for v in self.np)

    @property
    def capacity(self):
        return int(self.np[0]) * int(self.np[1]) * int(self.np[2])

    def as_dict(self):
        return {k: [float(v) for v in getattr(self, k)] for k in ("min", "extent", "cell", "cpe")} | {
            "np": [int(v) for v in self.np]
        }


MAX_LAYERS = 24


class Nested(C.Structure):
    _fields_ = [
        ("depth", C.c_uint32),
        ("layer", Grid * MAX_LAYERS),
        ("dec", (C.c_uint32 * 3) * MAX_LAYERS),
    ]


class Tree(C.Structure):
    _fields_ = [
        ("ng", Nested),
        ("pos", C.POINTER(C.c_float) * MAX_LAYERS),
        ("om", C.POINTER(C.c_float) * MAX_LAYERS),
        ("size", C.POINTER(C.c_float) * MAX_LAYERS),
        ("dip", C.POINTER(C.c_float) * MAX_LAYERS),
        ("cell_start", C.POINTER(C.c_uint32)),
        ("cell_count", C.POINTER(C.c_uint32)),
        ("sorted_idx", C.POINTER(C.c_uint32)),
        ("n", C.c_size_t),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path."""
    src_newer = (not os.path.exists(_LIB_PATH) or
                 os.path.getmtime(os.path.join(_HERE, "vg_oracle.c")) > os.path.getmtime(_LIB_PATH))
    if force or src_newer:
        args = ["make", "-s", "-C", _HERE] + (["-B"] if force else [])
        subprocess.run(args, check=True, capture_output=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.vgo_finvsqrtf.restype = C.c_float
        L.vgo_finvsqrtf.argtypes = [C.c_float]
        L.vgo_nearest_pow2.restype = C.c_uint32
        L.vgo_nearest_pow2.argtypes = [C.c_uint32]
        for name in ("vgo_accumulate_velocity", "vgo_accumulate_vecpot"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [f32p, f32p, f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_int]
        for name in ("vgo_direct_velocity", "vgo_direct_vecpot"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [f32p, f32p, f32p, C.c_size_t, f32p, C.c_size_t, C.c_float,
                           C.c_float, C.c_int, f32p, C.c_int]
        L.vgo_direct_velocity_f64.restype = None
        L.vgo_direct_velocity_f64.argtypes = [f32p, f32p, f32p, C.c_size_t, f32p, C.c_size_t,
                                              C.c_float, C.c_float, f64p, C.c_int]
        L.vgo_define_shape.restype = None
        L.vgo_define_shape.argtypes = [C.POINTER(Grid), C.c_size_t, f32p, f32p, C.c_int]
        L.vgo_decimate.argtypes = [C.POINTER(Grid), C.POINTER(Grid), C.c_uint32]
        L.vgo_grid_capacity.restype = C.c_size_t
        L.vgo_grid_capacity.argtypes = [C.POINTER(Grid)]
        L.vgo_indices_of_position.argtypes = [C.POINTER(Grid), f32p,
                                              np.ctypeslib.ndpointer(dtype=np.uint32)]
        L.vgo_find_bbox.argtypes = [f32p, C.c_size_t, C.c_size_t, f32p, f32p]
        L.vgo_finalize_bbox.argtypes = [f32p, f32p, C.c_float, C.c_size_t, C.POINTER(Grid)]
        L.vgo_nested_init.argtypes = [C.POINTER(Nested), C.POINTER(Grid)]
        L.vgo_tree_build.restype = C.c_int
        L.vgo_tree_build.argtypes = [C.POINTER(Tree), C.POINTER(Grid), f32p, f32p, f32p, C.c_size_t]
        L.vgo_tree_free.argtypes = [C.POINTER(Tree)]
        for name in ("vgo_tree_velocity", "vgo_tree_vecpot"):
            fn = getattr(L, name)
            fn.restype = None
            fn.argtypes = [C.POINTER(Tree), f32p, C.c_float, C.c_float, C.c_int, f32p]
        L.vgo_auto_mollify.argtypes = [C.POINTER(Grid), C.c_float, C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]
        L.vgo_velocity_grid.argtypes = [C.POINTER(Tree), C.POINTER(Grid), f32p, f32p, f32p,
                                        C.c_size_t, C.c_int, C.c_float, C.c_float, C.c_int,
                                        f32p, C.c_int]
        L.vgo_mac_velocity_grid.argtypes = [C.POINTER(Tree), C.POINTER(Grid), f32p, f32p, f32p,
                                            C.c_int, C.c_int, C.c_float, C.c_float, C.c_int,
                                            f32p, C.c_int]
        L.vgo_mac_velocity_grid_bricks.argtypes = [C.POINTER(Tree), C.POINTER(Grid), f32p, f32p,
                                                   f32p, C.c_int, C.c_int, C.c_float, C.c_float,
                                                   C.c_int, f32p, C.c_int, C.c_uint32,
                                                   C.c_uint32]
        L.vgo_mac_velocity_grid_ex.argtypes = [C.POINTER(Tree), C.POINTER(Grid), f32p, f32p,
                                               f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                               C.c_float, C.c_int, f32p, C.c_int, C.c_uint32,
                                               C.c_uint32]
        L.vgo_mac_velocity_grid_box.argtypes = [C.POINTER(Tree), C.POINTER(Grid), f32p, f32p,
                                                f32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                                C.c_float, C.c_int, f32p, C.c_int, C.c_uint32,
                                                C.c_uint32, C.c_uint32, C.c_uint32]
        for name in ("vgo_mac_velocity_grid", "vgo_mac_velocity_grid_bricks",
                     "vgo_mac_velocity_grid_ex", "vgo_mac_velocity_grid_box"):
            getattr(L, name).restype = C.c_int
        L.vgo_nested_has_orphans.restype = C.c_int
        L.vgo_nested_has_orphans.argtypes = [C.POINTER(Nested)]
        L.vgo_local_expansion.argtypes = [f32p, f32p, f32p, C.c_float, f32p, C.c_int, C.c_float,
                                          C.POINTER(C.c_double)]
        L.vgo_local_expansion.restype = None
        L.vgo_vecpot_boundary.argtypes = [C.POINTER(Tree), C.POINTER(Grid), f32p, f32p, f32p,
                                          C.c_size_t, C.c_float, C.c_float, C.c_int, C.c_int,
                                          f32p, C.c_int]
        L.vgo_interpolate.argtypes = [C.POINTER(Grid), f32p, f32p, f32p]
        L.vgo_accumulate.argtypes = [C.POINTER(Grid), f32p, C.c_int, f32p, f32p]
        L.vgo_splat_vorticity.argtypes = [C.POINTER(Grid), f32p, f32p, C.c_size_t, C.c_float,
                                          C.c_float, f32p]
        L.vgo_populate_velocity_grid.argtypes = [C.POINTER(Grid), f32p, f32p, C.c_size_t, f32p]
        L.vgo_downsample.argtypes = [C.POINTER(Grid), f32p, C.POINTER(Grid), f32p, C.c_int]
        L.vgo_upsample.argtypes = [C.POINTER(Grid), f32p, C.POINTER(Grid), f32p, C.c_int]
        L.vgo_sor_pass.argtypes = [C.POINTER(Grid), f32p, f32p, C.c_int]
        L.vgo_solve_vector_poisson.argtypes = [C.POINTER(Grid), f32p, f32p, C.c_size_t]
        L.vgo_mg_cascade.argtypes = [C.POINTER(Grid), f32p, f32p, C.c_size_t]
        L.vgo_curl.argtypes = [C.POINTER(Grid), f32p, f32p]
        L.vgo_jacobian.argtypes = [C.POINTER(Grid), f32p, f32p]
        L.vgo_brick_of_bytes.argtypes = [C.POINTER(Grid), f32p, C.c_void_p, f32p]
        L.vgo_pse_exp.restype = C.c_float
        L.vgo_pse_exp.argtypes = [C.c_float]
        L.vgo_pse_diffuse.argtypes = [C.POINTER(Grid), f32p, f32p, C.c_void_p, C.c_float,
                                      C.c_void_p, C.c_size_t, C.c_float, C.c_float]
        L.vgo_pse_diffuse_simultaneous.argtypes = L.vgo_pse_diffuse.argtypes
        L.vgo_stretch_tilt.argtypes = [C.POINTER(Grid), f32p, f32p, C.c_size_t, f32p, C.c_size_t,
                                       C.c_size_t, C.c_float]
        L.vgo_residual_rel.restype = C.c_double
        L.vgo_residual_rel.argtypes = [C.POINTER(Grid), f32p, f32p]
        L.vgo_residual_rel_pair.restype = C.c_double
        L.vgo_residual_rel_pair.argtypes = [C.POINTER(Grid), f32p, f32p, f32p]
        L.vgo_assign_velocity_from_field.argtypes = [C.POINTER(Grid), f32p, f32p, f32p,
                                                     C.c_size_t, C.c_size_t, C.c_float]
        L.vgo_evolve.argtypes = [f32p, f32p, C.c_size_t, C.c_size_t, C.c_float]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ---------------------------------------------------------------- primitives
def finvsqrtf(x: float) -> float:
    return float(lib().vgo_finvsqrtf(float(x)))


def nearest_pow2(v: int) -> int:
    return int(lib().vgo_nearest_pow2(int(v)))


def define_shape(n: int, vmin, vmax, pow2: bool = True) -> Grid:
    g = Grid()
    lib().vgo_define_shape(C.byref(g), n, _f32(vmin), _f32(vmax), int(pow2))
    return g


def decimate(src: Grid, dec: int) -> Grid:
    g = Grid()
    lib().vgo_decimate(C.byref(g), C.byref(src), dec)
    return g


def nested(src: Grid) -> Nested:
    ng = Nested()
    lib().vgo_nested_init(C.byref(ng), C.byref(src))
    return ng


def indices_of_position(g: Grid, p):
    out = np.zeros(3, np.uint32)
    lib().vgo_indices_of_position(C.byref(g), _f32(p), out)
    return out


def find_bbox(pos):
    pos = _f32(pos).reshape(-1, 3)
    mn = np.full(3, np.finfo(np.float32).max, np.float32)
    mx = -mn
    lib().vgo_find_bbox(pos.reshape(-1), 3, pos.shape[0], mn, mx)
    return mn, mx


def finalize_bbox(mn, mx, vorton_radius: float, n_vortons: int) -> Grid:
    return finalize_bbox_corners(mn, mx, vorton_radius, n_vortons)[0]


def finalize_bbox_corners(mn, mx, vorton_radius: float, n_vortons: int):
    """UpdateBoundingBox(bFinal): (grid template, padded min, padded max)."""
    mn = _f32(mn).copy()
    mx = _f32(mx).copy()
    g = Grid()
    lib().vgo_finalize_bbox(mn, mx, vorton_radius, n_vortons, C.byref(g))
    return g, mn, mx


def grid_for(pos, vorton_radius: float, extra_pos=None) -> Grid:
    """FindBoundingBox + UpdateBoundingBox(final) + DefineShape as the
    reference frame does (vortonSim.cpp:1001-1096)."""
    mn, mx = find_bbox(pos)
    if extra_pos is not None and len(extra_pos):
        mn2, mx2 = find_bbox(extra_pos)
        mn = np.minimum(mn, mn2)
        mx = np.maximum(mx, mx2)
    return finalize_bbox(mn, mx, vorton_radius, _f32(pos).reshape(-1, 3).shape[0])


def auto_mollify(g: Grid, vorton_radius: float):
    s = C.c_float()
    c = C.c_float()
    lib().vgo_auto_mollify(C.byref(g), vorton_radius, C.byref(s), C.byref(c))
    return s.value, c.value


def query_points(g: Grid) -> np.ndarray:
    """All grid points with the (1-2eps) spacing nudge, x fastest."""
    nudge = np.float32(1.0) - np.float32(2.0) * np.float32(np.finfo(np.float32).eps)
    axes = []
    for a in range(3):
        sp = np.float32(g.cell[a]) * nudge
        idx = np.arange(g.np[a], dtype=np.float32)
        axes.append((np.float32(g.min[a]) + idx * sp).astype(np.float32))
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([x, y, z], axis=-1).reshape(-1, 3).astype(np.float32)


# ---------------------------------------------------------------- direct sums
def _size_arr(size, n):
    size = np.asarray(size, dtype=np.float32)
    if size.ndim == 0:
        size = np.full(n, size, np.float32)
    return _f32(size)


def direct_velocity(pos, om, size, q, spread=1.0, circ=1.0, policy=0, threads=1):
    pos = _f32(pos).reshape(-1, 3)
    q = _f32(q).reshape(-1, 3)
    out = np.zeros_like(q)
    lib().vgo_direct_velocity(pos.reshape(-1), _f32(om).reshape(-1), _size_arr(size, len(pos)),
                              len(pos), q.reshape(-1), len(q), spread, circ, policy,
                              out.reshape(-1), threads)
    return out


def direct_vecpot(pos, om, size, q, spread=1.0, circ=1.0, policy=0, threads=1):
    pos = _f32(pos).reshape(-1, 3)
    q = _f32(q).reshape(-1, 3)
    out = np.zeros_like(q)
    lib().vgo_direct_vecpot(pos.reshape(-1), _f32(om).reshape(-1), _size_arr(size, len(pos)),
                            len(pos), q.reshape(-1), len(q), spread, circ, policy,
                            out.reshape(-1), threads)
    return out


def direct_velocity_f64(pos, om, size, q, spread=1.0, circ=1.0, threads=1):
    pos = _f32(pos).reshape(-1, 3)
    q = _f32(q).reshape(-1, 3)
    out = np.zeros(q.shape, np.float64)
    lib().vgo_direct_velocity_f64(pos.reshape(-1), _f32(om).reshape(-1),
                                  _size_arr(size, len(pos)), len(pos), q.reshape(-1), len(q),
                                  spread, circ, out.reshape(-1), threads)
    return out


# ---------------------------------------------------------------- tree
class InfluenceTree:
    """Owns a vgo_tree (NestedGrid<Vorton> built by MakeBaseVortonGrid +
    AggregateClusters)."""

    def __init__(self, grid: Grid, pos, om, size):
        self._t = Tree()
        pos = _f32(pos).reshape(-1)
        om = _f32(om).reshape(-1)
        n = pos.size // 3
        rc = lib().vgo_tree_build(C.byref(self._t), C.byref(grid), pos, om, _size_arr(size, n), n)
        if rc != 0:
            raise MemoryError("oracle tree build failed")

    def __del__(self):
        try:
            lib().vgo_tree_free(C.byref(self._t))
        except Exception:
            pass

    @property
    def depth(self):
        return int(self._t.ng.depth)

    def layer(self, l):
        g = self._t.ng.layer[l]
        cap = g.capacity
        pos = np.ctypeslib.as_array(self._t.pos[l], shape=(cap * 3,)).reshape(-1, 3).copy()
        om = np.ctypeslib.as_array(self._t.om[l], shape=(cap * 3,)).reshape(-1, 3).copy()
        size = np.ctypeslib.as_array(self._t.size[l], shape=(cap,)).copy()
        return g, pos, om, size

    def decimations(self, l):
        return tuple(int(v) for v in self._t.ng.dec[l])

    def dipoles(self, l):
        cap = self._t.ng.layer[l].capacity
        return np.ctypeslib.as_array(self._t.dip[l], shape=(cap * 9,)).reshape(-1, 3, 3).copy()

    def velocity(self, q, spread=1.0, circ=1.0, policy=0):
        out = np.zeros(3, np.float32)
        lib().vgo_tree_velocity(C.byref(self._t), _f32(q), spread, circ, policy, out)
        return out

    def vecpot(self, q, spread=1.0, circ=1.0, policy=0):
        out = np.zeros(3, np.float32)
        lib().vgo_tree_vecpot(C.byref(self._t), _f32(q), spread, circ, policy, out)
        return out

    def ptr(self):
        return C.byref(self._t)


def velocity_grid(tree, g: Grid, pos, om, size, technique=1, spread=1.0, circ=1.0,
                  policy=0, threads=1):
    pos = _f32(pos).reshape(-1)
    n = pos.size // 3
    out = np.zeros(g.capacity * 3, np.float32)
    lib().vgo_velocity_grid(tree.ptr() if tree is not None else None, C.byref(g), pos,
                            _f32(om).reshape(-1), _size_arr(size, n), n, technique, spread,
                            circ, policy, out, threads)
    return out.reshape(-1, 3)


def _walk_status(rc: int):
    if rc != 0:
        raise ValueError("the tree's grid leaves a cell without a parent cluster (a cell count "
                         "odd and above 3 on some layer); the near-field walk refuses it")


def nested_has_orphans(g: Grid) -> bool:
    """True when some NestedGrid layer of g leaves a cell without a parent."""
    ng = nested(g)
    return bool(lib().vgo_nested_has_orphans(C.byref(ng)))


def mac_velocity_grid(tree, g: Grid, pos, om, size, near_k=2, order=1, spread=1.0, circ=1.0,
                      policy=1, threads=1):
    """Near-field treecode at every grid point (bricks of 4^3 points)."""
    pos = _f32(pos).reshape(-1)
    n = pos.size // 3
    out = np.zeros(g.capacity * 3, np.float32)
    _walk_status(lib().vgo_mac_velocity_grid(tree.ptr(), C.byref(g), pos, _f32(om).reshape(-1),
                                             _size_arr(size, n), near_k, order, spread, circ,
                                             policy, out, threads))
    return out.reshape(-1, 3)


def mac_velocity_grid_bricks(tree, g: Grid, pos, om, size, near_k, order, spread, circ, policy,
                             threads, bz_begin, bz_end):
    """Near-field treecode on z-brick rows [bz_begin, bz_end) only."""
    pos = _f32(pos).reshape(-1)
    n = pos.size // 3
    out = np.zeros(g.capacity * 3, np.float32)
    _walk_status(lib().vgo_mac_velocity_grid_bricks(tree.ptr(), C.byref(g), pos,
                                                    _f32(om).reshape(-1), _size_arr(size, n),
                                                    near_k, order, spread, circ, policy, out,
                                                    threads, bz_begin, bz_end))
    return out.reshape(-1, 3)


def mac_velocity_grid_ex(tree, g: Grid, pos, om, size, near_k, order, expand_from, spread=1.0,
                         circ=1.0, policy=1, threads=1, bz_begin=0, bz_end=None):
    """Near-field treecode with brick-centred local expansions from layer
    expand_from (z-brick rows [bz_begin, bz_end))."""
    pos = _f32(pos).reshape(-1)
    n = pos.size // 3
    out = np.zeros(g.capacity * 3, np.float32)
    if bz_end is None:
        bz_end = (g.np[2] + 3) // 4
    _walk_status(lib().vgo_mac_velocity_grid_ex(tree.ptr(), C.byref(g), pos,
                                                _f32(om).reshape(-1), _size_arr(size, n), near_k,
                                                order, expand_from, spread, circ, policy, out,
                                                threads, bz_begin, bz_end))
    return out.reshape(-1, 3)


def mac_velocity_grid_box(tree, g: Grid, pos, om, size, near_k, order, expand_from, spread,
                          circ, policy, threads, bx, bz):
    """Near-field treecode on the bricks with x index in range(*bx) and z
    index in range(*bz) only (CPU timing samples)."""
    pos = _f32(pos).reshape(-1)
    n = pos.size // 3
    out = np.zeros(g.capacity * 3, np.float32)
    _walk_status(lib().vgo_mac_velocity_grid_box(tree.ptr(), C.byref(g), pos,
                                                 _f32(om).reshape(-1), _size_arr(size, n), near_k,
                                                 order, expand_from, spread, circ, policy, out,
                                                 threads, bx[0], bx[1], bz[0], bz[1]))
    return out.reshape(-1, 3)


def local_expansion(cb, cp, w, size, D, order=1, spread=1.0):
    """(c0 [3], c1 [3,3], c2 [3,3,3]) of one cluster's field about cb."""
    out = np.zeros(39, np.float64)
    lib().vgo_local_expansion(_f32(cb), _f32(cp), _f32(w), C.c_float(size),
                              _f32(np.asarray(D).reshape(-1)), int(order), C.c_float(spread),
                              out.ctypes.data_as(C.POINTER(C.c_double)))
    return out[:3], out[3:12].reshape(3, 3), out[12:].reshape(3, 3, 3)


def vecpot_boundary(tree, g: Grid, pos, om, size, spread=1.0, circ=1.0, policy=0,
                    copy_odd=True, threads=1, out=None):
    pos = _f32(pos).reshape(-1)
    n = pos.size // 3
    if out is None:
        out = np.zeros(g.capacity * 3, np.float32)
    out = out.reshape(-1)
    lib().vgo_vecpot_boundary(tree.ptr() if tree is not None else None, C.byref(g), pos,
                              _f32(om).reshape(-1), _size_arr(size, n), n, spread, circ, policy,
                              int(copy_odd), out, threads)
    return out.reshape(-1, 3)


# ---------------------------------------------------------------- grids
def interpolate(g: Grid, field, p):
    out = np.zeros(3, np.float32)
    lib().vgo_interpolate(C.byref(g), _f32(field).reshape(-1), _f32(p), out)
    return out


def splat_vorticity(g: Grid, pos, om, vorton_radius, scale=-1.0):
    pos = _f32(pos).reshape(-1)
    out = np.zeros(g.capacity * 3, np.float32)
    lib().vgo_splat_vorticity(C.byref(g), pos, _f32(om).reshape(-1), pos.size // 3,
                              vorton_radius, scale, out)
    return out.reshape(-1, 3)


def populate_velocity_grid(g: Grid, pos, vel):
    pos = _f32(pos).reshape(-1)
    out = np.zeros(g.capacity * 3, np.float32)
    lib().vgo_populate_velocity_grid(C.byref(g), pos, _f32(vel).reshape(-1), pos.size // 3, out)
    return out.reshape(-1, 3)


def downsample(lo: Grid, hi: Grid, hi_field, accurate=False):
    out = np.zeros(lo.capacity * 3, np.float32)
    lib().vgo_downsample(C.byref(lo), out, C.byref(hi), _f32(hi_field).reshape(-1), int(accurate))
    return out.reshape(-1, 3)


def upsample(hi: Grid, hi_field, lo: Grid, lo_field, interior_only=True):
    out = _f32(hi_field).reshape(-1).copy()
    lib().vgo_upsample(C.byref(hi), out, C.byref(lo), _f32(lo_field).reshape(-1),
                       int(interior_only))
    return out.reshape(-1, 3)


def sor_pass(g: Grid, soln, lap, color):
    s = _f32(soln).reshape(-1).copy()
    lib().vgo_sor_pass(C.byref(g), s, _f32(lap).reshape(-1), color)
    return s.reshape(-1, 3)


def solve_vector_poisson(g: Grid, soln, lap, steps):
    s = _f32(soln).reshape(-1).copy()
    lib().vgo_solve_vector_poisson(C.byref(g), s, _f32(lap).reshape(-1), steps)
    return s.reshape(-1, 3)


def mg_cascade(g: Grid, a0, negw, steps=16):
    a = _f32(a0).reshape(-1).copy()
    lib().vgo_mg_cascade(C.byref(g), a, _f32(negw).reshape(-1), steps)
    return a.reshape(-1, 3)


def curl(g: Grid, a):
    out = np.zeros(g.capacity * 3, np.float32)
    lib().vgo_curl(C.byref(g), _f32(a).reshape(-1), out)
    return out.reshape(-1, 3)


def jacobian(g: Grid, v):
    """ComputeJacobian: (capacity, 3, 3) with [o, a, b] = d v_b / d a."""
    out = np.zeros(g.capacity * 9, np.float32)
    lib().vgo_jacobian(C.byref(g), _f32(v).reshape(-1), out)
    return out.reshape(-1, 3, 3)


def stretch_tilt(g: Grid, jac, pos, angvel, dt):
    """StretchAndTiltVortons on packed (n, 3) arrays; returns the new angvel."""
    p = _f32(pos).reshape(-1)
    w = _f32(angvel).reshape(-1).copy()
    n = p.size // 3
    lib().vgo_stretch_tilt(C.byref(g), _f32(jac).reshape(-1), p, 3, w, 3, n, dt)
    return w.reshape(-1, 3)


def pse_exp(x: float) -> float:
    return float(lib().vgo_pse_exp(float(x)))


def pse_diffuse(g: Grid, pos, angvel, size, viscosity, dt, alive=None, simultaneous=False):
    """DiffuseAndDissipateVorticityPSE (threaded-build order, or all
    exchanges from the starting values with simultaneous=True); size: scalar
    diameter or per-vorton array; alive: optional bool array.  Returns the
    new angular velocities."""
    p = _f32(pos).reshape(-1)
    w = _f32(angvel).reshape(-1).copy()
    n = p.size // 3
    if np.ndim(size) == 0:
        sz_arr, uni = None, float(size)
    else:
        sz_arr, uni = _f32(size).reshape(-1), 0.0
    al = None if alive is None else np.ascontiguousarray(alive, np.uint8)
    fn = lib().vgo_pse_diffuse_simultaneous if simultaneous else lib().vgo_pse_diffuse
    fn(C.byref(g), p, w,
                          sz_arr.ctypes.data_as(C.c_void_p) if sz_arr is not None else None, uni,
                          al.ctypes.data_as(C.c_void_p) if al is not None else None, n,
                          float(viscosity), float(dt))
    return w.reshape(-1, 3)


def brick_of_bytes(g: Grid, field):
    """GenerateBrickOfBytes data: (planes (4, capacity) uint8, stats[7])."""
    planes = np.zeros((4, g.capacity), np.uint8)
    stats = np.zeros(7, np.float32)
    lib().vgo_brick_of_bytes(C.byref(g), _f32(field).reshape(-1),
                             planes.ctypes.data_as(C.c_void_p), stats)
    return planes, stats


def residual_rel(g: Grid, u, f, lo=None) -> float:
    """FP64 ||f - L u|| / ||f|| over the interior; u + lo when lo is given."""
    if lo is None:
        return float(lib().vgo_residual_rel(C.byref(g), _f32(u).reshape(-1), _f32(f).reshape(-1)))
    return float(lib().vgo_residual_rel_pair(C.byref(g), _f32(u).reshape(-1),
                                             _f32(lo).reshape(-1), _f32(f).reshape(-1)))


def advect(g: Grid, field, pos, vel, dt, gain=1.0):
    """AssignVelocityFromField then Evolve (pclOpAdvect.cpp, pclOpEvolve.cpp)."""
    p = _f32(pos).reshape(-1).copy()
    v = _f32(vel).reshape(-1).copy()
    n = p.size // 3
    lib().vgo_assign_velocity_from_field(C.byref(g), _f32(field).reshape(-1), p, v, 3, n, gain)
    lib().vgo_evolve(p, v, 3, n, dt)
    return p.reshape(-1, 3), v.reshape(-1, 3)
