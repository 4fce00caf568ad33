"""The C-ABI library loads and exports every entry point include/vortegrid.h
declares (no GPU needed: nothing here launches a kernel)."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "vortegrid.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vg_[a-z0-9_]+)\s*\(", text)))


def _lib():
    from vortegrid_amd import build, vg

    if not os.path.exists(vg.LIB_PATH):
        build.build()
    return vg, vg.lib()


def test_every_declared_symbol_is_exported():
    vg, L = _lib()
    names = _declared()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(vg.EXPORTS) == names


def test_host_geometry_matches_oracle(oracle):
    """vg_define_shape / vg_nested_layers are host code: identical to the
    oracle's DefineShape and NestedGrid on the same inputs."""
    vg, _ = _lib()
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = int(rng.integers(1, 2_000_000))
        mn = rng.uniform(-3, 1, 3).astype(np.float32)
        mx = (mn + rng.uniform(0.01, 5, 3)).astype(np.float32)
        a = vg.define_shape(n, mn, mx)
        b = oracle.define_shape(n, mn, mx)
        assert a.as_dict() == b.as_dict()
        layers, dec = vg.nested_layers(a)
        ng = oracle.nested(b)
        assert len(layers) == ng.depth
        for l, lay in enumerate(layers):
            assert lay.as_dict() == {k: [float(v) for v in getattr(ng.layer[l], k)] for k in
                                     ("min", "extent", "cell", "cpe")} | {"np": list(ng.layer[l].np)}
            assert dec[l] == tuple(ng.dec[l])


def test_bounding_box_finalize_and_mollify_match_oracle(oracle):
    vg, _ = _lib()
    mn = np.array([-1.2, -0.3, 0.1], np.float32)
    mx = np.array([1.1, 0.4, 2.0], np.float32)
    a = vg.finalize_bounding_box(mn, mx, 0.05, 4096)
    b = oracle.finalize_bbox(mn, mx, 0.05, 4096)
    assert a.as_dict() == b.as_dict()
    assert vg.auto_mollify(a, 0.05) == oracle.auto_mollify(b, 0.05)


def test_status_codes_without_gpu_work():
    vg, L = _lib()
    g = vg.GridDesc()
    assert L.vg_define_shape(None, 1, None, None, 1) == vg.VG_EINVAL
    assert L.vg_tree_depth(None, None) == vg.VG_EINVAL
    assert L.vg_decimate(C.byref(g), C.byref(g), 2) == vg.VG_EINVAL  # empty source
    assert b"empty" in L.vg_last_error()
    assert L.vg_version().startswith(b"vortegrid-amd")


def test_status_codes_of_the_vorticity_terms():
    """Argument checks of the stretching, partition and PSE entry points
    return before any device work."""
    vg, L = _lib()
    g = vg.GridDesc()
    buf = (C.c_float * 64)()
    assert L.vg_jacobian(None, buf, buf, None) == vg.VG_EINVAL
    assert L.vg_jacobian(C.byref(g), buf, buf, None) == vg.VG_EINVAL      # np < 2
    for a in range(3):
        g.np[a], g.extent[a], g.cell[a], g.cpe[a] = 3, 1.0, 0.5, 2.0
    assert L.vg_jacobian(C.byref(g), None, buf, None) == vg.VG_EINVAL
    # no vortons: nothing to do, no device work
    assert L.vg_stretch_tilt(C.byref(g), buf, buf, C.c_size_t(3), buf, C.c_size_t(3),
                             C.c_size_t(0), C.c_float(0.1), None, None) == vg.VG_OK
    assert L.vg_stretch_tilt_from_velocity(C.byref(g), buf, buf, C.c_size_t(2), buf,
                                           C.c_size_t(3), C.c_size_t(5), C.c_float(0.1), None,
                                           None) == vg.VG_EINVAL  # stride < 3
    assert L.vg_sim_stretch_tilt(None, C.c_float(0.1), buf, C.c_size_t(3)) == vg.VG_EINVAL
    assert L.vg_sim_pse_diffuse(None, C.c_float(0.1), C.c_float(0.1), buf, C.c_size_t(3), None,
                                C.c_size_t(18), 0) == vg.VG_EINVAL
    t = C.c_void_p()
    assert L.vg_tree_create(C.byref(t)) == vg.VG_OK
    try:
        p = C.c_void_p()
        cells = C.c_size_t()
        assert L.vg_tree_partition(t, None, C.byref(p), C.byref(p), C.byref(cells)) == vg.VG_EINVAL
        assert L.vg_tree_partition(t, C.byref(p), C.byref(p), C.byref(p),
                                   C.byref(cells)) == vg.VG_ESTATE   # not built
        assert L.vg_tree_pse_diffuse(t, None, C.c_size_t(3), None, C.c_size_t(18), C.c_float(0.1),
                                     C.c_float(0.1), 0, None) == vg.VG_EINVAL
        assert L.vg_tree_pse_diffuse(t, buf, C.c_size_t(3), None, C.c_size_t(18), C.c_float(20.0),
                                     C.c_float(0.1), 0, None) == vg.VG_EINVAL  # nu dt > 1
        assert b"flip" in L.vg_last_error()
        assert L.vg_tree_pse_diffuse(t, buf, C.c_size_t(3), None, C.c_size_t(18), C.c_float(0.1),
                                     C.c_float(0.1), 0, None) == vg.VG_ESTATE  # not built
    finally:
        L.vg_tree_destroy(t)
