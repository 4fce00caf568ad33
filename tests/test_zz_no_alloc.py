"""vg_direct_velocity_noalloc (fixed device scratch, no allocation) against
vg_direct_velocity through the C-ABI: bit for bit on two calls in a row, the
second smaller, with the source range split over workgroups in both; sizes
beyond the fixed scratch are refused."""
import numpy as np
import pytest

from vortegrid_amd import scenarios

pytestmark = pytest.mark.gpu
F32 = np.float32


def case(msg):
    print(f"[case] {msg}", flush=True)


def run_both(vg, pos, om, radius, q, policy):
    dp, dw = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
    v = vg.vortons_view(dp, dw, len(pos), uniform_size=2 * radius)
    dq = vg.DeviceArray.from_numpy(q)
    a = vg.DeviceArray.zeros((len(q), 3))
    b = vg.DeviceArray.zeros((len(q), 3))
    vg.direct_velocity(v, dq, len(q), a, policy=policy)
    vg.direct_velocity_noalloc(v, dq, len(q), b, policy=policy)
    vg.synchronize()
    return a.numpy(), b.numpy()


@pytest.mark.parametrize("policy", [0, 1])
def test_noalloc_equals_direct_two_calls(vg, policy):
    rng = np.random.default_rng(5)
    for n, nq, seed in ((65536, 2000, 1), (3000, 700, 2)):
        case(f"direct velocity without allocation: N={n}, {nq} queries, policy={policy}")
        pos, om, radius = scenarios.vortex_ring(n, seed=seed)
        q = (pos[rng.choice(n, nq, replace=False)] + rng.normal(0, 0.05, (nq, 3))).astype(F32)
        a, b = run_both(vg, pos, om, radius, q, policy)
        assert np.all(np.isfinite(a)) and np.any(a)
        assert np.array_equal(a, b)


def test_noalloc_refuses_sizes_beyond_scratch(vg):
    case("direct velocity without allocation: 1.2M vortons exceed the fixed scratch")
    n = 1_200_000
    pos, om, radius = scenarios.vortex_ring(n, seed=3)
    dp, dw = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
    v = vg.vortons_view(dp, dw, n, uniform_size=2 * radius)
    dq = vg.DeviceArray.from_numpy(pos[:1000])
    out = vg.DeviceArray.zeros((1000, 3))
    with pytest.raises(vg.VGError, match="scratch"):
        vg.direct_velocity_noalloc(v, dq, 1000, out)
    # the library still works after the refusal
    small = vg.DeviceArray.zeros((1000, 3))
    v2 = vg.vortons_view(dp, dw, 5000, uniform_size=2 * radius)
    vg.direct_velocity_noalloc(v2, dq, 1000, small)
    ref = vg.DeviceArray.zeros((1000, 3))
    vg.direct_velocity(v2, dq, 1000, ref)
    vg.synchronize()
    assert np.array_equal(small.numpy(), ref.numpy())
