"""pg_grain_stats_noalloc against pg_grain_stats, through the C ABI.

The no-allocation entry point stages the per-grain results in a static
device array instead of a buffer allocated at call time; the kernel and its
launch are those of pg_grain_stats, so the outputs must agree bit for bit.
Two calls in a row, the second on a smaller handle, check that the shared
array carries nothing over; an instance with more grains than the array
holds is refused with an error, before any device work.
"""
import ctypes

import numpy as np
import pytest

from perigrain_amd import Stepper, _ffi, setups

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _stats(g, noalloc, inst=0):
    _, ng, _ = g.sizes(inst)
    out = np.full((ng, 5 * g.dim + 1), np.nan)
    f = _ffi.lib().pg_grain_stats_noalloc if noalloc else _ffi.lib().pg_grain_stats
    _ffi.check(f(g._h, inst, out.ctypes.data_as(_dp)))
    return out


def test_noalloc_grain_stats_equal_the_allocating_entry_point():
    big, pb = setups.bulk_settle_2d(seed=3, n_grains=120, box=(40e-3, 40e-3), nodes_per_grain=120)
    small, ps = setups.compression_3d(seed=2, n_side=2, R=0.5e-3, h_div=5, cube=3e-3,
                                      wall_speed=-20.0)
    print(f"no-alloc stats: {big.n_grains} grains (2D), then {small.n_grains} grains (3D)",
          flush=True)
    with Stepper(big, pb, device=0) as g:
        g.step(37)
        a = _stats(g, noalloc=True)
        b = _stats(g, noalloc=False)
        assert np.array_equal(a, b, equal_nan=True)
        assert np.isfinite(a[:, :2]).all() and np.any(a[:, :2] != 0)
    with Stepper(small, ps, device=0) as g:  # second call: fewer grains, other dimension
        g.step(25)
        a = _stats(g, noalloc=True)
        b = _stats(g, noalloc=False)
        assert a.shape == (small.n_grains, 16)
        assert np.array_equal(a, b, equal_nan=True)


def test_noalloc_grain_stats_rejects_too_many_grains():
    # one node per grain: 2**18 / 11 + 1 grains exceed the staging array in 2D
    n = (1 << 18) // 11 + 1
    pos = np.stack([np.arange(n) * 1e-3, np.zeros(n)], axis=1)
    s = setups.Setup(
        dim=2, node_offset=np.arange(n + 1, dtype=np.int64), pos=pos, vol=np.full(n, 1e-12),
        disp=np.zeros_like(pos), vel=np.zeros_like(pos), acc=np.zeros_like(pos),
        extforce=np.zeros_like(pos), conn_offset=np.zeros(n + 1, np.int64),
        conn=np.zeros((0, 2), np.uint32), bdry_offset=np.zeros(n + 1, np.int64),
        bdry_nodes=np.zeros(0, np.uint32), clamped_offset=np.zeros(n + 1, np.int64),
        clamped_nodes=np.zeros(0, np.uint32), delta=np.ones(n), rho=np.full(n, 1e3),
        cnot=np.ones(n), snot=np.ones(n), movable=np.ones(n, np.int32),
        breakable=np.ones(n, np.int32), stoppable=np.ones(n, np.int32),
        torque_axis=np.full(n, 2, np.uint32), torque_val=np.zeros(n), contact_radius=0.0,
        normal_stiffness=0.0, damping_ratio=0.0, friction_coefficient=0.0, allow_wall=0,
        wall=np.zeros(6))
    with Stepper(s, setups.Params(dt=1e-8), device=0) as g:
        g.step(1)
        with pytest.raises(_ffi.PgError, match="static staging array"):
            _stats(g, noalloc=True)
        assert _stats(g, noalloc=False).shape == (n, 11)
