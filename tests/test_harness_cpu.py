"""CPU tests of the timing helper shared with the reference
(benchmark_functions, src/_utils.py:6-32)."""
from warpvbd_amd import benchmark_functions


def test_benchmark_functions_samples_and_skip():
    calls = []

    def f(x, y=0):
        calls.append((x, y))
        return x + y

    def g(x, y=0):
        return x * y

    t = benchmark_functions([f, g], 2, num_samples=3, y=5)
    assert set(t) == {"f", "g"} and len(t["f"]) == 3 and len(t["g"]) == 3
    assert all(v >= 0.0 for v in t["f"])
    assert calls == [(2, 5)] * 3
    calls.clear()
    t = benchmark_functions([f], 1, num_samples=2, skipOne=True)
    assert len(t["f"]) == 2 and len(calls) == 3


def test_parallel_grid_scene_numbers_each_environment_like_its_own_mesh():
    """The concatenated parallel scene is one padded voxel grid
    (examples/run_mass_spring_parallel.py:35-58): per environment, 1750 tets
    (the cube and column of one environment mesh), the same vertex set at the
    same grid coordinates, and the environment's vertices and tets in the same
    relative order as in the environment's own mesh -- so first-fit colouring
    and the sweep order of each environment are the same in both forms."""
    import numpy as np

    from warpvbd_amd.harness import _mass_spring_env, _mass_spring_grid

    nsim, nx = 6, 5
    V, T, env = _mass_spring_grid(nsim, nx)
    assert np.array_equal(np.bincount(env), np.full(nsim, len(_mass_spring_env(nx, 0, 0)[1])))
    nsimy = int(np.ceil(np.sqrt(nsim)))
    for k in range(nsim):
        Ve, Te = _mass_spring_env(nx, (k // nsimy) * 2 * nx, (k % nsimy) * 2 * nx)
        mine = T[env == k]
        used = np.unique(mine)  # ascending grid ids = the environment's vertices in grid order
        assert len(used) == len(Ve)
        assert np.array_equal(np.round(V[used] * nx), np.round(Ve * nx))
        # tets in grid order, renumbered to the environment's vertices, equal its own mesh's tets
        local = np.searchsorted(used, mine)
        assert np.array_equal(local, Te)
