"""Pins of the CPU oracle beyond the reference's golden vectors (CPU only).

Each check is computed independently of the oracle's code path, so a dropped term, a wrong sign or index,
or a transposed operand in the oracle's restatement fails at least one of them:
  * one diffusion solve equals the LOD composition of dense FP64 solves of the per-axis tridiagonal systems
    (host/diffusion_solver.h:359-373 defines the matrices) — numpy.linalg.solve on full matrices;
  * without Dirichlet conditions a sweep conserves mass up to the decay factor: the columns of each 1-D
    operator sum to 1 + decay*dt/dims, so sum(after) = sum(before) / (1 + decay*dt/dims)^dims;
  * the voxel index equals float32 truncation of (p - min)/dx at every voxel boundary and +-1 ulp;
  * the gradient equals numpy central/one-sided differences;
  * single-agent secretion equals the closed form of host/cell_solver.h:383-401 evaluated in FP64;
  * release at death equals D + I*f/vv at the agent's voxel with I cleared (host/cell_solver.cpp:301-346);
  * the bulk update equals its geometric closed forms (host/bulk_solver.cpp:23-35).
"""
import numpy as np
import pytest


def lod_matrix(n, D, decay, dt, dx, dims):
    """The 1-D implicit operator of one LOD sweep (host/diffusion_solver.h:359-373)."""
    if n == 1:
        return np.array([[1.0 + decay * dt / dims]])
    r = dt * D / (dx * dx)
    M = np.zeros((n, n))
    for i in range(n):
        M[i, i] = 1.0 + decay * dt / dims + (r if i in (0, n - 1) else 2.0 * r)
        if i > 0:
            M[i, i - 1] = -r
        if i < n - 1:
            M[i, i + 1] = -r
    return M


def dense_lod_solve(dens, grid, S, D, decay, dt, vs, dims):
    X, Y, Z = grid
    u = dens.reshape(Z, Y, X, S).astype(np.float64).copy()
    for axis in range(dims):
        n = grid[axis]
        ax = 2 - axis
        for s in range(S):
            M = lod_matrix(n, float(D[s]), float(decay[s]), dt, vs[axis], dims)
            us = np.moveaxis(u[..., s], ax, -1)
            sol = np.linalg.solve(M, us.reshape(-1, n).T).T.reshape(us.shape)
            u[..., s] = np.moveaxis(sol, -1, ax)
    return u.reshape(-1)


CASES = [
    (1, (17, 1, 1), 3),
    (2, (9, 13, 1), 2),
    (3, (7, 5, 6), 3),
    (3, (1, 4, 5), 2),
    (3, (2, 2, 2), 1),
]


def make(oracle_mod, dims, grid, S, dtype, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    vs = (20, 16, 25)
    maxs = [grid[d] * vs[d] for d in range(3)]
    D = np.exp(rng.uniform(np.log(1e2), np.log(1e5), size=S))
    decay = rng.uniform(0.01, 1.0, size=S)
    o = oracle_mod.Oracle(dims, [0, 0, 0], maxs, vs, S, 0.01, D, decay, np.zeros(S), dtype=dtype, threads=2)
    o.densities[:] = rng.uniform(0, 1, size=o.densities.size)
    o.initialize()
    return o, D.astype(dtype), decay.astype(dtype), vs


@pytest.mark.parametrize("dims,grid,S", CASES)
def test_diffusion_equals_dense_lod_solve(dims, grid, S, oracle_mod):
    o, D, decay, vs = make(oracle_mod, dims, grid, S, np.float64, seed=3)
    before = o.densities.copy()
    o.diffusion()
    exp = dense_lod_solve(before, grid, S, D, decay, 0.01, vs, dims)
    np.testing.assert_allclose(o.densities, exp, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dims,grid,S", CASES)
def test_float_diffusion_close_to_dense(dims, grid, S, oracle_mod):
    o, D, decay, vs = make(oracle_mod, dims, grid, S, np.float32, seed=5)
    before = o.densities.astype(np.float64)
    o.diffusion()
    exp = dense_lod_solve(before, grid, S, D.astype(np.float64), decay.astype(np.float64), np.float32(0.01), vs,
                          dims)
    np.testing.assert_allclose(o.densities, exp, rtol=2e-5, atol=1e-7)


@pytest.mark.parametrize("dims,grid,S", CASES)
def test_diffusion_conserves_mass_up_to_decay(dims, grid, S, oracle_mod):
    o, D, decay, vs = make(oracle_mod, dims, grid, S, np.float64, seed=7)
    before = o.densities.reshape(-1, S).sum(axis=0)
    o.diffusion()
    after = o.densities.reshape(-1, S).sum(axis=0)
    np.testing.assert_allclose(after, before / (1 + decay * 0.01 / dims) ** dims, rtol=1e-12)


def test_voxel_index_matches_float32_truncation(oracle_mod):
    o, _, _, vs = make(oracle_mod, 3, (50, 7, 3), 1, np.float32, seed=9)
    grid = (50, 7, 3)
    pts = []
    for d in range(3):
        e = (np.arange(-2, grid[d] + 3) * vs[d]).astype(np.float32)
        pts.append(np.concatenate([e, np.nextafter(e, np.float32(-np.inf)), np.nextafter(e, np.float32(np.inf))]))
    rng = np.random.Generator(np.random.PCG64(1))
    P = np.stack([rng.choice(pts[d], size=5000) for d in range(3)], axis=1).astype(np.float32)
    got = o.voxel_positions(P.reshape(-1))
    q = (P - np.float32(0)) / np.asarray(vs, dtype=np.float32)  # float32 arithmetic, IEEE division
    assert q.dtype == np.float32
    inside = (q > -1) & (q < np.asarray(grid, dtype=np.float32))
    exp = np.where(inside, np.trunc(q).astype(np.int64), -1)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("dims,grid,S", [c for c in CASES if min(c[1][:c[0]]) >= 2])
def test_gradient_matches_numpy_differences(dims, grid, S, oracle_mod):
    o, _, _, vs = make(oracle_mod, dims, grid, S, np.float64, seed=11)
    o.gradient()
    X, Y, Z = grid
    u = o.densities.reshape(Z, Y, X, S)
    for d in range(dims):
        ax = 2 - d
        g = np.empty_like(u)
        n = grid[d]
        idx = lambda i: tuple(i if a == ax else slice(None) for a in range(4))
        for i in range(n):
            if i == 0:
                g[idx(i)] = (u[idx(1)] - u[idx(0)]) * np.float32(1.0 / vs[d])
            elif i == n - 1:
                g[idx(i)] = (u[idx(n - 1)] - u[idx(n - 2)]) * np.float32(1.0 / vs[d])
            else:
                g[idx(i)] = (u[idx(i + 1)] - u[idx(i - 1)]) * np.float32(1.0 / (2 * vs[d]))
        V = X * Y * Z
        got = (o.gradients.reshape(3, V * S)[d] if dims == 3 else o.gradients.reshape(V * S, dims)[:, d])
        np.testing.assert_allclose(got, g.reshape(-1), rtol=1e-14, atol=0)


@pytest.mark.parametrize("internalized", [True, False])
def test_single_agent_secretion_closed_form(internalized, oracle_mod):
    """One agent per voxel: D' = (D + num)/(1 + den) + fac and I' = I - vv((-den D + num)/(1+den) + fac)."""
    o, _, _, vs = make(oracle_mod, 3, (6, 5, 4), 2, np.float64, seed=13)
    rng = np.random.Generator(np.random.PCG64(2))
    A, S = 40, 2
    vox = rng.choice(6 * 5 * 4, size=A, replace=False)
    x, y, z = vox % 6, (vox // 6) % 5, vox // 30
    pos = np.stack([(x + 0.5) * vs[0], (y + 0.5) * vs[1], (z + 0.5) * vs[2]], axis=1).reshape(-1)
    a = {k: rng.uniform(0, 1, size=A * S) for k in ("secretion_rates", "uptake_rates", "saturation_densities",
                                                     "net_export_rates", "internalized_substrates",
                                                     "fraction_released_at_death")}
    a["volumes"] = rng.uniform(500, 3000, size=A)
    a["positions"] = pos
    conv = o.set_agents(a)
    o.set_compute_internalized(internalized)
    o.initialize()
    before = o.densities.copy()
    I0 = conv["internalized_substrates"].copy()
    o.cell(True)
    vv = float(np.prod(vs))
    dt = 0.01
    for i in range(A):
        for s in range(S):
            k = i * S + s
            num = a["secretion_rates"][k] * a["saturation_densities"][k] * dt * a["volumes"][i] / vv
            den = (a["uptake_rates"][k] + a["secretion_rates"][k]) * dt * a["volumes"][i] / vv
            fac = a["net_export_rates"][k] * dt / vv
            D = before[vox[i] * S + s]
            np.testing.assert_allclose(o.densities[vox[i] * S + s], (D + num) / (1 + den) + fac, rtol=1e-13)
            if internalized:
                np.testing.assert_allclose(o.internalized[k], I0[k] - vv * ((-den * D + num) / (1 + den) + fac),
                                           rtol=1e-10, atol=1e-10)


def test_bulk_closed_forms(oracle_mod):
    """host/bulk_solver.cpp:23-35, D' = (D + dt*S*T) / (1 + dt*(U+S)), pinned by its closed forms in FP64:
    without uptake D_n - T = (D_0 - T) / (1 + dt*S)^n; without supply D_n = D_0 / (1 + dt*U)^n; and one
    general step satisfies D'(1 + dt*(U+S)) - dt*S*T = D."""
    o, _, _, _ = make(oracle_mod, 3, (7, 5, 6), 3, np.float64, seed=17)
    rng = np.random.Generator(np.random.PCG64(5))
    n = o.densities.size
    S, U, T = (rng.uniform(0, 50, size=n) for _ in range(3))
    dt = 0.01
    d0 = o.densities.copy()
    o.bulk(S, np.zeros(n), T)
    o.bulk(S, np.zeros(n), T)
    np.testing.assert_allclose(o.densities - T, (d0 - T) / (1 + dt * S) ** 2, rtol=1e-12, atol=1e-12)
    d0 = o.densities.copy()
    for _ in range(3):
        o.bulk(np.zeros(n), U, T)
    np.testing.assert_allclose(o.densities, d0 / (1 + dt * U) ** 3, rtol=1e-12)
    d0 = o.densities.copy()
    o.bulk(S, U, T)
    np.testing.assert_allclose(o.densities * (1 + dt * (U + S)) - dt * S * T, d0, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dims,grid", [(3, (6, 5, 4)), (2, (7, 6, 1)), (1, (15, 1, 1))])
def test_cell_release_closed_form(dims, grid, oracle_mod):
    """host/cell_solver.cpp:301-346, release at death, pinned in FP64 by its closed form: at the agent's voxel
    D' = D + I*f/vv for every substrate and I' = 0, nothing else changes; two agents sharing a voxel add their
    terms in turn; an agent outside the mesh changes nothing; with internalization off release is a no-op."""
    S = 3
    o, _, _, vs = make(oracle_mod, dims, grid, S, np.float64, seed=23)
    rng = np.random.Generator(np.random.PCG64(9))
    V = int(np.prod(grid))
    A = 12
    vox = rng.choice(V, size=A - 2, replace=False)
    vox = np.concatenate([vox, [vox[0]]])  # agent A-2 shares agent 0's voxel
    coords = np.stack([vox % grid[0], (vox // grid[0]) % grid[1], vox // (grid[0] * grid[1])], axis=1)
    pos = np.zeros((A, dims))
    for d in range(dims):
        pos[: A - 1, d] = (coords[:, d] + rng.uniform(0.05, 0.95, size=A - 1)) * vs[d]
    pos[A - 1, 0] = (grid[0] + 3) * vs[0]  # outside the mesh along x
    a = {k: rng.uniform(0, 1, size=A * S) for k in ("secretion_rates", "uptake_rates", "saturation_densities",
                                                     "net_export_rates", "fraction_released_at_death")}
    a["internalized_substrates"] = rng.uniform(-5e3, 5e4, size=A * S)
    a["volumes"] = rng.uniform(500, 3000, size=A)
    a["positions"] = pos.reshape(-1)
    vv = float(vs[0] * vs[1] * vs[2])  # the reference's integer voxel volume (mesh.cpp:26-32)

    conv = o.set_agents(a)
    o.set_compute_internalized(True)
    o.initialize()
    I0 = conv["internalized_substrates"].copy()
    expect = o.densities.copy()
    for i in (3, 0, A - 2):
        before_I = o.internalized.copy()
        o.release(i)
        for s in range(S):
            expect[vox[i] * S + s] = expect[vox[i] * S + s] + I0[i * S + s] * a["fraction_released_at_death"][i * S + s] / vv
        np.testing.assert_allclose(o.densities, expect, rtol=1e-15, atol=0)
        assert np.all(o.internalized[i * S:(i + 1) * S] == 0)
        keep = np.ones(A * S, dtype=bool)
        keep[i * S:(i + 1) * S] = False
        np.testing.assert_array_equal(o.internalized[keep], before_I[keep])
    # a released agent releases nothing a second time
    o.release(3)
    np.testing.assert_allclose(o.densities, expect, rtol=1e-15, atol=0)
    # outside the mesh: ignored
    before = o.densities.copy()
    o.release(A - 1)
    np.testing.assert_array_equal(o.densities, before)
    # an index past the agent count is an error
    with pytest.raises(Exception):
        o.release(A)
    # internalization off: no-op on densities and internalized amounts
    o.set_compute_internalized(False)
    before_I = o.internalized.copy()
    o.release(5)
    np.testing.assert_array_equal(o.densities, before)
    np.testing.assert_array_equal(o.internalized, before_I)
