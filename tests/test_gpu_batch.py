"""Batched microenvironments (bfvm_batch_*) against the oracle run instance by instance (needs an MI355X).

A batch advances every instance with one launch per sweep axis and one gradient launch over all instances
(work lists of (instance, range) items), replayed as one captured graph per step kind.  Each instance must
end where the oracle puts it when it steps that instance alone, to the partitioned sweep's bar: relative
error <= 1e-5 per voxel for densities, gradients relative to the density scale over dx, internalized
substrates to the cancellation bound of the cell solver tests.
"""
import numpy as np
import pytest

from test_gpu_parity import assert_internalized

pytestmark = pytest.mark.gpu

REL = 1e-5


@pytest.fixture(scope="module")
def pkg():
    import parabiofvm_amd

    parabiofvm_amd.load()
    return parabiofvm_amd


def args_of(cfg):
    return (cfg["dims"], cfg["mins"], cfg["maxs"], cfg["voxel_shape"], cfg["S"], cfg["dt"], cfg["D"], cfg["decay"],
            np.zeros(cfg["S"], dtype=np.float32))


def assert_close(got, exp, what):
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    err = np.abs(got - exp) / np.maximum(np.abs(exp), 1e-6)
    i = int(np.argmax(err))
    assert err.max() <= REL, f"{what}: worst {err.max():.3e} at {i} (gpu {got[i]!r} cpu {exp[i]!r})"


def run_batch(pkg, oracle_mod, cfgs, steps, storage="f32"):
    from parabiofvm_amd.configs import apply_config

    oracles, solvers = [], []
    for cfg in cfgs:
        o = oracle_mod.Oracle(*args_of(cfg), dtype=np.float32)
        g = pkg.DeviceSolver(*args_of(cfg))
        g.set_storage(storage)
        for be in (o, g):
            apply_config(cfg, be)
            be.initialize()
        g.store()
        oracles.append(o)
        solvers.append(g)
    batch = pkg.Batch(solvers)
    for i in range(steps):
        recompute = i % 10 == 0
        batch.timestep(recompute, True)
        for o in oracles:
            o.timestep(recompute, True)
    batch.wait()
    for g in solvers:
        g.load()
    return batch, oracles, solvers


def numpy_gradient(dens, grid, S, vs):
    """Central differences of a 3-D field in float32, one-sided at the mesh ends, in the reference's expression
    order (host/gradient_solver.cpp): the gradient of the densities as stored, [d][z][y][x][s]."""
    X, Y, Z = grid
    D = dens.reshape(Z, Y, X, S)
    out = np.empty((3, Z, Y, X, S), dtype=np.float32)
    for d, (axis, n, h) in enumerate(((2, X, vs[0]), (1, Y, vs[1]), (0, Z, vs[2]))):
        b = np.float32(1) / np.float32(h)
        c = np.float32(1) / np.float32(h * 2)
        A = np.moveaxis(D, axis, 0)
        G = np.moveaxis(out[d], axis, 0)
        G[1:n - 1] = (A[2:] - A[:n - 2]) * c
        G[0] = (A[1] - A[0]) * b
        G[n - 1] = (A[n - 1] - A[n - 2]) * b
    return out.reshape(-1)


def check_gradients_of_stored(cfgs, solvers, tag=""):
    """The batched gradient of 16-bit densities equals the gradient of those densities as loaded (the same FP32
    values the kernel computes with)."""
    for k, (cfg, g) in enumerate(zip(cfgs, solvers)):
        exp = numpy_gradient(g.densities, cfg["grid"], cfg["S"], cfg["voxel_shape"])
        scale = max(float(np.abs(g.densities).max()), 1e-6) / min(cfg["voxel_shape"])
        err = np.abs(g.gradients.astype(np.float64) - exp) / scale
        assert err.max() <= 1e-6, f"{tag} instance {k} grid {cfg['grid']} S {cfg['S']}: gradient {err.max():.2e}"


def check_instances(cfgs, oracles, solvers, steps):
    for k, (cfg, o, g) in enumerate(zip(cfgs, oracles, solvers)):
        tag = f"instance {k} grid {cfg['grid']} S {cfg['S']}"
        assert_close(g.densities, o.densities, f"{tag} densities")
        scale = max(float(np.abs(o.densities).max()), 1e-6) / min(cfg["voxel_shape"])
        gerr = np.abs(g.gradients.astype(np.float64) - o.gradients) / (np.abs(o.gradients) + scale)
        assert gerr.max() <= REL, f"{tag} gradients: worst {gerr.max():.3e}"
        dmax = max(1.0, float(np.abs(o.densities).max()), float(np.abs(cfg["init"]).max()))
        assert_internalized(g.internalized, o.internalized, cfg["agents"], cfg["S"], cfg["dt"], steps, dmax,
                            what=f"{tag} internalized")


def test_batch_matches_instances(pkg, oracle_mod):
    """Eight uneven 3-D instances (substrate counts 1..8, some multiples of 4), 12 steps with two recomputes."""
    from parabiofvm_amd.configs import make_batch_config

    cfgs = make_batch_config(seed=7, count=8, edge=(6, 40), substrates=(1, 8), agent_fraction=0.02)
    batch, oracles, solvers = run_batch(pkg, oracle_mod, cfgs, 12)
    check_instances(cfgs, oracles, solvers, 12)
    batch.close()


def test_batch_mixed_dimensions(pkg, oracle_mod):
    """A 2-D and a 1-D instance next to 3-D ones: their axes beyond dims and their gradients take the
    per-instance path inside the same graph."""
    from parabiofvm_amd.configs import make_batch_config, make_config

    cfgs = make_batch_config(seed=11, count=3, edge=(10, 30), substrates=(2, 5))
    for dims, shape in ((2, (33, 17, 1)), (1, (45, 1, 1))):
        c = make_config(0, seed=5 + dims, shape=shape, substrates=3, agents=20)
        c["dims"] = dims
        c["maxs"] = [shape[d] * 20 if d < dims else 0 for d in range(3)]
        c["agents"]["positions"] = c["agents"]["positions"].reshape(-1, 3)[:, :dims].copy().reshape(-1)
        cfgs.insert(1, c)
    batch, oracles, solvers = run_batch(pkg, oracle_mod, cfgs, 5)
    check_instances(cfgs, oracles, solvers, 5)
    batch.close()


def test_batch_refuses_reinitialized_instance(pkg):
    from parabiofvm_amd.configs import apply_config, make_batch_config

    cfgs = make_batch_config(seed=3, count=2, edge=(8, 12), substrates=(1, 3))
    solvers = []
    for cfg in cfgs:
        g = pkg.DeviceSolver(*args_of(cfg))
        apply_config(cfg, g)
        g.initialize()
        solvers.append(g)
    batch = pkg.Batch(solvers)
    batch.timestep(True, True)
    batch.wait()
    solvers[1].initialize()
    with pytest.raises(pkg.BfvmError) as e:
        batch.timestep(False, True)
    assert e.value.code == -2
    batch.close()


def test_batch_benchmark_entry_points(pkg):
    from parabiofvm_amd.configs import apply_config, make_batch_config

    cfgs = make_batch_config(seed=5, count=4, edge=(16, 48), substrates=(1, 16))
    solvers = []
    for cfg in cfgs:
        g = pkg.DeviceSolver(*args_of(cfg))
        apply_config(cfg, g)
        g.initialize()
        g.store()
        solvers.append(g)
    batch = pkg.Batch(solvers)
    for what in ("sweep_x", "sweep_y", "sweep_z", "gradient", "secretion", "secretion_recompute", "diffusion",
                 "gradient_phase", "timestep"):
        assert batch.benchmark(what, 2) > 0, what
    batch.close()


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_batch_16bit_storage(storage, pkg, oracle_mod):
    """Batches keep 16-bit densities too (one rounding counter for all instances): every instance stays within
    the storage form's error of the FP32 oracle stepping it alone (relative L2 per instance: FP16 1e-3,
    BF16 1e-2, the bounds of the single-instance precision tests)."""
    from parabiofvm_amd.configs import make_batch_config

    cfgs = make_batch_config(seed=13, count=5, edge=(8, 36), substrates=(1, 8), agent_fraction=0.02)
    batch, oracles, solvers = run_batch(pkg, oracle_mod, cfgs, 10, storage=storage)
    bound = 1e-3 if storage == "f16" else 1e-2
    for k, (cfg, o, g) in enumerate(zip(cfgs, oracles, solvers)):
        d = g.densities.astype(np.float64) - o.densities
        rel = np.linalg.norm(d) / max(np.linalg.norm(o.densities.astype(np.float64)), 1e-30)
        assert rel <= bound, f"instance {k} grid {cfg['grid']} S {cfg['S']}: rel L2 {rel:.2e}"
    check_gradients_of_stored(cfgs, solvers, storage)
    batch.close()


def test_batch_vector_widths_with_dirichlet(pkg, oracle_mod):
    """Instances whose rows take 4-, 2- and 1-element vectors per thread, with substrate counts those widths do
    not divide (the vectors of y and z then wrap over the substrates and cross voxels), Dirichlet conditions on
    every face with some substrates' conditions off, and interior Dirichlet voxels; 6 steps against the oracle."""
    from parabiofvm_amd.configs import make_config

    rng = np.random.default_rng(17)
    # (grid, S): x*S % 4 == 0 with odd S (y: 4-wide, wrapping), x*S % 4 == 2 (2-wide), odd x*S (1-wide),
    # S = 6 (x: 2-wide; y, z: 4-wide wrapping), S = 1 (4 voxels per vector), S = 4 and S = 2 (no wrapping)
    shapes = [((36, 20, 18), 3), ((10, 9, 21), 5), ((9, 11, 7), 3), ((8, 6, 19), 6), ((16, 5, 6), 1),
              ((7, 18, 9), 4), ((33, 5, 7), 2)]
    cfgs = []
    for k, (grid, S) in enumerate(shapes):
        c = make_config(1, seed=100 + k, shape=grid, substrates=S, agents=60)
        c["faces"] = {key: (vals, (rng.random(S) < 0.7).astype(np.uint8)) for key, (vals, _) in c["faces"].items()}
        cfgs.append(c)
    batch, oracles, solvers = run_batch(pkg, oracle_mod, cfgs, 6)
    check_instances(cfgs, oracles, solvers, 6)
    batch.close()


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_batch_vector_widths_with_dirichlet_16bit(storage, pkg, oracle_mod):
    """The wrapping and one-element layouts of the previous test in 16-bit storage, where the solves run on the
    stored (scaled) values and only Dirichlet values take the scale: every instance within the storage form's
    error of the FP32 oracle (relative L2: FP16 1e-3, BF16 1e-2)."""
    from parabiofvm_amd.configs import make_config

    rng = np.random.default_rng(19)
    shapes = [((36, 20, 18), 3), ((10, 9, 21), 5), ((9, 11, 7), 3), ((8, 6, 19), 6), ((16, 5, 6), 1),
              ((7, 18, 9), 4)]
    cfgs = []
    for k, (grid, S) in enumerate(shapes):
        c = make_config(1, seed=200 + k, shape=grid, substrates=S, agents=60)
        c["faces"] = {key: (vals, (rng.random(S) < 0.7).astype(np.uint8)) for key, (vals, _) in c["faces"].items()}
        cfgs.append(c)
    batch, oracles, solvers = run_batch(pkg, oracle_mod, cfgs, 6, storage=storage)
    bound = 1e-3 if storage == "f16" else 1e-2
    for k, (cfg, o, g) in enumerate(zip(cfgs, oracles, solvers)):
        d = g.densities.astype(np.float64) - o.densities
        rel = np.linalg.norm(d) / max(np.linalg.norm(o.densities.astype(np.float64)), 1e-30)
        assert rel <= bound, f"instance {k} grid {cfg['grid']} S {cfg['S']}: rel L2 {rel:.2e}"
    check_gradients_of_stored(cfgs, solvers, storage)
    batch.close()


def test_batch_agents_change_between_steps(pkg, oracle_mod):
    """Steps of a batch, then, on instance 1, fewer agents and then more agents than its device allocation
    holds (the batch's secretion plan is rebuilt while steps may still be in flight on the batch's stream, and
    the context calls must wait for them), then more steps: every instance against the oracle stepping it
    alone."""
    from parabiofvm_amd.configs import apply_config, make_batch_config, make_config

    cfgs = make_batch_config(seed=29, count=4, edge=(12, 30), substrates=(2, 8), agent_fraction=0.03)
    oracles, solvers = [], []
    for cfg in cfgs:
        o = oracle_mod.Oracle(*args_of(cfg), dtype=np.float32)
        g = pkg.DeviceSolver(*args_of(cfg))
        for be in (o, g):
            apply_config(cfg, be)
            be.initialize()
        g.store()
        oracles.append(o)
        solvers.append(g)
    batch = pkg.Batch(solvers)
    step = 0

    def steps(n, recompute_first=False):
        nonlocal step
        for i in range(n):
            rc = step % 3 == 0 or (recompute_first and i == 0)  # new agents need their terms computed
            batch.timestep(rc, True)
            for o in oracles:
                o.timestep(rc, True)
            step += 1

    steps(4)  # no wait: the agent changes below are issued while steps may still run
    A = cfgs[1]["agents"]["volumes"].size
    for count in (A // 2, 3 * A + 17):
        grown = make_config(0, seed=900 + count, shape=cfgs[1]["grid"], substrates=cfgs[1]["S"], agents=count)
        agents = {k: v.copy() for k, v in grown["agents"].items()}
        cfgs[1]["agents"] = agents
        for be in (oracles[1], solvers[1]):
            be.set_agents({k: v.copy() for k, v in agents.items()})
        solvers[1].store_agents()
        oracles[1].initialize()  # the oracle's ballots and reduced arrays follow the agent count
        steps(3, recompute_first=True)
    batch.wait()
    for g in solvers:
        g.load()
    for k, (cfg, o, g) in enumerate(zip(cfgs, oracles, solvers)):
        tag = f"instance {k} grid {cfg['grid']} S {cfg['S']}"
        assert_close(g.densities, o.densities, f"{tag} densities")
    batch.close()


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("rows", [4, 8])
def test_batch_stacked_gradient_tiles(storage, rows, pkg, oracle_mod, tmp_path):
    """The batched gradient with stacked tiles, which the batch autotuner need not pick: forced through a
    tuning-cache entry for this batch's shapes (applied by bfvm_batch_autotune), on 4-substrate instances (16-bit:
    with halo lanes) and one-element instances, with 4 and 8 rows per thread.  Each instance's gradient equals the
    gradient of its densities as stored, and in FP32 storage the densities stay within the oracle's tolerance."""
    from parabiofvm_amd import _native
    from parabiofvm_amd.configs import apply_config, make_config

    shapes = [((20, 12, 9), 4), ((11, 10, 8), 3), ((9, 14, 7), 8), ((13, 6, 10), 5)]
    cfgs = [make_config(1, seed=300 + k, shape=g, substrates=S, agents=40) for k, (g, S) in enumerate(shapes)]
    code = {"f32": 0, "bf16": 1, "f16": 2}[storage]
    key = f"{code}:{len(shapes)}" + "".join(f";{g[0]},{g[1]},{g[2]},{S},3" for g, S in shapes)
    line = f"batch {key} " + " ".join(["256"] * 6) + " " + " ".join(["2048"] * 6) + f" 1 4096 {rows} 16 0 1\n"
    path = tmp_path / "tune.txt"
    path.write_text(line)
    _native.tuning_load(str(path))
    oracles, solvers = [], []
    for cfg in cfgs:
        o = oracle_mod.Oracle(*args_of(cfg), dtype=np.float32)
        g = pkg.DeviceSolver(*args_of(cfg))
        g.set_storage(storage)
        for be in (o, g):
            apply_config(cfg, be)
            be.initialize()
        g.store()
        oracles.append(o)
        solvers.append(g)
    batch = pkg.Batch(solvers)
    batch.autotune()
    steps = 3
    for i in range(steps):
        batch.timestep(i == 0, True)
        for o in oracles:
            o.timestep(i == 0, True)
    batch.wait()
    for g in solvers:
        g.load()
    check_gradients_of_stored(cfgs, solvers, f"stacked rows {rows} {storage}")
    if storage == "f32":
        check_instances(cfgs, oracles, solvers, steps)
    batch.close()
