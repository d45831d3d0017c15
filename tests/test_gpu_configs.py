"""The benchmark configurations at their stated sizes, and the persistent sweep loop run over many passes (needs
an MI355X).

  * multi-pass sweeps: with the sweep workgroups capped at 3, every workgroup of every axis loops over dozens of
    line groups (the default launch runs one pass on small meshes, several on 400^3), so the persistent group
    loop and the 16-bit forms' cross-group prefetch run; FP32 must meet the 1e-5 per-voxel bar against the float
    oracle and every storage form must give the same bits as the single-pass launch of the same steps;
  * config 1 at full size (400^3 x 8, 1e6 clustered agents, 6 Dirichlet faces + 1 % interior voxels): 3 graph
    steps in FP32 storage against the float oracle (1e-5 per voxel), and in FP16/BF16 storage against the double
    oracle (per-substrate relative L2 reported; FP16 <= 1e-3, BF16 <= its 1e-2 quantisation floor);
  * config 3 at full size (the 64-instance batch, edges 32..200, 1..16 substrates): 3 batched steps, every
    instance against the float oracle (FP32 storage, 1e-5) and against the double oracle (FP16 storage, 1e-3);
  * config 4 at full size (640^3 x 16 = 4.2e9 densities, past 2^31 elements), the whole stress workload (1e7
    clustered agents, Dirichlet faces and interior voxels) in BF16 and FP16 storage against the double oracle after
    3 steps, and through properties that do not need a CPU run of the whole workload: diffusion-only mass per substrate equals the (1 + decay dt/3)^-3 closed form
    (FP32 and BF16 storage), sampled x/y/z lines equal a double-precision tridiagonal solve of the same line
    (every line independent, so the sampled lines pin the sweep wherever they lie, including element offsets
    beyond 2^31), gradient planes equal the float oracle's on a 3-plane slab around them, and the voxel index of
    the 1e7 clustered agents equals float32 truncation of (p - min)/dx.
"""
import ctypes
import math

import numpy as np
import pytest

from test_gpu_batch import check_instances
from test_gpu_parity import REL, assert_internalized

pytestmark = pytest.mark.gpu

CHUNK = 1 << 25


@pytest.fixture(scope="module")
def pkg():
    import parabiofvm_amd

    parabiofvm_amd.load()
    return parabiofvm_amd


def args_of(cfg):
    return (cfg["dims"], cfg["mins"], cfg["maxs"], cfg["voxel_shape"], cfg["S"], cfg["dt"], cfg["D"], cfg["decay"],
            np.zeros(cfg["S"], dtype=np.float32))


def max_rel(got, exp, tiny=1e-6):
    """Largest |got - exp| / max(|exp|, tiny), in chunks (arrays of billions of elements)."""
    worst, at = 0.0, -1
    for i in range(0, got.size, CHUNK):
        g = got[i:i + CHUNK].astype(np.float64)
        e = exp[i:i + CHUNK].astype(np.float64)
        err = np.abs(g - e) / np.maximum(np.abs(e), tiny)
        k = int(np.argmax(err))
        if not err[k] <= worst:
            worst, at = float(err[k]), i + k
    return worst, at


def rel_l2_by_substrate(got, ref, S):
    num = np.zeros(S)
    den = np.zeros(S)
    rows = CHUNK // S * S
    for i in range(0, got.size, rows):
        g = got[i:i + rows].astype(np.float64).reshape(-1, S)
        r = ref[i:i + rows].astype(np.float64).reshape(-1, S)
        num += ((g - r) ** 2).sum(axis=0)
        den += (r * r).sum(axis=0)
    return np.sqrt(num / np.maximum(den, 1e-300))


def gradient_err(got, exp, scale):
    """Gradient error relative to the density scale over dx (a difference of neighbours may cancel)."""
    worst = 0.0
    for i in range(0, got.size, CHUNK):
        g = got[i:i + CHUNK].astype(np.float64)
        e = exp[i:i + CHUNK].astype(np.float64)
        worst = max(worst, float((np.abs(g - e) / (np.abs(e) + scale)).max()))
    return worst


# ------------------------------------------------------------------------------------------------
# the persistent sweep loop over many passes
# ------------------------------------------------------------------------------------------------

def _run_small(pkg, oracle_mod, storage, workgroups, S, steps=4):
    from parabiofvm_amd.configs import apply_config, make_config

    cfg = make_config(1, seed=31, shape=(48, 40, 36), substrates=S, agents=2500)
    g = pkg.DeviceSolver(*args_of(cfg))
    apply_config(cfg, g)
    g.set_storage(storage)
    g.initialize()
    g.set_sweep_workgroups(workgroups)
    # the sorted reduction sums each voxel's agents in a fixed order, so two runs can be compared bit for bit
    g.set_secretion_reduction("sorted")
    g.store()
    for i in range(steps):
        g.timestep(i % 10 == 0, True)
    g.load()
    out = g.densities.copy(), g.gradients.copy(), g.internalized.copy(), [g.partition(a) for a in range(3)]
    g.close()
    return cfg, out


@pytest.mark.parametrize("S", [8, 3])
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_multipass_sweeps(storage, S, pkg, oracle_mod):
    from parabiofvm_amd.configs import apply_config

    cfg, (dens, grad, intern, parts) = _run_small(pkg, oracle_mod, storage, 3, S)
    # at most 3 workgroups per launch: each loops over all groups of its axis
    X, Y, Z = cfg["grid"]
    groups = []
    for axis, (P, M, lpg) in enumerate(parts):
        lines = X * Y * Z * S // (cfg["grid"][axis] * (4 if S % 4 == 0 else 1))
        groups.append(-(-lines // lpg))
    print(f"storage {storage} S {S}: groups per axis {groups}, partitions {parts}")
    assert min(groups) >= 9, groups  # >= 3 passes per workgroup on every axis
    _, (dens1, grad1, intern1, _) = _run_small(pkg, oracle_mod, storage, 0, S)
    np.testing.assert_array_equal(dens.view(np.uint32), dens1.view(np.uint32))
    np.testing.assert_array_equal(grad.view(np.uint32), grad1.view(np.uint32))
    if storage == "f32":
        o = oracle_mod.Oracle(*args_of(cfg), dtype=np.float32)
        apply_config(cfg, o)
        o.initialize()
        for i in range(4):
            o.timestep(i % 10 == 0, True)
        worst, at = max_rel(dens, o.densities)
        assert worst <= REL, f"densities: worst {worst:.3e} at {at}"
        scale = float(np.abs(o.densities).max()) / min(cfg["voxel_shape"])
        assert gradient_err(grad, o.gradients, scale) <= REL
        dmax = max(1.0, float(np.abs(o.densities).max()))
        assert_internalized(intern, o.internalized, cfg["agents"], S, cfg["dt"], 4, dmax, what="internalized")


# ------------------------------------------------------------------------------------------------
# config 1 at full size
# ------------------------------------------------------------------------------------------------

STEPS_FULL = 3


@pytest.fixture(scope="module")
def config1():
    from parabiofvm_amd.configs import make_config

    return make_config(1)


def _gpu_steps(pkg, cfg, storage, steps):
    from parabiofvm_amd.configs import apply_config

    g = pkg.DeviceSolver(*args_of(cfg))
    apply_config(cfg, g)
    g.set_storage(storage)
    g.initialize()
    g.store()
    for i in range(steps):
        g.timestep(i % 10 == 0, True)
    return g


def _oracle_steps(oracle_mod, cfg, dtype, steps):
    from parabiofvm_amd.configs import apply_config

    o = oracle_mod.Oracle(*args_of(cfg), dtype=dtype)
    apply_config(cfg, o)
    o.initialize()
    for i in range(steps):
        o.timestep(i % 10 == 0, True)
    return o


@pytest.mark.timeout(900)
def test_config1_full_size_f32(config1, pkg, oracle_mod):
    cfg = config1
    print(f"config 1: grid {cfg['grid']} x {cfg['S']}, {cfg['agents']['volumes'].size} agents, "
          f"{cfg['interior'][0].size // 3} interior Dirichlet voxels", flush=True)
    g = _gpu_steps(pkg, cfg, "f32", STEPS_FULL)
    parts = [g.partition(a) for a in range(3)]
    g.load()
    g.close()
    o = _oracle_steps(oracle_mod, cfg, np.float32, STEPS_FULL)
    print(f"partitions {parts}", flush=True)
    worst, at = max_rel(g.densities, o.densities)
    print(f"densities: worst relative error {worst:.3e}", flush=True)
    assert worst <= REL, f"densities: worst {worst:.3e} at {at}"
    scale = float(np.abs(o.densities).max()) / min(cfg["voxel_shape"])
    gw = gradient_err(g.gradients, o.gradients, scale)
    print(f"gradients: worst error relative to the density scale / dx {gw:.3e}", flush=True)
    assert gw <= REL
    dmax = max(1.0, float(np.abs(o.densities).max()))
    assert_internalized(g.internalized, o.internalized, cfg["agents"], cfg["S"], cfg["dt"], STEPS_FULL, dmax,
                        what="internalized")


@pytest.fixture(scope="module")
def config1_f64(config1, oracle_mod):
    o = _oracle_steps(oracle_mod, config1, np.float64, STEPS_FULL)
    dens = o.densities.copy()
    del o
    return dens


@pytest.mark.timeout(900)
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_config1_full_size_16bit(storage, config1, config1_f64, pkg):
    cfg = config1
    g = _gpu_steps(pkg, cfg, storage, STEPS_FULL)
    g.load_densities()
    g.close()
    l2 = rel_l2_by_substrate(g.densities, config1_f64, cfg["S"])
    print(f"config 1, {storage} storage, {STEPS_FULL} steps: relative L2 per substrate vs the double oracle "
          f"{np.array2string(l2, precision=2)}", flush=True)
    assert np.all(np.isfinite(g.densities[:CHUNK]))
    bound = 1e-3 if storage == "f16" else 1e-2
    assert l2.max() <= bound, l2


# ------------------------------------------------------------------------------------------------
# config 3 at full size
# ------------------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def config3():
    from parabiofvm_amd.configs import make_batch_config

    return make_batch_config()


def _batch_steps(pkg, cfgs, storage, steps):
    from parabiofvm_amd.configs import apply_config

    solvers = []
    for c in cfgs:
        g = pkg.DeviceSolver(*args_of(c))
        apply_config(c, g)
        g.set_storage(storage)
        g.initialize()
        g.store()
        solvers.append(g)
    batch = pkg.Batch(solvers)
    for i in range(steps):
        batch.timestep(i % 10 == 0, True)
    batch.wait()
    for g in solvers:
        g.load()
    batch.close()
    return solvers


@pytest.mark.timeout(900)
def test_config3_full_batch_f32(config3, pkg, oracle_mod):
    cfgs = config3
    vs = sum(c["updates_per_step"] for c in cfgs)
    print(f"config 3: {len(cfgs)} instances, {vs / 1e6:.0f}M voxel*substrates", flush=True)
    solvers = _batch_steps(pkg, cfgs, "f32", STEPS_FULL)
    oracles = [_oracle_steps(oracle_mod, c, np.float32, STEPS_FULL) for c in cfgs]
    check_instances(cfgs, oracles, solvers, STEPS_FULL)
    for g in solvers:
        g.close()


@pytest.mark.timeout(900)
def test_config3_full_batch_f16(config3, pkg, oracle_mod):
    cfgs = config3
    solvers = _batch_steps(pkg, cfgs, "f16", STEPS_FULL)
    worst = 0.0
    for k, (c, g) in enumerate(zip(cfgs, solvers)):
        o = _oracle_steps(oracle_mod, c, np.float64, STEPS_FULL)
        l2 = rel_l2_by_substrate(g.densities, o.densities, c["S"])
        worst = max(worst, float(l2.max()))
        assert l2.max() <= 1e-3, f"instance {k} grid {c['grid']} S {c['S']}: rel L2 {l2}"
        g.close()
    print(f"config 3, FP16 storage: worst per-substrate relative L2 vs the double oracle {worst:.2e}")


# ------------------------------------------------------------------------------------------------
# config 4 at full size: properties
# ------------------------------------------------------------------------------------------------

GRID4, S4 = (640, 640, 640), 16


def _hip():
    """The HIP runtime the solver library already loaded (for reading slices of device buffers)."""
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMemcpy.restype = ctypes.c_int
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib


def _lod_banded(n, D, decay, dt, dx, dims=3):
    """The 1-D implicit operator of one sweep (host/diffusion_solver.h:359-373) in solve_banded form."""
    r = dt * D / (dx * dx)
    ab = np.zeros((3, n))
    ab[0, 1:] = -r
    ab[2, :-1] = -r
    ab[1, :] = 1.0 + decay * dt / dims + 2.0 * r
    ab[1, 0] -= r
    ab[1, -1] -= r
    return ab


def _lines(axis, count, rng):
    """Sampled lines along `axis` as (fixed coordinates, substrate), always including lines at the far corner
    (element offsets near 4.2e9)."""
    X, Y, Z = GRID4
    out = [((X - 1, Y - 1, Z - 1), S4 - 1), ((0, Y - 1, Z - 1), 3)]
    for _ in range(count):
        out.append(((int(rng.integers(X)), int(rng.integers(Y)), int(rng.integers(Z))), int(rng.integers(S4))))
    return out


def _take_line(dens, axis, fixed, s):
    X, Y, Z = GRID4
    a = dens.reshape(Z, Y, X, S4)
    x, y, z = fixed
    if axis == 0:
        return a[z, y, :, s]
    if axis == 1:
        return a[z, :, x, s]
    return a[:, y, x, s]


@pytest.fixture(scope="module")
def config4_densities():
    rng = np.random.Generator(np.random.PCG64(404))
    n = GRID4[0] * GRID4[1] * GRID4[2] * S4
    dens = np.empty(n, dtype=np.float32)
    for i in range(0, n, 1 << 28):
        rng.random(out=dens[i:i + (1 << 28)], dtype=np.float32)
    D = np.exp(rng.uniform(np.log(1e3), np.log(1e5), size=S4)).astype(np.float32)
    decay = rng.uniform(0.01, 1.0, size=S4).astype(np.float32)
    return dens, D, decay


def _solver4(pkg, storage, D, decay, dens):
    g = pkg.DeviceSolver(3, [0, 0, 0], [v * 20 for v in GRID4], [20, 20, 20], S4, 0.01, D, decay,
                         np.zeros(S4, dtype=np.float32))
    g.set_storage(storage)
    g.initialize()
    g.densities = dens  # host array handed to the upload without a copy
    g.store_densities()
    return g


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_config4_full_size_properties(storage, config4_densities, pkg, oracle_mod):
    from scipy.linalg import solve_banded

    dens0, D, decay = config4_densities
    dt, dx = 0.01, 20.0
    print(f"config 4: {GRID4} x {S4} = {dens0.size} densities, {storage} storage", flush=True)
    g = _solver4(pkg, storage, D, decay, dens0)
    assert g.densities is dens0
    # the stored values (16-bit forms round at upload)
    g.densities = np.empty_like(dens0)
    g.load_densities()
    cur = g.densities
    mass0 = np.zeros(S4)
    for i in range(0, cur.size, CHUNK):
        mass0 += cur[i:i + CHUNK].reshape(-1, S4).sum(axis=0, dtype=np.float64)
    # one sweep at a time (the benchmark entry point runs its kernel twice: one untimed launch + one timed),
    # each axis checked on sampled lines against two double-precision solves of the line's input
    rng = np.random.Generator(np.random.PCG64(41))
    # BF16: the stochastic rounding of each stored value errs by less than one BF16 ulp (2^-7 relative to the
    # value's power of two), two sweeps per axis
    tol = REL if storage == "f32" else 4 * 2.0 ** -7
    kernels = ("sweep_x", "sweep_y", "sweep_z")
    nxt = np.empty_like(dens0)
    for axis in range(3):
        g.benchmark(kernels[axis], 1)
        g.densities = nxt
        g.load_densities()
        worst = 0.0
        for fixed, s in _lines(axis, 40, rng):
            u0 = _take_line(cur, axis, fixed, s).astype(np.float64)
            ab = _lod_banded(GRID4[axis], float(D[s]), float(decay[s]), dt, dx)
            exp = solve_banded((1, 1), ab, solve_banded((1, 1), ab, u0))
            got = _take_line(nxt, axis, fixed, s).astype(np.float64)
            if storage == "f32":
                err = float((np.abs(got - exp) / np.maximum(np.abs(exp), 1e-6)).max())
            else:
                err = float(np.abs(got - exp).max() / np.abs(exp).max())
            worst = max(worst, err)
        print(f"  axis {axis}: worst sampled-line error {worst:.3e}", flush=True)
        assert worst <= tol, (axis, worst)
        cur, nxt = nxt, cur
    # diffusion-only mass: 2 sweeps per axis, each scales the mass by 1 / (1 + decay dt / 3)
    mass = np.zeros(S4)
    for i in range(0, cur.size, CHUNK):
        mass += cur[i:i + CHUNK].reshape(-1, S4).sum(axis=0, dtype=np.float64)
    expect = mass0 / (1.0 + decay.astype(np.float64) * dt / 3.0) ** 6
    mrel = np.abs(mass - expect) / expect
    print(f"  mass after 6 sweeps: worst relative error {mrel.max():.3e}", flush=True)
    assert mrel.max() <= (1e-5 if storage == "f32" else 1e-4), mrel
    if storage == "f32":
        # gradient planes on both sides of the 2^31-element offset and at the last plane, against the float
        # oracle run on a 3-plane slab (its middle plane has the full mesh's x, y and z differences)
        g.gradient()
        g.wait()
        bufs = g.device_buffers()
        hip = _hip()
        X, Y, Z = GRID4
        plane = X * Y * S4
        for z in (1, 327, 328, Z - 2):
            o = oracle_mod.Oracle(3, [0, 0, 0], [X * 20, Y * 20, 3 * 20], [20, 20, 20], S4, dt, D, decay,
                                  np.zeros(S4), dtype=np.float32)
            o.densities[:] = cur[(z - 1) * plane:(z + 2) * plane]
            o.initialize()
            o.gradient()
            for d in range(3):
                got = np.empty(plane, dtype=np.float32)
                src = bufs.gradients + 4 * (d * (X * Y * Z * S4) + z * plane)
                assert hip.hipMemcpy(got.ctypes.data, src, 4 * plane, 2) == 0
                exp = o.gradients.reshape(3, 3 * plane)[d, plane:2 * plane]
                diff = np.nonzero(got.view(np.uint32) != exp.view(np.uint32))[0]
                assert diff.size == 0, f"gradient plane z={z} component {d}: {diff.size} differ"
        print("  gradient planes z = 1, 327, 328, 638 bit-identical to the oracle", flush=True)
    g.close()


@pytest.mark.timeout(600)
def test_config4_agent_voxels(pkg):
    """The 1e7 clustered agents of the stress workload (64 clusters, sigma 200 um, rejection-sampled into the
    domain): the GPU voxel index equals float32 truncation of (p - min) / dx (src/mesh.cpp:47-53)."""
    from parabiofvm_amd.configs import clustered_positions

    rng = np.random.Generator(np.random.PCG64(7919 * 4 + 1234))
    maxs = [v * 20 for v in GRID4]
    pos = clustered_positions(rng, 10_000_000, 64, 200.0, maxs).astype(np.float32)
    g = pkg.DeviceSolver(3, [0, 0, 0], maxs, [20, 20, 20], 1, 0.01, np.ones(1, np.float32),
                         np.ones(1, np.float32), np.zeros(1, np.float32))
    g.initialize()
    got = g.voxel_positions(pos.reshape(-1))
    g.close()
    exp = ((pos - np.float32(0)) / np.float32(20)).astype(np.int32)
    np.testing.assert_array_equal(got, exp)
    _, counts = np.unique((exp[:, 2].astype(np.int64) * 640 + exp[:, 1]) * 640 + exp[:, 0], return_counts=True)
    print(f"1e7 agents: most agents sharing one voxel {counts.max()}")


@pytest.fixture(scope="module")
def config4_double(oracle_mod):
    """The stress workload (make_config(4): 640^3 x 16, 1e7 agents in 64 clusters, 6 Dirichlet faces + 0.5 %
    interior voxels) and the double oracle (the reference's -DDOUBLE_PRECISION host path) after STEPS_FULL steps
    without the gradient (the double gradient of 4.2e9 densities would take 100 GB); the oracle is kept alive
    and its densities are a view onto its memory."""
    from parabiofvm_amd.configs import apply_config, make_config

    cfg = make_config(4)
    o = oracle_mod.Oracle(*args_of(cfg), dtype=np.float64)
    apply_config(cfg, o)
    o.initialize()
    for i in range(STEPS_FULL):
        o.timestep(i % 10 == 0, with_gradient=False)
    yield cfg, o
    del o


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_config4_full_size_vs_double(storage, config4_double, pkg):
    """The stress workload in 16-bit storage against the double oracle of the same workload after STEPS_FULL
    steps (the sort-by-voxel reduction, as the stress bench runs it): per-substrate relative L2 and largest
    relative error are reported; FP16 (scaled) meets 1e-3, BF16 its 1e-2 quantisation floor."""
    cfg, o = config4_double
    from parabiofvm_amd.configs import apply_config

    g = pkg.DeviceSolver(*args_of(cfg))
    g.set_secretion_reduction("sorted")
    apply_config(cfg, g)
    g.set_storage(storage)
    g.initialize()
    g.store()
    for i in range(STEPS_FULL):
        g.timestep(i % 10 == 0, True)
    g.load_densities()
    g.close()
    ref = o.densities
    l2 = rel_l2_by_substrate(g.densities, ref, cfg["S"])
    worst, at = max_rel(g.densities, ref, tiny=1e-3)
    l2s = " ".join(f"{v:.1e}" for v in l2)
    print(f"config 4, {storage} storage, {STEPS_FULL} steps vs the double oracle: relative L2 per substrate "
          f"[{l2s}] (max {l2.max():.2e}); largest relative error (|ref| floored at 1e-3) {worst:.2e}", flush=True)
    assert np.all(np.isfinite(g.densities[:CHUNK]))
    bound = 1e-3 if storage == "f16" else 1e-2
    assert l2.max() <= bound, l2
