"""16-bit density storage against the FP64 oracle (needs an MI355X).

Densities live in HBM between kernels as BF16 (the value itself) or FP16 (value * 2^k_s, one power of two
per substrate); every kernel computes in FP32 registers.  Checks:
  * the conversion at upload/download is round-to-nearest-even, bit for bit, on host and device uploads;
  * the FP16 scale puts the largest |density or Dirichlet value| of each substrate in [2^10, 2^11);
  * after N timesteps of the 100^3 x 4 substrate workload (10k agents, secretion -> x/y/z sweeps ->
    gradient, recompute every 10 steps) the per-substrate relative L2 error vs the FP64 oracle is <= 1e-3
    for FP16 storage and <= 1e-5 for FP32 storage.  BF16 keeps 8 significant bits: one rounding of values
    spread over an octave already has a relative RMS error of about 2^-8/sqrt(12)*1.5 ~ 1.7e-3, and every
    sweep rounds again, so BF16 is held to a sanity bound and its errors are reported;
  * the exact sweep refuses 16-bit storage with BFVM_ERR_UNSUPPORTED.
The full 200-step run prints the per-substrate max and L2 relative errors at steps 1, 50 and 200 and, when
BFVM_REPORT_DIR is set, writes them there as JSON.
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# The error bar of the 16-bit forms: relative L2 <= 1e-3 against the FP64 oracle.  FP16 storage (scaled) meets
# it and is the 16-bit form the benchmark reports.  BF16 does not: its 8 significant bits give a quantisation
# floor of about 2-6e-3 (one rounding of values spread over an octave alone is ~1.7e-3 RMS), so BF16 is checked
# only against that floor, 1e-2.  FP32 storage against the FP64 oracle: FP32 coefficient rounding biases every
# sweep by ~1 ulp the same way, so the error grows about linearly with the step count (~1e-7 per step); the
# <= 1e-5 per-voxel bar of the FP32 path is against the FP32 oracle (test_gpu_parity).
ERROR_BAR = 1e-3
BF16_FLOOR = 1e-2
L2_BOUND = {"f32": 1e-4, "f16": ERROR_BAR, "bf16": BF16_FLOOR}


@pytest.fixture(scope="module")
def pkg():
    import parabiofvm_amd

    parabiofvm_amd.load()
    return parabiofvm_amd


def bf16_round(x):
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def f16_scale(max_abs):
    if max_abs == 0 or not math.isfinite(max_abs):
        return 1.0
    _, e = math.frexp(max_abs)
    return math.ldexp(1.0, 11 - e)


def f16_round(x, scale):
    return ((np.asarray(x, dtype=np.float32) * np.float32(scale)).astype(np.float16).astype(np.float32)
            / np.float32(scale))


def solver_for(pkg, grid, S, storage, D=None, decay=None):
    vs = [20, 20, 20]
    D = np.full(S, 1e3, dtype=np.float32) if D is None else D
    decay = np.full(S, 0.1, dtype=np.float32) if decay is None else decay
    s = pkg.DeviceSolver(3, [0, 0, 0], [g * 20 for g in grid], vs, S, 0.01, D, decay, np.zeros(S, np.float32))
    s.set_storage(storage)
    return s


def magnitudes(grid, S, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    V = grid[0] * grid[1] * grid[2]
    x = rng.uniform(-1, 1, size=(V, S)) * (10.0 ** rng.integers(-6, 4, size=S))[None, :]
    x[rng.random(size=(V, S)) < 0.05] = 0
    return x.astype(np.float32).reshape(-1)


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_storage_roundtrip_host(storage, pkg):
    grid, S = (37, 19, 11), 5
    s = solver_for(pkg, grid, S, storage)
    x = magnitudes(grid, S, seed=3)
    s.densities[:] = x
    s.initialize()
    s.store()
    b = s.device_buffers()
    assert b.storage == pkg._native.STORAGE[storage]
    s.densities[:] = np.nan
    s.load()
    xs = x.reshape(-1, S)
    if storage == "bf16":
        exp = bf16_round(x)
    else:
        exp = np.empty_like(xs)
        for k in range(S):
            exp[:, k] = f16_round(xs[:, k], f16_scale(float(np.abs(xs[:, k]).max())))
        exp = exp.reshape(-1)
    assert np.array_equal(s.densities.view(np.uint32), exp.view(np.uint32)), storage


def hip_runtime():
    """The HIP runtime the solver library links (resolved by soname once the library is loaded)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_storage_roundtrip_device(storage, pkg):
    import ctypes

    grid, S = (64, 33, 9), 4
    s = solver_for(pkg, grid, S, storage)
    s.initialize()
    x = magnitudes(grid, S, seed=5)
    hip = hip_runtime()
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), x.nbytes) == 0
    try:
        assert hip.hipMemcpy(dev, x.ctypes.data, x.nbytes, 1) == 0  # host to device
        s.store_densities_device(dev.value)
        s.wait()
    finally:
        assert hip.hipFree(dev) == 0
    s.load()
    host = solver_for(pkg, grid, S, storage)
    host.densities[:] = x
    host.initialize()
    host.store()
    host.load()
    assert np.array_equal(s.densities.view(np.uint32), host.densities.view(np.uint32))


def test_f16_scale_includes_dirichlet_values(pkg):
    grid, S = (12, 10, 8), 2
    s = solver_for(pkg, grid, S, "f16")
    s.set_boundary(0, 0, np.array([3000.0, 0.0], np.float32), np.array([1, 0], np.uint8))
    s.set_boundary(2, 1, np.array([0.0, 7.0], np.float32), np.array([0, 1], np.uint8))
    s.densities[:] = 0.5
    s.initialize()
    s.store()
    s.dirichlet()
    s.load()
    d = s.densities.reshape(grid[2], grid[1], grid[0], S)
    assert np.all(d[:, :, 0, 0] == 3000.0)
    assert np.all(d[-1, :, :, 1] == 7.0)
    assert np.all(d[1:-1, 1:-1, 1:, 0] == 0.5)
    scales = np.zeros(S, np.float32)
    hip = hip_runtime()
    assert hip.hipMemcpy(scales.ctypes.data, s.device_buffers().density_scales, 4 * S, 2) == 0  # device to host
    assert scales.tolist() == [f16_scale(3000.0), f16_scale(7.0)]


def test_exact_sweep_refuses_16bit(pkg):
    s = solver_for(pkg, (10, 9, 8), 2, "bf16")
    s.initialize()
    s.set_sweep_algorithm(pkg.BFVM_SWEEP_EXACT, 0)
    with pytest.raises(pkg.BfvmError) as e:
        s.diffusion()
    assert e.value.code == -4


def rel_errors(got, exp, S):
    g = np.asarray(got, dtype=np.float64).reshape(-1, S)
    e = np.asarray(exp, dtype=np.float64).reshape(-1, S)
    l2 = np.linalg.norm(g - e, axis=0) / np.maximum(np.linalg.norm(e, axis=0), 1e-300)
    mx = (np.abs(g - e) / np.maximum(np.abs(e), 1e-6)).max(axis=0)
    return l2, mx


def run_precision(pkg, oracle_mod, cfg, storage, checkpoints, algorithm=None):
    from parabiofvm_amd.configs import apply_config

    args = (cfg["dims"], cfg["mins"], cfg["maxs"], cfg["voxel_shape"], cfg["S"], cfg["dt"], cfg["D"], cfg["decay"],
            np.zeros(cfg["S"], dtype=np.float32))
    o = oracle_mod.Oracle(*args, dtype=np.float64, threads=min(16, os.cpu_count() or 1))
    g = pkg.DeviceSolver(*args)
    g.set_storage(storage)
    for be in (o, g):
        apply_config(cfg, be)
        be.initialize()
    if algorithm is not None:
        g.set_sweep_algorithm(algorithm, 0)
    g.store()
    S = cfg["S"]
    rows = []
    for i in range(max(checkpoints)):
        o.timestep(i % 10 == 0, True)
        g.timestep(i % 10 == 0, True)
        if i + 1 in checkpoints:
            g.load()
            l2, mx = rel_errors(g.densities, o.densities, S)
            assert cfg["dims"] == 3  # gradients [d][voxel][s]: substrate is the fastest index
            gl2, gmx = rel_errors(g.gradients, o.gradients, S)
            rows.append({"step": i + 1, "storage": storage, "densities_l2": l2.tolist(), "densities_max": mx.tolist(),
                         "gradients_l2": gl2.tolist(), "gradients_max": gmx.tolist()})
            fmt = {"float_kind": lambda v: f"{v:.2e}"}
            print(f"{storage} step {i + 1}: densities L2 {np.array2string(l2, formatter=fmt)} "
                  f"max {np.array2string(mx, formatter=fmt)}; gradients L2 {np.array2string(gl2, formatter=fmt)}",
                  flush=True)
    return rows


def check_rows(rows, storage):
    bound = L2_BOUND[storage]
    for r in rows:
        assert max(r["densities_l2"]) <= bound, r


@pytest.mark.parametrize("algorithm", ["partitioned", "staged"])
@pytest.mark.parametrize("storage", ["bf16", "f16", "f32"])
def test_precision_small_clustered_dirichlet(storage, algorithm, pkg, oracle_mod):
    """Clustered agents with Dirichlet faces and interior voxels (a 40^3 x 8 version of the large instance)."""
    from parabiofvm_amd.configs import make_config

    cfg = make_config(1, seed=21, shape=(40, 40, 40), agents=4000)
    alg = pkg.BFVM_SWEEP_STAGED if algorithm == "staged" else pkg.BFVM_SWEEP_PARTITIONED
    check_rows(run_precision(pkg, oracle_mod, cfg, storage, [1, 20], alg), storage)


@pytest.mark.parametrize("storage", ["bf16", "f16", "f32"])
def test_precision_100cube_200_steps(storage, pkg, oracle_mod):
    """The 100^3 x 4 substrate workload with 10k uniform agents stored in 16 bits, 200 timesteps; per-substrate
    errors vs the FP64 oracle at steps 1, 50 and 200."""
    from parabiofvm_amd.configs import make_config

    cfg = make_config(0, seed=1234)
    rows = run_precision(pkg, oracle_mod, cfg, storage, [1, 50, 200])
    out = os.environ.get("BFVM_REPORT_DIR")
    if out:
        os.makedirs(out, exist_ok=True)
        with open(os.path.join(out, f"precision_100cube_{storage}.json"), "w") as f:
            json.dump({"workload": "100^3 mesh, 4 substrates, 10k agents, dt 0.01", "rows": rows}, f, indent=1)
    check_rows(rows, storage)


def _scales(s, S):
    scales = np.zeros(S, np.float32)
    hip = hip_runtime()
    assert hip.hipMemcpy(scales.ctypes.data, s.device_buffers().density_scales, 4 * S, 2) == 0  # device to host
    return scales


def _secreting_agents(grid, S, A, rng, export):
    pos = rng.uniform(0, 1, size=(A, 3)) * (np.array(grid) * 20.0)
    return {"secretion_rates": rng.uniform(0.5, 1.0, size=A * S), "uptake_rates": rng.uniform(0.0, 0.2, size=A * S),
            "saturation_densities": rng.uniform(50.0, 80.0, size=A * S),
            "net_export_rates": np.full(A * S, export), "internalized_substrates": np.zeros(A * S),
            "fraction_released_at_death": np.zeros(A * S), "volumes": np.full(A, 2494.0),
            "positions": pos.reshape(-1)}


@pytest.mark.parametrize("export", [0.0, 8e8])
def test_f16_headroom_near_zero_field_with_secretion(export, pkg, oracle_mod):
    """FP16 scale safety: densities uploaded near zero (the upload alone would pick a scale near 2^30) with
    agents secreting towards saturation densities of 50..80, and, in the second case, a net export that keeps
    raising the densities past what the first scale could hold.  The scale must come from the agents' targets
    and shrink as the export accumulates: every value finite and within the error bar of the FP64 oracle."""
    grid, S, A = (24, 20, 16), 2, 300
    rng = np.random.Generator(np.random.PCG64(47))
    D = np.array([2e3, 5e4], np.float32)
    decay = np.array([0.1, 0.5], np.float32)
    args = (3, [0, 0, 0], [g * 20 for g in grid], [20, 20, 20], S, 0.01, D, decay, np.zeros(S, np.float32))
    o = oracle_mod.Oracle(*args, dtype=np.float64)
    g = pkg.DeviceSolver(*args)
    g.set_storage("f16")
    agents = _secreting_agents(grid, S, A, rng, export)
    init = rng.uniform(0, 1e-6, size=grid[0] * grid[1] * grid[2] * S)
    for be in (o, g):
        be.densities[:] = init
        be.set_agents({k: v.copy() for k, v in agents.items()})
        be.set_compute_internalized(True)
        be.initialize()
    g.store()
    first = _scales(g, S)
    assert first.max() <= 2.0 ** 5, first  # from the saturation densities (< 2^7), not from the 1e-6 upload
    steps = 40
    for i in range(steps):
        o.timestep(i % 10 == 0, True)
        g.timestep(i % 10 == 0, True)
    g.load()
    assert np.all(np.isfinite(g.densities))
    l2, mx = rel_errors(g.densities, o.densities, S)
    last = _scales(g, S)
    print(f"export {export:g}: scales {first} -> {last}, max density {np.abs(o.densities).max():.3e}, "
          f"rel L2 {l2}")
    assert l2.max() <= ERROR_BAR, l2
    if export > 0:
        assert np.abs(o.densities).max() * first.max() > 65504.0  # the first scale would have overflowed
        assert np.all(last < first), (first, last)
