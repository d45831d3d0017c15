"""HIP path vs the CPU oracle and the reference test-suite's expectations (needs an MI355X).

Tolerances (north-star contract for the FP32 path):
  * exact sweep algorithm, Dirichlet, gradient, voxel indices: bit-identical to the float oracle;
  * partitioned sweep algorithm (FP64-derived coefficients, different elimination order): relative error
    <= 1e-5 per voxel, |gpu - cpu| <= 1e-5 * max(|cpu|, 1e-6);
  * secretion/uptake: <= 1e-5 relative (float atomics sum in a different order), voxel indices exact;
  * golden values: within 1e-5 relative + 4 float ULPs of the reference's printed expectation.
"""
import ctypes

import numpy as np
import pytest

from golden_runner import float_ulp_diff, load_cases, run_case

pytestmark = pytest.mark.gpu

FIXTURES, CASES = load_cases()
REL = 1e-5
TINY = 1e-6


@pytest.fixture(scope="module")
def pkg():
    import parabiofvm_amd

    parabiofvm_amd.load()
    return parabiofvm_amd


class GpuBackend:
    def __init__(self, pkg, algorithm, chunk, *args):
        self.s = pkg.DeviceSolver(*args)
        self._alg = (algorithm, chunk)
        self.grid = self.s.grid

    def initialize(self):
        self.s.initialize()
        self.s.set_sweep_algorithm(*self._alg)

    def __getattr__(self, name):
        return getattr(self.s, name)

    @property
    def densities(self):
        return self.s.densities

    @property
    def gradients(self):
        return self.s.gradients

    @property
    def internalized(self):
        return self.s.internalized


def assert_close(got, exp, rel=REL, tiny=TINY, what=""):
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    assert got.shape == exp.shape, what
    err = np.abs(got - exp) / np.maximum(np.abs(exp), tiny)
    bad = np.nonzero(~(err <= rel))[0]
    assert bad.size == 0, (f"{what}: {bad.size}/{got.size} beyond {rel:g}; worst rel {np.nanmax(err):.3e} at "
                           f"{bad[:5]} got {got[bad[:5]]} expected {exp[bad[:5]]}")


def internalized_scale(agents, S, dt, steps, dmax):
    """Sum of the magnitudes of the terms an agent's internalized substrate accumulates over `steps` calls:
    vv*(den*D + num + |fac|) = (U+S)*dt*vol*D + S*T*dt*vol + |N|*dt per call (host/cell_solver.cpp:137-152).
    I is the running difference of these terms and is often much smaller than they are, so its rounding error
    is bounded relative to this sum (the componentwise bound of a floating-point sum), not relative to I."""
    vol = np.repeat(np.asarray(agents["volumes"], dtype=np.float64), S)
    sec = np.asarray(agents["secretion_rates"], dtype=np.float64)
    upt = np.asarray(agents["uptake_rates"], dtype=np.float64)
    sat = np.asarray(agents["saturation_densities"], dtype=np.float64)
    nex = np.asarray(agents["net_export_rates"], dtype=np.float64)
    return steps * ((upt + sec) * dt * vol * dmax + sec * sat * dt * vol + np.abs(nex) * dt)


def assert_internalized(got, exp, agents, S, dt, steps, dmax, what=""):
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    err = np.abs(got - exp) / (np.abs(exp) + internalized_scale(agents, S, dt, steps, dmax))
    bad = np.nonzero(~(err <= REL))[0]
    assert bad.size == 0, (f"{what}: {bad.size}/{got.size} beyond {REL:g}; worst {np.nanmax(err):.3e} at "
                           f"{bad[:5]} got {got[bad[:5]]} expected {exp[bad[:5]]}")


def assert_bitwise(got, exp, what=""):
    g = np.asarray(got, dtype=np.float32)
    e = np.asarray(exp, dtype=np.float32)
    diff = np.nonzero(g.view(np.uint32) != e.view(np.uint32))[0]
    assert diff.size == 0, (f"{what}: {diff.size}/{g.size} differ, e.g. {diff[:5]} got {g[diff[:5]]} "
                            f"expected {e[diff[:5]]}")


# ------------------------------------------------------------------------------------------------
# the reference test-suite's cases (tests/golden/reference_expectations.json)
# ------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("algorithm", ["partitioned", "exact"])
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_matches_reference_expectations(case, algorithm, pkg):
    alg = pkg.BFVM_SWEEP_PARTITIONED if algorithm == "partitioned" else pkg.BFVM_SWEEP_EXACT
    make = lambda *a: GpuBackend(pkg, alg, 0, *a)
    n = 0
    for label, got, exp in run_case(case, FIXTURES, make):
        got = np.asarray(got, dtype=np.float64)
        exp = np.asarray(exp, dtype=np.float64)
        ok = (np.abs(got - exp) <= REL * np.abs(exp)) | (float_ulp_diff(got, exp) <= 4)
        bad = np.nonzero(~ok)[0]
        assert bad.size == 0, f"{label}: {bad.size} bad, e.g. got {got[bad[:5]]} expected {exp[bad[:5]]}"
        n += 1
    assert n > 0


# ------------------------------------------------------------------------------------------------
# random inputs vs the float oracle
# ------------------------------------------------------------------------------------------------

SHAPES = [
    ("3d_small", 3, (3, 4, 5), 2),
    ("3d_partitioned", 3, (70, 45, 33), 3),
    ("3d_unit_axis", 3, (17, 1, 9), 2),
    ("3d_two", 3, (2, 5, 2), 4),
    ("2d", 2, (150, 37, 1), 5),
    ("1d", 1, (129, 1, 1), 3),
    ("3d_wide_s", 3, (40, 12, 10), 16),
    ("3d_vec4", 3, (64, 40, 36), 8),
]


def random_problem(dims, shape, S, seed, faces=True, interior=True):
    rng = np.random.Generator(np.random.PCG64(seed))
    dx = (20, 16, 25)
    mins = (-40, 0, 100)
    maxs = [mins[d] + shape[d] * dx[d] for d in range(3)]
    D = np.exp(rng.uniform(np.log(1e2), np.log(1e5), size=S)).astype(np.float32)
    decay = rng.uniform(0.01, 1.0, size=S).astype(np.float32)
    init = rng.uniform(0, 1, size=S).astype(np.float32)
    V = shape[0] * shape[1] * shape[2]
    dens = rng.uniform(0, 1, size=V * S).astype(np.float32)
    fc = {}
    if faces:
        for d in range(dims):
            for mx in (0, 1):
                if rng.uniform() < 0.8:
                    fc[(d, mx)] = (rng.uniform(0, 2, size=S).astype(np.float32),
                                   (rng.uniform(size=S) < 0.7).astype(np.uint8))
    it = None
    if interior:
        k = max(1, V // 20)
        coords = np.stack([rng.integers(0, shape[d], size=k) for d in range(dims)], axis=1).astype(np.int32)
        vals = rng.uniform(0, 3, size=k * S).astype(np.float32)
        conds = (rng.uniform(size=k * S) < 0.6).astype(np.uint8)
        # distinct voxels (duplicates are applied in list order by the oracle but are racy in the reference)
        _, first = np.unique(coords, axis=0, return_index=True)
        first = np.sort(first)
        coords = coords[first]
        vals = vals.reshape(k, S)[first].reshape(-1)
        conds = conds.reshape(k, S)[first].reshape(-1)
        it = (coords.reshape(-1), vals, conds)
    return dict(dims=dims, mins=mins, maxs=maxs, vs=dx, S=S, dt=0.01, D=D, decay=decay, init=init, dens=dens,
                faces=fc, interior=it)


def build_pair(pkg, oracle_mod, p, algorithm=None, chunk=0, dtype=np.float32):
    args = (p["dims"], p["mins"], p["maxs"], p["vs"], p["S"], p["dt"], p["D"], p["decay"], p["init"])
    o = oracle_mod.Oracle(*args, dtype=dtype)
    g = pkg.DeviceSolver(*args)
    for be in (o, g):
        be.densities[:] = p["dens"]
        for (d, mx), (v, c) in p["faces"].items():
            be.set_boundary(d, mx, v, c)
        if p["interior"] is not None:
            be.set_interior(*p["interior"])
    o.initialize()
    g.initialize()
    if algorithm is not None:
        g.set_sweep_algorithm(algorithm, chunk)
    return o, g


@pytest.mark.parametrize("name,dims,shape,S", SHAPES, ids=[s[0] for s in SHAPES])
def test_gpu_exact_coefficients_bitwise(name, dims, shape, S, pkg, oracle_mod):
    p = random_problem(dims, shape, S, seed=11)
    o, g = build_pair(pkg, oracle_mod, p)
    for axis in range(dims):
        bo, co, eo = o.coefficients(axis)
        bg, cg, eg = g.exact_coefficients(axis)
        assert_bitwise(bg, bo, f"b axis {axis}")
        if shape[axis] > 1:
            assert_bitwise(cg, co, f"c axis {axis}")
            assert_bitwise(eg, eo, f"e axis {axis}")


@pytest.mark.parametrize("name,dims,shape,S", SHAPES, ids=[s[0] for s in SHAPES])
def test_gpu_exact_diffusion_bitwise(name, dims, shape, S, pkg, oracle_mod):
    p = random_problem(dims, shape, S, seed=23)
    o, g = build_pair(pkg, oracle_mod, p, algorithm=pkg.BFVM_SWEEP_EXACT)
    g.store()
    for _ in range(3):
        o.diffusion()
        g.diffusion()
    g.load()
    assert_bitwise(g.densities, o.densities, "densities after 3 diffusion solves")


@pytest.mark.parametrize("chunk", [0, 8, 16, 48, 64])
@pytest.mark.parametrize("name,dims,shape,S", SHAPES, ids=[s[0] for s in SHAPES])
def test_gpu_partitioned_diffusion(name, dims, shape, S, chunk, pkg, oracle_mod):
    p = random_problem(dims, shape, S, seed=37)
    o, g = build_pair(pkg, oracle_mod, p, algorithm=pkg.BFVM_SWEEP_PARTITIONED, chunk=chunk)
    g.store()
    for _ in range(3):
        o.diffusion()
        g.diffusion()
    g.load()
    assert_close(g.densities, o.densities, what=f"chunk {chunk} partition {[g.partition(a) for a in range(dims)]}")


STAGED_SHAPES = SHAPES + [
    ("3d_odd_s", 3, (37, 23, 19), 3),    # rows not 16-byte aligned: head/tail elements of every segment
    ("3d_s5_long", 3, (21, 130, 9), 5),  # long y lines, misaligned segments
    ("3d_s7", 3, (50, 31, 41), 7),
]


@pytest.mark.parametrize("name,dims,shape,S", STAGED_SHAPES, ids=[s[0] for s in STAGED_SHAPES])
def test_gpu_staged_diffusion(name, dims, shape, S, pkg, oracle_mod):
    """The partitioned algorithm on LDS-staged tiles (any substrate count, unaligned rows), Dirichlet faces
    and interior voxels included."""
    p = random_problem(dims, shape, S, seed=43)
    o, g = build_pair(pkg, oracle_mod, p, algorithm=pkg.BFVM_SWEEP_STAGED, chunk=0)
    g.store()
    for _ in range(3):
        o.diffusion()
        g.diffusion()
    g.load()
    assert_close(g.densities, o.densities, what=f"staged partition {[g.partition(a) for a in range(dims)]}")


@pytest.mark.parametrize("name,dims,shape,S", [s for s in SHAPES if min(s[2][:s[1]]) >= 2],
                         ids=[s[0] for s in SHAPES if min(s[2][:s[1]]) >= 2])
def test_gpu_gradient_bitwise(name, dims, shape, S, pkg, oracle_mod):
    p = random_problem(dims, shape, S, seed=41)
    o, g = build_pair(pkg, oracle_mod, p)
    g.store()
    o.gradient()
    g.gradient()
    g.load()
    assert_bitwise(g.gradients, o.gradients, "gradients")


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("shape,S", [((3, 4, 5), 2), ((70, 45, 33), 3), ((64, 40, 36), 8), ((100, 37, 9), 16),
                                     ((17, 2, 9), 4), ((130, 71, 6), 1), ((7, 9, 5), 124), ((9, 6, 5), 128)])
def test_gpu_gradient_stacked_tiles(shape, S, storage, pkg, oracle_mod):
    """The gradient's stacked tile layout (waves stacked along y, the rows where they meet passed through LDS) on
    meshes whose rows and columns do not fill whole tiles: bit-identical to the float oracle (FP32 storage) and
    to the other tile layout (FP16 storage), at several workgroup targets (z split into chunks of different
    lengths, so tiles start and end at every plane parity), in both workgroup orders (the plane order covers
    every tile exactly once: the output is NaN-filled before each launch).  16-bit stacked tiles give S / 4
    lanes at each end of a wave to halo columns: 124 substrates leave 2 computing lanes per wave, and at 128
    the launch takes the side-by-side layout instead."""
    p = random_problem(3, shape, S, seed=45)
    o, g = build_pair(pkg, oracle_mod, p)
    if storage != "f32":
        g.set_storage(storage)
        g.initialize()
    g.store()
    o.gradient()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    poison = np.full(g.gradients.size, np.nan, dtype=np.float32)
    results = []
    for order in (0, 1):
        for stacked in (False, True):
            for groups in (1, 3, 64, 4096):
                g.set_gradient_order(order)
                g.set_gradient_stacked(stacked)
                g.set_gradient_tuning(groups)
                # NaN in the device gradient buffer (the context is idle after load), so a tile left out shows
                assert hip.hipMemcpy(g.device_buffers().gradients, poison.ctypes.data, poison.nbytes, 1) == 0
                g.gradient()
                g.load()
                if storage == "f32":
                    assert_bitwise(g.gradients, o.gradients, f"order {order} stacked {stacked} groups {groups}")
                results.append(g.gradients.copy())
    for r in results[1:]:
        assert_bitwise(r, results[0], "tile layouts, workgroup orders and z chunkings")
    with pytest.raises(RuntimeError):
        g.set_gradient_order(2)


@pytest.mark.parametrize("name,dims,shape,S", SHAPES, ids=[s[0] for s in SHAPES])
def test_gpu_dirichlet_bitwise(name, dims, shape, S, pkg, oracle_mod):
    p = random_problem(dims, shape, S, seed=43)
    o, g = build_pair(pkg, oracle_mod, p)
    g.store()
    o.dirichlet()
    g.dirichlet()
    g.load()
    assert_bitwise(g.densities, o.densities, "densities after Dirichlet")


def test_gpu_voxel_index_boundaries(pkg, oracle_mod):
    """mesh.cpp:47-53 bit-exact: positions at every voxel boundary, one float ULP either side, and outside."""
    p = random_problem(3, (50, 7, 3), 1, seed=3, faces=False, interior=False)
    o, g = build_pair(pkg, oracle_mod, p)
    pts = []
    rng = np.random.Generator(np.random.PCG64(5))
    for d in range(3):
        n = (p["maxs"][d] - p["mins"][d]) // p["vs"][d]
        e = (p["mins"][d] + np.arange(-2, n + 3) * p["vs"][d]).astype(np.float32)
        cand = np.concatenate([e, np.nextafter(e, np.float32(-np.inf)), np.nextafter(e, np.float32(np.inf)),
                               e - np.float32(0.5) * p["vs"][d]])
        pts.append(cand)
    k = max(len(c) for c in pts) * 8
    P = np.stack([rng.choice(pts[d], size=k) for d in range(3)], axis=1).astype(np.float32)
    P = np.concatenate([P, np.array([[np.nan, 0, 0], [np.inf, 0, 0], [0, -np.inf, 0]], dtype=np.float32)])
    got = g.voxel_positions(P.reshape(-1))
    exp = o.voxel_positions(P.reshape(-1))
    assert np.array_equal(got, exp)


# ------------------------------------------------------------------------------------------------
# secretion / uptake and the full timestep
# ------------------------------------------------------------------------------------------------

def agents_for(p, A, seed, crowd=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    S, dims = p["S"], p["dims"]
    lo = np.array(p["mins"][:dims], dtype=np.float64)
    hi = np.array(p["maxs"][:dims], dtype=np.float64)
    if crowd:
        # many agents share few voxels
        pos = lo + (hi - lo) * rng.uniform(0, 1, size=(max(1, A // 10), dims))
        pos = pos[rng.integers(0, pos.shape[0], size=A)]
    else:
        pos = lo + (hi - lo) * rng.uniform(0, 1, size=(A, dims))
    return {
        "secretion_rates": rng.uniform(0, 1, size=A * S), "uptake_rates": rng.uniform(0, 1, size=A * S),
        "saturation_densities": rng.uniform(0, 1, size=A * S), "net_export_rates": rng.uniform(0, 0.1, size=A * S),
        "internalized_substrates": rng.uniform(-1, 1, size=A * S),
        "fraction_released_at_death": rng.uniform(0, 1, size=A * S),
        "volumes": rng.uniform(500, 4000, size=A), "positions": pos.reshape(-1),
    }


@pytest.mark.parametrize("reduction", ["atomic", "sorted"])
@pytest.mark.parametrize("ci", [True, False])
@pytest.mark.parametrize("crowd", [False, True])
@pytest.mark.parametrize("dims,shape", [(3, (20, 15, 10)), (2, (30, 20, 1)), (1, (40, 1, 1))])
def test_gpu_cell_solver(dims, shape, crowd, ci, reduction, pkg, oracle_mod):
    """Secretion/uptake with and without internalization, sparse and crowded voxels, both ways of combining
    agents that share a voxel (claim + atomics, or sorted by voxel with segmented sums)."""
    p = random_problem(dims, shape, 3, seed=53, faces=False, interior=False)
    o, g = build_pair(pkg, oracle_mod, p)
    g.set_secretion_reduction(reduction)
    a = agents_for(p, 400, seed=59, crowd=crowd)
    oa = o.set_agents(a)
    g.set_agents(a)
    for be in (o, g):
        be.set_compute_internalized(ci)
    o.initialize()
    g.initialize()
    g.store()
    dmax = max(1.0, float(np.abs(o.densities).max()))
    for step, rc in enumerate([True, False, True]):
        o.cell(rc)
        g.cell(rc)
        g.load()
        assert_close(g.densities, o.densities, what=f"densities step {step}")
        dmax = max(dmax, float(np.abs(o.densities).max()))
        if ci:
            assert_internalized(g.internalized, o.internalized, a, p["S"], p["dt"], step + 1, dmax,
                                what=f"internalized step {step}")
    # release at death (host/cell_solver.cpp:301-346)
    for idx in (0, 7):
        o.release(idx)
        g.release(idx)
    g.load()
    assert_close(g.densities, o.densities, what="densities after release")
    if ci:
        assert_internalized(g.internalized, o.internalized, a, p["S"], p["dt"], 3, dmax,
                            what="internalized after release")


@pytest.mark.parametrize("algorithm,reduction", [("partitioned", "atomic"), ("exact", "atomic"),
                                                 ("partitioned", "sorted")])
def test_gpu_timestep_graph(algorithm, reduction, pkg, oracle_mod):
    """Ten timesteps (secretion -> diffusion -> gradient, recompute every 10 steps as in main.cpp:100) replayed
    from the captured HIP graph, against the oracle stepping the same order."""
    from parabiofvm_amd.configs import apply_config, make_config

    cfg = make_config(1, seed=99, shape=(48, 40, 36), agents=3000)
    args = (cfg["dims"], cfg["mins"], cfg["maxs"], cfg["voxel_shape"], cfg["S"], cfg["dt"], cfg["D"], cfg["decay"],
            np.zeros(cfg["S"], dtype=np.float32))
    o = oracle_mod.Oracle(*args, dtype=np.float32)
    g = pkg.DeviceSolver(*args)
    for be in (o, g):
        apply_config(cfg, be)
        be.initialize()
    alg = pkg.BFVM_SWEEP_PARTITIONED if algorithm == "partitioned" else pkg.BFVM_SWEEP_EXACT
    g.set_sweep_algorithm(alg, 0)
    g.set_secretion_reduction(reduction)
    g.store()
    for i in range(10):
        o.timestep(i % 10 == 0, True)
        g.timestep(i % 10 == 0, True)
    g.load()
    assert_close(g.densities, o.densities, what="densities")
    # a gradient is a difference of neighbouring densities, so its error is bounded relative to the density
    # scale over dx, not relative to the (possibly cancelling) difference itself
    scale = np.abs(o.densities).max() / min(cfg["voxel_shape"])
    gerr = np.abs(g.gradients.astype(np.float64) - o.gradients) / (np.abs(o.gradients) + scale)
    assert gerr.max() <= REL, f"gradients: worst {gerr.max():.3e}"
    dmax = max(1.0, float(np.abs(o.densities).max()), float(np.abs(cfg["init"]).max()))
    assert_internalized(g.internalized, o.internalized, cfg["agents"], cfg["S"], cfg["dt"], 10, dmax,
                        what="internalized")


@pytest.mark.parametrize("axis,chunk,lines", [(0, 16, 1), (0, 32, 3), (0, 64, 2), (1, 16, 16), (1, 48, 64),
                                              (2, 32, 16), (2, 64, 32), (2, 8, 16)])
def test_gpu_partitioned_tuning(axis, chunk, lines, pkg, oracle_mod):
    """Every per-axis (chunk, lines per workgroup) the tuner may pick gives the same parity."""
    p = random_problem(3, (70, 45, 33), 3, seed=61)
    o, g = build_pair(pkg, oracle_mod, p)
    g.set_sweep_tuning(axis, chunk, lines)
    g.store()
    o.diffusion()
    g.diffusion()
    g.load()
    assert_close(g.densities, o.densities, what=f"axis {axis} chunk {chunk} lines {lines}: {g.partition(axis)}")


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_gpu_wide_workgroups(axis, pkg, oracle_mod):
    """4-substrate sweeps in workgroups of more than 512 threads (400-voxel lines, P = 25 blocks, 32 lines or x
    rows per workgroup: 800 threads), with Dirichlet faces and interior voxels, against the float oracle."""
    shape = [16, 12, 10]
    shape[axis] = 400
    p = random_problem(3, tuple(shape), 4, seed=67)
    o, g = build_pair(pkg, oracle_mod, p)
    g.set_sweep_tuning(axis, 16, 32)
    P, M, lpg = g.partition(axis)
    threads = lpg * P if axis else (lpg // 1) * P
    assert threads > 512, (P, M, lpg)
    g.store()
    for _ in range(2):
        o.diffusion()
        g.diffusion()
    g.load()
    assert_close(g.densities, o.densities, what=f"axis {axis}: partition {(P, M, lpg)}")


def test_gpu_sorted_reduction_agents_outside_and_reupload(pkg, oracle_mod):
    """Sorted reduction with agents outside the mesh (ignored), then a second upload that moves agents and
    switches back to the claim scheme: both uploads rebuild the agent -> voxel tables."""
    p = random_problem(3, (12, 9, 7), 2, seed=71, faces=False, interior=False)
    o, g = build_pair(pkg, oracle_mod, p)
    a = agents_for(p, 300, seed=73, crowd=True)
    pos = a["positions"].reshape(-1, 3)
    pos[:20] = np.array(p["mins"]) - 50.0  # outside
    for be in (o, g):
        be.set_agents(a)
        be.set_compute_internalized(True)
    g.set_secretion_reduction("sorted")
    o.initialize()
    g.initialize()
    g.store()
    for rc in (True, False):
        o.cell(rc)
        g.cell(rc)
    g.load()
    assert_close(g.densities, o.densities, what="sorted")
    a2 = agents_for(p, 250, seed=79, crowd=True)
    for be in (o, g):
        be.set_agents(a2)
    g.set_secretion_reduction("atomic")
    o.initialize()
    g.store_agents()
    g.densities[:] = o.densities
    g.store()
    o.cell(True)
    g.cell(True)
    g.load()
    assert_close(g.densities, o.densities, what="re-uploaded")


@pytest.mark.parametrize("dims,shape,S", [(3, (19, 11, 7), 3), (2, (33, 17, 1), 4), (1, (50, 1, 1), 2)])
def test_gpu_bulk_bitwise(dims, shape, S, pkg, oracle_mod):
    """bulk_solver (host/bulk_solver.cpp:23-35) with per-voxel supply/uptake/target fields: bit-identical to
    the float oracle over several solves."""
    p = random_problem(dims, shape, S, seed=83, faces=False, interior=False)
    o, g = build_pair(pkg, oracle_mod, p)
    rng = np.random.Generator(np.random.PCG64(89))
    n = o.densities.size
    sup, upt, tgt = (rng.uniform(0, 20, size=n).astype(np.float32) for _ in range(3))
    g.bulk_store_rates(sup, upt, tgt)
    g.store()
    for _ in range(3):
        o.bulk(sup, upt, tgt)
        g.bulk()
    g.load()
    assert_bitwise(g.densities, o.densities, "bulk")


def test_gpu_bulk_requires_rates(pkg, oracle_mod):
    p = random_problem(3, (5, 4, 3), 2, seed=97, faces=False, interior=False)
    _, g = build_pair(pkg, oracle_mod, p)
    with pytest.raises(pkg.BfvmError) as e:
        g.bulk()
    assert e.value.code == -2


@pytest.mark.parametrize("reduction", ["atomic", "sorted"])
def test_gpu_stress_shape_timestep(reduction, pkg, oracle_mod):
    """A reduced stress workload: 16 substrates (4-substrate vectors with 4 substrate groups per voxel), heavy
    same-voxel contention (64 Gaussian clusters), Dirichlet faces and interior voxels, 6 graph-replayed steps
    against the oracle."""
    from parabiofvm_amd.configs import apply_config, make_config

    cfg = make_config(4, seed=17, shape=(48, 40, 36), agents=60000)
    args = (cfg["dims"], cfg["mins"], cfg["maxs"], cfg["voxel_shape"], cfg["S"], cfg["dt"], cfg["D"], cfg["decay"],
            np.zeros(cfg["S"], dtype=np.float32))
    o = oracle_mod.Oracle(*args, dtype=np.float32)
    g = pkg.DeviceSolver(*args)
    for be in (o, g):
        apply_config(cfg, be)
        be.initialize()
    g.set_secretion_reduction(reduction)
    g.store()
    steps = 6
    for i in range(steps):
        o.timestep(i % 10 == 0, True)
        g.timestep(i % 10 == 0, True)
    g.load()
    assert_close(g.densities, o.densities, what=f"densities ({reduction})")
    dmax = max(1.0, float(np.abs(o.densities).max()), float(np.abs(cfg["init"]).max()))
    assert_internalized(g.internalized, o.internalized, cfg["agents"], cfg["S"], cfg["dt"], steps, dmax,
                        what="internalized")
