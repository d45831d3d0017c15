"""Deterministic synthetic workloads of the per-timestep solver loop.

Sizes follow the benchmark configurations of this project: a 100^3 x 4 substrate CPU-runnable case, a
400^3 x 8 substrate single large instance with 1e6 clustered agents and Dirichlet faces plus 1% interior
Dirichlet voxels, and reduced versions of the same shapes for parity tests.  All draws come from
numpy.random.Generator(PCG64(seed)) in a fixed order, so a (config, seed, overrides) triple always yields
the same arrays.
"""
import numpy as np

AGENT_VOLUME = 2494.0


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def clustered_positions(rng, count, clusters, sigma, lengths):
    """Agents drawn from `clusters` isotropic Gaussian clusters (centres uniform in the domain), rejection
    sampled into [0, L) on every axis (the reference does not clamp agent positions)."""
    lengths = np.asarray(lengths, dtype=np.float64)
    centres = rng.uniform(0.0, 1.0, size=(clusters, lengths.size)) * lengths
    out = np.empty((count, lengths.size), dtype=np.float64)
    filled = 0
    while filled < count:
        need = count - filled
        k = rng.integers(0, clusters, size=need * 2)
        p = centres[k] + rng.normal(0.0, sigma, size=(need * 2, lengths.size))
        ok = np.all((p >= 0.0) & (p < lengths), axis=1)
        p = p[ok][:need]
        out[filled:filled + p.shape[0]] = p
        filled += p.shape[0]
    return out


def make_config(index, seed=1234, shape=None, substrates=None, agents=None, steps=None, dirichlet=None):
    """Returns a dict describing one workload.

    index 0: 100x100x100, S=4, 10k uniform agents, no Dirichlet, 100 steps.
    index 1: 400x400x400, S=8, 1e6 agents in 16 Gaussian clusters (sigma 400), 6 Dirichlet faces + 1%
             interior Dirichlet voxels, 50 steps with the gradient.
    index 4: 640x640x640, S=16, 1e7 agents in 64 Gaussian clusters (sigma 200), 6 Dirichlet faces + 0.5%
             interior Dirichlet voxels, 20 steps with the gradient (the stress workload, 16-bit storage).
    `shape`, `substrates`, `agents`, `steps`, `dirichlet` override the sizes (parity tests use small ones).
    """
    dx = 20
    if index == 0:
        grid = (100, 100, 100)
        S, A, nsteps, dir_on = 4, 10_000, 100, False
    elif index == 1:
        grid = (400, 400, 400)
        S, A, nsteps, dir_on = 8, 1_000_000, 50, True
    elif index == 4:
        grid = (640, 640, 640)
        S, A, nsteps, dir_on = 16, 10_000_000, 20, True
    else:
        raise ValueError("unknown config")
    if shape is not None:
        grid = tuple(int(v) for v in shape)
    if substrates is not None:
        S = int(substrates)
    if agents is not None:
        A = int(agents)
    if steps is not None:
        nsteps = int(steps)
    if dirichlet is not None:
        dir_on = bool(dirichlet)

    rng = _rng(seed + 7919 * index)
    maxs = [g * dx for g in grid]
    D = np.exp(rng.uniform(np.log(1e3), np.log(1e5), size=S)).astype(np.float32)
    decay = rng.uniform(0.01, 1.0, size=S).astype(np.float32)
    V = grid[0] * grid[1] * grid[2]
    init = rng.random(V * S, dtype=np.float32)

    if index == 0:
        pos = rng.uniform(0.0, 1.0, size=(A, 3)) * np.asarray(maxs, dtype=np.float64)
    elif index == 1:
        sigma = 400.0 * (grid[0] / 400.0 if shape is not None else 1.0)
        pos = clustered_positions(rng, A, 16, max(sigma, dx), maxs)
    else:
        sigma = 200.0 * (grid[0] / 640.0 if shape is not None else 1.0)
        pos = clustered_positions(rng, A, 64, max(sigma, dx), maxs)
    agents_arr = {
        "secretion_rates": rng.uniform(0.0, 1.0, size=A * S).astype(np.float32),
        "uptake_rates": rng.uniform(0.0, 1.0, size=A * S).astype(np.float32),
        "saturation_densities": rng.uniform(0.0, 1.0, size=A * S).astype(np.float32),
        "net_export_rates": np.zeros(A * S, dtype=np.float32),
        "internalized_substrates": np.zeros(A * S, dtype=np.float32),
        "fraction_released_at_death": np.zeros(A * S, dtype=np.float32),
        "volumes": np.full(A, AGENT_VOLUME, dtype=np.float32),
        "positions": pos.astype(np.float32).reshape(-1),
    }

    faces = {}
    interior = None
    if dir_on:
        for d in range(3):
            for mx in (0, 1):
                faces[(d, mx)] = (rng.uniform(0.0, 1.0, size=S).astype(np.float32), np.ones(S, dtype=np.uint8))
        k = max(1, V // (200 if index == 4 else 100))
        lin = np.unique(rng.integers(0, V, size=k))
        x = lin % grid[0]
        y = (lin // grid[0]) % grid[1]
        z = lin // (grid[0] * grid[1])
        vox = np.stack([x, y, z], axis=1).astype(np.int32).reshape(-1)
        vals = rng.uniform(0.0, 1.0, size=lin.size * S).astype(np.float32)
        interior = (vox, vals, np.ones(lin.size * S, dtype=np.uint8))

    return {
        "index": index, "seed": seed, "dims": 3, "mins": [0, 0, 0], "maxs": maxs, "voxel_shape": [dx, dx, dx],
        "grid": grid, "S": S, "dt": 0.01, "D": D, "decay": decay, "init": init, "agents": agents_arr,
        "faces": faces, "interior": interior, "steps": nsteps, "compute_internalized": True,
        "updates_per_step": V * S,
    }


def make_batch_config(seed=4321, count=64, edge=(32, 200), substrates=(1, 16), agent_fraction=0.01):
    """Independent microenvironments of uneven shapes: each mesh edge ~ uniform integer in `edge` (independent
    per axis), substrate count ~ uniform integer in `substrates`, D/decay/initial densities and uniformly placed
    agents (agent_fraction of the voxel count, at least one) drawn as in the 100^3 workload (index 0), no
    Dirichlet conditions.  Returns a list of make_config dicts."""
    rng = _rng(seed)
    out = []
    for i in range(count):
        shape = tuple(int(v) for v in rng.integers(edge[0], edge[1] + 1, size=3))
        S = int(rng.integers(substrates[0], substrates[1] + 1))
        A = max(1, int(round(agent_fraction * shape[0] * shape[1] * shape[2])))
        out.append(make_config(0, seed=seed * 1000 + i, shape=shape, substrates=S, agents=A))
    return out


def apply_config(cfg, backend):
    """Loads a config into a backend with the DeviceSolver / oracle interface (host arrays)."""
    backend.densities[:] = cfg["init"]
    for (d, mx), (vals, conds) in cfg["faces"].items():
        backend.set_boundary(d, mx, vals, conds)
    if cfg["interior"] is not None:
        backend.set_interior(*cfg["interior"])
    backend.set_agents({k: v.copy() for k, v in cfg["agents"].items()})
    backend.set_compute_internalized(cfg["compute_internalized"])
