"""Drives the reference test-suite's cases (tests/golden/reference_expectations.json) through a backend.

A backend mirrors the reference's microenvironment + solver facade: it is built
from the mesh and substrate parameters, exposes host `densities` / `gradients`
arrays, `set_boundary` / `set_interior` / `set_agents`, `initialize()`,
`store()` / `load()` (the device solver's store_data_to_solver /
load_data_from_solver, no-ops on the CPU oracle) and the solve calls
`diffusion()`, `dirichlet()`, `gradient()`, `cell(recompute)`.

Each case runs the same calls the reference test makes (test/test.cpp and
test/test_device.cpp, cited per case in the JSON) and returns (got, expected)
arrays; the caller applies its tolerance.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "reference_expectations.json")


def load_cases():
    with open(GOLDEN) as f:
        data = json.load(f)
    return data["fixtures"], data["cases"]


def grid_of(mesh):
    g = [1, 1, 1]
    for d in range(mesh["dims"]):
        g[d] = (mesh["maxs"][d] - mesh["mins"][d] + mesh["voxel_shape"][d] - 1) // mesh["voxel_shape"][d]
    return g


def default_agent_arrays(specs, dims, S=2):
    """test/utils.cpp:204-222 set_default_agent_values (substrate 0 gets the rates, substrate 1 zeros)."""
    n = len(specs)
    a = {k: np.zeros(n * S) for k in ("secretion_rates", "saturation_densities", "uptake_rates",
                                      "net_export_rates", "internalized_substrates", "fraction_released_at_death")}
    a["volumes"] = np.zeros(n)
    a["positions"] = np.zeros(n * dims)
    for i, sp in enumerate(specs):
        o = sp["rates_offset"]
        a["secretion_rates"][i * S] = o + 100
        a["uptake_rates"][i * S] = o + 200
        a["saturation_densities"][i * S] = o + 300
        a["net_export_rates"][i * S] = o + 400
        a["volumes"][i] = sp["volume"]
        a["positions"][i * dims:(i + 1) * dims] = sp["position"][:dims]
    return a


def expand_agents(case):
    if "agents" in case:
        return case["agents"]
    spec = case["agents_per_voxel_center"]
    X = grid_of(case["mesh"])[0]
    vs = case["mesh"]["voxel_shape"][0]
    out = []
    for i in range(X):
        for _ in range(spec["per_voxel"]):
            # mesh.cpp:56-60 voxel_center: x*dx + dx/2.0 + min (computed in double, stored as real)
            out.append({"rates_offset": spec["rates_offset"], "volume": spec["volume"],
                        "position": [i * vs + vs / 2.0]})
    return out


def bruteforce_internalized(dens, a, S, dt, vv, dims, vs, mins):
    """test/utils.cpp:134-162 compute_expected_agent_internalized_1d: each agent with its own terms and the
    density of its voxel before the update."""
    n = a["volumes"].shape[0]
    out = np.zeros(n * S, dtype=np.float64)
    for i in range(n):
        x = int((a["positions"][i * dims] - mins[0]) / vs[0])
        for s in range(S):
            k = i * S + s
            num = a["secretion_rates"][k] * a["saturation_densities"][k] * dt * a["volumes"][i]
            den = (a["secretion_rates"][k] + a["uptake_rates"][k]) * dt * a["volumes"][i] / vv
            fac = a["net_export_rates"][k] * dt
            out[k] -= (vv * -den * dens[x * S + s] + num) / (1 + den) + fac
    return out


def bruteforce_densities(dens, a, S, dt, vv, dims, vs, mins, X):
    """The per-voxel reduction test/utils.cpp:164-202 intends, D = (D + sum num) / (1 + sum den) + sum fac,
    with every agent assigned to voxel floor((p - min) / dx).  (The reference helper compares the float
    quotient with the integer voxel index, which never holds for voxel-centre positions; this restatement
    uses the voxel the mesh assigns.)"""
    out = np.array(dens, dtype=np.float64)
    n = a["volumes"].shape[0]
    vox = [int((a["positions"][i * dims] - mins[0]) / vs[0]) for i in range(n)]
    for s in range(S):
        for x in range(X):
            num = den = fac = 0.0
            for i in range(n):
                if vox[i] != x:
                    continue
                k = i * S + s
                num += a["secretion_rates"][k] * a["saturation_densities"][k] * dt * a["volumes"][i] / vv
                den += (a["secretion_rates"][k] + a["uptake_rates"][k]) * dt * a["volumes"][i] / vv
                fac += a["net_export_rates"][k] * dt / vv
            out[x * S + s] = (dens[x * S + s] + num) / (1 + den) + fac
    return out


def gradient_expectation(case, dens, grid, S):
    """test/test.cpp:507-664: one-sided differences at the borders (divide by dx), central inside (2 dx)."""
    dims = case["mesh"]["dims"]
    vs = case["mesh"]["voxel_shape"]
    X, Y, Z = grid
    d4 = dens.reshape(Z, Y, X, S).astype(np.float64)
    exp = []
    for d in range(dims):
        n = grid[d]
        ax = 2 - d  # numpy axis of x/y/z in [z][y][x][s]
        g = np.empty_like(d4)
        idx = [slice(None)] * 4
        for i in range(n):
            up = i + (0 if i == n - 1 else 1)
            dn = i - (0 if i == 0 else 1)
            div = (vs[d] if (i == 0 or i == n - 1) else 2 * vs[d])
            a = list(idx); a[ax] = up
            b = list(idx); b[ax] = dn
            c = list(idx); c[ax] = i
            g[tuple(c)] = (d4[tuple(a)] - d4[tuple(b)]) / div
        exp.append(g)
    return exp


def gradient_component(case, grads, grid, S, d):
    dims = case["mesh"]["dims"]
    X, Y, Z = grid
    V = X * Y * Z
    if dims == 3:
        return grads.reshape(3, V * S)[d].reshape(Z, Y, X, S)
    return grads.reshape(V * S, dims)[:, d].reshape(Z, Y, X, S)


def fill_densities(case, dens, grid, S):
    X, Y, Z = grid
    fill = case.get("fill")
    if fill == "linear_index":
        dens[:] = np.arange(dens.size, dtype=np.float64)
    elif fill == "gradient_D1":
        # test/test.cpp:514-519: s0 = x^2, s1 mirrored
        d = dens.reshape(X, S)
        for x in range(X):
            d[x, 0] = x * x
            d[X - 1 - x, 1] = x * x
    elif fill == "gradient_square_sum":
        # test/test.cpp:553-561, 607-618: (x + y + z)^2 + s
        z, y, x, s = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), np.arange(S), indexing="ij")
        dens[:] = ((x + y + z) ** 2 + s).reshape(-1)


def run_case(case, fixtures, make_backend):
    """Run one golden case; yields (label, got, expected) triples."""
    fx = fixtures[case["microenv"]]
    dt = case.get("dt_override", fx["dt"])
    m = case["mesh"]
    S = len(fx["D"])
    be = make_backend(m["dims"], m["mins"], m["maxs"], m["voxel_shape"], S, dt, fx["D"], fx["decay"], fx["init"])
    grid = list(be.grid)
    assert grid == grid_of(m)
    fill_densities(case, be.densities, grid, S)
    for b in case.get("boundary", []):
        be.set_boundary(b["dim"], b["is_max"], b["values"], b["conds"])
    if "interior" in case:
        it = case["interior"]
        n = len(it["voxels"])
        vox = np.array([v[: m["dims"]] for v in it["voxels"]], dtype=np.int32).reshape(-1)
        vals = np.zeros(n * S)
        conds = np.zeros(n * S, dtype=np.uint8)
        for i, v in enumerate(it["values"]):
            vals[i * S] = v
            conds[i * S] = 1
        be.set_interior(vox, vals, conds)

    op = case["op"]
    if op in ("diffusion", "dirichlet", "gradient"):
        be.initialize()
        be.store()
        getattr(be, op)()
        be.load()
        dens = np.asarray(be.densities, dtype=np.float64)
        if "expect_uniform" in case:
            d = dens.reshape(-1, S)
            for s in range(S):
                yield f"s{s}", d[:, s], np.full(d.shape[0], case["expect_uniform"][s])
        if "expect_linear" in case:
            yield "linear", dens, np.array(case["expect_linear"])
        if "expect_points" in case:
            X, Y, Z = grid
            pts = np.array(case["expect_points"])
            idx = (((pts[:, 2] * Y + pts[:, 1]) * X + pts[:, 0]) * S + pts[:, 3]).astype(np.int64)
            yield "points", dens[idx], pts[:, 4]
        if op == "gradient":
            # the input densities are not modified by the gradient solve
            exp = gradient_expectation(case, dens, grid, S)
            grads = np.asarray(be.gradients, dtype=np.float64)
            for d in range(m["dims"]):
                yield f"grad{d}", gradient_component(case, grads, grid, S, d).reshape(-1), exp[d].reshape(-1)
        return

    # agents
    specs = expand_agents(case)
    dims = m["dims"]
    for ci in (True, False):
        for recompute in (True, False):
            be = make_backend(dims, m["mins"], m["maxs"], m["voxel_shape"], S, dt, fx["D"], fx["decay"], fx["init"])
            a = default_agent_arrays(specs, dims, S)
            be.set_agents(a)
            be.set_compute_internalized(ci)
            be.initialize()
            tag = f"ci={int(ci)},rc={int(recompute)}"
            vv = float(np.prod(m["voxel_shape"]))
            if op == "agents":
                be.store()
                be.cell(True)
                be.load()
                X = grid[0] * grid[1] * grid[2]
                dens = np.asarray(be.densities, dtype=np.float64).reshape(-1, S)
                lin = [((v + [0, 0, 0])[2] * grid[1] + (v + [0, 0, 0])[1]) * grid[0] + v[0]
                       for v in case["agent_voxels"]]
                yield tag + ":d1", dens[lin, 0], np.array(case["first"]["densities"])
                yield tag + ":d1s1", dens[:, 1], np.ones(X)
                if ci:
                    yield tag + ":i1", np.asarray(be.internalized).reshape(-1, S)[:, 0], np.array(case["first"]["internalized"])
                be.store()
                be.cell(recompute)
                be.load()
                dens = np.asarray(be.densities, dtype=np.float64).reshape(-1, S)
                yield tag + ":d2", dens[lin, 0], np.array(case["second"]["densities"])
                if ci:
                    yield tag + ":i2", np.asarray(be.internalized).reshape(-1, S)[:, 0], np.array(case["second"]["internalized"])
            else:
                X = grid[0]
                exp_i = np.zeros(len(specs) * S)
                for step, rc in ((0, True), (1, recompute)):
                    before = np.array(be.densities, dtype=np.float64)
                    exp_i = exp_i + bruteforce_internalized(before, a, S, dt, vv, dims, m["voxel_shape"], m["mins"])
                    exp_d = bruteforce_densities(before, a, S, dt, vv, dims, m["voxel_shape"], m["mins"], X)
                    be.store()
                    be.cell(rc)
                    be.load()
                    yield f"{tag}:d{step}", np.asarray(be.densities, dtype=np.float64), exp_d
                    if ci:
                        yield f"{tag}:i{step}", np.asarray(be.internalized, dtype=np.float64), exp_i


def float_ulp_diff(got, exp):
    """Distance in float32 ULPs (the EXPECT_FLOAT_EQ metric)."""
    g = np.asarray(got, dtype=np.float32).view(np.int32).astype(np.int64)
    e = np.asarray(exp, dtype=np.float32).view(np.int32).astype(np.int64)
    g = np.where(g < 0, np.int64(-(2 ** 31)) - g, g)
    e = np.where(e < 0, np.int64(-(2 ** 31)) - e, e)
    return np.abs(g - e)
