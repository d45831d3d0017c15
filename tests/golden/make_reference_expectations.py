"""Writes tests/golden/reference_expectations.json.

The expected values below are DATA transcribed from the reference's own
test-suite (test/test.cpp, test/test_device.cpp, test/utils.cpp); each case
names the lines it comes from.  Where the reference states an expectation as a
rule over voxels ("diagonally next to the boundary"), the rule is expanded here
into explicit (x, y, z, s) -> value entries so the JSON is self-contained.
The reference compares with EXPECT_FLOAT_EQ (4 float ULPs).

Run:  python tests/golden/make_reference_expectations.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

# test/utils.cpp:10-54
FIXTURES = {
    "default_microenv": {"dt": 5.0, "D": [4.0, 2.0], "decay": [5.0, 3.0], "init": [1.0, 1.0]},
    "biorobots_microenv": {"dt": 0.01, "D": [1000.0, 1000.0], "decay": [0.1, 0.4], "init": [0.0, 0.0]},
}


def mesh(dims, maxs, vs, mins=(0, 0, 0)):
    return {"dims": dims, "mins": list(mins), "maxs": list(maxs), "voxel_shape": list(vs)}


def grid(m):
    g = [1, 1, 1]
    for d in range(m["dims"]):
        g[d] = (m["maxs"][d] - m["mins"][d] + m["voxel_shape"][d] - 1) // m["voxel_shape"][d]
    return g


def voxels(m):
    X, Y, Z = grid(m)
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                yield x, y, z


cases = []

# ---- diffusion: uniform closed forms (test/test.cpp:38-114; test/test_device.cpp:35-92) ----
for name, m, exp, ref in [
    ("diffusion_D1_uniform", mesh(1, (80, 0, 0), (20, 0, 0)), [0.03846154, 0.0625], "test/test.cpp:38-62"),
    ("diffusion_D2_uniform", mesh(2, (800, 800, 0), (20, 20, 0)), [0.0054869675, 0.013840831], "test/test.cpp:64-88"),
    ("diffusion_D3_uniform", mesh(3, (200, 200, 200), (20, 20, 20)), [0.0012299563, 0.0046296306], "test/test.cpp:90-114"),
    ("diffusion_D2_uniform_200x20000", mesh(2, (200, 20000, 0), (20, 20, 0)), [0.0054869675, 0.013840831], "test/test_device.cpp:35-67"),
    ("diffusion_D2_uniform_20000x200", mesh(2, (20000, 200, 0), (20, 20, 0)), [0.0054869675, 0.013840831], "test/test_device.cpp:35-67"),
    ("diffusion_D2_uniform_20000x20000", mesh(2, (20000, 20000, 0), (20, 20, 0)), [0.0054869675, 0.013840831], "test/test_device.cpp:35-67"),
]:
    cases.append({"name": name, "ref": ref, "mesh": m, "microenv": "default_microenv", "op": "diffusion",
                  "expect_uniform": exp})

# ---- diffusion: known-answer vectors with density = linear index (test/test.cpp:116-244) ----
D1_random = [0.0486842592, 1.0444132121, 1.9980019980, 2.9880478088, 3.9473197368, 4.9316824055]
D2_random = [0.1948319355, 1.1899772978, 2.1441254507, 3.1335099015, 4.0934189658, 5.0770425053, 6.0427124809,
             7.0205751090, 7.9920082, 8.9641077127, 9.9412995111, 10.9076403164, 11.8905930262, 12.8511729202,
             13.8398865413, 14.7947055239, 15.7891800565, 16.7382381276]
D3_random = [
    0.6333066643, 1.6268066007, 2.5825920996, 3.5703051208, 4.5318775349, 5.5138036408, 6.4811629703,
    7.4573021609, 8.4304484056, 9.4008006809, 10.3797338410, 11.3442992010, 12.3290192763, 13.2877977210,
    14.2783047117, 15.2312962410, 16.2275901470, 17.1747947611, 18.1768755823, 19.1182932811, 20.1261610177,
    21.0617918012, 22.0754464530, 23.0052903212, 24.0247318884, 24.9487888412, 25.9740173237, 26.8922873613,
    27.9233027591, 28.8357858813, 29.8725881944, 30.7792844014, 31.8218736297, 32.7227829214, 33.7711590651,
    34.6662814414, 35.7204445004, 36.6097799615, 37.6697299358, 38.5532784815, 39.6190153711, 40.4967770016,
    41.5683008064, 42.4402755216, 43.5175862418, 44.3837740416, 45.4668716771, 46.3272725617, 47.4161571125,
    48.2707710817, 49.3654425478, 50.2142696018, 51.3147279832, 52.1577681218]
for name, m, exp, ref in [
    ("diffusion_D1_random", mesh(1, (60, 0, 0), (20, 20, 20)), D1_random, "test/test.cpp:116-148"),
    ("diffusion_D2_random", mesh(2, (60, 60, 0), (20, 20, 20)), D2_random, "test/test.cpp:150-189"),
    ("diffusion_D3_random", mesh(3, (60, 60, 60), (20, 20, 20)), D3_random, "test/test.cpp:191-244"),
]:
    # index = s + x*S + y*S*X + z*S*X*Y is exactly the [z][y][x][s] linear index
    cases.append({"name": name, "ref": ref, "mesh": m, "microenv": "biorobots_microenv", "op": "diffusion",
                  "fill": "linear_index", "expect_linear": exp})


# ---- interior Dirichlet (test/test.cpp:246-505); only substrate 0 has a condition (test/utils.cpp:88-109) ----
def dirichlet_case(name, ref, m, vox, vals, rule, s1_value):
    exp = []
    for x, y, z in voxels(m):
        exp.append([x, y, z, 0, rule(x, y, z)])
        exp.append([x, y, z, 1, s1_value])
    cases.append({"name": name, "ref": ref, "mesh": m, "microenv": "default_microenv", "op": "diffusion",
                  "interior": {"voxels": vox, "values": vals}, "expect_points": exp})


dirichlet_case("dirichlet_one_cond_D1", "test/test.cpp:246-271", mesh(1, (100, 0, 0), (20, 0, 0)), [[2, 0, 0]], [1.0],
               lambda x, y, z: 1.0 if x == 2 else 0.03846154, 0.0625)


def one_cond_D2(x, y, z):
    if x == 1 and y == 1:
        return 10.0
    if abs(x - 1) == 1 and abs(y - 1) == 1:
        return 0.0054876315
    if abs(x - 1) == 1 and y == 1:
        return 0.0056665326
    return 0.0081802057


dirichlet_case("dirichlet_one_cond_D2", "test/test.cpp:273-317", mesh(2, (60, 60, 0), (20, 20, 0)), [[1, 1, 0]], [10.0],
               one_cond_D2, 0.013840831)

_oc3 = {
    (0, 0): 0.0012301364, (1, 0): 0.001549035, (2, 0): 0.0012301364,
    (0, 1): 0.0012637526, (1, 1): 0.566124, (2, 1): 0.0012637526,
    (0, 2): 0.0012301364, (1, 2): 0.001549035, (2, 2): 0.0012301364,
}
_oc3_mid = {
    (0, 0): 0.0012637526, (1, 0): 0.061110392, (2, 0): 0.0012637526,
    (0, 1): 0.0075723953, (1, 1): 1000.0, (2, 1): 0.0075723953,
    (0, 2): 0.0012637526, (1, 2): 0.061110392, (2, 2): 0.0012637526,
}
dirichlet_case("dirichlet_one_cond_D3", "test/test.cpp:319-373", mesh(3, (60, 60, 60), (20, 20, 20)), [[1, 1, 1]], [1000.0],
               lambda x, y, z: _oc3_mid[(x, y)] if z == 1 else _oc3[(x, y)], 0.0046296306)
dirichlet_case("dirichlet_multiple_cond_D1", "test/test.cpp:375-402", mesh(1, (100, 0, 0), (20, 0, 0)),
               [[0, 0, 0], [4, 0, 0]], [1.0, 1.0], lambda x, y, z: 1.0 if x in (0, 4) else 0.03846154, 0.0625)
dirichlet_case("dirichlet_multiple_cond_D2", "test/test.cpp:404-443", mesh(2, (60, 60, 0), (20, 20, 0)),
               [[0, 0, 0], [1, 0, 0], [2, 0, 0]], [10.0] * 3,
               lambda x, y, z: 10.0 if y == 0 else (0.0081802057 if y == 1 else 0.0054969075), 0.013840831)
dirichlet_case("dirichlet_multiple_cond_D3", "test/test.cpp:445-505", mesh(3, (60, 60, 60), (20, 20, 20)),
               [[x, y, 1] for y in range(3) for x in range(3)], [1000.0] * 9,
               lambda x, y, z: 1000.0 if z == 1 else 0.566124, 0.0046296306)


# ---- boundary Dirichlet precedence (test/test.cpp:666-810); add_boundary_dirichlet: test/utils.cpp:111-132 ----
def boundary_case(name, ref, m, faces, rule, zrange=None):
    exp = []
    X, Y, Z = grid(m)
    for x, y, z in voxels(m):
        v = rule(x, y, z, X, Y, Z)
        if v is not None:
            exp.append([x, y, z, 0, v])
        exp.append([x, y, z, 1, 1.0])
    cases.append({"name": name, "ref": ref, "mesh": m, "microenv": "default_microenv", "op": "dirichlet",
                  "boundary": [{"dim": d, "is_max": mx, "values": [val, 42.0], "conds": [1, 0]} for d, mx, val in faces],
                  "expect_points": exp})


def bD1(x, y, z, X, Y, Z):
    return 4.0 if x == 0 else (5.0 if x == X - 1 else 1.0)


def bD2(x, y, z, X, Y, Z):
    if y == 0:
        return 6.0
    if y == Y - 1:
        return 7.0
    if x == 0:
        return 4.0
    if x == X - 1:
        return 5.0
    return 1.0


def bD3(x, y, z, X, Y, Z):
    if z == 0:
        return 8.0
    if z == Z - 1:
        return 9.0
    return bD2(x, y, z, X, Y, Z)


boundary_case("dirichlet_boundaries_D1", "test/test.cpp:666-693", mesh(1, (100, 100, 100), (20, 20, 20)),
              [(0, 0, 4.0), (0, 1, 5.0)], bD1)
boundary_case("dirichlet_boundaries_D2", "test/test.cpp:695-744", mesh(2, (100, 100, 100), (20, 20, 20)),
              [(0, 0, 4.0), (0, 1, 5.0), (1, 0, 6.0), (1, 1, 7.0)], bD2)
boundary_case("dirichlet_boundaries_D3", "test/test.cpp:746-810", mesh(3, (100, 100, 100), (20, 20, 20)),
              [(0, 0, 4.0), (0, 1, 5.0), (1, 0, 6.0), (1, 1, 7.0), (2, 0, 8.0), (2, 1, 9.0)], bD3)

# ---- gradient closed forms (test/test.cpp:507-664) ----
cases.append({"name": "gradient_D1", "ref": "test/test.cpp:507-542", "mesh": mesh(1, (100, 0, 0), (20, 0, 0)),
              "microenv": "default_microenv", "op": "gradient", "fill": "gradient_D1"})
cases.append({"name": "gradient_D2", "ref": "test/test.cpp:544-596", "mesh": mesh(2, (100, 100, 0), (20, 20, 0)),
              "microenv": "default_microenv", "op": "gradient", "fill": "gradient_square_sum"})
cases.append({"name": "gradient_D3", "ref": "test/test.cpp:598-664", "mesh": mesh(3, (100, 100, 100), (20, 20, 20)),
              "microenv": "default_microenv", "op": "gradient", "fill": "gradient_square_sum"})

# ---- agents, one agent per voxel (test/test.cpp:812-1041); values from test/utils.cpp:204-222 ----
first = {"internalized": [-216004.000000, -1469060.631579, -2927715.703704],
         "densities": [28.000500, 184.632579, 366.964463]}
second = {"internalized": [-216004.000000 + -157093.818182, -1469060.631579 + -618551.844632,
                           -2927715.703704 + -867471.319407],
          "densities": [47.637227, 261.951560, 475.398378]}
for dims, maxs, ref in [(1, (60, 20, 20), "test/test.cpp:818-877"), (2, (60, 60, 20), "test/test.cpp:879-959"),
                        (3, (60, 60, 60), "test/test.cpp:961-1041")]:
    agents = []
    for k, (off, pos) in enumerate([(0, 10), (400, 30), (800, 50)]):
        agents.append({"rates_offset": off, "volume": 1000, "position": [pos] * dims})
    cases.append({"name": f"agents_simple_D{dims}", "ref": ref, "mesh": mesh(dims, maxs, (20, 20, 20)),
                  "microenv": "default_microenv", "dt_override": 0.01, "op": "agents", "agents": agents,
                  "agent_voxels": [[0] * dims, [1] * dims, [2] * dims],
                  "first": first, "second": second})

# ---- agents with conflicts (test/test.cpp:1043-1214): expectations are brute force, see test_oracle_golden.py ----
cases.append({"name": "agents_conflict", "ref": "test/test.cpp:1043-1127", "mesh": mesh(1, (60, 20, 20), (20, 20, 20)),
              "microenv": "default_microenv", "dt_override": 0.01, "op": "agents_bruteforce",
              "agents": [{"rates_offset": o, "volume": v, "position": [p]} for o, v, p in
                         [(0, 500, 10), (600, 1000, 30), (1100, 1500, 30), (1600, 2000, 50), (2100, 2500, 50),
                          (2600, 3000, 50)]]})
cases.append({"name": "agents_conflict_big", "ref": "test/test.cpp:1129-1214",
              "mesh": mesh(1, (2000, 20, 20), (20, 20, 20)), "microenv": "default_microenv", "dt_override": 0.01,
              "op": "agents_bruteforce",
              # 50 agents at each of the 100 voxel centres (mesh.cpp:56-60: x*20 + 20/2.0 + 0)
              "agents_per_voxel_center": {"per_voxel": 50, "rates_offset": 0, "volume": 500}})

out = {"source": "asmelko/paraBioFVM test/test.cpp, test/test_device.cpp, test/utils.cpp (transcribed values)",
       "compare": "EXPECT_FLOAT_EQ: 4 float ULPs", "fixtures": FIXTURES, "cases": cases}
with open(os.path.join(HERE, "reference_expectations.json"), "w") as f:
    json.dump(out, f, separators=(",", ":"))
print(f"wrote {len(cases)} cases")
