"""TEST INFRASTRUCTURE ONLY — ctypes front-end of the CPU oracle (liboracle_f32/_f64.so).

The oracle restates the reference's host OpenMP solver path (see biofvm_oracle.h).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as
the parity checker and the timed CPU baseline; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_LIBS = {}


def build(force=False):
    """Compile the oracle (gcc -O3 -fopenmp -ffp-contract=off) in place."""
    args = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def _lib(dtype):
    key = np.dtype(dtype).name
    if key in _LIBS:
        return _LIBS[key]
    suffix = {"float32": "f32", "float64": "f64"}[key]
    path = os.path.join(HERE, f"liboracle_{suffix}.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    real = ctypes.c_float if suffix == "f32" else ctypes.c_double
    P = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)

    def fn(name, res, *args):
        f = getattr(lib, f"{name}_{suffix}")
        f.restype = res
        f.argtypes = list(args)
        return f

    api = {
        "create": fn("or_create", P, ctypes.c_int, ip, ip, ip, ctypes.c_int, real, P, P, P),
        "destroy": fn("or_destroy", None, P),
        "grid": fn("or_grid_shape", None, P, ip),
        "densities": fn("or_densities", P, P),
        "gradients": fn("or_gradients", P, P),
        "set_boundary": fn("or_set_boundary", None, P, ctypes.c_int, ctypes.c_int, P, P),
        "set_interior": fn("or_set_interior", None, P, ctypes.c_int, P, P, P),
        "set_agents": fn("or_set_agents", None, P, ctypes.c_int, P, P, P, P, P, P, P, P),
        "set_ci": fn("or_set_compute_internalized", None, P, ctypes.c_int),
        "set_threads": fn("or_set_threads", None, ctypes.c_int),
        "initialize": fn("or_initialize", ctypes.c_int, P),
        "coefs": fn("or_diffusion_coefficients", ctypes.c_int, P, ctypes.c_int, P, P, P),
        "diffusion": fn("or_diffusion_solve", ctypes.c_int, P),
        "dirichlet": fn("or_dirichlet_solve", ctypes.c_int, P),
        "gradient": fn("or_gradient_solve", ctypes.c_int, P),
        "cell": fn("or_cell_simulate", ctypes.c_int, P, ctypes.c_int),
        "bulk": fn("or_bulk_solve", ctypes.c_int, P, P, P, P),
        "release": fn("or_cell_release", ctypes.c_int, P, ctypes.c_int),
        "timestep": fn("or_timestep", ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P),
        "voxel_positions": fn("or_voxel_positions", None, P, ctypes.c_int, P, P),
        "last_error": fn("or_last_error", ctypes.c_char_p),
    }
    _LIBS[key] = api
    return api


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class OracleError(RuntimeError):
    pass


class Oracle:
    """One microenvironment + host solver state, mirroring microenvironment + solvers::host::solver."""

    def __init__(self, dims, mins, maxs, voxel_shape, substrates, dt, D, decay, init, dtype=np.float32,
                 threads=None):
        self.dtype = np.dtype(dtype)
        self.api = _lib(self.dtype)
        if threads:
            self.api["set_threads"](int(threads))
        self.S = int(substrates)
        self.dims = int(dims)
        D = np.ascontiguousarray(D, dtype=self.dtype)
        decay = np.ascontiguousarray(decay, dtype=self.dtype)
        init = np.ascontiguousarray(init, dtype=self.dtype)
        self.h = ctypes.c_void_p(self._create(dims, mins, maxs, voxel_shape, D, decay, init, dt))
        g = (ctypes.c_int * 3)()
        self.api["grid"](self.h, g)
        self.grid = tuple(g)
        self.V = self.grid[0] * self.grid[1] * self.grid[2]
        self._keep = []

    def _create(self, dims, mins, maxs, vs, D, decay, init, dt):
        h = self.api["create"](int(dims), _i3(mins), _i3(maxs), _i3(vs), self.S, float(dt), _ptr(D), _ptr(decay), _ptr(init))
        if not h:
            raise OracleError(self.api["last_error"]().decode())
        return h

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.api["destroy"](self.h)
                self.h = None
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise OracleError(self.api["last_error"]().decode())

    # -- arrays (views onto oracle-owned memory) --
    @property
    def densities(self):
        p = self.api["densities"](self.h)
        n = self.V * self.S
        buf = (ctypes.c_char * (n * self.dtype.itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=self.dtype, count=n)

    @property
    def gradients(self):
        p = self.api["gradients"](self.h)
        n = self.V * self.S * self.dims
        buf = (ctypes.c_char * (n * self.dtype.itemsize)).from_address(p)
        return np.frombuffer(buf, dtype=self.dtype, count=n)

    # -- configuration --
    def set_boundary(self, dim, is_max, values, conds):
        if values is None:
            self.api["set_boundary"](self.h, dim, int(is_max), None, None)
            return
        v = np.ascontiguousarray(values, dtype=self.dtype)
        c = np.ascontiguousarray(conds, dtype=np.uint8)
        self.api["set_boundary"](self.h, dim, int(is_max), _ptr(v), _ptr(c))

    def set_interior(self, voxels, values, conds):
        vox = np.ascontiguousarray(voxels, dtype=np.int32)
        val = np.ascontiguousarray(values, dtype=self.dtype)
        con = np.ascontiguousarray(conds, dtype=np.uint8)
        count = vox.size // self.dims if vox.size else 0
        self.api["set_interior"](self.h, count, _ptr(vox), _ptr(val), _ptr(con))

    def set_agents(self, agents):
        """agents: dict of numpy arrays (secretion_rates, saturation_densities, uptake_rates, net_export_rates,
        internalized_substrates, fraction_released_at_death, volumes, positions) in the reference SoA layout.
        Arrays are converted to the oracle dtype and kept alive; internalized_substrates is updated in place."""
        conv = {k: np.ascontiguousarray(v, dtype=self.dtype) for k, v in agents.items()}
        self._agents = conv
        n = conv["volumes"].shape[0]
        self.api["set_agents"](self.h, n, _ptr(conv["secretion_rates"]), _ptr(conv["saturation_densities"]),
                               _ptr(conv["uptake_rates"]), _ptr(conv["net_export_rates"]),
                               _ptr(conv["internalized_substrates"]), _ptr(conv["fraction_released_at_death"]),
                               _ptr(conv["volumes"]), _ptr(conv["positions"]))
        return conv

    @property
    def internalized(self):
        return self._agents["internalized_substrates"]

    def set_compute_internalized(self, flag):
        self.api["set_ci"](self.h, int(bool(flag)))

    # -- solvers --
    def initialize(self):
        self._check(self.api["initialize"](self.h))

    def coefficients(self, axis):
        n = self.grid[axis]
        b = np.zeros(n * self.S, dtype=self.dtype)
        c = np.zeros(self.S, dtype=self.dtype)
        e = np.zeros(max(n - 1, 1) * self.S, dtype=self.dtype)
        self._check(self.api["coefs"](self.h, axis, _ptr(b), _ptr(c), _ptr(e)))
        return b.reshape(n, self.S), c, e[: (n - 1) * self.S].reshape(max(n - 1, 0), self.S)

    def diffusion(self):
        self._check(self.api["diffusion"](self.h))

    def dirichlet(self):
        self._check(self.api["dirichlet"](self.h))

    def gradient(self):
        self._check(self.api["gradient"](self.h))

    def cell(self, recompute):
        self._check(self.api["cell"](self.h, int(bool(recompute))))

    def bulk(self, supply, uptake, target):
        """host/bulk_solver.cpp: the callbacks' values as [V][S] arrays."""
        arrs = [np.ascontiguousarray(a, dtype=self.dtype) for a in (supply, uptake, target)]
        self._check(self.api["bulk"](self.h, *[_ptr(a) for a in arrs]))

    def release(self, index):
        self._check(self.api["release"](self.h, int(index)))

    def timestep(self, recompute, with_gradient=True):
        ms = (ctypes.c_double * 3)()
        self._check(self.api["timestep"](self.h, int(bool(recompute)), int(bool(with_gradient)), ms))
        return tuple(ms)

    def voxel_positions(self, positions):
        pos = np.ascontiguousarray(positions, dtype=self.dtype)
        count = pos.size // self.dims
        out = np.zeros(count * 3, dtype=np.int32)
        self.api["voxel_positions"](self.h, count, _ptr(pos), _ptr(out))
        return out.reshape(count, 3)
