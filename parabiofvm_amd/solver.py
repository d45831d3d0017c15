"""Python mirror of the reference's microenvironment + device solver facade over the C ABI.

`Microenvironment` holds the host arrays the way `microenvironment`
(include/BioFVM/microenvironment.h:19-63) and `agent_data`
(include/BioFVM/agent_data.h:11-34) do; `DeviceSolver` drives one bfvm_ctx with
the call sequence of solvers::device::solver (CUDA/solver.h:12-31):
initialize, store_data_to_solver, diffusion/gradient/cell solves,
load_data_from_solver, wait_for_all.
"""
import ctypes

import numpy as np

from . import _native
from ._native import AgentsDesc, DeviceBuffers, MicroenvironmentDesc, check

AGENT_FIELDS = ("secretion_rates", "saturation_densities", "uptake_rates", "net_export_rates",
                "internalized_substrates", "fraction_released_at_death", "volumes", "positions")


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None


def grid_shape(dims, mins, maxs, voxel_shape):
    """cartesian_mesh grid shape (src/mesh.cpp:7-25)."""
    g = [1, 1, 1]
    for d in range(dims):
        g[d] = (maxs[d] - mins[d] + voxel_shape[d] - 1) // voxel_shape[d]
    return tuple(g)


class DeviceSolver:
    """One microenvironment on one GPU.

    Host-side state (numpy, float32): densities [V*S], gradients [dims*V*S], agent SoA arrays.
    """

    def __init__(self, dims, mins, maxs, voxel_shape, substrates, dt, D, decay, init, device=0):
        self.lib = _native.load()
        self.dims = int(dims)
        self.mins = [int(v) for v in mins]
        self.maxs = [int(v) for v in maxs]
        self.voxel_shape = [int(v) for v in voxel_shape]
        self.S = int(substrates)
        self.dt = float(dt)
        self.D = np.ascontiguousarray(D, dtype=np.float32)
        self.decay = np.ascontiguousarray(decay, dtype=np.float32)
        self.grid = grid_shape(self.dims, self.mins, self.maxs, self.voxel_shape)
        self.V = self.grid[0] * self.grid[1] * self.grid[2]
        # microenvironment.cpp:14-25: densities start at the initial condition of each substrate
        self.densities = np.empty(self.V * self.S, dtype=np.float32)
        self.densities.reshape(self.V, self.S)[:] = np.asarray(init, dtype=np.float32)[None, :]
        self.gradients = np.zeros(self.V * self.S * self.dims, dtype=np.float32)
        self._faces = {}
        self._interior = None
        self._agents = None
        self.compute_internalized = False
        self._algorithm = (_native.BFVM_SWEEP_PARTITIONED, 0)
        h = ctypes.c_void_p()
        check("bfvm_create", self.lib.bfvm_create(int(device), ctypes.byref(h)))
        self.h = h
        self.initialized = False

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            check("bfvm_destroy", self.lib.bfvm_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration (what microenvironment_builder / the test utilities set) --
    def set_boundary(self, dim, is_max, values, conds):
        if values is None:
            self._faces.pop((dim, int(is_max)), None)
        else:
            self._faces[(dim, int(is_max))] = (np.ascontiguousarray(values, dtype=np.float32),
                                               np.ascontiguousarray(conds, dtype=np.uint8))

    def set_interior(self, voxels, values, conds):
        self._interior = (np.ascontiguousarray(voxels, dtype=np.int32),
                          np.ascontiguousarray(values, dtype=np.float32),
                          np.ascontiguousarray(conds, dtype=np.uint8))

    def set_agents(self, agents):
        conv = {k: np.ascontiguousarray(agents[k], dtype=np.float32) for k in AGENT_FIELDS if k in agents}
        if "fraction_released_at_death" not in conv:
            conv["fraction_released_at_death"] = np.zeros_like(conv["secretion_rates"])
        self._agents = conv
        return conv

    @property
    def internalized(self):
        return self._agents["internalized_substrates"]

    @property
    def agents_count(self):
        return 0 if self._agents is None else int(self._agents["volumes"].shape[0])

    def set_compute_internalized(self, flag):
        self.compute_internalized = bool(flag)

    def set_storage(self, storage):
        """Density storage form in HBM ("f32", "bf16", "f16" or a BFVM_STORAGE_* value); applies at the next
        initialize.  Host arrays stay float32 either way."""
        code = _native.STORAGE[storage] if isinstance(storage, str) else int(storage)
        check("bfvm_set_storage", self.lib.bfvm_set_storage(self.h, code))

    def set_secretion_reduction(self, mode):
        """How agents sharing a voxel are combined: "atomic" (claim + float atomics) or "sorted" (agents ordered
        by voxel at upload, segmented sums); applies at the next store/store_agents."""
        code = _native.REDUCTION[mode] if isinstance(mode, str) else int(mode)
        check("bfvm_set_secretion_reduction", self.lib.bfvm_set_secretion_reduction(self.h, code))

    def set_sweep_algorithm(self, algorithm, chunk=0):
        self._algorithm = (int(algorithm), int(chunk))
        check("bfvm_set_sweep_algorithm", self.lib.bfvm_set_sweep_algorithm(self.h, *self._algorithm))

    # -- solver facade --
    def initialize(self):
        desc = MicroenvironmentDesc()
        desc.dims = self.dims
        for d in range(3):
            desc.bounding_box_mins[d] = self.mins[d]
            desc.bounding_box_maxs[d] = self.maxs[d]
            desc.voxel_shape[d] = self.voxel_shape[d]
        desc.substrates_count = self.S
        desc.diffusion_time_step = self.dt
        desc.diffusion_coefficients = _fp(self.D)
        desc.decay_rates = _fp(self.decay)
        keep = []
        for (dim, mx), (vals, conds) in self._faces.items():
            keep += [vals, conds]
            if mx:
                desc.dirichlet_max_boundary_values[dim] = _fp(vals)
                desc.dirichlet_max_boundary_conditions[dim] = conds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            else:
                desc.dirichlet_min_boundary_values[dim] = _fp(vals)
                desc.dirichlet_min_boundary_conditions[dim] = conds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        if self._interior is not None:
            vox, vals, conds = self._interior
            desc.dirichlet_interior_voxels_count = vox.size // self.dims
            desc.dirichlet_interior_voxels = vox.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            desc.dirichlet_interior_values = _fp(vals)
            desc.dirichlet_interior_conditions = conds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        desc.compute_internalized_substrates = int(self.compute_internalized)
        check("bfvm_initialize", self.lib.bfvm_initialize(self.h, ctypes.byref(desc)))
        if self._algorithm != (_native.BFVM_SWEEP_PARTITIONED, 0):
            check("bfvm_set_sweep_algorithm", self.lib.bfvm_set_sweep_algorithm(self.h, *self._algorithm))
        self.initialized = True

    def store(self):
        """store_data_to_solver: densities and agent data to the device."""
        check("bfvm_store_densities", self.lib.bfvm_store_densities(self.h, self.densities.ctypes.data))
        self.store_agents()

    def store_densities(self):
        """Densities only (agents untouched)."""
        check("bfvm_store_densities", self.lib.bfvm_store_densities(self.h, self.densities.ctypes.data))

    def store_densities_device(self, device_ptr):
        """Densities from device memory (a float32 [V*S] buffer on this context's device), enqueued on the
        context's stream."""
        check("bfvm_store_densities_device", self.lib.bfvm_store_densities_device(self.h, ctypes.c_void_p(device_ptr)))

    def store_agents(self):
        a = self._agents
        desc = AgentsDesc()
        if a is not None:
            desc.agents_count = a["volumes"].shape[0]
            for k in AGENT_FIELDS:
                setattr(desc, k, _fp(a[k]))
        check("bfvm_store_agents", self.lib.bfvm_store_agents(self.h, ctypes.byref(desc)))

    def load_densities(self):
        """Densities only (the gradients of a large mesh need not cross to the host)."""
        check("bfvm_load_densities", self.lib.bfvm_load_densities(self.h, self.densities.ctypes.data))

    def load(self):
        """load_data_from_solver: densities, gradients and internalized substrates to the host."""
        check("bfvm_load_densities", self.lib.bfvm_load_densities(self.h, self.densities.ctypes.data))
        check("bfvm_load_gradients", self.lib.bfvm_load_gradients(self.h, self.gradients.ctypes.data))
        if self._agents is not None and self.agents_count > 0:
            check("bfvm_load_internalized",
                  self.lib.bfvm_load_internalized(self.h, self._agents["internalized_substrates"].ctypes.data))

    def bulk_store_rates(self, supply, uptake, target):
        """Bulk supply/uptake fields ([V][S] each, the densities' layout): the values of the reference's
        supply_rate_func, uptake_rate_func and supply_target_densities_func."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (supply, uptake, target)]
        check("bfvm_bulk_store_rates", self.lib.bfvm_bulk_store_rates(self.h, *[a.ctypes.data for a in arrs]))

    def bulk(self):
        check("bfvm_bulk_solve", self.lib.bfvm_bulk_solve(self.h))

    def diffusion(self):
        check("bfvm_diffusion_solve", self.lib.bfvm_diffusion_solve(self.h))

    def dirichlet(self):
        check("bfvm_dirichlet_solve", self.lib.bfvm_dirichlet_solve(self.h))

    def gradient(self):
        check("bfvm_gradient_solve", self.lib.bfvm_gradient_solve(self.h))

    def cell(self, recompute):
        check("bfvm_cell_simulate", self.lib.bfvm_cell_simulate(self.h, int(bool(recompute))))

    def release(self, index):
        check("bfvm_cell_release", self.lib.bfvm_cell_release(self.h, int(index)))

    def timestep(self, recompute, with_gradient=True):
        check("bfvm_timestep", self.lib.bfvm_timestep(self.h, int(bool(recompute)), int(bool(with_gradient))))

    def wait(self):
        check("bfvm_wait_for_all", self.lib.bfvm_wait_for_all(self.h))

    def device_buffers(self):
        b = DeviceBuffers()
        check("bfvm_get_device_buffers", self.lib.bfvm_get_device_buffers(self.h, ctypes.byref(b)))
        return b

    def partition(self, axis):
        """(blocks, block length, lines per workgroup) of the partitioned sweep along `axis`."""
        p = ctypes.c_int32()
        m = ctypes.c_int32()
        lpg = ctypes.c_int32()
        check("bfvm_sweep_partition", self.lib.bfvm_sweep_partition(self.h, axis, ctypes.byref(p), ctypes.byref(m),
                                                                    ctypes.byref(lpg)))
        return p.value, m.value, lpg.value

    def set_gradient_stacked(self, stacked):
        """3-D gradient tiles with the workgroup's waves stacked along y (True) or side by side (False)."""
        check("bfvm_set_gradient_stacked", self.lib.bfvm_set_gradient_stacked(self.h, int(bool(stacked))))

    def set_gradient_order(self, order):
        """3-D gradient workgroup order: 0 = z chunks of a row band per XCD, 1 = plane order."""
        check("bfvm_set_gradient_order", self.lib.bfvm_set_gradient_order(self.h, int(order)))

    def set_gradient_tuning(self, workgroups):
        check("bfvm_set_gradient_tuning", self.lib.bfvm_set_gradient_tuning(self.h, int(workgroups)))

    def autotune(self):
        """Time candidate launch shapes of every sweep axis and the gradient on this mesh and keep the fastest
        (cached per shape for the process); the densities are unchanged."""
        check("bfvm_autotune", self.lib.bfvm_autotune(self.h))

    def set_sweep_workgroups(self, workgroups):
        """Most workgroups a sweep launches (0 = automatic); fewer workgroups run more passes each."""
        check("bfvm_set_sweep_workgroups", self.lib.bfvm_set_sweep_workgroups(self.h, int(workgroups)))

    def set_sweep_tuning(self, axis, chunk, lines_per_group):
        check("bfvm_set_sweep_tuning", self.lib.bfvm_set_sweep_tuning(self.h, int(axis), int(chunk),
                                                                      int(lines_per_group)))

    def exact_coefficients(self, axis):
        n = self.grid[axis]
        b = np.zeros(n * self.S, dtype=np.float32)
        c = np.zeros(self.S, dtype=np.float32)
        e = np.zeros(max(n - 1, 1) * self.S, dtype=np.float32)
        check("bfvm_exact_coefficients",
              self.lib.bfvm_exact_coefficients(self.h, axis, b.ctypes.data, c.ctypes.data, e.ctypes.data))
        return b.reshape(n, self.S), c, e[: (n - 1) * self.S].reshape(max(n - 1, 0), self.S)

    def benchmark(self, what, iterations):
        """Average ms per launch of one kernel/phase (names in _native.KERNELS), timed with HIP events on the
        context's stream."""
        ms = ctypes.c_float()
        check("bfvm_benchmark", self.lib.bfvm_benchmark(self.h, _native.KERNELS[what], int(iterations),
                                                        ctypes.byref(ms)))
        return ms.value

    def voxel_positions(self, positions):
        pos = np.ascontiguousarray(positions, dtype=np.float32)
        count = pos.size // self.dims
        out = np.zeros(count * 3, dtype=np.int32)
        check("bfvm_voxel_positions", self.lib.bfvm_voxel_positions(self.h, count, pos.ctypes.data, out.ctypes.data))
        return out.reshape(count, 3)


class Batch:
    """Many initialized DeviceSolvers advanced together (bfvm_batch_*): one launch per sweep axis and one for
    the gradient over all instances, the whole step replayed as one captured HIP graph.  Densities are
    stored and loaded through each solver as before; call wait() before loading."""

    def __init__(self, solvers):
        self.lib = _native.load()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        h = ctypes.c_void_p()
        check("bfvm_batch_create", self.lib.bfvm_batch_create(arr, len(self.solvers), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            check("bfvm_batch_destroy", self.lib.bfvm_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def timestep(self, recompute, with_gradient=True):
        check("bfvm_batch_timestep",
              self.lib.bfvm_batch_timestep(self.h, int(bool(recompute)), int(bool(with_gradient))))

    def wait(self):
        check("bfvm_batch_wait", self.lib.bfvm_batch_wait(self.h))

    def autotune(self):
        """Time candidate work-list launch shapes of the batch's sweeps and gradient and keep the fastest."""
        check("bfvm_batch_autotune", self.lib.bfvm_batch_autotune(self.h))

    def benchmark(self, what, iterations):
        ms = ctypes.c_float()
        check("bfvm_batch_benchmark", self.lib.bfvm_batch_benchmark(self.h, _native.KERNELS[what], int(iterations),
                                                                    ctypes.byref(ms)))
        return ms.value
