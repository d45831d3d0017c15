// Secretion and uptake of agents on the device (interface of src/solver/device/CUDA/cell_solver.h:39-64 and
// src/solver/host/cell_solver.h:39-64 of the reference; formulas in src/solver/host/cell_solver.h:1-26).
#pragma once

#include "device_context.h"

namespace biofvm {
namespace solvers {
namespace device {

class cell_solver : common_solver
{
	device_context& ctx_;

public:
	cell_solver(device_context& ctx);

	void initialize(microenvironment& m);

	void simulate_secretion_and_uptake(microenvironment& m, bool recompute);

	// D += I*F/vv at the agent's voxel, I = 0, on the device copies (the reference's device facade edits the
	// host arrays, CUDA/cell_solver.cpp:288-294); visible on the host after load_data_from_solver
	void release_internalized_substrates(microenvironment& m, index_t index);
};

} // namespace device
} // namespace solvers
} // namespace biofvm
