// Dirichlet conditions on the device (interface of src/solver/device/CUDA/dirichlet_solver.h:21-42 and
// src/solver/host/dirichlet_solver.h:12-20 of the reference).  Boundary faces then the interior list.
#pragma once

#include "device_context.h"

namespace biofvm {
namespace solvers {
namespace device {

class dirichlet_solver
{
	device_context& ctx_;

public:
	dirichlet_solver(device_context& ctx);

	void initialize(microenvironment& m);

	void solve_2d(microenvironment& m);
	void solve_3d(microenvironment& m);

	void solve(microenvironment& m);
};

} // namespace device
} // namespace solvers
} // namespace biofvm
