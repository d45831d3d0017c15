// LOD implicit diffusion on the device (interface of src/solver/device/CUDA/diffusion_solver.h:39-66 of the
// reference): Dirichlet, x sweep, Dirichlet, y sweep, Dirichlet, z sweep, Dirichlet.
#pragma once

#include "dirichlet_solver.h"

namespace biofvm {
namespace solvers {
namespace device {

class diffusion_solver
{
	device_context& ctx_;

public:
	diffusion_solver(device_context& ctx);

	dirichlet_solver dirichlet;

	// the Thomas coefficients are computed on the GPU by solver::initialize
	void initialize(microenvironment& m);

	void solve_2d(microenvironment& m);
	void solve_3d(microenvironment& m);

	void solve(microenvironment& m);
};

} // namespace device
} // namespace solvers
} // namespace biofvm
