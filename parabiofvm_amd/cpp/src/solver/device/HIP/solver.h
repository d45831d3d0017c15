// The device solver facade (interface of src/solver/device/CUDA/solver.h:12-31 of the reference).
#pragma once

#include "bulk_solver.h"
#include "cell_solver.h"
#include "device_context.h"
#include "diffusion_solver.h"
#include "gradient_solver.h"

namespace biofvm {
namespace solvers {
namespace device {

class solver
{
	device_context ctx;

public:
	solver();

	static constexpr bool is_device_solver = true;

	bulk_solver bulk; // the host solver's member (host/solver.h:17), here on the device
	cell_solver cell;
	diffusion_solver diffusion;
	gradient_solver gradient;

	void initialize(microenvironment& m);

	void store_data_to_solver(microenvironment& m);
	void load_data_from_solver(microenvironment& m);

	void wait_for_all();

	// one whole timestep (secretion, diffusion, gradient) replayed from a captured HIP graph
	void timestep(microenvironment& m, bool recompute, bool with_gradient = true);

	bfvm_ctx* context() { return ctx.ctx; }
};

} // namespace device
} // namespace solvers
} // namespace biofvm
