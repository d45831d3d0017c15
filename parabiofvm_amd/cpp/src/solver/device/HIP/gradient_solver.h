// Central-difference gradients on the device (host/gradient_solver.h:9-13 of the reference; the reference's
// device facade runs the host gradient, here it is a device kernel).  Results reach m.gradients with
// solver::load_data_from_solver.
#pragma once

#include "device_context.h"

namespace biofvm {
namespace solvers {
namespace device {

class gradient_solver
{
	device_context& ctx_;

public:
	gradient_solver(device_context& ctx);

	void solve(microenvironment& m);
};

} // namespace device
} // namespace solvers
} // namespace biofvm
