#include "gradient_solver.h"

using namespace biofvm;
using namespace biofvm::solvers::device;

gradient_solver::gradient_solver(device_context& ctx) : ctx_(ctx) {}

void gradient_solver::solve(microenvironment&)
{
	once([&] { hip_check(bfvm_gradient_solve(ctx_.ctx), "bfvm_gradient_solve"); });
}
