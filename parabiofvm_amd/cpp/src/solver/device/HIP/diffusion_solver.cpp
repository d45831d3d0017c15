#include "diffusion_solver.h"

using namespace biofvm;
using namespace biofvm::solvers::device;

diffusion_solver::diffusion_solver(device_context& ctx) : ctx_(ctx), dirichlet(ctx) {}

void diffusion_solver::initialize(microenvironment& m) { dirichlet.initialize(m); }

void diffusion_solver::solve(microenvironment&)
{
	once([&] { hip_check(bfvm_diffusion_solve(ctx_.ctx), "bfvm_diffusion_solve"); });
}

void diffusion_solver::solve_2d(microenvironment& m) { solve(m); }
void diffusion_solver::solve_3d(microenvironment& m) { solve(m); }
