#include "dirichlet_solver.h"

using namespace biofvm;
using namespace biofvm::solvers::device;

dirichlet_solver::dirichlet_solver(device_context& ctx) : ctx_(ctx) {}

// the Dirichlet description is uploaded by solver::initialize with the rest of the microenvironment
void dirichlet_solver::initialize(microenvironment&) {}

void dirichlet_solver::solve(microenvironment&)
{
	once([&] { hip_check(bfvm_dirichlet_solve(ctx_.ctx), "bfvm_dirichlet_solve"); });
}

void dirichlet_solver::solve_2d(microenvironment& m) { solve(m); }
void dirichlet_solver::solve_3d(microenvironment& m) { solve(m); }
