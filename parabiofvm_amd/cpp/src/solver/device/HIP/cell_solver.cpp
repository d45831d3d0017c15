#include "cell_solver.h"

using namespace biofvm;
using namespace biofvm::solvers::device;

cell_solver::cell_solver(device_context& ctx) : ctx_(ctx) {}

void cell_solver::initialize(microenvironment&) {}

void cell_solver::simulate_secretion_and_uptake(microenvironment&, bool recompute)
{
	once([&] { hip_check(bfvm_cell_simulate(ctx_.ctx, recompute ? 1 : 0), "bfvm_cell_simulate"); });
}

void cell_solver::release_internalized_substrates(microenvironment&, index_t index)
{
	once([&] { hip_check(bfvm_cell_release(ctx_.ctx, index), "bfvm_cell_release"); });
}
