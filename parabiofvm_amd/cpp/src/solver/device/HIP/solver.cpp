#include "solver.h"

using namespace biofvm;
using namespace biofvm::solvers::device;

solver::solver() : bulk(ctx), cell(ctx), diffusion(ctx), gradient(ctx) {}

void solver::initialize(microenvironment& m)
{
	once([&] {
		ctx.initialize(m);
		bulk.initialize(m); // host/solver.cpp:5-10 initializes bulk, cell, then diffusion
		cell.initialize(m);
		diffusion.initialize(m);
	});
}

void solver::store_data_to_solver(microenvironment& m)
{
	once([&] { ctx.copy_to_device(m); });
}

void solver::load_data_from_solver(microenvironment& m)
{
	once([&] { ctx.copy_to_host(m); });
}

void solver::wait_for_all()
{
	once([&] { hip_check(bfvm_wait_for_all(ctx.ctx), "bfvm_wait_for_all"); });
}

void solver::timestep(microenvironment&, bool recompute, bool with_gradient)
{
	once([&] { hip_check(bfvm_timestep(ctx.ctx, recompute, with_gradient), "bfvm_timestep"); });
}
