#include "bulk_solver.h"

#include <stdexcept>
#include <vector>

using namespace biofvm;
using namespace biofvm::solvers::device;

bulk_solver::bulk_solver(device_context& ctx) : ctx_(ctx) {}

void bulk_solver::initialize(microenvironment& m)
{
	if (m.supply_rate_func && m.uptake_rate_func && m.supply_target_densities_func)
		upload(m);
}

void bulk_solver::update_rates(microenvironment& m) { once([&] { upload(m); }); }

void bulk_solver::upload(microenvironment& m)
{
	{
		if (!m.supply_rate_func || !m.uptake_rate_func || !m.supply_target_densities_func)
			throw std::runtime_error("bulk_solver: the microenvironment has no bulk functions");
		const index_t S = m.substrates_count;
		const auto g = m.mesh.grid_shape;
		const std::size_t n = (std::size_t)m.mesh.voxel_count() * S;
		std::vector<real_t> supply(n), uptake(n), target(n);
		// [z][y][x][s], the densities' layout (traits.h:65-93)
		for (index_t z = 0; z < g[2]; z++)
			for (index_t y = 0; y < g[1]; y++)
				for (index_t x = 0; x < g[0]; x++)
				{
					const std::size_t at = (((std::size_t)z * g[1] + y) * g[0] + x) * S;
					m.supply_rate_func(m, { x, y, z }, supply.data() + at);
					m.uptake_rate_func(m, { x, y, z }, uptake.data() + at);
					m.supply_target_densities_func(m, { x, y, z }, target.data() + at);
				}
		hip_check(bfvm_bulk_store_rates(ctx_.ctx, supply.data(), uptake.data(), target.data()),
				  "bfvm_bulk_store_rates");
	}
	has_rates_ = true;
}

void bulk_solver::solve(microenvironment&)
{
	if (!has_rates_)
		throw std::runtime_error("bulk_solver: no rates (set the bulk functions before initialize)");
	once([&] { hip_check(bfvm_bulk_solve(ctx_.ctx), "bfvm_bulk_solve"); });
}
