#include "device_context.h"

#include <vector>

#include "agent_data.h"

using namespace biofvm;
using namespace biofvm::solvers::device;

device_context::device_context() { hip_check(bfvm_create(0, &ctx), "bfvm_create"); }

device_context::~device_context() { bfvm_destroy(ctx); }

// solver::initialize: the mesh, substrates and Dirichlet description of `m` (CUDA/device_context.cpp:19-26,
// CUDA/dirichlet_solver.cpp:17-88, CUDA/diffusion_solver.cpp:41-103 of the reference).
void device_context::initialize(microenvironment& m)
{
	static_assert(sizeof(bool) == 1, "Dirichlet conditions are passed as one byte per entry");
	bfvm_microenvironment_desc d {};
	d.dims = m.mesh.dims;
	for (int i = 0; i < 3; i++)
	{
		d.bounding_box_mins[i] = m.mesh.bounding_box_mins[i];
		d.bounding_box_maxs[i] = m.mesh.bounding_box_maxs[i];
		d.voxel_shape[i] = m.mesh.voxel_shape[i];
		d.dirichlet_min_boundary_values[i] = m.dirichlet_min_boundary_values[i].get();
		d.dirichlet_max_boundary_values[i] = m.dirichlet_max_boundary_values[i].get();
		d.dirichlet_min_boundary_conditions[i] = (const uint8_t*)m.dirichlet_min_boundary_conditions[i].get();
		d.dirichlet_max_boundary_conditions[i] = (const uint8_t*)m.dirichlet_max_boundary_conditions[i].get();
	}
	d.substrates_count = m.substrates_count;
	d.diffusion_time_step = m.diffusion_time_step;
	d.diffusion_coefficients = m.diffusion_coefficients.get();
	d.decay_rates = m.decay_rates.get();
	d.dirichlet_interior_voxels_count = m.dirichlet_interior_voxels_count;
	d.dirichlet_interior_voxels = m.dirichlet_interior_voxels.get();
	d.dirichlet_interior_values = m.dirichlet_interior_values.get();
	d.dirichlet_interior_conditions = (const uint8_t*)m.dirichlet_interior_conditions.get();
	d.compute_internalized_substrates = m.compute_internalized_substrates ? 1 : 0;
	hip_check(bfvm_initialize(ctx, &d), "bfvm_initialize");
	initialized = true;
}

// store_data_to_solver (CUDA/device_context.cpp:56-97 of the reference): densities and the agent SoA
void device_context::copy_to_device(microenvironment& m)
{
	hip_check(bfvm_store_densities(ctx, m.substrate_densities.get()), "bfvm_store_densities");
	auto& a = get_agent_data(m);
	bfvm_agents_desc ad {};
	ad.agents_count = a.agents_count;
	if (a.agents_count > 0)
	{
		ad.secretion_rates = a.secretion_rates.data();
		ad.saturation_densities = a.saturation_densities.data();
		ad.uptake_rates = a.uptake_rates.data();
		ad.net_export_rates = a.net_export_rates.data();
		ad.internalized_substrates = a.internalized_substrates.data();
		ad.fraction_released_at_death = a.fraction_released_at_death.data();
		ad.volumes = a.volumes.data();
		ad.positions = a.positions.data();
	}
	hip_check(bfvm_store_agents(ctx, &ad), "bfvm_store_agents");
	m.data_residency = data_residency::device;
}

// load_data_from_solver (CUDA/device_context.cpp:99-116 of the reference), plus the gradients, which the
// device computes here
void device_context::copy_to_host(microenvironment& m)
{
	hip_check(bfvm_load_densities(ctx, m.substrate_densities.get()), "bfvm_load_densities");
	hip_check(bfvm_load_gradients(ctx, m.gradients.get()), "bfvm_load_gradients");
	auto& a = get_agent_data(m);
	if (a.agents_count > 0)
		hip_check(bfvm_load_internalized(ctx, a.internalized_substrates.data()), "bfvm_load_internalized");
	m.data_residency = data_residency::host;
}
