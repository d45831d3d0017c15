#include "microenvironment_builder.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>

using namespace biofvm;

void microenvironment_builder::set_name(const std::string& n) { name = n; }
void microenvironment_builder::set_time_units(const std::string& u) { time_units = u; }
void microenvironment_builder::set_space_units(const std::string& u) { space_units = u; }
void microenvironment_builder::set_time_step(real_t dt) { time_step = dt; }

void microenvironment_builder::resize(index_t dims, point_t<index_t, 3> bounding_box_mins,
									  point_t<index_t, 3> bounding_box_maxs, point_t<index_t, 3> voxel_shape)
{
	mesh_ = cartesian_mesh(dims, bounding_box_mins, bounding_box_maxs, voxel_shape);
}

void microenvironment_builder::add_density(const std::string& n, const std::string& units,
										   real_t diffusion_coefficient, real_t decay_rate, real_t initial_condition)
{
	substrates_names.push_back(n);
	substrates_units.push_back(units);
	diffusion_coefficients.push_back(diffusion_coefficient);
	decay_rates.push_back(decay_rate);
	initial_conditions.push_back(initial_condition);
	boundary_dirichlet_mins_values.push_back({ 0, 0, 0 });
	boundary_dirichlet_maxs_values.push_back({ 0, 0, 0 });
	boundary_dirichlet_mins_conditions.push_back({ false, false, false });
	boundary_dirichlet_maxs_conditions.push_back({ false, false, false });
}

std::size_t microenvironment_builder::get_density_index(const std::string& n) const
{
	auto it = std::find(substrates_names.begin(), substrates_names.end(), n);
	if (it == substrates_names.end())
		throw std::runtime_error("Density " + n + " not found");
	return (std::size_t)(it - substrates_names.begin());
}

// src/microenvironment_builder.cpp:52-83 of the reference
void microenvironment_builder::add_dirichlet_node(point_t<index_t, 3> voxel_index, std::vector<real_t> values,
												  std::vector<bool> conditions)
{
	if (values.size() != substrates_names.size())
		throw std::runtime_error("Dirichlet node values size does not match the number of densities");
	if (!conditions.empty() && conditions.size() != substrates_names.size())
		throw std::runtime_error("Dirichlet node conditions size does not match the number of densities");
	if (!mesh_)
		throw std::runtime_error("Dirichlet node cannot be added without a mesh");
	for (index_t d = 0; d < mesh_->dims; d++)
		dirichlet_voxels.push_back(voxel_index[d]);
	dirichlet_values.insert(dirichlet_values.end(), values.begin(), values.end());
	if (conditions.empty())
		conditions.assign(values.size(), true);
	dirichlet_conditions.insert(dirichlet_conditions.end(), conditions.begin(), conditions.end());
}

void microenvironment_builder::add_boundary_dirichlet_conditions(std::size_t density_index,
																 point_t<real_t, 3> mins_values,
																 point_t<real_t, 3> maxs_values,
																 point_t<bool, 3> mins_conditions,
																 point_t<bool, 3> maxs_conditions)
{
	if (density_index >= substrates_names.size())
		throw std::runtime_error("Density index out of bounds");
	boundary_dirichlet_mins_values[density_index] = mins_values;
	boundary_dirichlet_maxs_values[density_index] = maxs_values;
	boundary_dirichlet_mins_conditions[density_index] = mins_conditions;
	boundary_dirichlet_maxs_conditions[density_index] = maxs_conditions;
}

void microenvironment_builder::set_bulk_functions(bulk_func_t supply, bulk_func_t uptake, bulk_func_t targets)
{
	supply_rate_func = std::move(supply);
	uptake_rate_func = std::move(uptake);
	supply_target_densities_func = std::move(targets);
}

void microenvironment_builder::do_compute_internalized_substrates() { compute_internalized_substrates = true; }

// src/microenvironment_builder.cpp:112-144 of the reference: a face array exists only when some substrate
// has a condition on it
void microenvironment_builder::fill_dirichlet_vectors(microenvironment& m)
{
	const index_t S = m.substrates_count;
	auto fill = [&](index_t d, const std::vector<point_t<real_t, 3>>& vals, const std::vector<point_t<bool, 3>>& conds,
					point_t<std::unique_ptr<real_t[]>, 3>& out_v, point_t<std::unique_ptr<bool[]>, 3>& out_c) {
		bool any = false;
		for (index_t s = 0; s < S; s++)
			any |= conds[s][d];
		if (!any)
			return;
		out_v[d] = std::make_unique<real_t[]>(S);
		out_c[d] = std::make_unique<bool[]>(S);
		for (index_t s = 0; s < S; s++)
		{
			out_v[d][s] = vals[s][d];
			out_c[d][s] = conds[s][d];
		}
	};
	for (index_t d = 0; d < m.mesh.dims; d++)
	{
		fill(d, boundary_dirichlet_mins_values, boundary_dirichlet_mins_conditions, m.dirichlet_min_boundary_values,
			 m.dirichlet_min_boundary_conditions);
		fill(d, boundary_dirichlet_maxs_values, boundary_dirichlet_maxs_conditions, m.dirichlet_max_boundary_values,
			 m.dirichlet_max_boundary_conditions);
	}
}

microenvironment microenvironment_builder::build()
{
	if (!mesh_)
		throw std::runtime_error("Microenvironment cannot be built without a mesh");
	if (substrates_names.empty())
		throw std::runtime_error("Microenvironment cannot be built with no densities");

	microenvironment m(*mesh_, (index_t)substrates_names.size(), time_step, initial_conditions.data());
	m.name = std::move(name);
	m.time_units = std::move(time_units);
	m.space_units = std::move(space_units);
	m.substrates_names = std::move(substrates_names);
	m.substrates_units = std::move(substrates_units);

	m.diffusion_coefficients = std::make_unique<real_t[]>(diffusion_coefficients.size());
	std::copy(diffusion_coefficients.begin(), diffusion_coefficients.end(), m.diffusion_coefficients.get());
	m.decay_rates = std::make_unique<real_t[]>(decay_rates.size());
	std::copy(decay_rates.begin(), decay_rates.end(), m.decay_rates.get());

	m.dirichlet_interior_voxels_count = (index_t)(dirichlet_voxels.size() / m.mesh.dims);
	m.dirichlet_interior_voxels = std::make_unique<index_t[]>(dirichlet_voxels.size());
	std::copy(dirichlet_voxels.begin(), dirichlet_voxels.end(), m.dirichlet_interior_voxels.get());
	m.dirichlet_interior_values = std::make_unique<real_t[]>(dirichlet_values.size());
	std::copy(dirichlet_values.begin(), dirichlet_values.end(), m.dirichlet_interior_values.get());
	m.dirichlet_interior_conditions = std::make_unique<bool[]>(dirichlet_conditions.size());
	std::copy(dirichlet_conditions.begin(), dirichlet_conditions.end(), m.dirichlet_interior_conditions.get());

	fill_dirichlet_vectors(m);

	m.supply_rate_func = std::move(supply_rate_func);
	m.uptake_rate_func = std::move(uptake_rate_func);
	m.supply_target_densities_func = std::move(supply_target_densities_func);
	m.compute_internalized_substrates = compute_internalized_substrates;
	return m;
}
