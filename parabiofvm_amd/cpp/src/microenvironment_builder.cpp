#include "microenvironment_builder.h"

#include <stdexcept>

using namespace biofvm;

void microenvironment_builder::set_name(const std::string& n) { name = n; }
void microenvironment_builder::set_time_units(const std::string& u) { time_units = u; }
void microenvironment_builder::set_space_units(const std::string& u) { space_units = u; }
void microenvironment_builder::set_time_step(real_t dt) { time_step = dt; }

void microenvironment_builder::resize(index_t dims, point_t<index_t, 3> bounding_box_mins,
									  point_t<index_t, 3> bounding_box_maxs, point_t<index_t, 3> voxel_shape)
{
	mesh_.emplace(dims, bounding_box_mins, bounding_box_maxs, voxel_shape);
}

void microenvironment_builder::add_density(const std::string& n, const std::string& units,
										   real_t diffusion_coefficient, real_t decay_rate, real_t initial_condition)
{
	substrate_spec spec;
	spec.name = n;
	spec.units = units;
	spec.diffusion_coefficient = diffusion_coefficient;
	spec.decay_rate = decay_rate;
	spec.initial_condition = initial_condition;
	substrates_.push_back(std::move(spec));
}

std::size_t microenvironment_builder::get_density_index(const std::string& n) const
{
	for (std::size_t i = 0; i < substrates_.size(); i++)
		if (substrates_[i].name == n)
			return i;
	throw std::runtime_error("no density named '" + n + "'");
}

void microenvironment_builder::add_dirichlet_node(point_t<index_t, 3> voxel_index, std::vector<real_t> values,
												  std::vector<bool> conditions)
{
	const std::size_t S = substrates_.size();
	if (values.size() != S)
		throw std::runtime_error("Dirichlet node: expected one value per density (" + std::to_string(S) + "), got " +
								 std::to_string(values.size()));
	if (conditions.empty())
		conditions = std::vector<bool>(S, true); // no conditions given: every substrate is fixed
	else if (conditions.size() != S)
		throw std::runtime_error("Dirichlet node: expected one condition per density (" + std::to_string(S) +
								 "), got " + std::to_string(conditions.size()));
	if (!mesh_.has_value())
		throw std::runtime_error("Dirichlet node: set the mesh (resize) first");
	nodes_.push_back(dirichlet_node { voxel_index, std::move(values), std::move(conditions) });
}

void microenvironment_builder::add_boundary_dirichlet_conditions(std::size_t density_index,
																 point_t<real_t, 3> mins_values,
																 point_t<real_t, 3> maxs_values,
																 point_t<bool, 3> mins_conditions,
																 point_t<bool, 3> maxs_conditions)
{
	if (density_index >= substrates_.size())
		throw std::runtime_error("boundary Dirichlet: density index " + std::to_string(density_index) +
								 " out of range");
	substrate_spec& spec = substrates_[density_index];
	for (int axis = 0; axis < 3; axis++)
	{
		spec.face_value[0][axis] = mins_values[axis];
		spec.face_value[1][axis] = maxs_values[axis];
		spec.face_condition[0][axis] = mins_conditions[axis];
		spec.face_condition[1][axis] = maxs_conditions[axis];
	}
}

void microenvironment_builder::set_bulk_functions(bulk_func_t supply, bulk_func_t uptake, bulk_func_t targets)
{
	supply_rate_func = std::move(supply);
	uptake_rate_func = std::move(uptake);
	supply_target_densities_func = std::move(targets);
}

void microenvironment_builder::do_compute_internalized_substrates() { compute_internalized_substrates = true; }

// The microenvironment holds a face's arrays only when at least one substrate has a condition on that face
// (a null array means "no Dirichlet face", include/BioFVM/microenvironment.h:47-50).
void microenvironment_builder::fill_dirichlet_vectors(microenvironment& m)
{
	const index_t S = m.substrates_count;
	for (int side = 0; side < 2; side++)
	{
		auto& values = side == 0 ? m.dirichlet_min_boundary_values : m.dirichlet_max_boundary_values;
		auto& conds = side == 0 ? m.dirichlet_min_boundary_conditions : m.dirichlet_max_boundary_conditions;
		for (index_t axis = 0; axis < m.mesh.dims; axis++)
		{
			bool used = false;
			for (const substrate_spec& spec : substrates_)
				used = used || spec.face_condition[side][axis];
			if (!used)
				continue;
			values[axis] = std::make_unique<real_t[]>(S);
			conds[axis] = std::make_unique<bool[]>(S);
			for (index_t s = 0; s < S; s++)
			{
				values[axis][s] = substrates_[s].face_value[side][axis];
				conds[axis][s] = substrates_[s].face_condition[side][axis];
			}
		}
	}
}

microenvironment microenvironment_builder::build()
{
	if (!mesh_.has_value())
		throw std::runtime_error("microenvironment: set the mesh (resize) before build");
	if (substrates_.empty())
		throw std::runtime_error("microenvironment: add at least one density before build");

	const index_t S = (index_t)substrates_.size();
	std::vector<real_t> initial(S);
	for (index_t s = 0; s < S; s++)
		initial[s] = substrates_[s].initial_condition;

	microenvironment m(*mesh_, S, time_step, initial.data());
	m.name = name;
	m.time_units = time_units;
	m.space_units = space_units;
	m.diffusion_coefficients = std::make_unique<real_t[]>(S);
	m.decay_rates = std::make_unique<real_t[]>(S);
	for (index_t s = 0; s < S; s++)
	{
		m.substrates_names.push_back(substrates_[s].name);
		m.substrates_units.push_back(substrates_[s].units);
		m.diffusion_coefficients[s] = substrates_[s].diffusion_coefficient;
		m.decay_rates[s] = substrates_[s].decay_rate;
	}

	// interior nodes flattened in insertion order: voxels [n][dims], values and conditions [n][S]
	const index_t dims = m.mesh.dims;
	const std::size_t n = nodes_.size();
	m.dirichlet_interior_voxels_count = (index_t)n;
	m.dirichlet_interior_voxels = std::make_unique<index_t[]>(n * dims);
	m.dirichlet_interior_values = std::make_unique<real_t[]>(n * S);
	m.dirichlet_interior_conditions = std::make_unique<bool[]>(n * S);
	for (std::size_t i = 0; i < n; i++)
	{
		for (index_t d = 0; d < dims; d++)
			m.dirichlet_interior_voxels[i * dims + d] = nodes_[i].voxel[d];
		for (index_t s = 0; s < S; s++)
		{
			m.dirichlet_interior_values[i * S + s] = nodes_[i].values[s];
			m.dirichlet_interior_conditions[i * S + s] = nodes_[i].conditions[s];
		}
	}

	fill_dirichlet_vectors(m);

	m.supply_rate_func = supply_rate_func;
	m.uptake_rate_func = uptake_rate_func;
	m.supply_target_densities_func = supply_target_densities_func;
	m.compute_internalized_substrates = compute_internalized_substrates;
	return m;
}
