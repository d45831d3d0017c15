// Step-by-step construction of a microenvironment with validation.  Public interface of the reference's
// builder (include/BioFVM/microenvironment_builder.h:9-72), so that setup code written against it compiles
// unchanged; the bookkeeping behind it is this facade's own: one record per substrate and one per interior
// Dirichlet node, flattened into the microenvironment's arrays only by build().
#pragma once

#include <array>
#include <optional>
#include <string>
#include <vector>

#include "mesh.h"
#include "microenvironment.h"
#include "types.h"

namespace biofvm {

class microenvironment_builder
{
	// everything known about one substrate, including its six boundary faces (index [is_max][axis])
	struct substrate_spec
	{
		std::string name, units;
		real_t diffusion_coefficient = 0, decay_rate = 0, initial_condition = 0;
		std::array<std::array<real_t, 3>, 2> face_value {};
		std::array<std::array<bool, 3>, 2> face_condition {};
	};

	// one interior Dirichlet voxel with a value and a condition per substrate
	struct dirichlet_node
	{
		point_t<index_t, 3> voxel;
		std::vector<real_t> values;
		std::vector<bool> conditions;
	};

	std::string name, time_units, space_units;
	real_t time_step = 0.01;
	std::optional<cartesian_mesh> mesh_;

	std::vector<substrate_spec> substrates_;
	std::vector<dirichlet_node> nodes_;

	bulk_func_t supply_rate_func, uptake_rate_func, supply_target_densities_func;

	bool compute_internalized_substrates = false;

	void fill_dirichlet_vectors(microenvironment& m);

public:
	void set_name(const std::string& name);
	void set_time_units(const std::string& units);
	void set_space_units(const std::string& units);

	void set_time_step(real_t time_step);

	void resize(index_t dims, point_t<index_t, 3> bounding_box_mins, point_t<index_t, 3> bounding_box_maxs,
				point_t<index_t, 3> voxel_shape);

	void add_density(const std::string& name, const std::string& units, real_t diffusion_coefficient = 0,
					 real_t decay_rate = 0, real_t initial_condition = 0);
	std::size_t get_density_index(const std::string& name) const;

	void add_dirichlet_node(point_t<index_t, 3> voxel_index, std::vector<real_t> values,
							std::vector<bool> conditions = {});
	void add_boundary_dirichlet_conditions(std::size_t density_index, point_t<real_t, 3> mins_values,
										   point_t<real_t, 3> maxs_values,
										   point_t<bool, 3> mins_conditions = { true, true, true },
										   point_t<bool, 3> maxs_conditions = { true, true, true });

	void set_bulk_functions(bulk_func_t supply_rate_func, bulk_func_t uptake_rate_func,
							bulk_func_t supply_target_densities_func);

	void do_compute_internalized_substrates();

	microenvironment build();
};

} // namespace biofvm
