#include "microenvironment.h"

#include <cmath>
#include <iostream>

#include "agent_container.h"

using namespace biofvm;

// src/microenvironment.cpp:27-53 of the reference: allocate the fields, densities start at the initial
// condition of each substrate, no Dirichlet conditions, agents in an empty container
microenvironment::microenvironment(cartesian_mesh mesh, index_t substrates_count, real_t time_step,
								   const real_t* initial_conditions)
	: mesh(mesh),
	  agents(std::make_unique<agent_container>(*this)),
	  substrates_count(substrates_count),
	  diffusion_time_step(time_step),
	  substrate_densities(std::make_unique<real_t[]>(substrates_count * mesh.voxel_count())),
	  gradients(std::make_unique<real_t[]>(mesh.dims * substrates_count * mesh.voxel_count())),
	  dirichlet_interior_voxels_count(0),
	  compute_internalized_substrates(false),
	  data_residency(solvers::data_residency::host)
{
	const std::size_t V = mesh.voxel_count();
#pragma omp parallel for
	for (long long v = 0; v < (long long)V; v++)
		for (index_t s = 0; s < substrates_count; s++)
			substrate_densities[(std::size_t)v * substrates_count + s] = initial_conditions[s];
}

index_t microenvironment::find_substrate_index(const std::string& n) const
{
	for (index_t i = 0; i < (index_t)substrates_names.size(); i++)
		if (substrates_names[i] == n)
			return i;
	return -1;
}

real_t microenvironment::substrate_density_value(point_t<index_t, 3> position, index_t substrate_index) const
{
	std::size_t v = position[0];
	if (mesh.dims >= 2)
		v += (std::size_t)position[1] * mesh.grid_shape[0];
	if (mesh.dims >= 3)
		v += (std::size_t)position[2] * mesh.grid_shape[0] * mesh.grid_shape[1];
	return substrate_densities[v * substrates_count + substrate_index];
}

void microenvironment::display_info()
{
	std::cout << "\nMicroenvironment " << name << ":" << std::endl;
	mesh.display_info();
	std::cout << "Densities (" << substrates_count << "):" << std::endl;
	for (index_t i = 0; i < (index_t)substrates_names.size(); i++)
	{
		std::cout << "  " << substrates_names[i] << " [" << substrates_units[i] << "]: D = " << diffusion_coefficients[i]
				  << " " << space_units << "^2/" << time_units << ", decay = " << decay_rates[i] << " 1/" << time_units
				  << ", length scale = " << std::sqrt(diffusion_coefficients[i] / (1e-12 + decay_rates[i])) << " "
				  << space_units << std::endl;
	}
}
