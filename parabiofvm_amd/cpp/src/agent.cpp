#include "agent.h"

#include "agent_data.h"
#include "microenvironment.h"

using namespace biofvm;

agent::agent(agent_id_t id, agent_data& data, index_t index) : index_(index), data_(data), id(id) {}

static real_t* row(std::vector<real_t>& v, index_t i, index_t S) { return v.data() + (std::size_t)i * S; }

real_t* agent::secretion_rates() { return row(data_.secretion_rates, index_, data_.m.substrates_count); }
real_t* agent::saturation_densities() { return row(data_.saturation_densities, index_, data_.m.substrates_count); }
real_t* agent::uptake_rates() { return row(data_.uptake_rates, index_, data_.m.substrates_count); }
real_t* agent::net_export_rates() { return row(data_.net_export_rates, index_, data_.m.substrates_count); }
real_t* agent::internalized_substrates()
{
	return row(data_.internalized_substrates, index_, data_.m.substrates_count);
}
real_t* agent::fraction_released_at_death()
{
	return row(data_.fraction_released_at_death, index_, data_.m.substrates_count);
}
real_t* agent::fraction_transferred_when_ingested()
{
	return row(data_.fraction_transferred_when_ingested, index_, data_.m.substrates_count);
}

real_t& agent::volume() { return data_.volumes[index_]; }

real_t* agent::position() { return data_.positions.data() + (std::size_t)index_ * data_.m.mesh.dims; }

point_t<real_t, 3> agent::get_position() const
{
	point_t<real_t, 3> p = { 0, 0, 0 };
	const index_t dims = data_.m.mesh.dims;
	for (index_t d = 0; d < dims; d++)
		p[d] = data_.positions[(std::size_t)index_ * dims + d];
	return p;
}

point_t<index_t, 3> agent::voxel_index() const
{
	const auto p = get_position();
	switch (data_.m.mesh.dims)
	{
	case 1:
		return data_.m.mesh.voxel_position<1>(p.data());
	case 2:
		return data_.m.mesh.voxel_position<2>(p.data());
	default:
		return data_.m.mesh.voxel_position<3>(p.data());
	}
}

static std::size_t linear_voxel(const cartesian_mesh& mesh, point_t<index_t, 3> v)
{
	return ((std::size_t)v[2] * mesh.grid_shape[1] + v[1]) * mesh.grid_shape[0] + v[0];
}

real_t* agent::nearest_density_vector()
{
	return data_.m.substrate_densities.get() + linear_voxel(data_.m.mesh, voxel_index()) * data_.m.substrates_count;
}

// gradient layouts: [d][z][y][x][s] in 3-D, [y][x][s][d] in 1-D/2-D (src/traits.h:27-33, 55-62, 86-92)
point_t<real_t, 3> agent::nearest_gradient(index_t substrate_index) const
{
	const auto& m = data_.m;
	const std::size_t v = linear_voxel(m.mesh, voxel_index());
	const std::size_t VS = m.mesh.voxel_count() * m.substrates_count;
	const std::size_t e = v * m.substrates_count + substrate_index;
	point_t<real_t, 3> g = { 0, 0, 0 };
	for (index_t d = 0; d < m.mesh.dims; d++)
		g[d] = m.mesh.dims == 3 ? m.gradients[d * VS + e] : m.gradients[e * m.mesh.dims + d];
	return g;
}

index_t agent::index() const { return index_; }
