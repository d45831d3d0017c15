#include "mesh.h"

#include <iostream>

using namespace biofvm;

// grid = ceil((max - min) / dx) per mesh axis, 1 on the others (src/mesh.cpp:7-25 of the reference)
cartesian_mesh::cartesian_mesh(index_t dims, point_t<index_t, 3> bounding_box_mins,
							   point_t<index_t, 3> bounding_box_maxs, point_t<index_t, 3> voxel_shape)
	: dims(dims), bounding_box_mins(bounding_box_mins), bounding_box_maxs(bounding_box_maxs), voxel_shape(voxel_shape)
{
	grid_shape = { 1, 1, 1 };
	for (index_t d = 0; d < dims && d < 3; d++)
		grid_shape[d] = (bounding_box_maxs[d] - bounding_box_mins[d] + voxel_shape[d] - 1) / voxel_shape[d];
}

std::size_t cartesian_mesh::voxel_count() const
{
	return (std::size_t)grid_shape[0] * (std::size_t)grid_shape[1] * (std::size_t)grid_shape[2];
}

index_t cartesian_mesh::voxel_volume() const { return voxel_shape[0] * voxel_shape[1] * voxel_shape[2]; }

template <index_t d>
static point_t<index_t, 3> position_to_voxel(const cartesian_mesh& m, const real_t* p)
{
	point_t<index_t, 3> r = { 0, 0, 0 };
	for (index_t i = 0; i < d; i++)
		r[i] = (index_t)((p[i] - m.bounding_box_mins[i]) / m.voxel_shape[i]);
	return r;
}

template <>
point_t<index_t, 3> cartesian_mesh::voxel_position<1>(const real_t* position) const
{
	return position_to_voxel<1>(*this, position);
}

template <>
point_t<index_t, 3> cartesian_mesh::voxel_position<2>(const real_t* position) const
{
	return position_to_voxel<2>(*this, position);
}

template <>
point_t<index_t, 3> cartesian_mesh::voxel_position<3>(const real_t* position) const
{
	return position_to_voxel<3>(*this, position);
}

point_t<real_t, 3> cartesian_mesh::voxel_center(point_t<index_t, 3> position) const
{
	point_t<real_t, 3> c;
	for (int d = 0; d < 3; d++)
		c[d] = (real_t)(position[d] * voxel_shape[d] + voxel_shape[d] / 2.0 + bounding_box_mins[d]);
	return c;
}

void cartesian_mesh::display_info()
{
	std::cout << "\nMesh: " << dims << "-D, domain";
	for (int d = 0; d < dims; d++)
		std::cout << " [" << bounding_box_mins[d] << "," << bounding_box_maxs[d] << "]";
	std::cout << ", voxel " << voxel_shape[0] << "x" << voxel_shape[1] << "x" << voxel_shape[2] << ", grid "
			  << grid_shape[0] << "x" << grid_shape[1] << "x" << grid_shape[2] << " (" << voxel_count()
			  << " voxels)" << std::endl;
}
