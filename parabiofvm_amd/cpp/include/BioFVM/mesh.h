// Cartesian mesh (include/BioFVM/mesh.h:7-29, src/mesh.cpp:7-80 of the reference).
#pragma once

#include <cstddef>

#include "types.h"

namespace biofvm {

struct cartesian_mesh
{
	index_t dims; // 1, 2 or 3

	point_t<index_t, 3> bounding_box_mins;
	point_t<index_t, 3> bounding_box_maxs;

	point_t<index_t, 3> voxel_shape; // [dx, dy, dz]
	point_t<index_t, 3> grid_shape;	 // voxels per axis, 1 on axes beyond dims

	cartesian_mesh(index_t dims, point_t<index_t, 3> bounding_box_mins, point_t<index_t, 3> bounding_box_maxs,
				   point_t<index_t, 3> voxel_shape);

	std::size_t voxel_count() const;
	index_t voxel_volume() const;

	// voxel containing `position` (dims coordinates): (index_t)((p - min) / dx) per axis, no clamping
	template <index_t dims>
	point_t<index_t, 3> voxel_position(const real_t* position) const;

	point_t<real_t, 3> voxel_center(point_t<index_t, 3> position) const;

	void display_info();
};

} // namespace biofvm
