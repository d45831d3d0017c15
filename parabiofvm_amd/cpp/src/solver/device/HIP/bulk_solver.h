// Bulk supply and uptake on the device (host/bulk_solver.h:21-36 of the reference, which only has a host
// version): per voxel and substrate D = (D + dt*S*T) / (1 + dt*(U + S)).  The reference calls its three
// per-voxel callbacks at every solve; here initialize() and update_rates() evaluate them for every voxel on
// the host and upload the values as fields, and solve() runs on the device with the last uploaded fields.
#pragma once

#include "device_context.h"

namespace biofvm {
namespace solvers {
namespace device {

class bulk_solver
{
	device_context& ctx_;
	bool has_rates_ = false;

	void upload(microenvironment& m);

public:
	bulk_solver(device_context& ctx);

	// samples m.supply_rate_func, m.uptake_rate_func and m.supply_target_densities_func (if all are set);
	// called by solver::initialize, once per team
	void initialize(microenvironment& m);
	// samples them again (after the callbacks' inputs changed)
	void update_rates(microenvironment& m);

	void solve(microenvironment& m);
};

} // namespace device
} // namespace solvers
} // namespace biofvm
