#pragma once

namespace biofvm {
namespace solvers {

// Where the current densities live (include/BioFVM/solver_data_residency.h of the reference).
enum class data_residency
{
	host,
	device
};

} // namespace solvers
} // namespace biofvm
