// Solver selection (include/BioFVM/solver.h:1-31 of the reference).  This library provides one backend,
// BioFVM_SOLVER_IMPL_HIP: biofvm::solver is the MI355X device solver.
#pragma once

#ifndef BioFVM_SOLVER_IMPL_HIP
	#define BioFVM_SOLVER_IMPL_HIP 1
#endif

#if BioFVM_SOLVER_IMPL_HIP
	#include "../../src/solver/device/HIP/solver.h"

namespace biofvm {
using solver = solvers::device::solver;
using cell_solver = solvers::device::cell_solver;
} // namespace biofvm

#endif
