#pragma once

namespace biofvm {

struct agent_data;
struct microenvironment;

namespace solvers {

// Friend accessor of the agent container's SoA (src/solver/common_solver.h of the reference).
class common_solver
{
protected:
	agent_data& get_agent_data(microenvironment& m);
};

} // namespace solvers
} // namespace biofvm
