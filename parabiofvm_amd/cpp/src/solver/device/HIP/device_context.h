// Device context of the MI355X solver: one bfvm_ctx (include/parabiofvm_hip.h) shared by the sub-solvers,
// the counterpart of the reference's device_context (src/solver/device/CUDA/device_context.h:23-56).
#pragma once

#include <stdexcept>
#include <string>

#include "../../common_solver.h"
#include "microenvironment.h"
#include "parabiofvm_hip.h"

namespace biofvm {
namespace solvers {
namespace device {

// Status -> exception (the C ABI never aborts; the facade reports failures as std::runtime_error).
inline void hip_check(int status, const char* what)
{
	if (status != BFVM_OK)
		throw std::runtime_error(std::string(what) + ": " + bfvm_last_error());
}

// Host solvers are called by every thread of an OpenMP team (orphaned worksharing, src/main.cpp:71); the
// device solvers enqueue each call exactly once: inside a parallel region one thread runs `f` and the team
// waits for it, then every thread sees its outcome.
template <typename F>
void once(F&& f);

struct device_context : common_solver
{
	bfvm_ctx* ctx = nullptr;
	bool initialized = false;

	device_context();
	~device_context();
	device_context(const device_context&) = delete;
	device_context& operator=(const device_context&) = delete;

	void initialize(microenvironment& m);

	void copy_to_device(microenvironment& m);
	void copy_to_host(microenvironment& m);
};

} // namespace device
} // namespace solvers
} // namespace biofvm

#include "once.inl"
