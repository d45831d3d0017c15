#pragma once

#include <exception>

#ifdef _OPENMP
	#include <omp.h>
#endif

namespace biofvm {
namespace solvers {
namespace device {

template <typename F>
void once(F&& f)
{
#ifdef _OPENMP
	if (omp_in_parallel())
	{
		static std::exception_ptr failure;
	#pragma omp single
		{
			failure = nullptr;
			try
			{
				f();
			}
			catch (...)
			{
				failure = std::current_exception();
			}
		}
		// implicit barrier of `single`: every thread now sees the outcome
		std::exception_ptr seen = failure;
	#pragma omp barrier
		if (seen)
			std::rethrow_exception(seen);
		return;
	}
#endif
	f();
}

} // namespace device
} // namespace solvers
} // namespace biofvm
