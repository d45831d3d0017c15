// Element moves between slots of the agent arrays (the helpers of include/BioFVM/data_utils.h), for the
// trivially copyable element types the SoA holds.
#pragma once

#include <algorithm>
#include <type_traits>

#include "types.h"

template <typename T>
void move_scalar(T* dst, const T* src)
{
	static_assert(std::is_trivially_copyable_v<T>);
	*dst = *src;
}

template <typename T>
void move_vector(T* dst, const T* src, biofvm::index_t size)
{
	static_assert(std::is_trivially_copyable_v<T>);
	std::copy_n(src, size, dst);
}
