// Basic types of the BioFVM API (include/BioFVM/types.h:8-18 of the reference).
#pragma once

#include <array>
#include <cstdint>

namespace biofvm {

using agent_id_t = uint32_t;
using index_t = int32_t;

// The MI355X solver library computes in FP32 (the reference's default build); a double build of the
// reference (USE_DOUBLES) has no device counterpart here.
using real_t = float;

template <typename T, int dims>
using point_t = std::array<T, dims>;

} // namespace biofvm
