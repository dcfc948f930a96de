// Helpers shared by the plain (mps_gpu.hip) and batched (mps_batch.hip) device workers.
#pragma once

#include <string>
#include <vector>

#include "mps_gpu.h"
#include "mps_kernels.h"

namespace mps {

// Kernel constants of one item (the conversions of Gpu.hpp:862-875 and gpu_classes,
// Math.hpp:324-337), the largest baby-step count m_max (plain-mode m is capped at mcap) and the
// class-mode size table csz[qp * 32 + s] = (m, giant steps per class).
// Returns false, with the reason, when the BSGS table would not fit.
bool item_params(const mps_item* it, bool classes_on, uint32_t mcap, uint32_t waves, ItemParams& ip,
                 uint64_t& m_max, std::vector<uint2>& csz, std::string& why);

}  // namespace mps
