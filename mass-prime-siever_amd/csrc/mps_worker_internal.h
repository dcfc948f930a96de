// Helpers shared by the plain (mps_gpu.hip) and batched (mps_batch.hip) device workers.
#pragma once

#include <string>

#include "mps_gpu.h"
#include "mps_kernels.h"

namespace mps {

// Kernel constants of one item (the conversions of Gpu.hpp:862-875 and gpu_classes,
// Math.hpp:324-337).  Returns false, with the reason, when the BSGS table would not fit.
bool item_params(const mps_item* it, bool classes_on, ItemParams& ip, uint32_t& tbits, std::string& why);

}  // namespace mps
