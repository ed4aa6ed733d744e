#pragma once

#include <neonufft/config.h>
#include <neonufft/enums.h>
#include <neonufft/types.hpp>

namespace neonufft {

/* Plan options, same members, order and defaults as the reference
 * (include/neonufft/plan.hpp:19-29). Passed by value, so the layout is ABI.
 * The GPU plans read tol, upsampfac, recenter_threshold and order; the other
 * members are accepted and ignored exactly as the reference GPU plan does. */
struct NEONUFFT_EXPORT Options {
  double tol = 0.001;
  double upsampfac = 2.0;
  double recenter_threshold = 0.1;
  int num_threads = 0;
  bool sort_input = true;
  bool sort_output = true;
  bool kernel_approximation = false;
  NeonufftModeOrder order = NEONUFFT_MODE_ORDER_CMCL;
  NeonufftKernelType kernel_type = NEONUFFT_ES_KERNEL;
};

/* The host (CPU) neonufft::Plan / PlanT3 of the reference
 * (include/neonufft/plan.hpp:31-85) are outside this library's scope: only
 * the GPU plans in <neonufft/gpu/plan.hpp> are provided. */

}  // namespace neonufft
