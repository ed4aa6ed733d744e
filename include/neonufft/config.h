#pragma once
/* Static configuration of the MI355X (gfx950) NeoNUFFT drop-in.
 * Replaces the CMake-generated config.h of the reference (cmake/config.h.in:1-11):
 * this build has exactly one backend, ROCm/HIP: the host-memory plans of
 * <neonufft/plan.hpp> also run on the GPU. */

#define NEONUFFT_VERSION "0.1.0-mi355x"
#define NEONUFFT_ROCM

#ifndef NEONUFFT_EXPORT
#define NEONUFFT_EXPORT __attribute__((visibility("default")))
#endif
