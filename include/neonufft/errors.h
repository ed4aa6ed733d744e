#pragma once

#include <neonufft/config.h>

/* Error codes, same values as the reference (include/neonufft/errors.h:5-38). */
enum NeonufftError {
  NEONUFFT_SUCCESS,
  NEONUFFT_GENERIC_ERROR,
  NEONUFFT_MEMORY_ALLOC_ERROR,
  NEONUFFT_UNKNOWN_ERROR,
  NEONUFFT_INTERNAL_ERROR,
  NEONUFFT_INPUT_ERROR,
  NEONUFFT_GPU_ERROR,
  NEONUFFT_NOT_IMPLEMENTED_ERROR,
};

#ifndef __cplusplus
typedef enum NeonufftError NeonufftError;
#endif
