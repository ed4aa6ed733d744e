#pragma once

#include <neonufft/config.h>

/* Same enumerators and values as the reference (include/neonufft/enums.h:6-11). */
enum NeonufftModeOrder {
  NEONUFFT_MODE_ORDER_FFT,
  NEONUFFT_MODE_ORDER_CMCL,
};

enum NeonufftKernelType { NEONUFFT_ES_KERNEL };

#ifndef __cplusplus
typedef enum NeonufftModeOrder NeonufftModeOrder;
typedef enum NeonufftKernelType NeonufftKernelType;
#endif
