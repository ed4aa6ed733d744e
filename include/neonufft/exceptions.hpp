#pragma once

#include <neonufft/config.h>
#include <neonufft/errors.h>

#include <exception>

namespace neonufft {

/* Exception hierarchy with the reference's class names and error codes
 * (include/neonufft/exceptions.hpp:16-103). Messages must be string literals. */
class NEONUFFT_EXPORT GenericError : public std::exception {
public:
  GenericError() : msg_("Neonufft:: Generic error") {}
  explicit GenericError(const char* msg) : msg_(msg) {}
  const char* what() const noexcept override { return msg_; }
  virtual NeonufftError error_code() const noexcept { return NEONUFFT_GENERIC_ERROR; }

private:
  const char* msg_;
};

class NEONUFFT_EXPORT InputError : public GenericError {
public:
  InputError() : GenericError("Neonufft:: Input error") {}
  explicit InputError(const char* msg) : GenericError(msg) {}
  NeonufftError error_code() const noexcept override { return NEONUFFT_INPUT_ERROR; }
};

class NEONUFFT_EXPORT InternalError : public GenericError {
public:
  InternalError() : GenericError("Neonufft:: Internal error") {}
  explicit InternalError(const char* msg) : GenericError(msg) {}
  NeonufftError error_code() const noexcept override { return NEONUFFT_INTERNAL_ERROR; }
};

class NEONUFFT_EXPORT MemoryAllocError : public GenericError {
public:
  MemoryAllocError() : GenericError("Neonufft:: Memory allocation error") {}
  explicit MemoryAllocError(const char* msg) : GenericError(msg) {}
  NeonufftError error_code() const noexcept override { return NEONUFFT_MEMORY_ALLOC_ERROR; }
};

class NEONUFFT_EXPORT NotImplementedError : public GenericError {
public:
  NotImplementedError() : GenericError("Neonufft:: Not implemented") {}
  explicit NotImplementedError(const char* msg) : GenericError(msg) {}
  NeonufftError error_code() const noexcept override { return NEONUFFT_NOT_IMPLEMENTED_ERROR; }
};

class NEONUFFT_EXPORT GPUError : public GenericError {
public:
  GPUError() : GenericError("Neonufft:: GPU error") {}
  explicit GPUError(const char* msg) : GenericError(msg) {}
  NeonufftError error_code() const noexcept override { return NEONUFFT_GPU_ERROR; }
};

class NEONUFFT_EXPORT GPUFFTError : public GenericError {
public:
  GPUFFTError() : GenericError("Neonufft:: GPU FFT error") {}
  explicit GPUFFTError(const char* msg) : GenericError(msg) {}
  NeonufftError error_code() const noexcept override { return NEONUFFT_GPU_ERROR; }
};

}  // namespace neonufft
