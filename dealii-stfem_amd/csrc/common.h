// Shared definitions of the stfem HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include "grid.h"
#include "status.h"

#include <cstdint>
#include <stdexcept>
#include <string>

namespace stfem
{
#define STFEM_HIP(call)                                                                           \
  do                                                                                              \
    {                                                                                             \
      hipError_t e_ = (call);                                                                     \
      if (e_ != hipSuccess)                                                                       \
        throw ::stfem::Error(::stfem::STATUS_ERR_HIP, std::string(#call) + ": " +                 \
                                                        hipGetErrorString(e_) + " (" __FILE__     \
                                                        ":" + std::to_string(__LINE__) + ")");   \
    }                                                                                             \
  while (0)

} // namespace stfem
