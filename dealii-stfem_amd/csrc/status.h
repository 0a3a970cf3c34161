// Status codes and the error type of the stfem library (host C++ only, no
// HIP dependency, so that host-only parts build with a plain C++ compiler).
#pragma once

#include <stdexcept>
#include <string>

namespace stfem
{
  struct Error : public std::runtime_error
  {
    int code;
    Error(int c, const std::string &m)
      : std::runtime_error(m)
      , code(c)
    {}
  };

  // status codes of the C ABI (include/stfem.h)
  enum Status : int
  {
    STATUS_OK               = 0,
    STATUS_ERR_ARG          = -1,
    STATUS_ERR_HIP          = -2,
    STATUS_ERR_NO_CONVERGE  = -3,
    STATUS_ERR_UNSUPPORTED  = -4,
    STATUS_ERR_NOT_FINITE   = -5,
    STATUS_ERR_INTERNAL     = -6
  };

#define STFEM_REQUIRE(cond, msg)                                                                  \
  do                                                                                              \
    {                                                                                             \
      if (!(cond))                                                                                \
        throw ::stfem::Error(::stfem::STATUS_ERR_ARG, msg);                                       \
    }                                                                                             \
  while (0)

#define STFEM_UNSUPPORTED(cond, msg)                                                              \
  do                                                                                              \
    {                                                                                             \
      if (!(cond))                                                                                \
        throw ::stfem::Error(::stfem::STATUS_ERR_UNSUPPORTED, msg);                               \
    }                                                                                             \
  while (0)
} // namespace stfem
