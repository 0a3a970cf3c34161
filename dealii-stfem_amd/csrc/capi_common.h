// Error plumbing shared by the extern "C" entry points: every entry point
// runs its body through guarded(), which maps exceptions to the status codes
// of include/stfem.h and keeps the message for stfem_last_error().
#pragma once

#include <exception>
#include <string>

#include "status.h"

namespace stfem
{
  // message of the last failure on the calling thread
  inline std::string &last_error_message()
  {
    static thread_local std::string m;
    return m;
  }

  template <typename F>
  int guarded(F &&f)
  {
    try
      {
        f();
        return STATUS_OK;
      }
    catch (const Error &e)
      {
        last_error_message() = e.what();
        return e.code;
      }
    catch (const std::exception &e)
      {
        last_error_message() = e.what();
        return STATUS_ERR_INTERNAL;
      }
    catch (...)
      {
        last_error_message() = "unknown error";
        return STATUS_ERR_INTERNAL;
      }
  }
} // namespace stfem
