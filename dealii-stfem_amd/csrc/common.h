// Shared definitions of the stfem HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

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

  // Largest 1D node count (space degree + 1) with compiled kernels.
  constexpr int MAX_N = 6;

  // Structured level description passed to kernels.  Nodes of the
  // Gauss-Lobatto grid are numbered lexicographically (x fastest); a vector
  // stores the nt time components of a node contiguously ([node][t]).
  struct GridDesc
  {
    int  dim;
    int  p;            // space degree
    int  nc[3];        // cells per direction
    int  nn[3];        // nodes per direction (nc*p + 1)
    long n_cells;
    long n_nodes;
    int  dirichlet;    // bit f set: face f (x0,x1,y0,y1,z0,z1) carries zero Dirichlet values
  };
} // namespace stfem
