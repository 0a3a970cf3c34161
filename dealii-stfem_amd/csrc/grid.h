// Structured grid description shared by the host code and the kernels
// (plain C++, no HIP dependency).
#pragma once

namespace stfem
{
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

  // host helpers: lattice position of cell c, and whether lattice node
  // (gx, gy, gz) lies on a Dirichlet face
  inline void cell_xyz(const GridDesc &g, long c, int &cx, int &cy, int &cz)
  {
    cx = int(c % g.nc[0]);
    cy = int((c / g.nc[0]) % g.nc[1]);
    cz = (g.dim == 3) ? int(c / ((long)g.nc[0] * g.nc[1])) : 0;
  }
  inline bool node_constrained(const GridDesc &g, long gx, long gy, long gz)
  {
    const int d = g.dirichlet;
    bool      c = (gx == 0 && (d & 1)) || (gx == g.nn[0] - 1 && (d & 2)) ||
                 (gy == 0 && (d & 4)) || (gy == g.nn[1] - 1 && (d & 8));
    if (g.dim == 3)
      c = c || (gz == 0 && (d & 16)) || (gz == g.nn[2] - 1 && (d & 32));
    return c;
  }
} // namespace stfem
