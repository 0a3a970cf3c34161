// Host-side mesh levels, shape data and patch matrices (mesh.h).
#include "mesh.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "status.h"

namespace stfem
{
  // =============================================================== Shape1D
  Shape1D::Shape1D(int p_)
    : p(p_)
    , n(p_ + 1)
  {
    gll           = quad_gauss_lobatto(n).x;
    const auto gq = quad_gauss(n);
    qx            = gq.x;
    qw            = gq.w;
    L.nodes       = gll;
    LagrangeBasis Lq{qx};
    S.resize(size_t(n) * n);
    D.resize(size_t(n) * n);
    Sd.resize(size_t(n) * n);
    for (int q = 0; q < n; ++q)
      for (int a = 0; a < n; ++a)
        {
          S[q * n + a]  = L.value(a, qx[q]);
          Sd[q * n + a] = L.deriv(a, qx[q]);
          D[q * n + a]  = Lq.deriv(a, qx[q]);
        }
  }

  // ============================================================ mesh levels
  std::vector<MeshLevel> build_mesh_levels(int dim, const int subdiv[3], int refine, int p,
                                           const double *fine_vertices, const double *lo,
                                           const double *hi, const double *fine_kappa,
                                           int dirichlet_mask, int coarse_coef)
  {
    STFEM_REQUIRE(coarse_coef >= 0 && coarse_coef <= 3, "mesh: coarse_coef must be in [0, 3]");
    STFEM_REQUIRE(dim == 2 || dim == 3, "mesh: dim must be 2 or 3");
    STFEM_REQUIRE(p >= 1 && p < MAX_N, "mesh: space degree must be in [1, 5]");
    STFEM_REQUIRE(refine >= 0 && refine < 12, "mesh: refinement out of range");
    int fnc[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d)
      {
        STFEM_REQUIRE(subdiv[d] >= 1, "mesh: subdivisions must be positive");
        fnc[d] = subdiv[d] << refine;
      }
    const long fnv0 = fnc[0] + 1, fnv1 = fnc[1] + 1, fnv2 = (dim == 3) ? fnc[2] + 1 : 1;
    std::vector<double> FX(size_t(fnv0 * fnv1 * fnv2) * dim);
    for (long k = 0; k < fnv2; ++k)
      for (long j = 0; j < fnv1; ++j)
        for (long i = 0; i < fnv0; ++i)
          {
            const long id = i + fnv0 * (j + fnv1 * k);
            const long ii[3] = {i, j, k};
            for (int d = 0; d < dim; ++d)
              FX[id * dim + d] = fine_vertices ?
                                   fine_vertices[id * dim + d] :
                                   lo[d] + (hi[d] - lo[d]) * double(ii[d]) / fnc[d];
          }
    const Shape1D          sh(p);
    const int              n = p + 1, nq = n;
    std::vector<MeshLevel> out(refine + 1);
    for (int s = 0; s <= refine; ++s)
      {
        MeshLevel &m   = out[s];
        const int  fac = 1 << (refine - s);
        m.g.dim        = dim;
        m.g.p          = p;
        m.g.n_cells    = 1;
        m.g.n_nodes    = 1;
        for (int d = 0; d < 3; ++d)
          {
            m.g.nc[d] = (d < dim) ? fnc[d] / fac : 1;
            m.g.nn[d] = (d < dim) ? m.g.nc[d] * p + 1 : 1;
            m.g.n_cells *= m.g.nc[d];
            m.g.n_nodes *= m.g.nn[d];
          }
        m.g.dirichlet = dirichlet_mask;
        // the kernels split node and cell indices with 32-bit arithmetic
        STFEM_REQUIRE(m.g.n_nodes < (1L << 31) && m.g.n_cells < (1L << 31),
                      "mesh: at most 2^31 nodes per level");
        const long nv0 = m.g.nc[0] + 1, nv1 = m.g.nc[1] + 1, nv2 = (dim == 3) ? m.g.nc[2] + 1 : 1;
        m.X.resize(size_t(nv0 * nv1 * nv2) * dim);
        for (long k = 0; k < nv2; ++k)
          for (long j = 0; j < nv1; ++j)
            for (long i = 0; i < nv0; ++i)
              {
                const long fid = (i * fac) + fnv0 * ((j * fac) + fnv1 * (k * fac));
                const long id  = i + nv0 * (j + nv1 * k);
                for (int d = 0; d < dim; ++d)
                  m.X[id * dim + d] = FX[fid * dim + d];
              }
        if (!fine_kappa)
          {
            m.coef_mode  = 0;
            m.coef_value = 1.0;
          }
        else if (s == refine)
          {
            m.coef_mode = 1;
            m.coef.assign(fine_kappa, fine_kappa + m.g.n_cells);
          }
        else
          {
            m.coef_mode     = 2;
            const long nqc  = (dim == 3) ? nq * nq * nq : nq * nq;
            m.coef.resize(size_t(m.g.n_cells * nqc));
            // per direction: the fine sub-cells met by each point's share
            // of [0,1] and their overlap weights
            std::vector<std::vector<std::pair<int, double>>> share(nq);
            for (int q = 0; q < nq; ++q)
              {
                if (coarse_coef == 0)
                  {
                    int sub = int(std::floor(sh.qx[q] * fac));
                    share[q].push_back({std::min(std::max(sub, 0), fac - 1), 1.0});
                    continue;
                  }
                const double lo_q = q == 0 ? 0.0 : 0.5 * (sh.qx[q - 1] + sh.qx[q]);
                const double hi_q = q == nq - 1 ? 1.0 : 0.5 * (sh.qx[q] + sh.qx[q + 1]);
                for (int sub = 0; sub < fac; ++sub)
                  {
                    const double ov = std::min(hi_q, double(sub + 1) / fac) -
                                      std::max(lo_q, double(sub) / fac);
                    if (ov > 0)
                      share[q].push_back({sub, ov});
                  }
              }
            for (long c = 0; c < m.g.n_cells; ++c)
              {
                const int cx = int(c % m.g.nc[0]);
                const int cy = int((c / m.g.nc[0]) % m.g.nc[1]);
                const int cz = (dim == 3) ? int(c / ((long)m.g.nc[0] * m.g.nc[1])) : 0;
                for (long q = 0; q < nqc; ++q)
                  {
                    const int qi[3] = {int(q % nq), int((q / nq) % nq),
                                       dim == 3 ? int(q / (nq * nq)) : 0};
                    const auto &sz = dim == 3 ? share[qi[2]] : share[0];
                    double      acc = 0.0, wsum = 0.0;
                    for (size_t kz = 0; kz < (dim == 3 ? sz.size() : 1); ++kz)
                      for (const auto &ey : share[qi[1]])
                        for (const auto &ex : share[qi[0]])
                          {
                            const long f0 = long(cx) * fac + ex.first;
                            const long f1 = long(cy) * fac + ey.first;
                            const long f2 = dim == 3 ? long(cz) * fac + sz[kz].first : 0;
                            const double w  = ex.second * ey.second * (dim == 3 ? sz[kz].second : 1.0);
                            const double kv = fine_kappa[f0 + (long)fnc[0] * (f1 + (long)fnc[1] * f2)];
                            acc += w * (coarse_coef == 1 ? std::log(kv) :
                                        coarse_coef == 2 ? 1.0 / kv : kv);
                            wsum += w;
                          }
                    acc /= wsum;
                    m.coef[c * nqc + q] = coarse_coef <= 1 ? (coarse_coef == 0 ? acc : std::exp(acc)) :
                                          coarse_coef == 2 ? 1.0 / acc : acc;
                  }
              }
          }
      }
    return out;
  }

  namespace
  {
    // element mass and stiffness matrices of one cell
    void element_matrices(const MeshLevel &m, const Shape1D &sh, long c, double *Me, double *Ke)
    {
      const int dim = m.g.dim, n = sh.n, nq = n;
      const int nl  = (dim == 3) ? n * n * n : n * n;
      int       cx, cy, cz;
      cell_xyz(m.g, c, cx, cy, cz);
      double    V[8][3] = {};
      const long nv0 = m.g.nc[0] + 1, nv1 = m.g.nc[1] + 1;
      for (int v = 0; v < (1 << dim); ++v)
        {
          const long id = (cx + (v & 1)) + nv0 * ((cy + ((v >> 1) & 1)) + nv1 * (cz + ((v >> 2) & 1)));
          for (int d = 0; d < dim; ++d)
            V[v][d] = m.X[id * dim + d];
        }
      std::fill(Me, Me + size_t(nl) * nl, 0.0);
      std::fill(Ke, Ke + size_t(nl) * nl, 0.0);
      std::vector<double> phi(nl), gr(size_t(nl) * 3);
      const int           nqz = (dim == 3) ? nq : 1;
      for (int qz = 0; qz < nqz; ++qz)
        for (int qy = 0; qy < nq; ++qy)
          for (int qx = 0; qx < nq; ++qx)
            {
              const double xi[3] = {sh.qx[qx], sh.qx[qy], dim == 3 ? sh.qx[qz] : 0.0};
              double       J[3][3] = {};
              for (int v = 0; v < (1 << dim); ++v)
                for (int e = 0; e < dim; ++e)
                  {
                    double g = 1;
                    for (int f = 0; f < dim; ++f)
                      {
                        const int bit = (v >> f) & 1;
                        g *= (f == e) ? (bit ? 1.0 : -1.0) : (bit ? xi[f] : 1 - xi[f]);
                      }
                    for (int d = 0; d < dim; ++d)
                      J[d][e] += g * V[v][d];
                  }
              double Ji[3][3], det;
              if (dim == 2)
                {
                  det      = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                  Ji[0][0] = J[1][1] / det;
                  Ji[0][1] = -J[0][1] / det;
                  Ji[1][0] = -J[1][0] / det;
                  Ji[1][1] = J[0][0] / det;
                }
              else
                {
                  double a[3][3];
                  a[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
                  a[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
                  a[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
                  a[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
                  a[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
                  a[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
                  a[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
                  a[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
                  a[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
                  det     = J[0][0] * a[0][0] + J[0][1] * a[1][0] + J[0][2] * a[2][0];
                  for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                      Ji[i][j] = a[i][j] / det;
                }
              const double w   = sh.qw[qx] * sh.qw[qy] * (dim == 3 ? sh.qw[qz] : 1.0);
              const double JxW = std::abs(det) * w;
              double       kap = m.coef_value;
              if (m.coef_mode == 1)
                kap = m.coef[c];
              else if (m.coef_mode == 2)
                {
                  const long nqc = (dim == 3) ? nq * nq * nq : nq * nq;
                  kap            = m.coef[c * nqc + qx + nq * (qy + nq * qz)];
                }
              int l = 0;
              for (int c3 = 0; c3 < ((dim == 3) ? n : 1); ++c3)
                for (int b = 0; b < n; ++b)
                  for (int a = 0; a < n; ++a, ++l)
                    {
                      const double sx = sh.S[qx * n + a], sy = sh.S[qy * n + b];
                      const double sz = (dim == 3) ? sh.S[qz * n + c3] : 1.0;
                      const double dx = sh.Sd[qx * n + a], dy = sh.Sd[qy * n + b];
                      const double dz = (dim == 3) ? sh.Sd[qz * n + c3] : 0.0;
                      phi[l]          = sx * sy * sz;
                      const double g[3] = {dx * sy * sz, sx * dy * sz, sx * sy * dz};
                      for (int d = 0; d < dim; ++d)
                        {
                          double s = 0;
                          for (int e = 0; e < dim; ++e)
                            s += Ji[e][d] * g[e];
                          gr[l * 3 + d] = s;
                        }
                    }
              for (int i = 0; i < nl; ++i)
                {
                  const double pi = phi[i] * JxW;
                  for (int j = 0; j < nl; ++j)
                    {
                      Me[size_t(i) * nl + j] += pi * phi[j];
                      double s = 0;
                      for (int d = 0; d < dim; ++d)
                        s += gr[i * 3 + d] * gr[j * 3 + d];
                      Ke[size_t(i) * nl + j] += kap * JxW * s;
                    }
                }
            }
    }
  } // namespace

  void cell_patch_matrices(const MeshLevel &m, const Shape1D &sh, std::vector<double> &Kp,
                           std::vector<double> &Mp)
  {
    const GridDesc &g   = m.g;
    const int       dim = g.dim, n = sh.n, p = sh.p;
    const int       nl  = (dim == 3) ? n * n * n : n * n;
    const long      nc  = g.n_cells;
    const size_t    bs  = size_t(nl) * nl;
    std::vector<double> Ke(size_t(nc) * bs), Me(size_t(nc) * bs);
#pragma omp parallel for schedule(dynamic, 16)
    for (long c = 0; c < nc; ++c)
      element_matrices(m, sh, c, Me.data() + c * bs, Ke.data() + c * bs);
    Kp.assign(size_t(nc) * bs, 0.0);
    Mp.assign(size_t(nc) * bs, 0.0);
#pragma omp parallel for schedule(dynamic, 16)
    for (long c = 0; c < nc; ++c)
      {
        int cx, cy, cz;
        cell_xyz(g, c, cx, cy, cz);
        double *K = Kp.data() + c * bs, *M = Mp.data() + c * bs;
        std::vector<int> li, lj;
        for (int dz = (dim == 3 ? -1 : 0); dz <= (dim == 3 ? 1 : 0); ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx)
              {
                const int ox = cx + dx, oy = cy + dy, oz = cz + dz;
                if (ox < 0 || oy < 0 || oz < 0 || ox >= g.nc[0] || oy >= g.nc[1] ||
                    (dim == 3 && oz >= g.nc[2]))
                  continue;
                const long o = ox + (long)g.nc[0] * (oy + (long)g.nc[1] * oz);
                // local nodes of c shared with o, and their index in o
                li.clear();
                lj.clear();
                const int dd[3] = {dx, dy, dz};
                for (int c3 = 0; c3 < ((dim == 3) ? n : 1); ++c3)
                  for (int b = 0; b < n; ++b)
                    for (int a = 0; a < n; ++a)
                      {
                        const int loc[3] = {a, b, c3};
                        int       oth[3];
                        bool      ok = true;
                        for (int d = 0; d < dim; ++d)
                          {
                            if (dd[d] == 0)
                              oth[d] = loc[d];
                            else if (dd[d] == -1)
                              {
                                ok     = ok && loc[d] == 0;
                                oth[d] = p;
                              }
                            else
                              {
                                ok     = ok && loc[d] == p;
                                oth[d] = 0;
                              }
                          }
                        if (!ok)
                          continue;
                        li.push_back(a + n * (b + n * c3));
                        lj.push_back(oth[0] + n * (oth[1] + n * (dim == 3 ? oth[2] : 0)));
                      }
                const double *Ko = Ke.data() + o * bs, *Mo = Me.data() + o * bs;
                for (size_t i = 0; i < li.size(); ++i)
                  for (size_t j = 0; j < li.size(); ++j)
                    {
                      K[size_t(li[i]) * nl + li[j]] += Ko[size_t(lj[i]) * nl + lj[j]];
                      M[size_t(li[i]) * nl + li[j]] += Mo[size_t(lj[i]) * nl + lj[j]];
                    }
              }
        // constrained rows/columns keep only their diagonal
        for (int l = 0; l < nl; ++l)
          {
            const int  a = l % n, b = (l / n) % n, c3 = (dim == 3) ? l / (n * n) : 0;
            const long gx = (long)cx * p + a, gy = (long)cy * p + b, gz = (long)cz * p + c3;
            if (!node_constrained(g, gx, gy, gz))
              continue;
            for (int j = 0; j < nl; ++j)
              if (j != l)
                {
                  K[size_t(l) * nl + j] = K[size_t(j) * nl + l] = 0.0;
                  M[size_t(l) * nl + j] = M[size_t(j) * nl + l] = 0.0;
                }
          }
      }
  }

} // namespace stfem
