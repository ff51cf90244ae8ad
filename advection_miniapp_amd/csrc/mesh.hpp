// Structured hexahedral (quadrilateral in 2D) mesh with the cell order and
// DoF numbering of deal.II's serial Triangulation + DoFHandler for the
// reference's meshes:
//   - GridGenerator::subdivided_hyper_rectangle / hyper_cube coarse cells in
//     lexicographic order (x fastest)          advection_solver.cc:1080-1083
//   - refine_global(r): every coarse cell is followed by its 2^(dim r)
//     descendants in child order (child c at (c&1, c>>1&1, c>>2&1)), i.e. a
//     Morton order with x in the least significant bit
//                                                advection_solver.cc:1128
//   - periodic pairs (2d, 2d+1) in every direction  advection_solver.cc:1085-1099
//   - FE_DGQ dofs cell-contiguous, lexicographic inside the cell
// Interior vertices can be perturbed by a seeded uniform random offset of
// up to `perturbation * h` per coordinate: mt19937_64, raw output >> 11
// times 2^-53, one draw per coordinate of every interior vertex in
// lexicographic vertex order.
#pragma once

#include <cstdint>
#include <random>
#include <stdexcept>
#include <vector>

namespace dgops
{
  struct MeshDesc
  {
    int      dim;
    int      n_coarse[3];  // coarse cells per direction
    int      refine;       // global refinements
    double   lo[3], hi[3]; // bounding box
    int      periodic[3];  // per direction
    double   perturbation; // fraction of h
    uint64_t seed;
  };

  struct StructuredMesh
  {
    MeshDesc desc;
    int      dim;
    int      N[3];    // cells per direction
    int64_t  n_cells;
    double   h[3];
    // cell index -> integer coordinates
    std::vector<int> cell_ijk; // [cell][3]
    // neighbor cell per face f = 2*d + side, -1 at a non-periodic boundary
    std::vector<int> neighbors; // [cell][2*dim]
    // vertex grid (N+1)^dim, lexicographic, dim coordinates each
    std::vector<double> vertices;
    // corner coordinates per cell, [cell][2^dim][dim] (corner v at
    // (v&1, v>>1&1, v>>2&1))
    std::vector<double> cell_vertices;
    double              min_vertex_distance;

    int64_t
    cell_index(int i, int j, int k) const
    {
      const int r  = desc.refine;
      const int cx = i >> r, cy = j >> r, cz = k >> r;
      const int lx = i - (cx << r), ly = j - (cy << r), lz = k - (cz << r);
      int64_t   coarse = cx + (int64_t)desc.n_coarse[0] *
                              (cy + (int64_t)desc.n_coarse[1] * cz);
      int64_t   m = 0;
      for (int b = 0; b < r; ++b)
        {
          m |= (int64_t)((lx >> b) & 1) << (dim * b);
          m |= (int64_t)((ly >> b) & 1) << (dim * b + 1);
          if (dim == 3)
            m |= (int64_t)((lz >> b) & 1) << (dim * b + 2);
        }
      return (coarse << (dim * r)) + m;
    }
  };

  inline StructuredMesh
  make_mesh(const MeshDesc &d)
  {
    // vertex coordinates are index data shared with other implementations:
    // no fused multiply-add contraction, so every compiler rounds alike
#pragma clang fp contract(off)
    if (d.dim != 2 && d.dim != 3)
      throw std::runtime_error("dim must be 2 or 3");
    if (d.refine < 0 || d.refine > 12)
      throw std::runtime_error("refinement out of range");
    StructuredMesh m;
    m.desc = d;
    m.dim  = d.dim;
    m.n_cells = 1;
    for (int e = 0; e < 3; ++e)
      {
        if (e < d.dim)
          {
            if (d.n_coarse[e] < 1)
              throw std::runtime_error("n_coarse must be >= 1");
            m.N[e] = d.n_coarse[e] << d.refine;
            m.h[e] = (d.hi[e] - d.lo[e]) / m.N[e];
            if (!(m.h[e] > 0))
              throw std::runtime_error("empty bounding box");
          }
        else
          {
            m.N[e] = 1;
            m.h[e] = 1;
          }
        m.n_cells *= m.N[e];
      }
    const int r  = d.refine;
    const int nv = 1 << d.dim;
    m.cell_ijk.resize(3 * m.n_cells);
    const int64_t per_coarse = (int64_t)1 << (d.dim * r);
    for (int64_t c = 0; c < m.n_cells; ++c)
      {
        const int64_t coarse = c / per_coarse;
        const int64_t mm     = c % per_coarse;
        int           ijk[3] = {0, 0, 0};
        for (int b = 0; b < r; ++b)
          for (int e = 0; e < d.dim; ++e)
            ijk[e] |= (int)((mm >> (d.dim * b + e)) & 1) << b;
        const int cx = (int)(coarse % d.n_coarse[0]);
        const int cy = (int)((coarse / d.n_coarse[0]) % d.n_coarse[1]);
        const int cz =
          d.dim == 3 ? (int)(coarse / ((int64_t)d.n_coarse[0] * d.n_coarse[1])) :
                       0;
        m.cell_ijk[3 * c + 0] = (cx << r) + ijk[0];
        m.cell_ijk[3 * c + 1] = (cy << r) + ijk[1];
        m.cell_ijk[3 * c + 2] = d.dim == 3 ? (cz << r) + ijk[2] : 0;
      }

    // vertices
    const int V[3] = {m.N[0] + 1, m.N[1] + 1, d.dim == 3 ? m.N[2] + 1 : 1};
    const int64_t n_vert = (int64_t)V[0] * V[1] * V[2];
    m.vertices.resize(n_vert * d.dim);
    std::mt19937_64 rng(d.seed);
    for (int k = 0; k < V[2]; ++k)
      for (int j = 0; j < V[1]; ++j)
        for (int i = 0; i < V[0]; ++i)
          {
            const int64_t v      = i + (int64_t)V[0] * (j + (int64_t)V[1] * k);
            const int     idx[3] = {i, j, k};
            bool          interior = true;
            for (int e = 0; e < d.dim; ++e)
              if (idx[e] == 0 || idx[e] == m.N[e])
                interior = false;
            for (int e = 0; e < d.dim; ++e)
              {
                double x = d.lo[e] + (d.hi[e] - d.lo[e]) * idx[e] / m.N[e];
                if (idx[e] == m.N[e])
                  x = d.hi[e];
                m.vertices[v * d.dim + e] = x;
              }
            if (interior && d.perturbation != 0.)
              for (int e = 0; e < d.dim; ++e)
                {
                  const double u = (double)(rng() >> 11) * 0x1.0p-53;
                  m.vertices[v * d.dim + e] +=
                    d.perturbation * m.h[e] * (2. * u - 1.);
                }
          }

    m.cell_vertices.resize(m.n_cells * nv * d.dim);
    m.neighbors.resize(m.n_cells * 2 * d.dim);
    double min_dist = 1e300;
    for (int64_t c = 0; c < m.n_cells; ++c)
      {
        const int i = m.cell_ijk[3 * c], j = m.cell_ijk[3 * c + 1],
                  k = m.cell_ijk[3 * c + 2];
        double    corners[8][3];
        for (int v = 0; v < nv; ++v)
          {
            const int64_t gv = (i + (v & 1)) +
                               (int64_t)V[0] * ((j + ((v >> 1) & 1)) +
                                                (int64_t)V[1] *
                                                  (k + ((v >> 2) & 1)));
            for (int e = 0; e < d.dim; ++e)
              {
                corners[v][e] = m.vertices[gv * d.dim + e];
                m.cell_vertices[(c * nv + v) * d.dim + e] = corners[v][e];
              }
          }
        // minimum distance between the cell's vertices (as
        // CellAccessor::minimum_vertex_distance)
        for (int a = 0; a < nv; ++a)
          for (int b = a + 1; b < nv; ++b)
            {
              double s = 0;
              for (int e = 0; e < d.dim; ++e)
                s += (corners[a][e] - corners[b][e]) *
                     (corners[a][e] - corners[b][e]);
              if (s < min_dist * min_dist)
                min_dist = std::sqrt(s);
            }
        const int ijk[3] = {i, j, k};
        for (int e = 0; e < d.dim; ++e)
          for (int side = 0; side < 2; ++side)
            {
              int nb[3] = {ijk[0], ijk[1], ijk[2]};
              nb[e] += side ? 1 : -1;
              int64_t nbc = -1;
              if (nb[e] < 0 || nb[e] >= m.N[e])
                {
                  if (d.periodic[e])
                    {
                      nb[e] = (nb[e] + m.N[e]) % m.N[e];
                      nbc   = m.cell_index(nb[0], nb[1], nb[2]);
                    }
                }
              else
                nbc = m.cell_index(nb[0], nb[1], nb[2]);
              m.neighbors[c * 2 * d.dim + 2 * e + side] = (int)nbc;
            }
      }
    m.min_vertex_distance = min_dist;
    return m;
  }
} // namespace dgops
