// mesh.cpp -- axis-aligned quadtree forest: the meshes of the shallow-water test cases.
//
// Replaces, for Cartesian coarse grids, the reference's mesh path: GridIn::read_msh of a
// structured Gmsh file (tests/shallowWaterVortex/shallowWaterVortex_20x10cells.msh),
// Triangulation::refine_global (main.cc:617-663) and MatrixFree's inner/boundary face
// lists (OceanDG.h:438-470).  Active cells are kept in forest order: coarse cells
// lexicographic (x fastest), children of a cell in z-order at the parent's position,
// the order p4est gives a parallel::distributed::Triangulation.
//
// Inner faces are listed once, from the cell on the lower-x / lower-y side ("minus" side,
// outward normal +x or +y).  A face between a coarse cell and its two finer neighbours is
// listed as two subfaces; the coarse side carries the subface index (0 = lower half along
// the face, 1 = upper half), as deal.II's MatrixFree does for hanging faces.

#include "oceano_dg.h"

#include <cmath>
#include <cstdint>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

namespace oceano {
void set_error(const std::string &msg);
}

struct oc_mesh
{
  double x0, x1, y0, y1;
  int32_t nx0, ny0;
  bool per_x, per_y;
  int32_t bid[4];
  struct Cell
  {
    int32_t level, ix, iy;
  };
  std::vector<Cell> cells;
  // derived topology
  bool built = false;
  std::vector<double> geom;
  std::vector<int32_t> inner, bdry;
  int64_t n_hanging = 0;
};

namespace {

inline uint64_t key(int32_t l, int32_t ix, int32_t iy)
{
  return (uint64_t(uint32_t(l)) << 56) | (uint64_t(uint32_t(ix)) << 28) | uint64_t(uint32_t(iy));
}

bool build(oc_mesh *m)
{
  using oceano::set_error;
  const int64_t N = (int64_t)m->cells.size();
  std::unordered_map<uint64_t, int32_t> index;
  index.reserve(N * 2);
  for (int64_t c = 0; c < N; ++c)
    index[key(m->cells[c].level, m->cells[c].ix, m->cells[c].iy)] = (int32_t)c;
  auto find = [&](int32_t l, int32_t ix, int32_t iy) -> int32_t {
    auto it = index.find(key(l, ix, iy));
    return it == index.end() ? -1 : it->second;
  };
  m->geom.resize(N * 4);
  for (int64_t c = 0; c < N; ++c)
    {
      const auto &C = m->cells[c];
      const double nx = double(int64_t(m->nx0) << C.level), ny = double(int64_t(m->ny0) << C.level);
      m->geom[4 * c + 0] = m->x0 + (m->x1 - m->x0) * C.ix / nx;
      m->geom[4 * c + 1] = m->y0 + (m->y1 - m->y0) * C.iy / ny;
      m->geom[4 * c + 2] = (m->x1 - m->x0) / nx;
      m->geom[4 * c + 3] = (m->y1 - m->y0) / ny;
    }
  m->inner.clear();
  m->bdry.clear();
  m->n_hanging = 0;
  // coverage count per (cell, face): 1 for whole faces, 2 for a coarse face split in two
  std::vector<int8_t> cover(N * 4, 0);
  for (int64_t c = 0; c < N; ++c)
    {
      const auto &C = m->cells[c];
      const int32_t l = C.level;
      const int32_t nxl = m->nx0 << l, nyl = m->ny0 << l;
      for (int f = 0; f < 4; ++f)
        {
          const bool xdir = f < 2;
          const bool plus = (f & 1);
          int32_t nix = C.ix + (xdir ? (plus ? 1 : -1) : 0);
          int32_t niy = C.iy + (xdir ? 0 : (plus ? 1 : -1));
          bool outside = false;
          if (nix < 0 || nix >= nxl)
            {
              if (m->per_x)
                nix = (nix + nxl) % nxl;
              else
                outside = true;
            }
          if (niy < 0 || niy >= nyl)
            {
              if (m->per_y)
                niy = (niy + nyl) % nyl;
              else
                outside = true;
            }
          if (outside)
            {
              m->bdry.insert(m->bdry.end(), {(int32_t)c, f, m->bid[f]});
              cover[4 * c + f] += 1;
              continue;
            }
          if (!plus)
            continue;  // inner faces are listed from their minus side
          const int32_t fo = f - 1;  // opposite face number on the neighbour
          int32_t nb = find(l, nix, niy);
          if (nb >= 0)
            {
              m->inner.insert(m->inner.end(), {(int32_t)c, f, -1, nb, fo, -1});
              cover[4 * c + f] += 1;
              cover[4 * nb + fo] += 1;
              continue;
            }
          if (l > 0)
            {
              nb = find(l - 1, nix >> 1, niy >> 1);
              if (nb >= 0)
                {
                  const int32_t sub = xdir ? (C.iy & 1) : (C.ix & 1);
                  m->inner.insert(m->inner.end(), {(int32_t)c, f, -1, nb, fo, sub});
                  cover[4 * c + f] += 1;
                  cover[4 * nb + fo] += 1;
                  m->n_hanging += 1;
                  continue;
                }
            }
          int32_t ch[2];
          for (int s = 0; s < 2; ++s)
            ch[s] = xdir ? find(l + 1, 2 * nix, 2 * C.iy + s) : find(l + 1, 2 * C.ix + s, 2 * niy);
          if (ch[0] >= 0 && ch[1] >= 0)
            {
              for (int s = 0; s < 2; ++s)
                {
                  m->inner.insert(m->inner.end(), {(int32_t)c, f, s, ch[s], fo, -1});
                  cover[4 * ch[s] + fo] += 1;
                  m->n_hanging += 1;
                }
              cover[4 * c + f] += 2;
              continue;
            }
          set_error("mesh is not 2:1 balanced at cell " + std::to_string(c));
          return false;
        }
    }
  // each face of each cell is covered once (whole) or twice (coarse side of two subfaces)
  for (int64_t c = 0; c < N; ++c)
    for (int f = 0; f < 4; ++f)
      if (cover[4 * c + f] < 1 || cover[4 * c + f] > 2)
        {
          set_error("face " + std::to_string(f) + " of cell " + std::to_string(c) +
                    " is not covered exactly once");
          return false;
        }
  m->built = true;
  return true;
}

}  // namespace

extern "C" {

int oc_mesh_create_rectangle(double x0, double x1, double y0, double y1, int32_t nx, int32_t ny,
                             int32_t periodic_x, int32_t periodic_y,
                             const int32_t boundary_id[4], oc_mesh **mesh)
{
  if (!mesh || nx <= 0 || ny <= 0 || !(x1 > x0) || !(y1 > y0))
    {
      oceano::set_error("oc_mesh_create_rectangle: bad arguments");
      return OC_ERR_ARGUMENT;
    }
  oc_mesh *m = new oc_mesh;
  m->x0 = x0;
  m->x1 = x1;
  m->y0 = y0;
  m->y1 = y1;
  m->nx0 = nx;
  m->ny0 = ny;
  m->per_x = periodic_x != 0;
  m->per_y = periodic_y != 0;
  for (int i = 0; i < 4; ++i)
    m->bid[i] = boundary_id ? boundary_id[i] : 0;
  m->cells.reserve((size_t)nx * ny);
  for (int32_t iy = 0; iy < ny; ++iy)
    for (int32_t ix = 0; ix < nx; ++ix)
      m->cells.push_back({0, ix, iy});
  if (!build(m))
    {
      delete m;
      return OC_ERR_MESH;
    }
  *mesh = m;
  return OC_OK;
}

static int refine_marked(oc_mesh *m, const std::vector<char> &mark)
{
  std::vector<oc_mesh::Cell> out;
  out.reserve(m->cells.size() * 2);
  for (size_t c = 0; c < m->cells.size(); ++c)
    {
      const auto C = m->cells[c];
      if (!mark[c])
        {
          out.push_back(C);
          continue;
        }
      for (int s = 0; s < 4; ++s)
        out.push_back({C.level + 1, 2 * C.ix + (s & 1), 2 * C.iy + (s >> 1)});
    }
  m->cells.swap(out);
  return build(m) ? OC_OK : OC_ERR_MESH;
}

int oc_mesh_refine_global(oc_mesh *m, int32_t times)
{
  if (!m || times < 0)
    return OC_ERR_ARGUMENT;
  for (int t = 0; t < times; ++t)
    {
      std::vector<char> mark(m->cells.size(), 1);
      const int rc = refine_marked(m, mark);
      if (rc)
        return rc;
    }
  return OC_OK;
}

int oc_mesh_refine_fraction(oc_mesh *m, double fraction, uint64_t seed)
{
  if (!m || !(fraction >= 0.0) || fraction > 1.0)
    return OC_ERR_ARGUMENT;
  const size_t N = m->cells.size();
  const size_t K = (size_t)std::floor(fraction * (double)N + 0.5);
  std::vector<size_t> perm(N);
  for (size_t i = 0; i < N; ++i)
    perm[i] = i;
  std::mt19937_64 engine(seed);
  for (size_t i = N - 1; i > 0; --i)
    {
      const size_t j = (size_t)(engine() % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
  std::vector<char> mark(N, 0);
  for (size_t k = 0; k < K; ++k)
    mark[perm[k]] = 1;
  return refine_marked(m, mark);
}

int oc_mesh_info(const oc_mesh *m, int64_t info[4])
{
  if (!m || !info)
    return OC_ERR_ARGUMENT;
  info[0] = (int64_t)m->cells.size();
  info[1] = (int64_t)m->inner.size() / 6;
  info[2] = (int64_t)m->bdry.size() / 3;
  info[3] = m->n_hanging;
  return OC_OK;
}

int oc_mesh_export(const oc_mesh *m, double *cell_geom, int32_t *cell_level, int32_t *inner,
                   int32_t *boundary)
{
  if (!m)
    return OC_ERR_ARGUMENT;
  if (cell_geom)
    std::copy(m->geom.begin(), m->geom.end(), cell_geom);
  if (cell_level)
    for (size_t c = 0; c < m->cells.size(); ++c)
      cell_level[c] = m->cells[c].level;
  if (inner)
    std::copy(m->inner.begin(), m->inner.end(), inner);
  if (boundary)
    std::copy(m->bdry.begin(), m->bdry.end(), boundary);
  return OC_OK;
}

void oc_mesh_destroy(oc_mesh *m) { delete m; }

}  // extern "C"
