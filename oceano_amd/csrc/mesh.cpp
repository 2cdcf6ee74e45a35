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

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <array>
#include <map>
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

// Gmsh 4.1 ASCII reader for the Cartesian coarse meshes of the test decks
// (tests/shallowWaterVortex/shallowWaterVortex_20x10cells.msh,
// tests/channelFlow/channelFlow_10x2cells.msh), standing in for GridIn::read_msh
// (main.cc:621-637).  Quadrilaterals (type 3) must be axis-aligned and tile a uniform
// nx x ny grid; boundary ids are the first physical tag of the curve entity that holds each
// boundary line (type 1), as deal.II assigns them, and must be constant along each side.
// The cells come out lexicographic (x fastest) rather than in file order: cell order
// changes no result of the operator (every cell's arithmetic is independent of it).
int oc_mesh_read_msh(const char *path, int32_t periodic_x, int32_t periodic_y, oc_mesh **mesh)
{
  using oceano::set_error;
  if (!path || !mesh)
    {
      set_error("oc_mesh_read_msh: null argument");
      return OC_ERR_ARGUMENT;
    }
  std::ifstream in(path);
  if (!in)
    {
      set_error(std::string("oc_mesh_read_msh: cannot open ") + path);
      return OC_ERR_ARGUMENT;
    }
  auto fail = [&](const std::string &msg) {
    set_error(std::string("oc_mesh_read_msh: ") + path + ": " + msg);
    return OC_ERR_MESH;
  };
  std::map<int64_t, int> curve_tag;  // curve entity -> first physical tag (0 if none)
  std::map<int64_t, std::pair<double, double>> node;
  std::vector<std::array<int64_t, 4>> quads;
  std::vector<std::array<int64_t, 3>> lines;  // node, node, curve entity
  std::string tok;
  bool version_ok = false;
  while (in >> tok)
    {
      if (tok == "$MeshFormat")
        {
          double version;
          int file_type, dsize;
          in >> version >> file_type >> dsize;
          if (file_type != 0)
            return fail("only ASCII files are read");
          version_ok = version >= 4.1 - 1e-9 && version < 5.0;
          if (!version_ok)
            return fail("format version 4.1 expected");
        }
      else if (tok == "$Entities")
        {
          int64_t np, nc, ns, nv;
          in >> np >> nc >> ns >> nv;
          for (int64_t i = 0; i < np; ++i)
            {
              int64_t tag, nphys, t;
              double x, y, z;
              in >> tag >> x >> y >> z >> nphys;
              for (int64_t k = 0; k < nphys; ++k)
                in >> t;
            }
          for (int64_t i = 0; i < nc; ++i)
            {
              int64_t tag, nphys, nb, t;
              double bb[6];
              in >> tag;
              for (double &v : bb)
                in >> v;
              in >> nphys;
              int first = 0;
              for (int64_t k = 0; k < nphys; ++k)
                {
                  in >> t;
                  if (k == 0)
                    first = (int)t;
                }
              curve_tag[tag] = first;
              in >> nb;
              for (int64_t k = 0; k < nb; ++k)
                in >> t;
            }
          for (int64_t i = 0; i < ns + nv; ++i)
            {
              int64_t tag, nphys, nb, t;
              double bb[6];
              in >> tag;
              for (double &v : bb)
                in >> v;
              in >> nphys;
              for (int64_t k = 0; k < nphys; ++k)
                in >> t;
              in >> nb;
              for (int64_t k = 0; k < nb; ++k)
                in >> t;
            }
        }
      else if (tok == "$Nodes")
        {
          int64_t nblocks, nnodes, mn, mx;
          in >> nblocks >> nnodes >> mn >> mx;
          for (int64_t b = 0; b < nblocks; ++b)
            {
              int64_t dim, ent, param, n;
              in >> dim >> ent >> param >> n;
              std::vector<int64_t> tags(n);
              for (auto &t : tags)
                in >> t;
              for (int64_t i = 0; i < n; ++i)
                {
                  double x, y, z, u;
                  in >> x >> y >> z;
                  for (int64_t k = 0; param && k < dim; ++k)
                    in >> u;
                  node[tags[i]] = {x, y};
                }
            }
        }
      else if (tok == "$Elements")
        {
          int64_t nblocks, nel, mn, mx;
          in >> nblocks >> nel >> mn >> mx;
          for (int64_t b = 0; b < nblocks; ++b)
            {
              int64_t dim, ent, type, n;
              in >> dim >> ent >> type >> n;
              const int nn = type == 1 ? 2 : type == 3 ? 4 : type == 15 ? 1 : -1;
              if (nn < 0)
                return fail("element type " + std::to_string(type) +
                            " (only points, lines and quadrilaterals)");
              for (int64_t i = 0; i < n; ++i)
                {
                  int64_t tag, v[4];
                  in >> tag;
                  for (int k = 0; k < nn; ++k)
                    in >> v[k];
                  if (type == 3)
                    quads.push_back({v[0], v[1], v[2], v[3]});
                  else if (type == 1)
                    lines.push_back({v[0], v[1], ent});
                }
            }
        }
      if (!in)
        return fail("truncated or malformed section " + tok);
    }
  if (!version_ok || quads.empty())
    return fail("no $MeshFormat 4.1 header or no quadrilaterals");
  // the grid lines
  std::vector<double> xs, ys;
  for (const auto &q : quads)
    for (int64_t t : q)
      {
        auto it = node.find(t);
        if (it == node.end())
          return fail("element refers to an unknown node");
        xs.push_back(it->second.first);
        ys.push_back(it->second.second);
      }
  // grid lines: the decks' node coordinates scatter around them by ~1e-12 (mesher
  // round-off); coordinates within 1e-9 of the domain size are one line, and the nodes are
  // snapped to the uniform lines (cells the reference sees deviate from these rectangles
  // by that round-off)
  auto lines_of = [](std::vector<double> v) {
    std::sort(v.begin(), v.end());
    const double tol = 1e-9 * (v.back() - v.front());
    std::vector<double> out;
    for (double x : v)
      if (out.empty() || x - out.back() > tol)
        out.push_back(x);
    return out;
  };
  xs = lines_of(xs);
  ys = lines_of(ys);
  const int nx = (int)xs.size() - 1, ny = (int)ys.size() - 1;
  if (nx < 1 || ny < 1 || (int64_t)quads.size() != (int64_t)nx * ny)
    return fail("the quadrilaterals do not tile a tensor grid");
  const double dx = (xs.back() - xs.front()) / nx, dy = (ys.back() - ys.front()) / ny;
  for (int i = 0; i <= nx; ++i)
    if (std::fabs(xs[i] - (xs.front() + i * dx)) > 1e-9 * dx)
      return fail("grid lines in x are not uniform");
  for (int j = 0; j <= ny; ++j)
    if (std::fabs(ys[j] - (ys.front() + j * dy)) > 1e-9 * dy)
      return fail("grid lines in y are not uniform");
  auto snap = [](double v, double v0, double d, long &k) {
    k = std::lround((v - v0) / d);
    return std::fabs(v - (v0 + k * d)) <= 1e-9 * d;
  };
  std::vector<char> seen((size_t)nx * ny, 0);
  for (const auto &q : quads)
    {
      long ix[4], iy[4];
      for (int k = 0; k < 4; ++k)
        if (!snap(node[q[k]].first, xs.front(), dx, ix[k]) ||
            !snap(node[q[k]].second, ys.front(), dy, iy[k]))
          return fail("a node is off the grid lines");
      const long x0 = *std::min_element(ix, ix + 4), x1 = *std::max_element(ix, ix + 4);
      const long y0 = *std::min_element(iy, iy + 4), y1 = *std::max_element(iy, iy + 4);
      int corners = 0;
      for (int k = 0; k < 4; ++k)
        corners |= 1 << ((ix[k] == x1) + 2 * (iy[k] == y1));
      if (x1 - x0 != 1 || y1 - y0 != 1 || corners != 15)
        return fail("a quadrilateral is not an axis-aligned grid cell");
      if (x0 < 0 || x0 >= nx || y0 < 0 || y0 >= ny || seen[(size_t)y0 * nx + x0]++)
        return fail("the quadrilaterals do not tile a tensor grid");
    }
  // boundary ids per side: 0 x = x0, 1 x = x1, 2 y = y0, 3 y = y1
  int32_t bid[4] = {0, 0, 0, 0};
  bool have[4] = {false, false, false, false};
  const double tx = 1e-6 * dx, ty = 1e-6 * dy;
  for (const auto &l : lines)
    {
      if (!node.count(l[0]) || !node.count(l[1]))
        return fail("line refers to an unknown node");
      const double xa = node[l[0]].first, ya = node[l[0]].second;
      const double xb = node[l[1]].first, yb = node[l[1]].second;
      int side = -1;
      if (std::fabs(xa - xs.front()) < tx && std::fabs(xb - xs.front()) < tx)
        side = 0;
      else if (std::fabs(xa - xs.back()) < tx && std::fabs(xb - xs.back()) < tx)
        side = 1;
      else if (std::fabs(ya - ys.front()) < ty && std::fabs(yb - ys.front()) < ty)
        side = 2;
      else if (std::fabs(ya - ys.back()) < ty && std::fabs(yb - ys.back()) < ty)
        side = 3;
      if (side < 0)
        continue;  // not on the outer boundary
      const int id = curve_tag.count(l[2]) ? curve_tag[l[2]] : 0;
      if (have[side] && bid[side] != id)
        return fail("boundary ids vary along a side of the rectangle");
      bid[side] = id;
      have[side] = true;
    }
  return oc_mesh_create_rectangle(xs.front(), xs.back(), ys.front(), ys.back(), nx, ny,
                                  periodic_x, periodic_y, bid, mesh);
}

}  // extern "C"
