// Orbits of the Q1 element matrix under the symmetry group of the cube.
//
// For an isotropic material on a cubic element, KE is invariant under the 48 symmetries
// of the cube (3 axis reflections x 6 axis permutations), which act on an entry
// ((node a, component i), (node b, component j)) by moving both nodes and relabelling
// both components, with a sign (-1) per reflected component; KE is also symmetric.
// The 576 entries then fall into 12 orbits: entry (x, y) equals sign(x, y) * v[idx(x, y)]
// with v the values at 12 representative entries (up to the rounding of the quadrature
// that formed KE).  The table is built at compile time, so a kernel indexing it with
// constant (x, y) holds only the 12 values, in scalar registers.
#pragma once

namespace tofe {

constexpr int kKeOrbitsMax = 16;

struct KeOrbits {
  signed char idx[24][24];   // orbit of entry (x, y)
  bool neg[24][24];          // KE(x, y) = -v[idx] (else +v[idx])
  unsigned char rep[kKeOrbitsMax][2];  // representative entry of each orbit
  int n;                     // number of orbits
};

// element-local node order (0,0,0),(1,0,0),(1,1,0),(0,1,0), then the +z layer
constexpr int ko_node(int x, int y, int z) { return z * 4 + (y ? (x ? 2 : 3) : (x ? 1 : 0)); }
constexpr int ko_coord(int l, int axis)
{
  return axis == 0 ? ((l & 3) == 1 || (l & 3) == 2) : axis == 1 ? ((l & 3) >= 2) : (l >= 4);
}

constexpr KeOrbits ke_cube_orbits()
{
  KeOrbits o{};
  constexpr int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int x = 0; x < 24; ++x)
    for (int y = 0; y < 24; ++y) o.idx[x][y] = -1;
  o.n = 0;
  for (int x = 0; x < 24; ++x)
    for (int y = 0; y < 24; ++y) {
      if (o.idx[x][y] >= 0) continue;
      const int k = o.n++;
      o.rep[k][0] = (unsigned char)x;
      o.rep[k][1] = (unsigned char)y;
      const int a = x / 3, i = x % 3, b = y / 3, j = y % 3;
      for (int p = 0; p < 6; ++p)
        for (int r = 0; r < 8; ++r) {
          // new coordinate m of a node is its old coordinate perms[p][m], reflected if bit m of r
          int ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
          int ni = 0, nj = 0;
          for (int m = 0; m < 3; ++m) {
            const int flip = (r >> m) & 1;
            ca[m] = ko_coord(a, perms[p][m]) ^ flip;
            cb[m] = ko_coord(b, perms[p][m]) ^ flip;
            if (perms[p][m] == i) ni = m;
            if (perms[p][m] == j) nj = m;
          }
          const bool s = (((r >> ni) & 1) ^ ((r >> nj) & 1)) != 0;
          const int xa = 3 * ko_node(ca[0], ca[1], ca[2]) + ni;
          const int yb = 3 * ko_node(cb[0], cb[1], cb[2]) + nj;
          if (o.idx[xa][yb] < 0) {
            o.idx[xa][yb] = (signed char)k;
            o.neg[xa][yb] = s;
          }
          if (o.idx[yb][xa] < 0) {
            o.idx[yb][xa] = (signed char)k;
            o.neg[yb][xa] = s;
          }
        }
    }
  return o;
}

constexpr KeOrbits kKeCube = ke_cube_orbits();
static_assert(kKeCube.n <= kKeOrbitsMax, "orbit table");

}  // namespace tofe
