// geometry.cpp -- synthetic porous media and lattice helpers for the engine.
//
// The reference builds its media with felbm_local's DomainInitializer_* (an
// un-vendored tree) and scripts/geometry/voxelize_spheres.py.  These are this
// build's own generators, seeded and deterministic on every host:
//   - random non-overlapping cylinders along z (RSA with rejection);
//   - random overlapping spheres added until the pore fraction drops to a
//     target;
//   - the largest periodic pore cluster under D3Q19 18-connectivity, the rule
//     of voxelize_spheres.py:158-248 (pore_connectivity).
// Solid masks are uint8 [Nz][Ny][Nx] (x fastest), 1 = solid.
#include <cstdint>
#include <cstdlib>
#include <cmath>
#include <vector>
#include <numeric>
#include "felbm_mi355x.h"

namespace {

// splitmix64 (Steele, Lea & Flood 2014), a counter-style 64-bit generator
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

inline double pdist(double a, double b, int L) {   // periodic distance on one axis
  double d = std::fabs(a - b);
  return std::fmin(d, (double)L - d);
}

struct Dsu {
  std::vector<int64_t> p;
  explicit Dsu(int64_t n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  int64_t find(int64_t a) { while (p[a] != a) { p[a] = p[p[a]]; a = p[a]; } return a; }
  void unite(int64_t a, int64_t b) { a = find(a); b = find(b); if (a != b) { if (a < b) p[b] = a; else p[a] = b; } }
};

const int DX[19] = { 0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1 };
const int DY[19] = { 0, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 1,-1, 1,-1, 0, 0, 0, 0 };
const int DZ[19] = { 0, 0, 0, 0, 0, 1,-1, 0, 0, 0, 0, 1,-1,-1, 1, 1,-1,-1, 1 };

}  // namespace

extern "C" {

int felbm_geom_cylinders(int Nx, int Ny, int Nz, int count, double rmin, double rmax,
                         double gap, uint64_t seed, int max_tries, uint8_t *solid)
{
  if (Nx <= 0 || Ny <= 0 || Nz <= 0 || !solid || rmin <= 0 || rmax < rmin) return FELBM_EINVAL;
  Rng rng(seed);
  std::vector<double> cx, cy, cr;
  int tries = 0;
  while ((int)cx.size() < count && tries < max_tries) {
    ++tries;
    double r = rng.uniform(rmin, rmax);
    double x = rng.uniform(0.0, (double)Nx), y = rng.uniform(0.0, (double)Ny);
    bool ok = true;
    for (size_t q = 0; q < cx.size() && ok; ++q) {
      double dx = pdist(x, cx[q], Nx), dy = pdist(y, cy[q], Ny);
      if (std::sqrt(dx*dx + dy*dy) < r + cr[q] + gap) ok = false;
    }
    if (ok) { cx.push_back(x); cy.push_back(y); cr.push_back(r); }
  }
  for (int64_t v = 0; v < (int64_t)Nx*Ny*Nz; ++v) solid[v] = 0;
  for (int y = 0; y < Ny; ++y)
    for (int x = 0; x < Nx; ++x) {
      bool s = false;
      for (size_t q = 0; q < cx.size() && !s; ++q) {
        double dx = pdist((double)x, cx[q], Nx), dy = pdist((double)y, cy[q], Ny);
        if (dx*dx + dy*dy <= cr[q]*cr[q]) s = true;
      }
      if (s) for (int z = 0; z < Nz; ++z) solid[(int64_t)x + (int64_t)Nx*((int64_t)y + (int64_t)Ny*z)] = 1;
    }
  return (int)cx.size();
}

int felbm_geom_spheres(int Nx, int Ny, int Nz, double rmin, double rmax, double porosity,
                       uint64_t seed, uint8_t *solid)
{
  if (Nx <= 0 || Ny <= 0 || Nz <= 0 || !solid || rmin <= 0 || rmax < rmin || porosity <= 0 || porosity >= 1)
    return FELBM_EINVAL;
  const int64_t V = (int64_t)Nx*Ny*Nz;
  for (int64_t v = 0; v < V; ++v) solid[v] = 0;
  int64_t nsolid = 0;
  const int64_t target_solid = (int64_t)std::ceil((1.0 - porosity) * (double)V);
  Rng rng(seed);
  int count = 0;
  while (nsolid < target_solid) {
    double r = rng.uniform(rmin, rmax);
    double x0 = rng.uniform(0.0, (double)Nx), y0 = rng.uniform(0.0, (double)Ny), z0 = rng.uniform(0.0, (double)Nz);
    int R = (int)std::ceil(r);
    for (int dz = -R - 1; dz <= R + 1; ++dz) {
      int z = ((int)std::floor(z0) + dz) % Nz; if (z < 0) z += Nz;
      double ddz = pdist((double)z, z0, Nz);
      for (int dy = -R - 1; dy <= R + 1; ++dy) {
        int y = ((int)std::floor(y0) + dy) % Ny; if (y < 0) y += Ny;
        double ddy = pdist((double)y, y0, Ny);
        for (int dx = -R - 1; dx <= R + 1; ++dx) {
          int x = ((int)std::floor(x0) + dx) % Nx; if (x < 0) x += Nx;
          double ddx = pdist((double)x, x0, Nx);
          if (ddx*ddx + ddy*ddy + ddz*ddz <= r*r) {
            uint8_t &s = solid[(int64_t)x + (int64_t)Nx*((int64_t)y + (int64_t)Ny*z)];
            if (!s) { s = 1; ++nsolid; }
          }
        }
      }
    }
    ++count;
  }
  return count;
}

// Keep only the largest pore cluster (D3Q19 18-connectivity, periodic on the
// flagged axes); every other pore voxel becomes solid.  Returns the kept size.
int64_t felbm_keep_largest_cluster(int Nx, int Ny, int Nz, int px, int py, int pz, uint8_t *solid)
{
  if (Nx <= 0 || Ny <= 0 || Nz <= 0 || !solid) return FELBM_EINVAL;
  const int64_t V = (int64_t)Nx*Ny*Nz;
  Dsu d(V);
  for (int z = 0; z < Nz; ++z)
    for (int y = 0; y < Ny; ++y)
      for (int x = 0; x < Nx; ++x) {
        int64_t v = (int64_t)x + (int64_t)Nx*((int64_t)y + (int64_t)Ny*z);
        if (solid[v]) continue;
        for (int m = 1; m < 19; ++m) {
          int a = x + DX[m], b = y + DY[m], c = z + DZ[m];
          if (a < 0 || a >= Nx) { if (!px) continue; a = (a + Nx) % Nx; }
          if (b < 0 || b >= Ny) { if (!py) continue; b = (b + Ny) % Ny; }
          if (c < 0 || c >= Nz) { if (!pz) continue; c = (c + Nz) % Nz; }
          int64_t u = (int64_t)a + (int64_t)Nx*((int64_t)b + (int64_t)Ny*c);
          if (!solid[u]) d.unite(v, u);
        }
      }
  std::vector<int64_t> size(V, 0);
  for (int64_t v = 0; v < V; ++v) if (!solid[v]) size[d.find(v)]++;
  int64_t best = -1, bs = 0;
  for (int64_t v = 0; v < V; ++v) if (size[v] > bs) { bs = size[v]; best = v; }
  for (int64_t v = 0; v < V; ++v) if (!solid[v] && d.find(v) != best) solid[v] = 1;
  return bs;
}

}  // extern "C"
