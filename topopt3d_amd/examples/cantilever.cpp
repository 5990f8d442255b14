// A short design loop through the C++ facade: cantilever, optimality-criteria update of
// the densities with a volume constraint, compliance printed per iteration.  Mirrors the
// reference's Main.cpp:131-174 loop (MatIntFnc -> FEAssemble -> FESolve -> functions ->
// update) with an OC step in place of MMA.
//
// usage: cantilever NX NY NZ ITERATIONS [precision 0|1]
// prints one line per iteration: "it c vol its reason", and a final "checksum" line.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tofe_topopt.hpp"

int main(int argc, char **argv)
{
  const int nx = argc > 1 ? std::atoi(argv[1]) : 24;
  const int ny = argc > 2 ? std::atoi(argv[2]) : 12;
  const int nz = argc > 3 ? std::atoi(argv[3]) : 12;
  const int iters = argc > 4 ? std::atoi(argv[4]) : 5;
  const int prec = argc > 5 ? std::atoi(argv[5]) : 0;
  tofe::TopOpt t;
  t.precision = prec;
  const double dims[6] = {0, (double)nx, 0, (double)ny, 0, (double)nz};
  const int nel[3] = {nx, ny, nz};
  t.CreateMesh(dims, nel, 1.5);
  const double big = 1e10;
  const double lim[6] = {-big, big, -big, big, -big, big};
  {  // clamp the x = 0 face
    const double c[3] = {0, ny / 2.0, nz / 2.0}, r[3] = {1e-10, big, big}, v[3] = {1, 1, 1};
    t.Set_BC(c, r, lim, v, tofe::SUPPORT);
  }
  {  // unit downward load at the bottom-centre node of the x = max face
    const double c[3] = {(double)nx, (double)(ny / 2), 0}, r[3] = {1e-10, 1e-10, 1e-10},
                 v[3] = {0, 0, -1};
    t.Set_BC(c, r, lim, v, tofe::LOAD);
  }
  if (t.FEInitialize()) {
    std::fprintf(stderr, "FEInitialize: %s\n", t.LastError().c_str());
    return 1;
  }
  const double volfrac = 0.3;
  std::vector<double> x(t.nElem, volfrac);
  tofe::Compliance comp;
  tofe::Volume vol;
  double checksum = 0;
  for (int it = 0; it < iters; ++it) {
    if (t.MatIntFnc(x) || t.FEAssemble() || t.FESolve() || comp.Function(&t) || vol.Function(&t)) {
      std::fprintf(stderr, "iteration %d: %s\n", it, t.LastError().c_str());
      return 1;
    }
    std::printf("%d %.15e %.15e %d %d\n", it, comp.values[0], vol.values[0], t.its, t.KUF_reason);
    checksum += comp.values[0];
    // optimality criteria update with bisection on the volume multiplier
    double l1 = 0, l2 = 1e9;
    std::vector<double> xn(x.size());
    while ((l2 - l1) / (l1 + l2) > 1e-4) {
      const double lm = 0.5 * (l1 + l2);
      double s = 0;
      for (size_t e = 0; e < x.size(); ++e) {
        const double b = -comp.gradients[e] / (lm * vol.gradients[e]);
        double v = x[e] * std::sqrt(std::max(b, 0.0));
        v = std::min({v, x[e] + 0.2, 1.0});
        v = std::max({v, x[e] - 0.2, 0.001});
        xn[e] = v;
        s += v;
      }
      if (s / x.size() > volfrac) l1 = lm;
      else l2 = lm;
    }
    x = xn;
  }
  std::printf("checksum %.15e\n", checksum);
  return 0;
}
