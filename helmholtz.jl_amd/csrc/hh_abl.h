// Host-side 1-D ramps of the absorbing layer getABL (src/GetHelmholtz.jl:97-220).
// Plain C++ (no HIP): the device kernel combines the per-axis profiles into
// gamma = amp (p1 + p2 - p1 p2) (2-D) or min(amp, amp (g1 + g2 + g3)) (3-D).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace hh {

// prof[d] has nodes(d) values (1 value for d >= dim).  faces (3-D only, may be
// null = every face): faces[2 d] switches the low face of axis d on, faces[2 d + 1]
// the high face, as the reference's code[d, 1] / code[d, 2]; the 2-D branch
// (the reference's active impl = 1) has no face mask.  A Neumann top removes
// the low face of the last axis.  Throws std::invalid_argument on bad input.
inline std::vector<std::vector<double>> abl_profiles(int dim, const int64_t nnodes[3], int neumann_at_first_dim,
                                                     const int64_t pad[3], const int32_t* faces) {
  std::vector<std::vector<double>> prof(3);
  for (int d = 0; d < 3; ++d) {
    const int64_t n = d < dim ? nnodes[d] : 1;
    if (n < 1 || n > (1 << 24)) throw std::invalid_argument("bad node count");
    prof[d].assign(n, 0.0);
    if (d >= dim) continue;
    const int64_t pd = pad[d];
    if (pd < 1 || pd > n) throw std::invalid_argument("pad must be in 1..n");
    std::vector<double>& P = prof[d];
    if (dim == 2) {
      // quadratic ramps ((pad - k)/pad)^2 at the low end, ((k + 1)/pad)^2 at the high end
      const bool low_on = !(d == 1 && neumann_at_first_dim);
      for (int64_t k = 0; k < pd; ++k) {
        if (low_on) P[k] += (double)((pd - k) * (pd - k)) / (double)(pd * pd);
        P[n - pd + k] += (double)((k + 1) * (k + 1)) / (double)(pd * pd);
      }
    } else {
      // squared distance to the inner edge of the pad on a [-1,1] (x, y) or [0,1] (z) axis
      const double lo = d == 2 ? 0.0 : -1.0, len = d == 2 ? 1.0 : 2.0;
      auto xc = [&](int64_t k) { return lo + len * (double)k / (double)(n - 1); };
      const bool low_on = !(d == 2 && neumann_at_first_dim) && (!faces || faces[2 * d] != 0);
      const bool high_on = !faces || faces[2 * d + 1] != 0;
      for (int64_t k = 0; k < n; ++k) {
        double v = 0.0;
        if (low_on && k < pd) v += (xc(k) - xc(pd - 1)) * (xc(k) - xc(pd - 1));
        if (high_on && k >= n - pd) v += (xc(k) - xc(n - pd)) * (xc(k) - xc(n - pd));
        P[k] = v;
      }
      double mx = 0.0;
      for (double v : P) mx = v > mx ? v : mx;
      for (double& v : P) v /= (mx + 1e-5);
    }
  }
  return prof;
}

}  // namespace hh
