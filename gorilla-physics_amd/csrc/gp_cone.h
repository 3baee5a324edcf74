// gp_cone.h — pieces of the velocity-stepping impulse program shared by the
// runtime-topology kernel (gp_vstep.hip) and the single-body kernel
// (gp_vstep_body.hip): the program's constants, second-order-cone step
// lengths and the Nesterov-Todd scaling of one cone.
#pragma once

#include <hip/hip_runtime.h>

namespace gp {

constexpr double kConeMu = 0.95;  // dynamics.rs:674
constexpr int kConeMaxIter = 50;
constexpr int kConeStall = 3;
constexpr double kConeGapTol = 1e-16;
constexpr double kConeResTol = 1e-14;
constexpr int kCcdMaxRounds = 64;

__device__ inline double soc_det(double a0, double a1, double a2) { return a0 * a0 - a1 * a1 - a2 * a2; }

// largest step t with u + t du in the cone: the first positive root of det,
// and never past the apex (u0 + t du0 = 0): a direction along the cone's axis
// has a double root there, which rounding can turn into no real root
__device__ inline double soc_max_step(const double* u, const double* du) {
  const double a = soc_det(du[0], du[1], du[2]);
  const double b = u[0] * du[0] - u[1] * du[1] - u[2] * du[2];
  const double c = soc_det(u[0], u[1], u[2]);
  double best = du[0] < 0.0 ? -u[0] / du[0] : 1e300;
  if (a == 0.0) {
    if (b < 0.0) best = fmin(best, -c / (2.0 * b));
    return best;
  }
  const double disc = b * b - a * c;
  if (disc < 0.0) return best;
  const double q = -(b + (b < 0.0 ? -sqrt(disc) : sqrt(disc)));
  const double r1 = q / a, r2 = q != 0.0 ? c / q : 1e300;
  if (r1 > 0.0 && r1 < best) best = r1;
  if (r2 > 0.0 && r2 < best) best = r2;
  return best;
}

// Nesterov-Todd scaling of one cone (W z = W^-1 s), row-major 3x3. RCP: each
// group of divisions by one denominator as one reciprocal and products (5
// divisions instead of 19; the results differ in the last bits)
template <bool RCP = false>
__device__ inline void nt_scaling(const double* s, const double* z, double* W, double* Wi) {
  const double ds = sqrt(soc_det(s[0], s[1], s[2])), dz = sqrt(soc_det(z[0], z[1], z[2]));
  auto over = [](double x, double d, double rd) { return RCP ? x * rd : x / d; };
  const double rds = RCP ? 1.0 / ds : 0.0, rdz = RCP ? 1.0 / dz : 0.0;
  const double sb0 = over(s[0], ds, rds), sb1 = over(s[1], ds, rds), sb2 = over(s[2], ds, rds);
  const double zb0 = over(z[0], dz, rdz), zb1 = over(z[1], dz, rdz), zb2 = over(z[2], dz, rdz);
  const double gam = sqrt((1.0 + sb0 * zb0 + sb1 * zb1 + sb2 * zb2) / 2.0);
  const double g2 = 2.0 * gam, rg2 = RCP ? 1.0 / g2 : 0.0;
  const double w0 = over(sb0 + zb0, g2, rg2), w1 = over(sb1 - zb1, g2, rg2), w2 = over(sb2 - zb2, g2, rg2);
  const double eta = sqrt(over(ds, dz, rdz)), f = 1.0 / (1.0 + w0), reta = RCP ? 1.0 / eta : 0.0;
  const double B[9] = {w0, w1, w2, w1, 1.0 + f * w1 * w1, f * w1 * w2, w2, f * w1 * w2, 1.0 + f * w2 * w2};
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int i = k / 3, j = k % 3;
    W[k] = eta * B[k];
    Wi[k] = over((i == 0) != (j == 0) ? -B[k] : B[k], eta, reta);
  }
}

}  // namespace gp
