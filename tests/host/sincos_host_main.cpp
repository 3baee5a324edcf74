// CPU build of the kernels' FP64 joint-angle sin/cos (sincos_q in
// gp_device.h) through the host stand-in for the HIP runtime: reads doubles
// from argv[1], writes sin then cos of each to argv[2].
#include <cstdio>
#include <vector>

#include "../../gorilla-physics_amd/csrc/gp_device.h"

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<double> x;
  double v;
  while (std::fread(&v, sizeof v, 1, f) == 1) x.push_back(v);
  std::fclose(f);
  std::vector<double> out(2 * x.size());
  for (size_t i = 0; i < x.size(); ++i) gp::sincos_q(x[i], &out[i], &out[x.size() + i]);
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) return 2;
  std::fwrite(out.data(), sizeof(double), out.size(), o);
  std::fclose(o);
  return 0;
}
