// CPU build of the library's runtime and its kernels for AddressSanitizer and
// UBSan: the runtime-topology kernels (plain dynamics, velocity stepping with
// and without continuous collision detection, joint-space terms, energy,
// transfers) and, built with GP_HOST_FAST, the compiled-topology kernels that
// run one instance per lane (chains and welded-base chains). Every device
// array is a host allocation of its exact size between poisoned guards, every
// lane runs in turn. Reads a batch written by the host-build tests, runs the
// entry points over it through the C ABI, prints a checksum and, given an
// output file, writes the final q, v and vdot (doubles) and the last step's
// collision-loop rounds per instance (int32) there.
//
//   kernels_host <batch file> [out file] [kernel] [precision]
//
// file: int32 n_models, integrator, n_steps, steps_per_launch; double dt;
// per model: int32 n_bodies, n_contacts, n_halfspaces, n_spring_contacts,
// n_spheres; double gravity; the five POD arrays (gorilla_amd.h layouts);
// then int64 counts[n_models], and q, v, tau (the batch's flattening).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/gorilla_amd.h"
#include "../../gorilla-physics_amd/csrc/gp_internal.h"

namespace gp {
#ifdef GP_HOST_FAST
// the grouped launch is not built here (batches use separate launches)
int launch_group_f32(const GroupTable*, int, double, int, int, hipStream_t) { return hipErrorNotSupported; }
int launch_group_f64(const GroupTable*, int, double, int, int, hipStream_t) { return hipErrorNotSupported; }
#else
// the compiled-topology kernels are not built here: no bucket is given a
// compiled topology
int fast_topology_id(int, const int*, const int*, int) { return -1; }
int launch_fast_step(int, int, const BucketArgs&, double, int, int, void*) { return hipErrorNotSupported; }
int launch_fast_rk_step(int, int, int, const BucketArgs&, double, int, int, void*) { return hipErrorNotSupported; }
int launch_grouped_step(const GroupTable*, int, int, double, int, int, void*) { return hipErrorNotSupported; }
#endif
}  // namespace gp

static void need(bool c, const char* what) {
  if (!c) {
    std::fprintf(stderr, "host kernels: %s (%s)\n", what, gp_last_error());
    std::exit(2);
  }
}

template <class T> static std::vector<T> rd(FILE* f, size_t n) {
  std::vector<T> v(n ? n : 1);
  if (n) need(std::fread(v.data(), sizeof(T), n, f) == n, "short read");
  return v;
}

int main(int argc, char** argv) {
  need(argc >= 2, "usage: kernels_host <batch file> [out file] [kernel] [precision]");
  const char* out_path = argc > 2 ? argv[2] : nullptr;
  const int kernel = argc > 3 ? std::atoi(argv[3]) : GP_KERNEL_PLAIN;
  const int precision = argc > 4 ? std::atoi(argv[4]) : GP_PREC_F64;
  FILE* f = std::fopen(argv[1], "rb");
  need(f != nullptr, "open");
  const std::vector<int32_t> hdr = rd<int32_t>(f, 4);
  const int n_models = hdr[0], integrator = hdr[1], n_steps = hdr[2], spl = hdr[3];
  const double dt = rd<double>(f, 1)[0];
  struct Owned {
    std::vector<gp_body_desc> b;
    std::vector<gp_contact_desc> c;
    std::vector<gp_halfspace_desc> h;
    std::vector<gp_spring_contact_desc> s;
    std::vector<gp_sphere_desc> sp;
  };
  std::vector<Owned> own(n_models);
  std::vector<gp_model*> models(n_models);
  for (int m = 0; m < n_models; ++m) {
    const std::vector<int32_t> n = rd<int32_t>(f, 5);
    const double g = rd<double>(f, 1)[0];
    Owned& o = own[m];
    o.b = rd<gp_body_desc>(f, n[0]);
    o.c = rd<gp_contact_desc>(f, n[1]);
    o.h = rd<gp_halfspace_desc>(f, n[2]);
    o.s = rd<gp_spring_contact_desc>(f, n[3]);
    o.sp = rd<gp_sphere_desc>(f, n[4]);
    gp_model_desc d{};
    d.n_bodies = n[0]; d.n_contacts = n[1]; d.n_halfspaces = n[2]; d.gravity = g;
    d.bodies = o.b.data(); d.contacts = o.c.data(); d.halfspaces = o.h.data();
    d.n_spring_contacts = n[3]; d.spring_contacts = o.s.data();
    d.n_spheres = n[4]; d.spheres = o.sp.data();
    need(gp_model_create(&d, &models[m]) == GP_OK, "gp_model_create");
  }
  const std::vector<int64_t> counts = rd<int64_t>(f, n_models);
  gp_batch_opts opts{};
  opts.precision = precision;
  opts.kernel = kernel;
  opts.steps_per_launch = spl;
  opts.use_graph = 0;
  opts.separate_launches = 1;
  opts.block_size = 64;
  opts.lanes_per_mech = 1;  // one lane per instance: no kernel that shuffles between lanes
  opts.lanes_per_tree = 1;
  gp_batch* b = nullptr;
  need(gp_batch_create(models.data(), counts.data(), n_models, &opts, &b) == GP_OK, "gp_batch_create");
  int64_t ni = 0, nq = 0, nv = 0;
  int32_t nb = 0;
  need(gp_batch_info(b, &ni, &nq, &nv, &nb) == GP_OK, "gp_batch_info");
  std::vector<double> q = rd<double>(f, nq), v = rd<double>(f, nv), tau = rd<double>(f, nv);
  std::fclose(f);
  need(gp_batch_set_state(b, q.data(), v.data()) == GP_OK, "set_state");
  need(gp_batch_set_tau(b, tau.data()) == GP_OK, "set_tau");
  int64_t msum = 0;
  std::vector<int64_t> qo(ni + 1), vo(ni + 1);
  need(gp_batch_offsets(b, qo.data(), vo.data()) == GP_OK, "offsets");
  for (int64_t i = 0; i < ni; ++i) msum += (vo[i + 1] - vo[i]) * (vo[i + 1] - vo[i]);
  std::vector<double> M(msum + 1), c(nv + 1), vdot(nv + 1), ke(ni), pe(ni);
  if (!out_path) {  // the FP64 runtime-topology entry points
    need(gp_mass_matrix(b, M.data()) == GP_OK, "gp_mass_matrix");
    need(gp_dynamics_bias(b, c.data()) == GP_OK, "gp_dynamics_bias");
    need(gp_dynamics(b, nullptr, vdot.data()) == GP_OK, "gp_dynamics");
  }
  need(gp_step(b, dt, integrator, n_steps, nullptr) == GP_OK, "gp_step");
  need(gp_batch_sync(b) == GP_OK, "sync");
  need(gp_batch_get_state(b, q.data(), v.data()) == GP_OK, "get_state");
  need(gp_batch_get_vdot(b, vdot.data()) == GP_OK, "get_vdot");
  if (out_path) {
    FILE* o = std::fopen(out_path, "wb");
    need(o != nullptr, "open output");
    std::fwrite(q.data(), sizeof(double), nq, o);
    std::fwrite(v.data(), sizeof(double), nv, o);
    std::fwrite(vdot.data(), sizeof(double), nv, o);
    std::vector<int32_t> rr(ni);
    need(gp_batch_collision_rounds(b, rr.data()) == GP_OK, "rounds");
    std::fwrite(rr.data(), sizeof(int32_t), ni, o);
    std::fclose(o);
  }
  need(gp_batch_energy(b, ke.data(), pe.data()) == GP_OK, "energy");
  std::vector<uint8_t> faults(ni);
  std::vector<int32_t> rounds(ni);
  need(gp_batch_faults(b, faults.data()) == GP_OK, "faults");
  need(gp_batch_collision_rounds(b, rounds.data()) == GP_OK, "rounds");
  double sq = 0, sv = 0;
  int nf = 0, max_rounds = 0;
  for (double x : q) sq += x;
  for (double x : v) sv += x;
  for (int64_t i = 0; i < ni; ++i) {
    nf += faults[i];
    max_rounds = rounds[i] > max_rounds ? rounds[i] : max_rounds;
  }
  std::printf("instances %lld steps %d faults %d max_rounds %d sum_q %.17g sum_v %.17g\n", (long long)ni, n_steps, nf,
              max_rounds, sq, sv);
  gp_batch_destroy(b);
  for (gp_model* m : models) gp_model_destroy(m);
  return 0;
}
