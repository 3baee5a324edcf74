// gp_runtime.h — host-side objects behind the opaque C handles.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gorilla_amd.h"
#include "gp_internal.h"

// A validated, immutable mechanism with its parameters already in aligned
// body frames (see gp_internal.h).
struct gp_model {
  int n = 0, nq = 0, nv = 0, nc = 0, nh = 0;
  double g = 9.81;
  std::vector<int> parent, dofcls, kind, qoff, voff, cstart;
  std::vector<double> bp;  // [n][NBF]
  std::vector<double> cp;  // [nc][NCF], sorted by body
  std::vector<double> hp;  // [nh][NHF]
  std::vector<int> contact_order;  // sorted slot -> index in the description
  int nsp = 0;
  std::vector<int> sstart;         // [n + 1] spring contacts sorted by body
  std::vector<double> sp;          // [nsp][NSF]
  std::vector<int> spring_slot;    // index in the description -> sorted slot
  int ns = 0;
  std::vector<int> sphere_of;      // [n] sphere collider slot of the body or -1
  std::vector<double> shp;         // [ns][NSH]
  uint64_t sig = 0;
  uint64_t serial = 0;  // unique per model created in the process (never reused)
  int fast_topo = -1;
  bool all_rev = true;  // no prismatic joint
};

namespace gp {

struct Bucket {
  int nb = 0, nq = 0, nv = 0, nc = 0, nh = 0, nsp = 0, ns = 0, B = 0, Bp = 0, fast_topo = -1;
  bool all_rev = false;  // every 1-DOF joint of every lane is revolute
  double g = 9.81;
  std::vector<int> parent, qoff, voff, dofcls, cstart, sstart;
  std::vector<int> spring_slot;
  std::vector<int> sphere_of;
  std::vector<int64_t> inst;  // lane -> instance index
  int *d_parent = nullptr, *d_qoff = nullptr, *d_voff = nullptr, *d_cstart = nullptr, *d_kind = nullptr;
  int* d_sstart = nullptr;
  double *d_sp = nullptr, *d_ss = nullptr;
  int* d_sphere_of = nullptr;
  double* d_shp = nullptr;
  double *d_bp = nullptr, *d_cp = nullptr, *d_hp = nullptr;
  float *d_bpf = nullptr, *d_cpf = nullptr, *d_hpf = nullptr;
  double *d_q = nullptr, *d_v = nullptr, *d_tau = nullptr, *d_tau_tmp = nullptr, *d_vdot = nullptr;
  int* d_fault = nullptr;
  int* d_rounds = nullptr;
  int* d_active = nullptr;
  double* d_scratch = nullptr;
  size_t scratch_doubles = 0;
  void* d_arena = nullptr;  // one allocation holding every array above except scratch and active
  size_t arena_bytes = 0;
  int64_t* d_flat_off = nullptr;  // [2][B]: each lane's q and v offsets in the flattened arrays
  Bucket() = default;
  Bucket(const Bucket&) = delete;
  Bucket& operator=(const Bucket&) = delete;
  Bucket(Bucket&& o) noexcept { *this = std::move(o); }
  Bucket& operator=(Bucket&& o) noexcept {
    if (this != &o) {
      release();
      nb = o.nb; nq = o.nq; nv = o.nv; nc = o.nc; nh = o.nh; nsp = o.nsp; ns = o.ns; B = o.B; Bp = o.Bp;
      fast_topo = o.fast_topo;
      g = o.g;
      sstart = std::move(o.sstart); spring_slot = std::move(o.spring_slot); sphere_of = std::move(o.sphere_of);
      d_sstart = o.d_sstart; d_sp = o.d_sp; d_ss = o.d_ss; d_sphere_of = o.d_sphere_of; d_shp = o.d_shp;
      o.d_sstart = nullptr; o.d_sp = o.d_ss = nullptr; o.d_sphere_of = nullptr; o.d_shp = nullptr;
      scratch_doubles = o.scratch_doubles;
      all_rev = o.all_rev;
      parent = std::move(o.parent); qoff = std::move(o.qoff); voff = std::move(o.voff); dofcls = std::move(o.dofcls);
      cstart = std::move(o.cstart); inst = std::move(o.inst);
      d_parent = o.d_parent; d_qoff = o.d_qoff; d_voff = o.d_voff; d_cstart = o.d_cstart; d_kind = o.d_kind;
      d_bp = o.d_bp; d_cp = o.d_cp; d_hp = o.d_hp; d_bpf = o.d_bpf; d_cpf = o.d_cpf; d_hpf = o.d_hpf;
      d_q = o.d_q; d_v = o.d_v; d_tau = o.d_tau; d_tau_tmp = o.d_tau_tmp; d_vdot = o.d_vdot; d_fault = o.d_fault;
      d_rounds = o.d_rounds;
      o.d_rounds = nullptr;
      d_active = o.d_active; d_scratch = o.d_scratch; d_arena = o.d_arena; arena_bytes = o.arena_bytes;
      d_flat_off = o.d_flat_off;
      o.d_arena = nullptr;
      o.d_flat_off = nullptr;
      o.d_parent = o.d_qoff = o.d_voff = o.d_cstart = o.d_kind = o.d_fault = o.d_active = nullptr;
      o.d_bp = o.d_cp = o.d_hp = o.d_q = o.d_v = o.d_tau = o.d_tau_tmp = o.d_vdot = o.d_scratch = nullptr;
      o.d_bpf = o.d_cpf = o.d_hpf = nullptr;
    }
    return *this;
  }
  ~Bucket();
  void release();
};

int fail(int code, const char* fmt, ...);
int model_create(const gp_model_desc* d, gp_model** out);
int bucket_build(Bucket& b, const std::vector<const gp_model*>& lane_models, bool fill_lanes = true);
int bucket_ensure_scratch(Bucket& b);
BucketArgs bucket_args(const Bucket& b, int lanes, const double* tau);
// Most lanes one bucket of this model can hold: the kernels address each
// per-bucket array ([rows][Bp]) with 32-bit byte offsets, so the largest one
// stays below 2^31 bytes.
int64_t soa_max_lanes(const gp_model& m);
// Relative time of one step of one wave of a bucket's compiled kernel.
double bucket_step_cost(const Bucket& b, int precision, int lanes_per_tree);

}  // namespace gp

struct gp_batch {
  std::vector<gp::Bucket> buckets;
  int64_t n_inst = 0;
  std::vector<int64_t> qoff, voff;
  std::vector<int> inst_bucket, inst_lane;
  gp_batch_opts opts{};
  hipStream_t stream = nullptr;
  // a caller's own stream has carried this batch's work since the last
  // gp_batch_sync (which then waits for the whole device)
  bool foreign_stream = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timing_valid = false;
  int64_t last_launches = 0;
  // chunk graphs by (dt, integrator, steps)
  std::map<std::tuple<double, int, int>, hipGraphExec_t> graphs;
  std::map<std::tuple<double, int, int>, int64_t> graph_launches;
  // grouped launch over the FAST buckets (built lazily, rebuilt on set_opts)
  gp::GroupTable* d_group = nullptr;
  int group_blocks = 0, group_block = 0;
  bool group_valid = false;
};
