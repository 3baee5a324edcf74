// gp_fast.hip — host dispatch of the fast kernels: topology matching and the
// per-precision launchers (the kernels live in gp_fast_f32/f64.hip and
// gp_group_f32/f64.hip, their templates in gp_fast_impl.h).
#include <hip/hip_runtime.h>

#include "../../include/gorilla_amd.h"
#include "gp_fast_impl.h"

namespace gp {

int launch_fast_f32(int topo, const BucketArgs& a, double dt, int n_steps, int block, hipStream_t s);
int launch_fast_f64(int topo, const BucketArgs& a, double dt, int n_steps, int block, hipStream_t s);
int launch_group_f32(const GroupTable* d_table, int n_blocks, double dt, int n_steps, int block, hipStream_t s);
int launch_group_f64(const GroupTable* d_table, int n_blocks, double dt, int n_steps, int block, hipStream_t s);
int launch_fast_rk_f32(int topo, const BucketArgs& a, double dt, int n_steps, int integrator, int block, hipStream_t s);
int launch_fast_rk_f64(int topo, const BucketArgs& a, double dt, int n_steps, int integrator, int block, hipStream_t s);

int fast_topology_id(int nb, const int* parent, const int* dofcls, int nc) {
  bool chain = nb >= 1 && nb <= 12 && nc == 0;
  for (int i = 0; i < nb && chain; ++i) chain = parent[i] == i - 1 && dofcls[i] == 1;
  if (chain) return nb;
  // a welded base and a chain behind it
  bool welded = nb >= 2 && nb - 1 <= kMaxWeldedJoints && nc == 0 && parent[0] == -1 && dofcls[0] == 0;
  for (int i = 1; i < nb && welded; ++i) welded = parent[i] == i - 1 && dofcls[i] == 1;
  if (welded) return TOPO_WELDED + (nb - 1);
  if (nb == 1 && parent[0] == -1 && dofcls[0] == 6) return TOPO_FREE;
  bool quad = nb == 13 && nc == 0;
  for (int i = 0; i < nb && quad; ++i) quad = parent[i] == TopoQuad::parent(i) && dofcls[i] == TopoQuad::dof(i);
  if (quad) return TOPO_QUAD;
  if (tree_topology_ok(nb, parent)) return TOPO_TREE;
  return -1;
}

int launch_fast_step(int topo_id, int precision, const BucketArgs& a, double dt, int n_steps, int block,
                     void* stream) {
  if (a.B == 0 || n_steps <= 0) return 0;
  if (block <= 0) block = 256;
  hipStream_t s = (hipStream_t)stream;
  if (precision == GP_PREC_MIXED) return launch_fast_f32(topo_id, a, dt, n_steps, block, s);
  return launch_fast_f64(topo_id, a, dt, n_steps, block, s);
}

int launch_fast_rk_step(int topo_id, int precision, int integrator, const BucketArgs& a, double dt, int n_steps,
                        int block, void* stream) {
  if (a.B == 0 || n_steps <= 0) return 0;
  if (block <= 0) block = 256;
  hipStream_t s = (hipStream_t)stream;
  if (precision == GP_PREC_MIXED) return launch_fast_rk_f32(topo_id, a, dt, n_steps, integrator, block, s);
  return launch_fast_rk_f64(topo_id, a, dt, n_steps, integrator, block, s);
}

int launch_grouped_step(const GroupTable* d_table, int n_blocks, int precision, double dt, int n_steps, int block,
                        void* stream) {
  if (n_blocks <= 0 || n_steps <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (precision == GP_PREC_MIXED) return launch_group_f32(d_table, n_blocks, dt, n_steps, block, s);
  return launch_group_f64(d_table, n_blocks, dt, n_steps, block, s);
}

}  // namespace gp
