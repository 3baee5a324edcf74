// gp_internal.h — device data layout of one topology bucket and the kernel
// launch interface shared by gp_kernels.hip and the host runtime.
//
// A bucket holds B instances that share a topology signature (body count,
// parents, DOF class per body, contact points per body, half-space count).
// Every per-instance array is structure-of-arrays with a lane pitch Bp
// (B rounded up to 64): element (field f, item i, lane l) lives at
// [(f * n_items + i) * Bp + l], so a wavefront reads 64 consecutive values.
//
// Body parameters are stored in each body's *aligned* frame: a 1-DOF joint's
// axis is the local z axis (the host rotates the reference's body frame by a
// fixed rotation A_i with A_i e_z = axis). The motion subspace is then e_z
// (revolute: angular z, prismatic: linear z), U = column 2 or 5 of I^A and
// D = its diagonal entry.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gp {

enum Kind { K_FIXED = 0, K_REV = 1, K_PRI = 2, K_FLOAT = 3 };

// body parameter fields
enum BodyField {
  BF_RT = 0,   // 9: tree rotation, child-aligned -> parent-aligned (row-major)
  BF_PT = 9,   // 3: tree translation, parent-aligned coords
  BF_J = 12,   // 6: moment about origin xx,yy,zz,xy,xz,yz
  BF_C = 18,   // 3: m * com
  BF_M = 21,   // 1: mass
  BF_SK = 22,  // 1: prismatic spring k
  BF_SL = 23,  // 1: spring rest length
  NBF = 24
};
enum ContactField { CF_P = 0, CF_K = 3, NCF = 4 };
enum HalfspaceField { HF_POINT = 0, HF_NORMAL = 3, HF_ALPHA = 6, HF_MU = 7, NHF = 8 };
// spring contacts: parameters (direction in the aligned body frame, rest
// length, stiffness, A^T of the body's frame alignment, row-major, to turn a
// world direction into the stored one) and registration state (registered
// flag, registered world point, half-space index); both [field][nsp][Bp]
enum SpringField { SF_DIR = 0, SF_L = 3, SF_K = 4, SF_AT = 5, NSF = 14 };
enum SpringState { SS_REG = 0, SS_P = 1, SS_H = 4, NSS = 5 };
// sphere colliders (at most one per body): centre in the aligned body frame,
// radius, enabled (1 / 0); [field][ns][Bp]
enum SphereField { SH_C = 0, SH_R = 3, SH_ON = 4, NSH = 5 };

// per-lane scratch fields of the runtime-topology (plain) kernel, per body
enum PlainField {
  PF_E = 0,     // 9
  PF_R = 9,     // 3
  PF_VEL = 12,  // 6
  PF_CB = 18,   // 6
  PF_PA = 24,   // 6
  PF_IA = 30,   // 21: A(6) B(9) C(6)
  PF_U = 51,    // 6
  PF_D = 57,    // 1
  PF_UU = 58,   // 6 (1-DOF: u in [0]; floating: u[6])
  PF_ACC = 64,  // 6
  PF_RW = 70,   // 9 world rotation
  PF_TW = 79,   // 3 world translation
  NPF = 82
};

// Per-lane scratch of the velocity-stepping kernel, after the plain kernel's
// per-body fields; offsets in units of Bp doubles ([k][Bp] columns). cmax =
// every (contact point or sphere, half-space) pair; m = 3 cmax impulse rows.
struct VsLayout {
  int cmax, m;
  size_t vfree, zero, col, out, qc, vc, vn, gv, mm, y, j, g, l, x, z, s, rd, lt, xb, dx, dz, ds, xi, rhs, w, wi, cn, cp,
      cb, total;
};
__host__ __device__ inline VsLayout vs_layout(int nb, int nq, int nv, int nc, int nh, int ns) {
  VsLayout L;
  L.cmax = (nc + ns) * nh;
  L.m = 3 * L.cmax;
  const size_t m = (size_t)L.m, c = (size_t)L.cmax;
  size_t at = (size_t)NPF * nb;
  auto take = [&](size_t n) { size_t o = at; at += n; return o; };
  L.vfree = take(nv); L.zero = take(nv); L.col = take(nv); L.out = take(nv);
  L.qc = take(nq); L.vc = take(nv); L.vn = take(nv); L.gv = take(m);
  L.mm = take((size_t)nv * nv);
  L.y = take((size_t)nv * m); L.j = take(m * nv); L.g = take(m * m); L.l = take(m * m);
  L.x = take(m); L.z = take(m); L.s = take(m); L.rd = take(m); L.lt = take(m); L.xb = take(m);
  L.dx = take(m); L.dz = take(m); L.ds = take(m); L.xi = take(m); L.rhs = take(m);
  L.w = take(9 * c); L.wi = take(9 * c); L.cn = take(3 * c); L.cp = take(3 * c); L.cb = take(c);
  L.total = at;
  return L;
}

constexpr int kMaxBodies = 64;
constexpr double kInsideMargin = 1e-8;  // halfspace.rs:42
constexpr double kHalfspaceK = 50e3;    // contact.rs:328
constexpr double kSlipTol = 1e-3;       // contact.rs:287

struct BucketArgs {
  int B;      // instances
  int Bp;     // lane pitch
  int nb, nq, nv, nc, nh;
  int lanes;    // lanes per instance (FAST kernels)
  int all_rev;  // every 1-DOF joint of every lane is revolute
  int n_steps;  // steps of this bucket in a grouped launch (serving pools); 0 = the launch's
  double g;
  double dt;    // step of this bucket (serving pools); 0 = the launch's dt
  // topology (device, uniform over the bucket)
  const int* parent;  // [nb]
  const int* qoff;    // [nb]
  const int* voff;    // [nb]
  const int* cstart;  // [nb + 1] contacts sorted by body
  // parameters
  const int* kind;     // [nb][Bp]
  const double* bp;    // [NBF][nb][Bp]
  const float* bpf;    // same, FP32
  const double* cp;    // [NCF][nc][Bp]
  const float* cpf;
  const double* hp;    // [NHF][nh][Bp]
  int nsp;             // spring contacts per instance
  const int* sstart;   // [nb + 1] spring contacts sorted by body
  double* sp;          // [NSF][nsp][Bp] (direction and rest length change when a leg detaches)
  double* ss;          // [NSS][nsp][Bp]
  const float* hpf;
  int ns;              // sphere colliders per instance
  const int* sphere_of;  // [nb] the body's sphere slot or -1
  const double* shp;   // [NSH][ns][Bp]
  int no_springs;      // plain dynamics without the prismatic springs (M^-1 products)
  // state
  double* q;           // [nq][Bp]
  double* v;           // [nv][Bp]
  const double* tau;   // [nv][Bp]
  double* vdot;        // [nv][Bp] acceleration of the last step
  int* fault;          // [Bp]
  int* rounds;         // [Bp] collision-loop rounds of the last velocity-stepping step
  int* active;         // [Bp] or null: per-lane remaining steps (serving);
                       // the FAST kernels count it down by the steps they run
  double* scratch;     // plain kernel: [NPF][nb][Bp] + RK stages
};

// All buckets of a batch in one launch (fast topologies only): bucket b owns
// workgroups [block_start[b], block_start[b + 1]).
constexpr int kMaxGroup = 32;
struct GroupTable {
  int n;
  int block_start[kMaxGroup + 1];
  int topo[kMaxGroup];
  BucketArgs args[kMaxGroup];
};

// Compile-time topologies besides the chains (whose id is their body count);
// TOPO_TREE: any tree of 13 to 64 bodies, one body per lane (gp_tree_impl.h);
// TOPO_WELDED + k: a body welded to the world and a chain of k one-DOF joints
// behind it, k = 1..7 (a fixed-base manipulator).
enum { TOPO_FREE = 100, TOPO_QUAD = 101, TOPO_TREE = 102, TOPO_WELDED = 200 };
constexpr int kMaxWeldedJoints = 7;

// Lanes per instance of a FAST bucket: free bodies with contacts split their
// contact points over lanes_per_mech lanes (0 = 4); the quadruped tree runs one
// lane per leg unless lanes_per_tree is 1.
// A TOPO_TREE bucket runs one body per lane on 16, 32 or 64 lanes per
// instance (the power of two that holds its bodies).
inline int fast_lanes(int fast_topo, int lanes_per_mech, int lanes_per_tree, int nb = 0) {
  if (fast_topo == TOPO_FREE) return lanes_per_mech > 0 ? lanes_per_mech : 4;
  if (fast_topo == TOPO_QUAD) return lanes_per_tree == 1 ? 1 : 4;
  if (fast_topo == TOPO_TREE) return nb <= 16 ? 16 : nb <= 32 ? 32 : 64;
  return 1;
}

// Host-side launchers (gp_plain.hip, gp_fast.hip). Each returns a hipError_t as int.
int launch_grouped_step(const GroupTable* d_table, int n_blocks, int precision, double dt, int n_steps, int block,
                        void* stream);
int launch_plain_step(const BucketArgs& a, double dt, int integrator, int dynamics_only, int n_steps, void* stream);
// Fast kernels: topo_id from fast_topology_id(); precision GP_PREC_*.
int fast_topology_id(int nb, const int* parent, const int* dofcls, int nc);
int launch_fast_step(int topo_id, int precision, const BucketArgs& a, double dt, int n_steps, int block,
                     void* stream);
// RK2 / RK4 (integrator GP_RUNGE_KUTTA_*) on the fast kernels; one lane per
// quadruped instance (a.lanes is used for free bodies only).
int launch_fast_rk_step(int topo_id, int precision, int integrator, const BucketArgs& a, double dt, int n_steps,
                        int block, void* stream);
size_t plain_scratch_doubles(int nb, int nq, int nv, int Bp);
// velocity stepping (ccd = 1: with continuous collision detection), FP64, one
// step per launch; scratch as vs_layout
int launch_vs_step(const BucketArgs& a, double dt, int ccd, int n_steps, void* stream);
// mass matrix and bias c(q, v) - tau_c per lane: out is [nv * nv + nv][Bp]
// (M row-major, then c); needs the bucket's plain scratch
int launch_terms(const BucketArgs& a, double* out, void* stream);
// kinetic and gravitational energy per lane (ke[lane], pe[lane])
int launch_energy(const BucketArgs& a, double* ke, double* pe, void* stream);
size_t vs_scratch_doubles(int nb, int nq, int nv, int nc, int nh, int ns, int Bp);
// velocity stepping of a lone floating body, `lanes` (1 or 8) lanes per
// instance sharing its `candidates` contact candidates (contact points x
// half-spaces, plus the sphere's half-spaces), the working set in registers
// and LDS. hipErrorInvalidValue when the candidates do not fit (more than
// kVsBodyMaxCandidates). block: lanes per workgroup at 8 lanes per body (64,
// 128 or 256, the same bits; 0 = kVsBodyBlock: one wave per workgroup, so a
// SIMD whose wave ends takes the next one without waiting for a workgroup's
// slowest, 1.8 % faster than 256 on the thrown cubes).
constexpr int kVsBodyMaxCandidates = 8;
constexpr int kVsBodyBlock = 64;
int launch_vs_body(const BucketArgs& a, double dt, int ccd, int n_steps, int lanes, int candidates, int block,
                   void* stream);
// rows x B values between a bucket's SoA rows and a flat instance-major array
// (lane l's values start at flat + flat_off[l]); to_soa = 1: flat -> SoA.
int launch_flat_xfer(double* soa, int Bp, int B, int rows, const int64_t* flat_off, double* flat, int to_soa,
                     void* stream);

}  // namespace gp
