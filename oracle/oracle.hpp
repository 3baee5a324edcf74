// oracle.hpp — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference's forward-dynamics path, used as the parity
// checker of the HIP kernels. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it; the product path never does.
//
// Two independent algorithms, both templated on the scalar type T so they can
// run in double (the oracle), long double / __float128 (to measure the
// oracle's own rounding) and an op-counting type (algorithmic flops):
//
//  1. crba_dynamics: the reference's own algorithm, restated step by step —
//     world-frame kinematics (mechanism.rs:153-170, twist.rs:176-204),
//     composite-rigid-body mass matrix (mechanism.rs:592-696, momentum.rs:17-48,
//     inertia.rs:106-134), RNEA bias with gravity as base acceleration
//     (dynamics.rs:41-251, wrench.rs:96-126), point/half-space Hunt-Crossley
//     contact (contact.rs:97-129,260-338, halfspace.rs:39-44), prismatic springs
//     (dynamics.rs:298-315) and a partial-pivot LU solve (dynamics.rs:255-276,
//     nalgebra 0.33.2 `lu()`).
//  2. aba_dynamics: Featherstone's articulated-body algorithm in body frames
//     whose z axis is each 1-DOF joint's axis — the formulation the GPU kernels
//     use. Exact in real arithmetic, so it agrees with (1) to rounding.
//
// Integrators: semi-implicit Euler (integrators.rs:276-319, util.rs:83-102),
// RK2 / RK4 (integrators.rs:177-271), and velocity stepping with and without
// continuous collision detection (integrators.rs:42-175, dynamics.rs:366-742):
// contact impulses from a second-order-cone program (cone_program).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/gorilla_amd.h"

namespace oracle {

// Operation counters of the op-counting scalar `Cnt` (defined at the end).
struct OpCounter {
  static inline thread_local uint64_t add = 0, mul = 0, div = 0, sqrt_ = 0, trig = 0, pow_ = 0;
  static inline thread_local uint64_t add64 = 0, mul64 = 0, div64 = 0, sqrt64 = 0, trig64 = 0, pow64 = 0;
  static inline thread_local int f64 = 0;       // > 0 inside an F64Scope
  static inline thread_local bool loading = false;
  static void reset() { add = mul = div = sqrt_ = trig = pow_ = 0; add64 = mul64 = div64 = sqrt64 = trig64 = pow64 = 0; }
  static void count(uint64_t& c32, uint64_t& c64) { ++(f64 > 0 ? c64 : c32); }
};
struct F64Scope {
  F64Scope() { ++OpCounter::f64; }
  ~F64Scope() { --OpCounter::f64; }
};

using std::size_t;

// ---------------------------------------------------------------- small algebra
template <class T>
struct V3 {
  T x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(T a, T b, T c) : x(a), y(b), z(c) {}
  T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  const T& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> V3<T> operator+(const V3<T>& a, const V3<T>& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> V3<T> operator-(const V3<T>& a, const V3<T>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> V3<T> operator-(const V3<T>& a) { return {-a.x, -a.y, -a.z}; }
template <class T> V3<T> operator*(const T& s, const V3<T>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> V3<T> operator*(const V3<T>& a, const T& s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> V3<T> operator/(const V3<T>& a, const T& s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> V3<T>& operator+=(V3<T>& a, const V3<T>& b) { a.x = a.x + b.x; a.y = a.y + b.y; a.z = a.z + b.z; return a; }
template <class T> V3<T>& operator-=(V3<T>& a, const V3<T>& b) { a.x = a.x - b.x; a.y = a.y - b.y; a.z = a.z - b.z; return a; }
template <class T> T dot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> V3<T> cross(const V3<T>& a, const V3<T>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

template <class T>
struct M3 {  // row-major
  T m[3][3];
  M3() { for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] = T(0); }
  static M3 identity() { M3 r; r.m[0][0] = r.m[1][1] = r.m[2][2] = T(1); return r; }
  static M3 diag(T a, T b, T c) { M3 r; r.m[0][0] = a; r.m[1][1] = b; r.m[2][2] = c; return r; }
  M3 transpose() const { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = m[j][i]; return r; }
  T trace() const { return m[0][0] + m[1][1] + m[2][2]; }
};
template <class T> V3<T> operator*(const M3<T>& A, const V3<T>& v) {
  return {A.m[0][0] * v.x + A.m[0][1] * v.y + A.m[0][2] * v.z,
          A.m[1][0] * v.x + A.m[1][1] * v.y + A.m[1][2] * v.z,
          A.m[2][0] * v.x + A.m[2][1] * v.y + A.m[2][2] * v.z};
}
template <class T> M3<T> operator*(const M3<T>& A, const M3<T>& B) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
  return r;
}
template <class T> M3<T> operator+(const M3<T>& A, const M3<T>& B) { M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = A.m[i][j] + B.m[i][j]; return r; }
template <class T> M3<T> operator-(const M3<T>& A, const M3<T>& B) { M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = A.m[i][j] - B.m[i][j]; return r; }
template <class T> M3<T> operator*(const T& s, const M3<T>& A) { M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = s * A.m[i][j]; return r; }
template <class T> M3<T> outer(const V3<T>& a, const V3<T>& b) { M3<T> r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a[i] * b[j]; return r; }
template <class T> M3<T> skew(const V3<T>& v) {  // util.rs:105
  M3<T> r;
  r.m[0][1] = -v.z; r.m[0][2] = v.y;
  r.m[1][0] = v.z;  r.m[1][2] = -v.x;
  r.m[2][0] = -v.y; r.m[2][1] = v.x;
  return r;
}

// Quaternion stored as (x, y, z, w) like nalgebra's coords.
template <class T>
struct Q {
  T x, y, z, w;
  Q() : x(0), y(0), z(0), w(1) {}
  Q(T a, T b, T c, T d) : x(a), y(b), z(c), w(d) {}
};
// Hamilton product (nalgebra UnitQuaternion * UnitQuaternion)
template <class T> Q<T> qmul(const Q<T>& a, const Q<T>& b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
          a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class T> Q<T> qconj(const Q<T>& a) { return {-a.x, -a.y, -a.z, a.w}; }
// Rotation of a vector by a unit quaternion: v + w t + u x t, t = 2 u x v
template <class T> V3<T> qrot(const Q<T>& q, const V3<T>& v) {
  V3<T> u(q.x, q.y, q.z);
  V3<T> t = cross(u, v) * T(2);
  return t * q.w + cross(u, t) + v;
}
template <class T> M3<T> qmat(const Q<T>& q) {  // UnitQuaternion::to_rotation_matrix
  T i = q.x, j = q.y, k = q.z, w = q.w;
  T ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  T ij = i * j * T(2), wk = w * k * T(2), wj = w * j * T(2), ik = i * k * T(2), jk = j * k * T(2), wi = w * i * T(2);
  M3<T> r;
  r.m[0][0] = ww + ii - jj - kk; r.m[0][1] = ij - wk;            r.m[0][2] = wj + ik;
  r.m[1][0] = wk + ij;            r.m[1][1] = ww - ii + jj - kk; r.m[1][2] = jk - wi;
  r.m[2][0] = ik - wj;            r.m[2][1] = wi + jk;            r.m[2][2] = ww - ii - jj + kk;
  return r;
}
template <class T> Q<T> q_axis_angle(const V3<T>& axis, const T& angle) {  // from_axis_angle
  using std::sin; using std::cos;
  T h = angle / T(2);
  T s = sin(h);
  return {axis.x * s, axis.y * s, axis.z * s, cos(h)};
}

// Isometry: x_to = R x_from + t (nalgebra Isometry3)
template <class T>
struct Iso {
  Q<T> r;
  V3<T> t;
};
template <class T> Iso<T> imul(const Iso<T>& a, const Iso<T>& b) { return {qmul(a.r, b.r), a.t + qrot(a.r, b.t)}; }
template <class T> Iso<T> iinv(const Iso<T>& a) { Q<T> ri = qconj(a.r); return {ri, -qrot(ri, a.t)}; }

// Spatial inertia in the reference's parametrisation (inertia.rs:32): moment
// about the frame origin, cross part = m * com, mass.
template <class T>
struct Inertia {
  M3<T> J;
  V3<T> c;
  T m;
};

// ---------------------------------------------------------------- model
template <class T>
struct Contact {
  V3<T> p;
  T k;
};
// SpringContact (contact.rs:74-94): its description values; the mutable part
// (registration, direction, rest length) lives in SpringSt, owned by the caller
template <class T>
struct SpringC {
  V3<T> dir;
  T l_rest, k;
};
template <class T>
struct SpringSt {
  bool registered = false;
  V3<T> point;       // registered contact, world frame
  int halfspace = 0;
  V3<T> dir;         // body frame, unit
  T l_rest = T(0);
};
template <class T>
struct HalfSpace {
  V3<T> point, normal;
  T alpha, mu;
};
// Sphere collider (collision/sphere.rs:6-39): centre in the body frame
// (base_translation), radius; Collider::enabled (rigid_body.rs:11-37)
template <class T>
struct SphereC {
  bool present = false, enabled = false;
  V3<T> c;
  T r = T(0);
};

template <class T>
struct Model {
  int n = 0;
  std::vector<int> parent;  // -1 = world
  std::vector<int> jtype;
  std::vector<V3<T>> axis;
  std::vector<Iso<T>> tree;
  std::vector<Inertia<T>> I;
  std::vector<T> spring_k, spring_l;
  std::vector<std::vector<Contact<T>>> contacts;  // per body
  std::vector<std::vector<SpringC<T>>> springs;   // per body, description order within a body
  std::vector<HalfSpace<T>> halfspaces;
  std::vector<SphereC<T>> spheres;                // per body
  T g = T(9.81);
  std::vector<int> qoff, voff;  // per body
  int nq = 0, nv = 0;
};

inline int jdof_q(int jt) { return jt == GP_JOINT_FLOATING ? 7 : (jt == GP_JOINT_FIXED ? 0 : 1); }
inline int jdof_v(int jt) { return jt == GP_JOINT_FLOATING ? 6 : (jt == GP_JOINT_FIXED ? 0 : 1); }

template <class T>
Model<T> load_model(const gp_model_desc& d) {
  Model<T> M;
  M.n = d.n_bodies;
  M.g = T(d.gravity);
  M.parent.resize(M.n); M.jtype.resize(M.n); M.axis.resize(M.n); M.tree.resize(M.n); M.I.resize(M.n);
  M.spring_k.resize(M.n); M.spring_l.resize(M.n); M.contacts.resize(M.n); M.qoff.resize(M.n); M.voff.resize(M.n);
  M.springs.resize(M.n);
  M.spheres.resize(M.n);
  for (int i = 0; i < M.n; ++i) {
    const gp_body_desc& b = d.bodies[i];
    M.parent[i] = b.parent;
    M.jtype[i] = b.jtype;
    M.axis[i] = V3<T>(T(b.axis[0]), T(b.axis[1]), T(b.axis[2]));
    M.tree[i].r = Q<T>(T(b.x_tree_quat[0]), T(b.x_tree_quat[1]), T(b.x_tree_quat[2]), T(b.x_tree_quat[3]));
    M.tree[i].t = V3<T>(T(b.x_tree_trans[0]), T(b.x_tree_trans[1]), T(b.x_tree_trans[2]));
    M3<T> J;
    J.m[0][0] = T(b.moment[0]); J.m[1][1] = T(b.moment[1]); J.m[2][2] = T(b.moment[2]);
    J.m[0][1] = J.m[1][0] = T(b.moment[3]);
    J.m[0][2] = J.m[2][0] = T(b.moment[4]);
    J.m[1][2] = J.m[2][1] = T(b.moment[5]);
    M.I[i].J = J;
    M.I[i].c = V3<T>(T(b.mc[0]), T(b.mc[1]), T(b.mc[2]));
    M.I[i].m = T(b.mass);
    M.spring_k[i] = T(b.spring_k);
    M.spring_l[i] = T(b.spring_l);
    M.qoff[i] = M.nq; M.voff[i] = M.nv;
    M.nq += jdof_q(b.jtype); M.nv += jdof_v(b.jtype);
  }
  for (int c = 0; c < d.n_contacts; ++c) {
    const gp_contact_desc& cd = d.contacts[c];
    M.contacts[cd.body].push_back({V3<T>(T(cd.p[0]), T(cd.p[1]), T(cd.p[2])), T(cd.k)});
  }
  for (int c = 0; c < d.n_spring_contacts; ++c) {
    const gp_spring_contact_desc& sd = d.spring_contacts[c];
    M.springs[sd.body].push_back({V3<T>(T(sd.direction[0]), T(sd.direction[1]), T(sd.direction[2])), T(sd.l_rest),
                                  T(sd.k)});
  }
  for (int s = 0; s < d.n_spheres; ++s) {
    const gp_sphere_desc& sd = d.spheres[s];
    SphereC<T>& S = M.spheres[sd.body];
    S.present = true;
    S.enabled = sd.enabled != 0;
    S.c = V3<T>(T(sd.center[0]), T(sd.center[1]), T(sd.center[2]));
    S.r = T(sd.radius);
  }
  for (int h = 0; h < d.n_halfspaces; ++h) {
    const gp_halfspace_desc& hd = d.halfspaces[h];
    M.halfspaces.push_back({V3<T>(T(hd.point[0]), T(hd.point[1]), T(hd.point[2])),
                            V3<T>(T(hd.normal[0]), T(hd.normal[1]), T(hd.normal[2])), T(hd.alpha), T(hd.mu)});
  }
  return M;
}

// Joint transform successor -> predecessor for the joint's current q
// (revolute.rs:97-102, prismatic.rs:82-87, floating.rs:26-31, pose.rs:30-33).
template <class T>
Iso<T> joint_transform(const Model<T>& M, int i, const T* q) {
  const Iso<T>& X0 = M.tree[i];
  const T* qi = q + M.qoff[i];
  switch (M.jtype[i]) {
    case GP_JOINT_REVOLUTE: {
      Iso<T> rot{q_axis_angle(M.axis[i], qi[0]), V3<T>()};
      return imul(X0, rot);
    }
    case GP_JOINT_PRISMATIC: {
      Iso<T> tr{Q<T>(), M.axis[i] * qi[0]};
      return imul(X0, tr);
    }
    case GP_JOINT_FLOATING: {
      Iso<T> pose{Q<T>(qi[0], qi[1], qi[2], qi[3]), V3<T>(qi[4], qi[5], qi[6])};
      return imul(X0, pose);
    }
    default:
      return X0;
  }
}

// A twist / acceleration / wrench as [angular; linear].
template <class T>
struct SV {
  V3<T> a, l;
};
template <class T> SV<T> operator+(const SV<T>& x, const SV<T>& y) { return {x.a + y.a, x.l + y.l}; }
template <class T> SV<T> operator-(const SV<T>& x, const SV<T>& y) { return {x.a - y.a, x.l - y.l}; }

// Twist::transform (twist.rs:74-93): express a motion vector of frame `from`
// in frame `to` of the isometry.
template <class T> SV<T> twist_transform(const SV<T>& s, const Iso<T>& X) {
  V3<T> a = qrot(X.r, s.a);
  V3<T> l = qrot(X.r, s.l) + cross(X.t, a);
  return {a, l};
}

// SpatialInertia::transform (inertia.rs:106-134)
template <class T> Inertia<T> inertia_transform(const Inertia<T>& I, const Iso<T>& X) {
  M3<T> R = qmat(X.r);
  V3<T> p = X.t;
  V3<T> Rmc = R * I.c;
  V3<T> mp = I.m * p;
  V3<T> mcnew = Rmc + mp;
  M3<T> Xm = outer(Rmc, p);
  M3<T> Y = Xm + Xm.transpose() + outer(mp, p);
  M3<T> Jnew = R * I.J * R.transpose() - Y + Y.trace() * M3<T>::identity();
  return {Jnew, mcnew, I.m};
}

// Motion subspace columns in the body frame (revolute.rs:87, prismatic.rs:72,
// floating.rs:37, fixed.rs:17).
template <class T> int motion_subspace(const Model<T>& M, int i, SV<T>* S) {
  switch (M.jtype[i]) {
    case GP_JOINT_REVOLUTE: S[0] = {M.axis[i], V3<T>()}; return 1;
    case GP_JOINT_PRISMATIC: S[0] = {V3<T>(), M.axis[i]}; return 1;
    case GP_JOINT_FLOATING:
      for (int k = 0; k < 6; ++k) {
        SV<T> s;
        if (k < 3) s.a[k] = T(1); else s.l[k - 3] = T(1);
        S[k] = s;
      }
      return 6;
    default: return 0;
  }
}

// GeometricJacobian::transform (geometric_jacobian.rs:87-107): uses rot() as a
// matrix and colwise cross with the translation.
template <class T> SV<T> jac_col_transform(const SV<T>& s, const Iso<T>& X) {
  M3<T> R = qmat(X.r);
  V3<T> a = R * s.a;
  V3<T> l = R * s.l + cross(X.t, a);
  return {a, l};
}

// mul_inertia (util.rs:18-28)
template <class T> SV<T> mul_inertia(const Inertia<T>& I, const SV<T>& v) {
  return {I.J * v.a + cross(I.c, v.l), I.m * v.l - cross(I.c, v.a)};
}

// Hunt-Crossley normal force + regularised Coulomb friction
// (contact.rs:260-302); the half-space side has k = 50e3 (contact.rs:321-338).
template <class T>
V3<T> contact_force(T z, const V3<T>& normal, const V3<T>& vel, T kA, T kB, T alpha, T mu) {
  using std::pow; using std::sqrt; using std::fmax;
  T z_dot = -dot(vel, normal);
  T zn = pow(z, T(1.5));
  T k = kA * kB / (kA + kB);
  T lam = T(1.5) * alpha * k;
  T pi = fmax(lam * zn * z_dot + k * zn, T(0));  // NaN.max(0.0) = 0.0 (F9)
  V3<T> f_normal = pi * normal;
  V3<T> v_t = vel + z_dot * normal;
  T v_t_norm = sqrt(dot(v_t, v_t));
  V3<T> f_fric;
  if (!(v_t_norm == T(0))) {
    T v_s = T(1e-3);
    T s = v_t_norm / v_s;
    T mu_eff = (s > T(1)) ? mu : mu * s;
    f_fric = -(mu_eff * pi) * (v_t / v_t_norm);
  }
  return f_normal + f_fric;
}

constexpr double kHalfspaceK = 50e3;     // contact.rs:328
constexpr double kInsideMargin = 1e-8;   // halfspace.rs:42

// ---------------------------------------------------------------- CRBA + RNEA + LU
// Workspace-free restatement of dynamics_continuous (dynamics.rs:322-364).
// Returns 0 on success, 1 if LU finds a zero pivot (the reference panics).
template <class T> T springs_k(const Model<T>& M, int body, size_t s) { return M.springs[body][s].k; }

// Spring contacts visited in body order (description order within a body):
// the state index of every spring contact. ss may be null when the model has
// none. Returns 2 where the reference panics on a spring pulling into its
// half-space (contact.rs:161-163).
template <class T>
int crba_dynamics(const Model<T>& M, const T* q, const T* v, const T* tau, T* vdot,
                  T* Mout = nullptr, T* cout = nullptr, SpringSt<T>* ss = nullptr, bool contact_wrenches = true) {
  const int n = M.n, nv = M.nv;
  std::vector<SpringSt<T>> fresh;  // no caller state: the springs as described, unregistered
  if (!ss) {
    for (int i = 0; i < n; ++i)
      for (const SpringC<T>& sc : M.springs[i]) {
        SpringSt<T> st;
        st.dir = sc.dir;
        st.l_rest = sc.l_rest;
        fresh.push_back(st);
      }
    ss = fresh.data();
  }
  std::vector<Iso<T>> b2r(n);        // bodies_to_root (without the world entry)
  std::vector<SV<T>> jtw(n), tw(n);  // joint twists (body frame), body twists (world)
  for (int i = 0; i < n; ++i) {      // mechanism.rs:153-170
    Iso<T> XJ = joint_transform(M, i, q);
    b2r[i] = M.parent[i] < 0 ? XJ : imul(b2r[M.parent[i]], XJ);
  }
  for (int i = 0; i < n; ++i) {  // twist.rs:36-61, 176-204
    const T* vi = v + M.voff[i];
    SV<T> j;
    switch (M.jtype[i]) {
      case GP_JOINT_REVOLUTE: j.a = M.axis[i] * vi[0]; break;
      case GP_JOINT_PRISMATIC: j.l = M.axis[i] * vi[0]; break;
      case GP_JOINT_FLOATING: j.a = V3<T>(vi[0], vi[1], vi[2]); j.l = V3<T>(vi[3], vi[4], vi[5]); break;
      default: break;
    }
    jtw[i] = j;
    SV<T> parent_tw = M.parent[i] < 0 ? SV<T>() : tw[M.parent[i]];
    tw[i] = parent_tw + twist_transform(j, b2r[i]);
  }

  // ---- mass matrix (mechanism.rs:637-696)
  std::vector<Inertia<T>> Iw(n), Ic(n);
  std::vector<std::array<SV<T>, 6>> Sw(n);
  std::vector<int> dof(n);
  for (int i = 0; i < n; ++i) {
    Iw[i] = inertia_transform(M.I[i], b2r[i]);
    SV<T> S[6];
    dof[i] = motion_subspace(M, i, S);
    for (int k = 0; k < dof[i]; ++k) Sw[i][k] = jac_col_transform(S[k], b2r[i]);
  }
  for (int i = n - 1; i >= 0; --i) {  // compute_crb_inertias (mechanism.rs:606-625)
    Inertia<T> ic = Iw[i];
    for (int c = 0; c < n; ++c)
      if (M.parent[c] == i) {
        ic.J = ic.J + Ic[c].J; ic.c = ic.c + Ic[c].c; ic.m = ic.m + Ic[c].m;
      }
    Ic[i] = ic;
  }
  std::vector<T> H(size_t(nv) * nv, T(0));
  auto supports = [&](int j, int i) {  // joint j supports body i
    for (int b = i; b >= 0; b = M.parent[b]) if (b == j) return true;
    return false;
  };
  for (int i = 0; i < n; ++i) {
    if (dof[i] == 0) continue;
    SV<T> F[6];  // MomentumMatrix::mul (momentum.rs:17-32)
    for (int k = 0; k < dof[i]; ++k) F[k] = mul_inertia(Ic[i], Sw[i][k]);
    for (int j = 0; j <= i; ++j) {
      if (dof[j] == 0 || !supports(j, i)) continue;
      for (int a = 0; a < dof[i]; ++a)
        for (int b = 0; b < dof[j]; ++b) {  // transpose_mul (momentum.rs:35-46)
          T h = dot(F[a].a, Sw[j][b].a) + dot(F[a].l, Sw[j][b].l);
          H[size_t(M.voff[i] + a) * nv + M.voff[j] + b] = h;
        }
    }
  }
  for (int r = 0; r < nv; ++r)
    for (int c = r + 1; c < nv; ++c) H[size_t(r) * nv + c] = H[size_t(c) * nv + r];

  // ---- contact wrenches (contact.rs:97-129), world frame about world origin;
  // none on the velocity-stepping path (dynamics.rs:377-383)
  std::vector<SV<T>> wc(n);
  for (int i = 0; i < n && contact_wrenches; ++i) {
    M3<T> R = qmat(b2r[i].r);
    for (const Contact<T>& cp : M.contacts[i]) {
      V3<T> pw = R * cp.p + b2r[i].t;              // ContactPoint::transform
      V3<T> vel = tw[i].l + cross(tw[i].a, pw);  // Twist::point_velocity
      for (const HalfSpace<T>& h : M.halfspaces) {
        if (!(dot(pw - h.point, h.normal) <= T(kInsideMargin))) continue;  // has_inside
        T z = -dot(pw - h.point, h.normal);
        V3<T> F = contact_force(z, h.normal, vel, cp.k, T(kHalfspaceK), h.alpha, h.mu);
        wc[i].a += cross(pw, F);
        wc[i].l += F;
      }
    }
    // spring contacts with half-spaces (contact.rs:131-186), state changed here
    for (size_t s = 0; s < M.springs[i].size(); ++s) {
      SpringSt<T>& st = *(ss++);
      const V3<T> body_location = b2r[i].t;
      if (!st.registered) {
        const V3<T> spring_direction = R * st.dir;
        const V3<T> contact_location = body_location + spring_direction * st.l_rest;
        for (size_t h = 0; h < M.halfspaces.size(); ++h) {
          const HalfSpace<T>& hs = M.halfspaces[h];
          if (!(dot(contact_location - hs.point, hs.normal) <= T(kInsideMargin))) continue;  // has_inside
          st.registered = true;
          st.point = contact_location;
          st.halfspace = int(h);
          break;
        }
      } else {
        const V3<T> diff = st.point - body_location;
        const T dist = std::sqrt(dot(diff, diff));
        const V3<T> spring_direction = diff / dist;  // UnitVector3::new_normalize
        if (dot(spring_direction, M.halfspaces[st.halfspace].normal) > T(0)) return 2;
        if (dist < st.l_rest) {
          const T force = -springs_k(M, i, s) * (dist - st.l_rest);  // spring_force (energy_control.rs:13)
          const V3<T> F = -spring_direction * force;
          wc[i].a += cross(body_location, F);  // Wrench::from_force at the body's origin
          wc[i].l += F;
        } else {
          st.registered = false;
          const T nn = std::sqrt(dot(spring_direction, spring_direction));
          st.dir = spring_direction / nn;
        }
      }
    }
  }

  // ---- bias accelerations (dynamics.rs:143-224)
  std::vector<SV<T>> cor(n), acc(n);
  for (int i = 0; i < n; ++i) {
    Iso<T> r2b = iinv(b2r[i]);
    SV<T> btw = twist_transform(tw[i], r2b);  // body twist in body frame
    SV<T> cb;  // se3_commutator(body_twist, joint_twist)
    cb.a = cross(btw.a, jtw[i].a);
    cb.l = cross(btw.a, jtw[i].l) + cross(btw.l, jtw[i].a);
    SV<T> cw = twist_transform(cb, b2r[i]);
    SV<T> cp = M.parent[i] < 0 ? SV<T>() : cor[M.parent[i]];
    cor[i] = cp + cw;
    SV<T> g;
    g.l = V3<T>(T(0), T(0), M.g);
    acc[i] = g + cor[i];
  }
  // ---- newton_euler (dynamics.rs:41-100)
  std::vector<SV<T>> w(n);
  for (int i = 0; i < n; ++i) {
    SV<T> f = mul_inertia(Iw[i], acc[i]);
    SV<T> hm = mul_inertia(Iw[i], tw[i]);
    f.a += cross(tw[i].a, hm.a) + cross(tw[i].l, hm.l);
    f.l += cross(tw[i].a, hm.l);
    w[i] = f - wc[i];  // dynamics.rs:244-246
  }
  // ---- compute_torques (wrench.rs:96-126)
  std::vector<T> c(nv, T(0));
  for (int i = n - 1; i >= 0; --i) {
    if (M.parent[i] >= 0) { w[M.parent[i]].a += w[i].a; w[M.parent[i]].l += w[i].l; }
    for (int k = 0; k < dof[i]; ++k) c[M.voff[i] + k] = dot(Sw[i][k].a, w[i].a) + dot(Sw[i][k].l, w[i].l);
  }
  // ---- prismatic springs (dynamics.rs:298-315) and the solve (dynamics.rs:255-276)
  std::vector<T> rhs(nv);
  for (int k = 0; k < nv; ++k) rhs[k] = tau[k];
  for (int i = 0; i < n; ++i)
    if (M.jtype[i] == GP_JOINT_PRISMATIC && !(M.spring_k[i] == T(0)))
      rhs[M.voff[i]] = rhs[M.voff[i]] + (-M.spring_k[i] * (q[M.qoff[i]] - M.spring_l[i]));
  for (int k = 0; k < nv; ++k) rhs[k] = rhs[k] - c[k];
  if (Mout) for (size_t k = 0; k < H.size(); ++k) Mout[k] = H[k];
  if (cout) for (int k = 0; k < nv; ++k) cout[k] = c[k];
  if (nv == 0) return 0;
  // partial-pivot LU (nalgebra: pivot = max |a| in the column), then solve
  std::vector<T> A = H;
  std::vector<int> piv(nv);
  for (int k = 0; k < nv; ++k) piv[k] = k;
  for (int col = 0; col < nv; ++col) {
    int p = col;
    T best = std::fabs(A[size_t(col) * nv + col]);
    for (int r = col + 1; r < nv; ++r) {
      T a = std::fabs(A[size_t(r) * nv + col]);
      if (a > best) { best = a; p = r; }
    }
    if (best == T(0)) return 1;
    if (p != col) {
      for (int k = 0; k < nv; ++k) std::swap(A[size_t(p) * nv + k], A[size_t(col) * nv + k]);
      std::swap(rhs[p], rhs[col]);
    }
    T d = A[size_t(col) * nv + col];
    for (int r = col + 1; r < nv; ++r) {
      T f = A[size_t(r) * nv + col] / d;
      A[size_t(r) * nv + col] = f;
      for (int k = col + 1; k < nv; ++k) A[size_t(r) * nv + k] = A[size_t(r) * nv + k] - f * A[size_t(col) * nv + k];
      rhs[r] = rhs[r] - f * rhs[col];
    }
  }
  for (int r = nv - 1; r >= 0; --r) {
    T s = rhs[r];
    for (int k = r + 1; k < nv; ++k) s = s - A[size_t(r) * nv + k] * vdot[k];
    vdot[r] = s / A[size_t(r) * nv + r];
  }
  return 0;
}

// ---------------------------------------------------------------- integrators
// compute_new_q (integrators.rs:296-319) with quaternion_derivative
// (util.rs:83-102) and UnitQuaternion::from_quaternion (normalisation).
template <class T>
void advance_q(const Model<T>& M, const T* q, const T* vq, T dt, T* q_out) {
  using std::sqrt;
  for (int i = 0; i < M.n; ++i) {
    const T* qi = q + M.qoff[i];
    const T* vi = vq + M.voff[i];
    T* qo = q_out + M.qoff[i];
    switch (M.jtype[i]) {
      case GP_JOINT_REVOLUTE:
      case GP_JOINT_PRISMATIC: qo[0] = qi[0] + vi[0] * dt; break;
      case GP_JOINT_FLOATING: {
        T x = qi[0], y = qi[1], z = qi[2], w = qi[3];
        T ox = vi[0], oy = vi[1], oz = vi[2];
        T half = T(2);
        T dw = (-x * ox - y * oy - z * oz) / half;
        T dx = (w * ox - z * oy + y * oz) / half;
        T dy = (z * ox + w * oy - x * oz) / half;
        T dz = (-y * ox + x * oy + w * oz) / half;
        Q<T> rot(x, y, z, w);
        V3<T> tdot = qrot(rot, V3<T>(vi[3], vi[4], vi[5]));
        qo[4] = qi[4] + tdot.x * dt;
        qo[5] = qi[5] + tdot.y * dt;
        qo[6] = qi[6] + tdot.z * dt;
        T nx = x + dx * dt, ny = y + dy * dt, nz = z + dz * dt, nw = w + dw * dt;
        T nrm = sqrt(nx * nx + ny * ny + nz * nz + nw * nw);
        qo[0] = nx / nrm; qo[1] = ny / nrm; qo[2] = nz / nrm; qo[3] = nw / nrm;
        break;
      }
      default: break;
    }
  }
}

template <class T, class Dyn>
int step_integrate(const Model<T>& M, int integrator, const T* q, const T* v, const T* tau, T dt,
                   T* q_out, T* v_out, T* vdot_out, Dyn dyn) {
  const int nq = M.nq, nv = M.nv;
  std::vector<T> a1(nv);
  int st = dyn(M, q, v, tau, a1.data());
  if (vdot_out) for (int k = 0; k < nv; ++k) vdot_out[k] = a1[k];
  if (st) return st;
  if (integrator == GP_SEMI_IMPLICIT_EULER) {  // integrators.rs:276-292
    for (int k = 0; k < nv; ++k) v_out[k] = v[k] + a1[k] * dt;
    std::vector<T> qn(nq);
    advance_q(M, q, v_out, dt, qn.data());
    for (int k = 0; k < nq; ++k) q_out[k] = qn[k];
    return 0;
  }
  // euler_step (integrators.rs:230-271): v' = v + dt vdot; q' = q (+) dt v
  auto euler = [&](const T* a, T h, T* qe, T* ve) {
    for (int k = 0; k < nv; ++k) ve[k] = v[k] + a[k] * h;
    advance_q(M, q, v, h, qe);
  };
  std::vector<T> q1(nq), v1(nv), a2(nv);
  if (integrator == GP_RUNGE_KUTTA_2) {  // integrators.rs:177-193
    euler(a1.data(), dt / T(2), q1.data(), v1.data());
    if ((st = dyn(M, q1.data(), v1.data(), tau, a2.data()))) return st;
    euler(a2.data(), dt, q_out, v_out);
    return 0;
  }
  // RK4 (integrators.rs:195-225)
  std::vector<T> a3(nv), a4(nv), af(nv);
  euler(a1.data(), dt / T(2), q1.data(), v1.data());
  if ((st = dyn(M, q1.data(), v1.data(), tau, a2.data()))) return st;
  euler(a2.data(), dt / T(2), q1.data(), v1.data());
  if ((st = dyn(M, q1.data(), v1.data(), tau, a3.data()))) return st;
  euler(a3.data(), dt, q1.data(), v1.data());
  if ((st = dyn(M, q1.data(), v1.data(), tau, a4.data()))) return st;
  for (int k = 0; k < nv; ++k) af[k] = (a1[k] + a2[k] * T(2) + a3[k] * T(2) + a4[k]) / T(6);
  euler(af.data(), dt, q_out, v_out);
  return 0;
}

// ---------------------------------------------------------------- velocity stepping
// Partial-pivot LU (nalgebra 0.33.2 `lu()`: the pivot is the largest |a| of
// the column; the row swaps are replayed on every right-hand side).
template <class T>
struct LUF {
  int n = 0;
  std::vector<T> a;      // L below the diagonal (unit diagonal), U on and above
  std::vector<int> swp;  // row swapped with row k at step k
  bool ok = true;
};
template <class T> LUF<T> lu_factor(const std::vector<T>& A, int n) {
  LUF<T> f;
  f.n = n;
  f.a = A;
  f.swp.assign(n, 0);
  for (int col = 0; col < n; ++col) {
    int p = col;
    T best = std::fabs(f.a[size_t(col) * n + col]);
    for (int r = col + 1; r < n; ++r) {
      T v = std::fabs(f.a[size_t(r) * n + col]);
      if (v > best) { best = v; p = r; }
    }
    f.swp[col] = p;
    if (best == T(0)) { f.ok = false; return f; }
    if (p != col)
      for (int k = 0; k < n; ++k) std::swap(f.a[size_t(p) * n + k], f.a[size_t(col) * n + k]);
    const T d = f.a[size_t(col) * n + col];
    for (int r = col + 1; r < n; ++r) {
      T m = f.a[size_t(r) * n + col] / d;
      f.a[size_t(r) * n + col] = m;
      for (int k = col + 1; k < n; ++k) f.a[size_t(r) * n + k] = f.a[size_t(r) * n + k] - m * f.a[size_t(col) * n + k];
    }
  }
  return f;
}
template <class T> void lu_solve(const LUF<T>& f, const T* b, T* x) {
  const int n = f.n;
  std::vector<T> y(b, b + n);
  for (int k = 0; k < n; ++k) std::swap(y[k], y[f.swp[k]]);
  for (int r = 0; r < n; ++r)
    for (int k = 0; k < r; ++k) y[r] = y[r] - f.a[size_t(r) * n + k] * y[k];
  for (int r = n - 1; r >= 0; --r) {
    T s = y[r];
    for (int k = r + 1; k < n; ++k) s = s - f.a[size_t(r) * n + k] * x[k];
    x[r] = s / f.a[size_t(r) * n + r];
  }
}

// The contact impulse program of dynamics.rs:668-704 (Clarabel 0.11.1 with
// default settings there): minimise 1/2 x^T G x + g^T x with
// (mu x_3c, x_3c+1, x_3c+2) in the 3-dimensional second-order cone for every
// contact c — Clarabel's A = -diag(mu, 1, 1), b = 0, SecondOrderConeT(3),
// mu = 0.95 (dynamics.rs:674). Clarabel is a Rust crate absent here; it is
// restated as the method family it implements: a primal-dual interior-point
// method with Nesterov-Todd scaling and Mehrotra's predictor-corrector
// (L. Vandenberghe, "The CVXOPT linear and quadratic cone program solvers",
// 2010, sections 2-4), without Clarabel's homogeneous embedding, run to a
// relative gap of 1e-16 and residual of 1e-14 (Clarabel stops at 1e-8), or,
// once rounding stalls the progress for 3 iterations, to the iterate with the
// smallest of the larger of the two. G is positive
// semidefinite and the optimum in x may be a set; the velocities it produces
// are unique. Returns 0, or 1 if the iterations produced no finite point.
constexpr double kConeMu = 0.95;
constexpr int kConeMaxIter = 50;
constexpr int kConeStall = 3;  // iterations without a better iterate
constexpr double kConeGapTol = 1e-16;
constexpr double kConeResTol = 1e-14;

template <class T> struct Soc3 {
  T a[3];
};
template <class T> T soc_det(const T* u) { return u[0] * u[0] - u[1] * u[1] - u[2] * u[2]; }
// Nesterov-Todd scaling of one cone: W z = W^-1 s (Vandenberghe 2010, 4.3):
// W = eta [w0 w1^T; w1 I + w1 w1^T / (1 + w0)], w = (s/|s| + R z/|z|) / (2 gamma)
template <class T> void nt_scaling(const T* s, const T* z, T W[3][3], T Wi[3][3]) {
  using std::sqrt;
  T ds = sqrt(soc_det(s)), dz = sqrt(soc_det(z));
  T sb[3] = {s[0] / ds, s[1] / ds, s[2] / ds}, zb[3] = {z[0] / dz, z[1] / dz, z[2] / dz};
  T gam = sqrt((T(1) + sb[0] * zb[0] + sb[1] * zb[1] + sb[2] * zb[2]) / T(2));
  T w[3] = {(sb[0] + zb[0]) / (T(2) * gam), (sb[1] - zb[1]) / (T(2) * gam), (sb[2] - zb[2]) / (T(2) * gam)};
  T eta = sqrt(ds / dz);
  T f = T(1) / (T(1) + w[0]);
  T B[3][3] = {{w[0], w[1], w[2]}, {w[1], T(1) + f * w[1] * w[1], f * w[1] * w[2]}, {w[2], f * w[1] * w[2], T(1) + f * w[2] * w[2]}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      W[i][j] = eta * B[i][j];
      T sg = ((i == 0) != (j == 0)) ? T(-1) : T(1);
      Wi[i][j] = sg * B[i][j] / eta;
    }
}
// u o v and the y solving u o y = w in the Jordan algebra of the cone
template <class T> void soc_prod(const T* u, const T* v, T* out) {
  out[0] = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
  out[1] = u[0] * v[1] + v[0] * u[1];
  out[2] = u[0] * v[2] + v[0] * u[2];
}
template <class T> void soc_div(const T* u, const T* w, T* y) {
  y[0] = (u[0] * w[0] - u[1] * w[1] - u[2] * w[2]) / soc_det(u);
  y[1] = (w[1] - y[0] * u[1]) / u[0];
  y[2] = (w[2] - y[0] * u[2]) / u[0];
}
// largest step a with u + a du still in the cone: the first positive root of
// det, and never past the apex (u0 + a du0 = 0): a direction along the cone's
// axis has a double root there, which rounding can turn into no real root
template <class T> T soc_max_step(const T* u, const T* du) {
  using std::sqrt;
  T a = soc_det(du);
  T b = u[0] * du[0] - u[1] * du[1] - u[2] * du[2];
  T c = soc_det(u);
  T best = du[0] < T(0) ? -u[0] / du[0] : T(1e300);
  if (a == T(0)) {
    if (b < T(0)) best = std::fmin(best, -c / (T(2) * b));
    return best;
  }
  T disc = b * b - a * c;
  if (disc < T(0)) return best;
  T q = -(b + (b < T(0) ? -sqrt(disc) : sqrt(disc)));
  T r1 = q / a, r2 = q != T(0) ? c / q : T(1e300);
  if (r1 > T(0) && r1 < best) best = r1;
  if (r2 > T(0) && r2 < best) best = r2;
  return best;
}

template <class T>
int cone_program(const std::vector<T>& G, const std::vector<T>& g, int nc, std::vector<T>& x, int* iters = nullptr) {
  using std::fabs; using std::sqrt;
  const int m = 3 * nc;
  const T mu = T(kConeMu);
  std::vector<T> d(m), z(m), s(m), rd(m), lt(m), xbest, dx(m), dz(m), ds(m), xi(m), rhs(m), L(size_t(m) * m);
  std::vector<std::array<std::array<T, 3>, 3>> W(nc), Wi(nc);
  x.assign(m, T(0));
  for (int c = 0; c < nc; ++c) {
    d[3 * c] = mu; d[3 * c + 1] = T(1); d[3 * c + 2] = T(1);
    x[3 * c] = T(1) / mu;
    z[3 * c] = T(1); z[3 * c + 1] = T(0); z[3 * c + 2] = T(0);
  }
  T gnorm = T(0);
  for (int k = 0; k < m; ++k) gnorm = std::fmax(gnorm, fabs(g[k]));
  T best = T(1e300);
  xbest = x;
  int it = 0, since_best = 0;
  auto solve = [&](const std::vector<T>& xi_, std::vector<T>& dx_, std::vector<T>& dz_) {
    for (int c = 0; c < nc; ++c)
      for (int i = 0; i < 3; ++i) {
        T t = T(0);
        for (int j = 0; j < 3; ++j) t = t + Wi[c][i][j] * xi_[3 * c + j];
        rhs[3 * c + i] = -rd[3 * c + i] + d[3 * c + i] * t;
      }
    for (int r = 0; r < m; ++r) {  // L y = rhs
      T t = rhs[r];
      for (int k = 0; k < r; ++k) t = t - L[size_t(r) * m + k] * dx_[k];
      dx_[r] = t / L[size_t(r) * m + r];
    }
    for (int r = m - 1; r >= 0; --r) {  // L^T dx = y
      T t = dx_[r];
      for (int k = r + 1; k < m; ++k) t = t - L[size_t(k) * m + r] * dx_[k];
      dx_[r] = t / L[size_t(r) * m + r];
    }
    for (int c = 0; c < nc; ++c) {
      T a[3], b[3];
      for (int i = 0; i < 3; ++i) {
        a[i] = T(0);
        for (int j = 0; j < 3; ++j) a[i] = a[i] + Wi[c][i][j] * d[3 * c + j] * dx_[3 * c + j];
      }
      for (int i = 0; i < 3; ++i) {
        b[i] = T(0);
        for (int j = 0; j < 3; ++j) b[i] = b[i] + Wi[c][i][j] * (xi_[3 * c + j] - a[j]);
      }
      for (int i = 0; i < 3; ++i) dz_[3 * c + i] = b[i];
    }
  };
  for (it = 0; it < kConeMaxIter; ++it) {
    T gap = T(0), rmax = T(0), xmax = T(0);
    for (int k = 0; k < m; ++k) s[k] = d[k] * x[k];
    for (int r = 0; r < m; ++r) {
      T t = g[r] - d[r] * z[r];
      for (int k = 0; k < m; ++k) t = t + G[size_t(r) * m + k] * x[k];
      rd[r] = t;
      rmax = std::fmax(rmax, fabs(t));
      gap = gap + s[r] * z[r];
      xmax = std::fmax(xmax, fabs(x[r]));
    }
    bool finite = std::isfinite(gap) && std::isfinite(rmax);
    if (!finite) break;
    // relative gap (impulse x velocity scale) and residual (velocity scale)
    const T gap_rel = gap / ((T(1) + gnorm) * (T(1) + xmax));
    const T res_rel = rmax / (T(1) + gnorm);
    const T merit = std::fmax(gap_rel, res_rel);
    if (merit < best) { best = merit; xbest = x; since_best = 0; }
    if (gap_rel <= T(kConeGapTol) && res_rel <= T(kConeResTol)) break;
    if (++since_best > kConeStall) break;  // rounding has stopped the progress
    const T mug = gap / T(nc);
    for (int c = 0; c < nc; ++c) {
      T Wc[3][3], Wic[3][3];
      nt_scaling(&s[3 * c], &z[3 * c], Wc, Wic);
      for (int i = 0; i < 3; ++i) {
        lt[3 * c + i] = Wc[i][0] * z[3 * c] + Wc[i][1] * z[3 * c + 1] + Wc[i][2] * z[3 * c + 2];
        for (int j = 0; j < 3; ++j) { W[c][i][j] = Wc[i][j]; Wi[c][i][j] = Wic[i][j]; }
      }
    }
    // K = G + D W^-2 D (block diagonal), Cholesky K = L L^T
    for (size_t k = 0; k < L.size(); ++k) L[k] = G[k];
    for (int c = 0; c < nc; ++c)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          T t = T(0);
          for (int k = 0; k < 3; ++k) t = t + Wi[c][i][k] * Wi[c][k][j];
          L[size_t(3 * c + i) * m + 3 * c + j] = L[size_t(3 * c + i) * m + 3 * c + j] + d[3 * c + i] * t * d[3 * c + j];
        }
    bool pd = true;
    for (int j = 0; j < m && pd; ++j) {
      T t = L[size_t(j) * m + j];
      for (int k = 0; k < j; ++k) t = t - L[size_t(j) * m + k] * L[size_t(j) * m + k];
      if (!(t > T(0))) { pd = false; break; }
      T ljj = sqrt(t);
      L[size_t(j) * m + j] = ljj;
      for (int r = j + 1; r < m; ++r) {
        T u = L[size_t(r) * m + j];
        for (int k = 0; k < j; ++k) u = u - L[size_t(r) * m + k] * L[size_t(j) * m + k];
        L[size_t(r) * m + j] = u / ljj;
      }
    }
    if (!pd) break;
    // predictor
    for (int k = 0; k < m; ++k) xi[k] = -lt[k];
    solve(xi, dx, dz);
    T alpha = T(1);
    for (int k = 0; k < m; ++k) ds[k] = d[k] * dx[k];
    for (int c = 0; c < nc; ++c)
      alpha = std::fmin(alpha, std::fmin(soc_max_step(&s[3 * c], &ds[3 * c]), soc_max_step(&z[3 * c], &dz[3 * c])));
    T sz = T(0);
    for (int k = 0; k < m; ++k) sz = sz + (s[k] + alpha * ds[k]) * (z[k] + alpha * dz[k]);
    T ratio = sz / gap;
    T sigma = ratio * ratio * ratio;
    // corrector: xi = lt \ (sigma mu e - lt o lt - (W^-1 ds) o (W dz))
    for (int c = 0; c < nc; ++c) {
      T a[3], b[3], ll[3], cc[3], w[3];
      for (int i = 0; i < 3; ++i) {
        a[i] = Wi[c][i][0] * ds[3 * c] + Wi[c][i][1] * ds[3 * c + 1] + Wi[c][i][2] * ds[3 * c + 2];
        b[i] = W[c][i][0] * dz[3 * c] + W[c][i][1] * dz[3 * c + 1] + W[c][i][2] * dz[3 * c + 2];
      }
      soc_prod(&lt[3 * c], &lt[3 * c], ll);
      soc_prod(a, b, cc);
      w[0] = sigma * mug - ll[0] - cc[0];
      w[1] = -ll[1] - cc[1];
      w[2] = -ll[2] - cc[2];
      soc_div(&lt[3 * c], w, &xi[3 * c]);
    }
    solve(xi, dx, dz);
    for (int k = 0; k < m; ++k) ds[k] = d[k] * dx[k];
    alpha = T(1);
    for (int c = 0; c < nc; ++c)
      alpha = std::fmin(alpha, T(0.99) * std::fmin(soc_max_step(&s[3 * c], &ds[3 * c]), soc_max_step(&z[3 * c], &dz[3 * c])));
    for (int k = 0; k < m; ++k) { x[k] = x[k] + alpha * dx[k]; z[k] = z[k] + alpha * dz[k]; }
  }
  if (iters) *iters = it;
  x = xbest;
  for (int k = 0; k < m; ++k)
    if (!std::isfinite(x[k])) return 1;
  return best < T(1e300) ? 0 : 1;
}

// One contact of the velocity-stepping path: normal, world point, body
// (collision_detection, dynamics.rs:546-606: point contacts inside a
// half-space, then sphere colliders touching one)
template <class T>
struct VsContact {
  V3<T> n, p;
  int body;
};

template <class T> std::vector<Iso<T>> bodies_to_root(const Model<T>& M, const T* q) {
  std::vector<Iso<T>> b2r(M.n);
  for (int i = 0; i < M.n; ++i) {
    Iso<T> XJ = joint_transform(M, i, q);
    b2r[i] = M.parent[i] < 0 ? XJ : imul(b2r[M.parent[i]], XJ);
  }
  return b2r;
}

// dynamics_discrete (dynamics.rs:366-424): free velocity without contact
// forces, then the impulses of the cone program over the touching contacts
// (solve_constraint_and_contact, dynamics.rs:647-742). Returns 0, 1 on a
// singular mass matrix, 3 when the cone program gives no finite impulse
// (the reference panics on NaN, dynamics.rs:731-733).
// forced_body / forced_h: a sphere the continuous-collision loop has just
// moved onto half-space forced_h; it is in contact (its distance is zero in
// exact arithmetic, and rounding may leave it some ulps above the plane, where
// the reference's loop would find the same collision again without moving).
template <class T>
int vs_dynamics(const Model<T>& M, const T* q, const T* v, const T* tau, T dt, T* v_next,
                std::vector<VsContact<T>>& contacts, int* cone_iters = nullptr, int forced_body = -1,
                int forced_h = -1) {
  using std::sqrt;
  const int n = M.n, nv = M.nv;
  std::vector<T> Mm(size_t(nv) * nv), c(nv), scratch(nv), zero(nv, T(0));
  crba_dynamics(M, q, v, zero.data(), scratch.data(), Mm.data(), c.data(), (SpringSt<T>*)nullptr, false);
  std::vector<T> rhs(nv), vfree(nv);
  for (int k = 0; k < nv; ++k) rhs[k] = tau[k];
  for (int i = 0; i < n; ++i)  // add_prismatic_joint_spring_force (dynamics.rs:298-315)
    if (M.jtype[i] == GP_JOINT_PRISMATIC && !(M.spring_k[i] == T(0)))
      rhs[M.voff[i]] = rhs[M.voff[i]] + (-M.spring_k[i] * (q[M.qoff[i]] - M.spring_l[i]));
  for (int k = 0; k < nv; ++k) rhs[k] = rhs[k] * dt - c[k] * dt;  // dynamics_solve(M, c dt, tau dt)
  LUF<T> lu = lu_factor(Mm, nv);
  if (nv > 0 && !lu.ok) return 1;
  if (nv > 0) lu_solve(lu, rhs.data(), vfree.data());
  for (int k = 0; k < nv; ++k) vfree[k] = v[k] + vfree[k];
  // world motion subspaces (build_jacobian_blocks, dynamics.rs:432-471)
  std::vector<Iso<T>> b2r = bodies_to_root(M, q);
  std::vector<std::array<SV<T>, 6>> Sw(n);
  std::vector<int> dof(n);
  for (int i = 0; i < n; ++i) {
    SV<T> S[6];
    dof[i] = motion_subspace(M, i, S);
    for (int k = 0; k < dof[i]; ++k) Sw[i][k] = jac_col_transform(S[k], b2r[i]);
  }
  // collision detection
  contacts.clear();
  for (int i = 0; i < n; ++i) {
    M3<T> R = qmat(b2r[i].r);
    for (const Contact<T>& cp : M.contacts[i]) {
      V3<T> pw = R * cp.p + b2r[i].t;
      for (const HalfSpace<T>& h : M.halfspaces)
        if (dot(pw - h.point, h.normal) <= T(kInsideMargin)) contacts.push_back({h.normal, pw, i});
    }
  }
  for (int i = 0; i < n; ++i) {
    const SphereC<T>& S = M.spheres[i];
    if (!S.present || !S.enabled) continue;
    V3<T> center = b2r[i].t + qrot(b2r[i].r, S.c);  // Sphere::center
    for (size_t hh = 0; hh < M.halfspaces.size(); ++hh) {
      const HalfSpace<T>& h = M.halfspaces[hh];
      T dist = dot(center - h.point, h.normal);
      if (dist <= S.r || (i == forced_body && int(hh) == forced_h))
        contacts.push_back({h.normal, center - dist * h.normal, i});  // contact_halfspace
    }
  }
  const int nc = (int)contacts.size();
  if (cone_iters) *cone_iters = 0;
  if (nc == 0) {
    for (int k = 0; k < nv; ++k) v_next[k] = vfree[k];
    return 0;
  }
  // contact Jacobians C X H (compose_contact_jacobian, dynamics.rs:747-820)
  const int m = 3 * nc;
  std::vector<T> J(size_t(m) * nv, T(0));
  for (int ci = 0; ci < nc; ++ci) {
    const V3<T> nn = contacts[ci].n, p = contacts[ci].p;
    V3<T> t = cross(nn, V3<T>(T(1), T(0), T(0)));
    if (sqrt(dot(t, t)) == T(0)) t = cross(nn, V3<T>(T(0), T(1), T(0)));
    t = t / sqrt(dot(t, t));
    V3<T> b = cross(nn, t);
    b = b / sqrt(dot(b, b));
    for (int j = contacts[ci].body; j >= 0; j = M.parent[j])
      for (int k = 0; k < dof[j]; ++k) {
        V3<T> pv = Sw[j][k].l - cross(p, Sw[j][k].a);  // [-skew(p) I] (w; v)
        const int col = M.voff[j] + k;
        J[size_t(3 * ci) * nv + col] = dot(nn, pv);
        J[size_t(3 * ci + 1) * nv + col] = dot(t, pv);
        J[size_t(3 * ci + 2) * nv + col] = dot(b, pv);
      }
  }
  // G = J M^-1 J^T (mass_inv from the LU, dynamics.rs:663-667), g = J v_free
  std::vector<T> Minv(size_t(nv) * nv), e(nv), col(nv);
  for (int k = 0; k < nv; ++k) {
    for (int r = 0; r < nv; ++r) e[r] = T(r == k ? 1 : 0);
    lu_solve(lu, e.data(), col.data());
    for (int r = 0; r < nv; ++r) Minv[size_t(r) * nv + k] = col[r];
  }
  std::vector<T> JM(size_t(m) * nv, T(0)), G(size_t(m) * m, T(0)), g(m, T(0));
  for (int r = 0; r < m; ++r)
    for (int k = 0; k < nv; ++k) {
      T s = T(0);
      for (int l = 0; l < nv; ++l) s = s + J[size_t(r) * nv + l] * Minv[size_t(l) * nv + k];
      JM[size_t(r) * nv + k] = s;
    }
  for (int r = 0; r < m; ++r) {
    for (int k = 0; k < m; ++k) {
      T s = T(0);
      for (int l = 0; l < nv; ++l) s = s + JM[size_t(r) * nv + l] * J[size_t(k) * nv + l];
      G[size_t(r) * m + k] = s;
    }
    T s = T(0);
    for (int l = 0; l < nv; ++l) s = s + J[size_t(r) * nv + l] * vfree[l];
    g[r] = s;
  }
  std::vector<T> lam;
  if (cone_program(G, g, nc, lam, cone_iters)) return 3;
  std::vector<T> imp(nv, T(0)), dv(nv);
  for (int k = 0; k < nv; ++k)
    for (int r = 0; r < m; ++r) imp[k] = imp[k] + J[size_t(r) * nv + k] * lam[r];
  lu_solve(lu, imp.data(), dv.data());
  for (int k = 0; k < nv; ++k) v_next[k] = vfree[k] + dv[k];
  return 0;
}

// velocity_stepping / ccd_velocity_stepping (integrators.rs:42-175). The
// continuous-collision loop cuts the step at the earliest time a separated
// contact point, or a sphere whose body has no contact, reaches a half-space
// with the new velocities; a sphere that ended a cut is in contact with that
// half-space in the next round; the loop gives up (returns 4) after
// kCcdMaxRounds rounds, where the reference would keep looping.
constexpr int kCcdMaxRounds = 64;

template <class T>
int vs_step(const Model<T>& M, bool ccd, const T* q, const T* v, const T* tau, T dt, T* q_out, T* v_out,
            int* rounds_out = nullptr) {
  const int nq = M.nq, nv = M.nv, n = M.n;
  std::vector<T> qc(q, q + nq), vc(v, v + nv), vn(nv), qn(nq);
  std::vector<VsContact<T>> contacts;
  T t_remain = dt;
  int forced_body = -1, forced_h = -1;
  for (int round = 0; round < kCcdMaxRounds; ++round) {
    if (rounds_out) *rounds_out = round + 1;
    int st = vs_dynamics(M, qc.data(), vc.data(), tau, t_remain, vn.data(), contacts, nullptr, forced_body, forced_h);
    if (st) return st;
    T earliest = T(1e300);
    int hit_body = -1, hit_h = -1;
    if (ccd) {
      std::vector<Iso<T>> b2r = bodies_to_root(M, qc.data());
      std::vector<SV<T>> tw(n);
      for (int i = 0; i < n; ++i) {  // compute_twists_wrt_world with the new joint twists
        const T* vi = vn.data() + M.voff[i];
        SV<T> j;
        switch (M.jtype[i]) {
          case GP_JOINT_REVOLUTE: j.a = M.axis[i] * vi[0]; break;
          case GP_JOINT_PRISMATIC: j.l = M.axis[i] * vi[0]; break;
          case GP_JOINT_FLOATING: j.a = V3<T>(vi[0], vi[1], vi[2]); j.l = V3<T>(vi[3], vi[4], vi[5]); break;
          default: break;
        }
        SV<T> parent_tw = M.parent[i] < 0 ? SV<T>() : tw[M.parent[i]];
        tw[i] = parent_tw + twist_transform(j, b2r[i]);
      }
      auto hit = [&](const V3<T>& p, const V3<T>& vel, T radius, int sphere_body) {
        for (size_t hh = 0; hh < M.halfspaces.size(); ++hh) {
          const HalfSpace<T>& h = M.halfspaces[hh];
          T vdn = dot(vel, h.normal);
          if (vdn == T(0)) continue;
          T dist = dot(p - h.point, h.normal) - radius;
          if (dist <= T(0)) continue;
          T t = dist / -vdn;
          if (t < earliest && t > T(0)) {
            earliest = t;
            hit_body = sphere_body;
            hit_h = int(hh);
          }
        }
      };
      for (int i = 0; i < n; ++i) {
        M3<T> R = qmat(b2r[i].r);
        for (const Contact<T>& cp : M.contacts[i]) {
          V3<T> pw = R * cp.p + b2r[i].t;
          hit(pw, tw[i].l + cross(tw[i].a, pw), T(0), -1);
        }
      }
      for (int i = 0; i < n; ++i) {
        const SphereC<T>& S = M.spheres[i];
        if (!S.present || !S.enabled) continue;
        V3<T> center = b2r[i].t + qrot(b2r[i].r, S.c);
        bool touching = false;
        for (const VsContact<T>& c : contacts) touching = touching || c.body == i;
        if (touching) continue;
        hit(center, tw[i].l + cross(tw[i].a, center), S.r, i);
      }
    }
    if (earliest < t_remain) {
      advance_q(M, qc.data(), vn.data(), earliest, qn.data());
      qc = qn;
      vc = vn;
      t_remain = t_remain - earliest;
      forced_body = hit_body;
      forced_h = hit_body >= 0 ? hit_h : -1;
    } else {
      advance_q(M, qc.data(), vn.data(), t_remain, q_out);
      for (int k = 0; k < nv; ++k) v_out[k] = vn[k];
      return 0;
    }
  }
  return 4;
}

// ---------------------------------------------------------------- energy (metric)
// KE (inertia.rs:182-202) and potential energy from the centre-of-mass height
// (sum m g z_com), not the frame-origin height of mechanism.rs:352-362.
template <class T>
void energy(const Model<T>& M, const T* q, const T* v, T* ke, T* pe) {
  const int n = M.n;
  std::vector<Iso<T>> b2r(n);
  std::vector<SV<T>> tw(n);
  T K = T(0), P = T(0);
  for (int i = 0; i < n; ++i) {
    Iso<T> XJ = joint_transform(M, i, q);
    b2r[i] = M.parent[i] < 0 ? XJ : imul(b2r[M.parent[i]], XJ);
    const T* vi = v + M.voff[i];
    SV<T> j;
    switch (M.jtype[i]) {
      case GP_JOINT_REVOLUTE: j.a = M.axis[i] * vi[0]; break;
      case GP_JOINT_PRISMATIC: j.l = M.axis[i] * vi[0]; break;
      case GP_JOINT_FLOATING: j.a = V3<T>(vi[0], vi[1], vi[2]); j.l = V3<T>(vi[3], vi[4], vi[5]); break;
      default: break;
    }
    SV<T> pt = M.parent[i] < 0 ? SV<T>() : tw[M.parent[i]];
    tw[i] = pt + twist_transform(j, b2r[i]);
    Inertia<T> Iw = inertia_transform(M.I[i], b2r[i]);
    const V3<T>& w = tw[i].a;
    const V3<T>& vl = tw[i].l;
    K = K + (dot(w, Iw.J * w) + dot(vl, Iw.m * vl + T(2) * cross(w, Iw.c))) / T(2);
    P = P + M.g * Iw.c.z;
  }
  *ke = K;
  *pe = P;
}

}  // namespace oracle

namespace oracle {

// ---------------------------------------------------------------- ABA (body frames)
// Rotation A with A e_z = a for a unit axis a. Signed coordinate axes give an
// exact signed permutation; any other axis a Gram-Schmidt frame.
template <class T>
M3<T> align_z(const V3<T>& a) {
  M3<T> A;
  auto setcols = [&](const V3<T>& b1, const V3<T>& b2, const V3<T>& b3) {
    for (int r = 0; r < 3; ++r) { A.m[r][0] = b1[r]; A.m[r][1] = b2[r]; A.m[r][2] = b3[r]; }
  };
  const T one(1), zero(0);
  // exact cases: a = +-e_x, +-e_y, +-e_z
  int nz = (a.x != zero) + (a.y != zero) + (a.z != zero);
  if (nz == 1) {
    if (a.z != zero) {
      T s = a.z > zero ? one : -one;
      setcols(V3<T>(one, zero, zero), V3<T>(zero, s, zero), V3<T>(zero, zero, s));
    } else if (a.x != zero) {
      T s = a.x > zero ? one : -one;
      setcols(V3<T>(zero, s, zero), V3<T>(zero, zero, one), V3<T>(s, zero, zero));
    } else {
      T s = a.y > zero ? one : -one;
      setcols(V3<T>(zero, zero, s), V3<T>(one, zero, zero), V3<T>(zero, s, zero));
    }
    return A;
  }
  using std::fabs; using std::sqrt;
  V3<T> t = fabs(a.x) < T(0.9) ? V3<T>(one, zero, zero) : V3<T>(zero, one, zero);
  V3<T> b1 = t - dot(t, a) * a;
  b1 = b1 / sqrt(dot(b1, b1));
  V3<T> b2 = cross(a, b1);
  setcols(b1, b2, a);
  return A;
}

// Joint kinds after alignment: 1-DOF joints move along/about the local z axis.
enum { K_FIXED = 0, K_REV = 1, K_PRI = 2, K_FLOAT = 3 };

template <class T>
struct AbaModel {
  int n = 0, nq = 0, nv = 0;
  std::vector<int> parent, kind, qoff, voff;
  std::vector<M3<T>> Rt;         // tree rotation, child-internal -> parent-internal
  std::vector<V3<T>> pt;         // tree translation in parent-internal coords
  std::vector<M3<T>> J;          // moment about origin, internal coords
  std::vector<V3<T>> c;          // m * com, internal coords
  std::vector<T> m, sk, sl;      // mass, spring k, rest length
  std::vector<std::vector<Contact<T>>> contacts;  // internal coords
  std::vector<HalfSpace<T>> halfspaces;
  T g = T(9.81);
};

template <class T>
AbaModel<T> make_aba_model(const Model<T>& M) {
  AbaModel<T> A;
  A.n = M.n; A.nq = M.nq; A.nv = M.nv; A.g = M.g;
  A.parent = M.parent; A.qoff = M.qoff; A.voff = M.voff;
  A.kind.resize(M.n); A.Rt.resize(M.n); A.pt.resize(M.n); A.J.resize(M.n); A.c.resize(M.n);
  A.m.resize(M.n); A.sk.resize(M.n); A.sl.resize(M.n); A.contacts.resize(M.n);
  std::vector<M3<T>> Al(M.n);
  for (int i = 0; i < M.n; ++i) {
    int jt = M.jtype[i];
    A.kind[i] = jt == GP_JOINT_REVOLUTE ? K_REV : jt == GP_JOINT_PRISMATIC ? K_PRI : jt == GP_JOINT_FLOATING ? K_FLOAT : K_FIXED;
    Al[i] = (A.kind[i] == K_REV || A.kind[i] == K_PRI) ? align_z(M.axis[i]) : M3<T>::identity();
    M3<T> Ap = M.parent[i] < 0 ? M3<T>::identity() : Al[M.parent[i]];
    M3<T> Rtree = qmat(M.tree[i].r);
    A.Rt[i] = Ap.transpose() * Rtree * Al[i];
    A.pt[i] = Ap.transpose() * M.tree[i].t;
    A.J[i] = Al[i].transpose() * M.I[i].J * Al[i];
    A.c[i] = Al[i].transpose() * M.I[i].c;
    A.m[i] = M.I[i].m;
    A.sk[i] = M.spring_k[i]; A.sl[i] = M.spring_l[i];
    for (const Contact<T>& cp : M.contacts[i]) A.contacts[i].push_back({Al[i].transpose() * cp.p, cp.k});
  }
  A.halfspaces = M.halfspaces;
  return A;
}

// Symmetric 6x6 stored dense (row-major) for clarity of the restatement.
template <class T> struct S6 { T a[6][6]; };
template <class T> struct V6 { T a[6]; };

template <class T> V6<T> sv_to6(const SV<T>& s) { return {{s.a.x, s.a.y, s.a.z, s.l.x, s.l.y, s.l.z}}; }
template <class T> SV<T> v6_to_sv(const V6<T>& s) { return {V3<T>(s.a[0], s.a[1], s.a[2]), V3<T>(s.a[3], s.a[4], s.a[5])}; }

// Plucker transform parent -> child given E (parent->child rotation) and r
// (child origin in parent coords): motion vectors.
template <class T> SV<T> xmotion(const M3<T>& E, const V3<T>& r, const SV<T>& m) {
  V3<T> w = E * m.a;
  V3<T> v = E * (m.l - cross(r, m.a));
  return {w, v};
}
// X^T f: child force -> parent force
template <class T> SV<T> xtforce(const M3<T>& E, const V3<T>& r, const SV<T>& f) {
  M3<T> Et = E.transpose();
  V3<T> fl = Et * f.l;
  V3<T> fa = Et * f.a + cross(r, fl);
  return {fa, fl};
}
// X^T I X for a symmetric 6x6 I = [[A, B], [B^T, C]] in child coords, to
// parent coords, in block form. With A' = E^T A E, B' = E^T B E, C' = E^T C E
// and rx = skew(r):  TR = B' + rx C',  TL = A' + rx B'^T - TR rx,  BR = C'.
template <class T> M3<T> rot_congruence(const M3<T>& E, const M3<T>& X) {  // E^T X E
  return E.transpose() * (X * E);
}
template <class T> S6<T> xtix(const M3<T>& E, const V3<T>& r, const S6<T>& I) {
  M3<T> A, B, C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) { A.m[i][j] = I.a[i][j]; B.m[i][j] = I.a[i][j + 3]; C.m[i][j] = I.a[i + 3][j + 3]; }
  M3<T> A1 = rot_congruence(E, A), B1 = rot_congruence(E, B), C1 = rot_congruence(E, C);
  M3<T> TR, rxC, TL;
  for (int j = 0; j < 3; ++j) {  // rx C' column by column
    V3<T> col(C1.m[0][j], C1.m[1][j], C1.m[2][j]);
    V3<T> x = cross(r, col);
    for (int i = 0; i < 3; ++i) rxC.m[i][j] = x[i];
  }
  TR = B1 + rxC;
  for (int j = 0; j < 3; ++j) {  // rx B'^T: columns r x (row j of B')
    V3<T> row(B1.m[j][0], B1.m[j][1], B1.m[j][2]);
    V3<T> x = cross(r, row);
    for (int i = 0; i < 3; ++i) TL.m[i][j] = A1.m[i][j] + x[i];
  }
  for (int i = 0; i < 3; ++i) {  // - TR rx: row i of TR rx = (row i of TR) x r
    V3<T> row(TR.m[i][0], TR.m[i][1], TR.m[i][2]);
    V3<T> x = cross(row, r);
    for (int j = 0; j < 3; ++j) TL.m[i][j] = TL.m[i][j] - x[j];
  }
  S6<T> out;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      out.a[i][j] = TL.m[i][j];
      out.a[i][j + 3] = TR.m[i][j];
      out.a[j + 3][i] = TR.m[i][j];
      out.a[i + 3][j + 3] = C1.m[i][j];
    }
  return out;
}

// Cholesky factor of a 6x6 SPD system (the floating joint's D); false if not SPD.
template <class T> bool chol6_factor(const S6<T>& A, T (&L)[6][6]) {
  using std::sqrt;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) L[i][j] = T(0);
  for (int j = 0; j < 6; ++j) {
    T s = A.a[j][j];
    for (int k = 0; k < j; ++k) s = s - L[j][k] * L[j][k];
    if (!(s > T(0))) return false;
    L[j][j] = sqrt(s);
    for (int i = j + 1; i < 6; ++i) {
      T t = A.a[i][j];
      for (int k = 0; k < j; ++k) t = t - L[i][k] * L[j][k];
      L[i][j] = t / L[j][j];
    }
  }
  return true;
}
// Cholesky solve of a 6x6 SPD system (the floating joint's D)
template <class T> int chol6_solve(const S6<T>& A, const V6<T>& b, V6<T>& x) {
  T L[6][6];
  if (!chol6_factor(A, L)) return 1;
  T y[6];
  for (int i = 0; i < 6; ++i) {
    T s = b.a[i];
    for (int k = 0; k < i; ++k) s = s - L[i][k] * y[k];
    y[i] = s / L[i][i];
  }
  for (int i = 5; i >= 0; --i) {
    T s = y[i];
    for (int k = i + 1; k < 6; ++k) s = s - L[k][i] * x.a[k];
    x.a[i] = s / L[i][i];
  }
  return 0;
}

// Featherstone ABA (RBDA Table 7.1) over the aligned model. Returns 0, or 1
// if a joint-space inertia D is not positive.
template <class T>
int aba_dynamics(const AbaModel<T>& A, const T* q, const T* v, const T* tau, T* vdot) {
  using std::sin; using std::cos;
  const int n = A.n;
  std::vector<M3<T>> E(n), Rw(n);
  std::vector<V3<T>> r(n), tw(n);
  std::vector<SV<T>> vel(n), cb(n), pA(n), aa(n);
  std::vector<S6<T>> IA(n);
  std::vector<S6<T>> U(n);
  std::vector<S6<T>> Dm(n);
  std::vector<V6<T>> uu(n);
  bool any_contact = false;
  for (const auto& cs : A.contacts) any_contact = any_contact || !cs.empty();
  // pass 1: kinematics, velocities, bias terms
  for (int i = 0; i < n; ++i) {
    const T* qi = q + A.qoff[i];
    const T* vi = v + A.voff[i];
    M3<T> R = A.Rt[i];
    V3<T> p = A.pt[i];
    SV<T> vj;
    switch (A.kind[i]) {
      case K_REV: {
        T cq = cos(qi[0]), sq = sin(qi[0]);
        M3<T> Rz; Rz.m[0][0] = cq; Rz.m[0][1] = -sq; Rz.m[1][0] = sq; Rz.m[1][1] = cq; Rz.m[2][2] = T(1);
        R = R * Rz;
        vj.a.z = vi[0];
        break;
      }
      case K_PRI:
        p = p + V3<T>(A.Rt[i].m[0][2], A.Rt[i].m[1][2], A.Rt[i].m[2][2]) * qi[0];
        vj.l.z = vi[0];
        break;
      case K_FLOAT: {
        F64Scope f64;  // the mixed-precision kernels keep the floating pose in FP64
        Q<T> qq(qi[0], qi[1], qi[2], qi[3]);
        M3<T> Rq = qmat(qq);
        p = p + A.Rt[i] * V3<T>(qi[4], qi[5], qi[6]);
        R = R * Rq;
        vj.a = V3<T>(vi[0], vi[1], vi[2]);
        vj.l = V3<T>(vi[3], vi[4], vi[5]);
        break;
      }
      default: break;
    }
    E[i] = R.transpose();
    r[i] = p;
    int pa = A.parent[i];
    SV<T> vp = pa < 0 ? SV<T>() : vel[pa];
    vel[i] = xmotion(E[i], r[i], vp) + vj;
    // c = v x_m vj (se3_commutator)
    cb[i].a = cross(vel[i].a, vj.a);
    cb[i].l = cross(vel[i].a, vj.l) + cross(vel[i].l, vj.a);
    // world pose (for contacts)
    if (any_contact) {
      F64Scope f64;  // contact distances are FP64 in every kernel
      if (pa < 0) { Rw[i] = R; tw[i] = p; }
      else { Rw[i] = Rw[pa] * R; tw[i] = tw[pa] + Rw[pa] * p; }
    }
    // rigid-body inertia and bias force v x* (I v) - f_ext
    S6<T>& I6 = IA[i];
    M3<T> cx = skew(A.c[i]);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        I6.a[a][b] = A.J[i].m[a][b];
        I6.a[a][b + 3] = cx.m[a][b];
        I6.a[a + 3][b] = cx.m[b][a];
        I6.a[a + 3][b + 3] = a == b ? A.m[i] : T(0);
      }
    Inertia<T> Ii{A.J[i], A.c[i], A.m[i]};
    SV<T> h = mul_inertia(Ii, vel[i]);
    SV<T> f;
    f.a = cross(vel[i].a, h.a) + cross(vel[i].l, h.l);
    f.l = cross(vel[i].a, h.l);
    // contacts: world-frame force from the half-spaces, applied in body coords
    for (const Contact<T>& cp : A.contacts[i]) {
      V3<T> pw;
      {
        F64Scope f64;
        pw = Rw[i] * cp.p + tw[i];
      }
      V3<T> vb = vel[i].l + cross(vel[i].a, cp.p);
      V3<T> vw = Rw[i] * vb;
      for (const HalfSpace<T>& hs : A.halfspaces) {
        T d;
        {
          F64Scope f64;
          d = dot(pw - hs.point, hs.normal);
        }
        if (!(d <= T(kInsideMargin))) continue;
        V3<T> Fw = contact_force(-d, hs.normal, vw, cp.k, T(kHalfspaceK), hs.alpha, hs.mu);
        V3<T> Fb = Rw[i].transpose() * Fw;
        f.a -= cross(cp.p, Fb);
        f.l -= Fb;
      }
    }
    pA[i] = f;
  }
  // pass 2: articulated inertias
  for (int i = n - 1; i >= 0; --i) {
    const S6<T>& I6 = IA[i];
    V6<T> p6 = sv_to6(pA[i]);
    V6<T> c6 = sv_to6(cb[i]);
    S6<T> Ia;
    V6<T> pa6;
    int k = A.kind[i];
    if (k == K_REV || k == K_PRI) {
      int s = k == K_REV ? 2 : 5;
      V6<T> Ui;
      for (int a = 0; a < 6; ++a) Ui.a[a] = I6.a[a][s];
      T D = Ui.a[s];
      if (!(D > T(0))) return 1;
      T tq = tau[A.voff[i]], u;
      {
        F64Scope f64;  // u and the joint-space inertia's solve are FP64 in every kernel
        if (k == K_PRI && !(A.sk[i] == T(0)))
          tq = tq + (-A.sk[i] * (q[A.qoff[i]] - A.sl[i]));
        u = tq - p6.a[s];
      }
      U[i].a[0][0] = D;  // keep D, U and u for pass 3
      for (int a = 0; a < 6; ++a) U[i].a[1][a] = Ui.a[a];
      uu[i].a[0] = u;
      V6<T> UD;
      for (int a = 0; a < 6; ++a) UD.a[a] = Ui.a[a] / D;
      for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) { Ia.a[a][b] = I6.a[a][b] - UD.a[a] * Ui.a[b]; Ia.a[b][a] = Ia.a[a][b]; }
      // pa = pA + Ia c + U u / D
      for (int a = 0; a < 6; ++a) {
        T s2(0);
        for (int b = 0; b < 6; ++b) s2 = s2 + Ia.a[a][b] * c6.a[b];
        pa6.a[a] = p6.a[a] + s2 + UD.a[a] * u;
      }
    } else if (k == K_FLOAT) {
      Dm[i] = I6;
      F64Scope f64;
      for (int a = 0; a < 6; ++a) uu[i].a[a] = tau[A.voff[i] + a] - p6.a[a];
      // Ia = IA - IA IA^-1 IA = 0; pa = pA + 0 c + IA IA^-1 u = pA + u = tau
      for (int a = 0; a < 6; ++a) { for (int b = 0; b < 6; ++b) Ia.a[a][b] = T(0); pa6.a[a] = p6.a[a] + uu[i].a[a]; }
    } else {
      Ia = I6;
      for (int a = 0; a < 6; ++a) {
        T s2(0);
        for (int b = 0; b < 6; ++b) s2 = s2 + I6.a[a][b] * c6.a[b];
        pa6.a[a] = p6.a[a] + s2;
      }
    }
    int pr = A.parent[i];
    if (pr >= 0) {
      S6<T> add = xtix(E[i], r[i], Ia);
      for (int a = 0; a < 6; ++a) for (int b = 0; b < 6; ++b) IA[pr].a[a][b] = IA[pr].a[a][b] + add.a[a][b];
      SV<T> fp = xtforce(E[i], r[i], v6_to_sv(pa6));
      pA[pr] = pA[pr] + fp;
    }
  }
  // pass 3: accelerations
  SV<T> a0;
  a0.l = V3<T>(T(0), T(0), A.g);
  for (int i = 0; i < n; ++i) {
    int pr = A.parent[i];
    SV<T> ap = xmotion(E[i], r[i], pr < 0 ? a0 : aa[pr]) + cb[i];
    int k = A.kind[i];
    if (k == K_REV || k == K_PRI) {
      int s = k == K_REV ? 2 : 5;
      V6<T> a6 = sv_to6(ap);
      T Ua(0);
      for (int a = 0; a < 6; ++a) Ua = Ua + U[i].a[1][a] * a6.a[a];
      T qdd;
      {
        F64Scope f64;
        qdd = (uu[i].a[0] - Ua) / U[i].a[0][0];
      }
      vdot[A.voff[i]] = qdd;
      a6.a[s] = a6.a[s] + qdd;
      aa[i] = v6_to_sv(a6);
    } else if (k == K_FLOAT) {
      F64Scope f64;  // the floating joint's 6x6 solve
      V6<T> a6 = sv_to6(ap);
      V6<T> rhs;
      for (int a = 0; a < 6; ++a) {
        T s2(0);
        for (int b = 0; b < 6; ++b) s2 = s2 + Dm[i].a[a][b] * a6.a[b];
        rhs.a[a] = uu[i].a[a] - s2;
      }
      V6<T> x;
      if (chol6_solve(Dm[i], rhs, x)) return 1;
      for (int a = 0; a < 6; ++a) { vdot[A.voff[i] + a] = x.a[a]; a6.a[a] = a6.a[a] + x.a[a]; }
      aa[i] = v6_to_sv(a6);
    } else {
      aa[i] = ap;
    }
  }
  return 0;
}

// Scalar that counts arithmetic operations (algorithmic flop count per step).
// An operation with a structural zero or one of the formulation (a literal
// T(0) / T(1): the zero blocks of a rigid-body inertia, the identity of a
// motion subspace, a velocity's unused components) is not counted, since an
// implementation specialised to the formulation's structure never performs
// it; values loaded from the model or state (OpCounter::loading) count even
// when they happen to be 0 or 1. Operations run in FP64 by the mixed-precision
// kernels (F64Scope) are counted apart from the rest.
struct Cnt {
  double v;
  signed char s;  // 0: a value; 1: structural zero; 2: structural one
  Cnt() : v(0), s(1) {}
  Cnt(double x) : v(x), s(OpCounter::loading ? 0 : (x == 0.0 ? 1 : (x == 1.0 ? 2 : 0))) {}
  static Cnt val(double x) { Cnt c; c.v = x; c.s = 0; return c; }
  explicit operator double() const { return v; }
  bool zero() const { return s == 1; }
  bool one() const { return s == 2; }
};
inline Cnt operator+(Cnt a, Cnt b) {
  if (a.zero()) return b;
  if (b.zero()) return a;
  OpCounter::count(OpCounter::add, OpCounter::add64);
  return Cnt::val(a.v + b.v);
}
inline Cnt operator-(Cnt a) { Cnt r = a; r.v = -a.v; if (a.one()) r.s = 0; return r; }
inline Cnt operator-(Cnt a, Cnt b) {
  if (b.zero()) return a;
  if (a.zero()) return -b;
  OpCounter::count(OpCounter::add, OpCounter::add64);
  return Cnt::val(a.v - b.v);
}
inline Cnt operator*(Cnt a, Cnt b) {
  if (a.zero() || b.zero()) return Cnt();
  if (a.one()) return b;
  if (b.one()) return a;
  OpCounter::count(OpCounter::mul, OpCounter::mul64);
  return Cnt::val(a.v * b.v);
}
inline Cnt operator/(Cnt a, Cnt b) {
  if (a.zero()) return Cnt();
  if (b.one()) return a;
  OpCounter::count(OpCounter::div, OpCounter::div64);
  return Cnt::val(a.v / b.v);
}
inline Cnt& operator+=(Cnt& a, Cnt b) { return a = a + b; }
inline Cnt& operator-=(Cnt& a, Cnt b) { return a = a - b; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { return a.v <= b.v; }
inline bool operator==(Cnt a, Cnt b) { return a.v == b.v; }
inline bool operator!=(Cnt a, Cnt b) { return a.v != b.v; }
inline Cnt sqrt(Cnt a) { OpCounter::count(OpCounter::sqrt_, OpCounter::sqrt64); return Cnt::val(std::sqrt(a.v)); }
inline Cnt sin(Cnt a) { OpCounter::count(OpCounter::trig, OpCounter::trig64); return Cnt::val(std::sin(a.v)); }
inline Cnt cos(Cnt a) { OpCounter::count(OpCounter::trig, OpCounter::trig64); return Cnt::val(std::cos(a.v)); }
inline Cnt fabs(Cnt a) { Cnt r = a; r.v = std::fabs(a.v); return r; }
inline Cnt fmax(Cnt a, Cnt b) { return Cnt::val(std::fmax(a.v, b.v)); }
inline Cnt pow(Cnt a, Cnt b) { OpCounter::count(OpCounter::pow_, OpCounter::pow64); return Cnt::val(std::pow(a.v, b.v)); }

}  // namespace oracle
