// oracle.cpp — TEST INFRASTRUCTURE ONLY: C entry points of the CPU oracle
// (oracle.hpp) for the pytest parity suite, smoke() and bench.py's cpu_baseline.
// The product library never links or loads this file.
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <thread>

#include "oracle.hpp"

using namespace oracle;

namespace {

enum { M_CRBA = 0, M_ABA = 1, M_CRBA_LD = 2, M_ABA_LD = 3 };

template <class T>
int dyn_any(const Model<T>& M, const AbaModel<T>* A, int method, const T* q, const T* v, const T* tau, T* vdot) {
  if (method == M_ABA || method == M_ABA_LD) return aba_dynamics(*A, q, v, tau, vdot);
  return crba_dynamics(M, q, v, tau, vdot);
}

template <class T>
int step_t(const gp_model_desc* d, int integrator, int method, const double* q, const double* v,
           const double* tau, double dt, double* q_out, double* v_out, double* vdot_out) {
  Model<T> M = load_model<T>(*d);
  AbaModel<T> A = make_aba_model(M);
  std::vector<T> qt(M.nq), vt(M.nv), tt(M.nv), qo(M.nq), vo(M.nv), ao(M.nv);
  for (int k = 0; k < M.nq; ++k) qt[k] = T(q[k]);
  for (int k = 0; k < M.nv; ++k) { vt[k] = T(v[k]); tt[k] = tau ? T(tau[k]) : T(0); }
  auto dyn = [&](const Model<T>& MM, const T* qq, const T* vv, const T* ta, T* out) {
    return dyn_any(MM, &A, method, qq, vv, ta, out);
  };
  int st = step_integrate(M, integrator, qt.data(), vt.data(), tt.data(), T(dt), qo.data(), vo.data(), ao.data(), dyn);
  for (int k = 0; k < M.nq; ++k) if (q_out) q_out[k] = double(qo[k]);
  for (int k = 0; k < M.nv; ++k) { if (v_out) v_out[k] = double(vo[k]); if (vdot_out) vdot_out[k] = double(ao[k]); }
  return st;
}

}  // namespace

extern "C" {

int oracle_dims(const gp_model_desc* d, int* nq, int* nv) {
  Model<double> M = load_model<double>(*d);
  *nq = M.nq; *nv = M.nv;
  return 0;
}

// dynamics_continuous (dynamics.rs:322). method: 0 CRBA f64, 1 ABA f64,
// 2 CRBA long double, 3 ABA long double.
int oracle_dynamics(const gp_model_desc* d, const double* q, const double* v, const double* tau,
                    double* vdot, int method) {
  if (method >= M_CRBA_LD) {
    Model<long double> M = load_model<long double>(*d);
    AbaModel<long double> A = make_aba_model(M);
    std::vector<long double> qt(M.nq), vt(M.nv), tt(M.nv), ao(M.nv);
    for (int k = 0; k < M.nq; ++k) qt[k] = q[k];
    for (int k = 0; k < M.nv; ++k) { vt[k] = v[k]; tt[k] = tau ? tau[k] : 0.0; }
    int st = dyn_any(M, &A, method, qt.data(), vt.data(), tt.data(), ao.data());
    for (int k = 0; k < M.nv; ++k) vdot[k] = double(ao[k]);
    return st;
  }
  Model<double> M = load_model<double>(*d);
  AbaModel<double> A = make_aba_model(M);
  std::vector<double> tt(M.nv, 0.0);
  if (tau) for (int k = 0; k < M.nv; ++k) tt[k] = tau[k];
  return dyn_any(M, &A, method, q, v, tt.data(), vdot);
}

// One semi-implicit Euler step (simulate.rs:20) of a model with spring
// contacts, CRBA, FP64. state: per spring contact in body order (description
// order within a body) 9 values [registered, point x y z, half-space, direction
// x y z (body frame), rest length], read and written (the reference mutates
// its spring contacts inside the dynamics, contact.rs:131-186). Returns 0, 1
// (singular mass matrix) or 2 (a spring pulling into its half-space).
int oracle_step_spring(const gp_model_desc* d, const double* q, const double* v, const double* tau, double dt,
                       double* q_out, double* v_out, double* vdot_out, double* state) {
  Model<double> M = load_model<double>(*d);
  std::vector<SpringSt<double>> ss;
  size_t k = 0;
  for (int i = 0; i < M.n; ++i)
    for (size_t s = 0; s < M.springs[i].size(); ++s, ++k) {
      const double* x = state + 9 * k;
      SpringSt<double> st;
      st.registered = x[0] != 0.0;
      st.point = V3<double>(x[1], x[2], x[3]);
      st.halfspace = int(x[4]);
      st.dir = V3<double>(x[5], x[6], x[7]);
      st.l_rest = x[8];
      ss.push_back(st);
    }
  std::vector<double> tt(M.nv, 0.0), qo(M.nq), vo(M.nv), ao(M.nv);
  if (tau) for (int j = 0; j < M.nv; ++j) tt[j] = tau[j];
  auto dyn = [&](const Model<double>& MM, const double* qq, const double* vv, const double* ta, double* out) {
    return crba_dynamics(MM, qq, vv, ta, out, (double*)nullptr, (double*)nullptr, ss.data());
  };
  int st = step_integrate(M, GP_SEMI_IMPLICIT_EULER, q, v, tt.data(), dt, qo.data(), vo.data(), ao.data(), dyn);
  for (int j = 0; j < M.nq; ++j) q_out[j] = qo[j];
  for (int j = 0; j < M.nv; ++j) { v_out[j] = vo[j]; if (vdot_out) vdot_out[j] = ao[j]; }
  for (size_t j = 0; j < ss.size(); ++j) {
    double* x = state + 9 * j;
    x[0] = ss[j].registered ? 1.0 : 0.0;
    x[1] = ss[j].point.x; x[2] = ss[j].point.y; x[3] = ss[j].point.z;
    x[4] = double(ss[j].halfspace);
    x[5] = ss[j].dir.x; x[6] = ss[j].dir.y; x[7] = ss[j].dir.z;
    x[8] = ss[j].l_rest;
  }
  return st;
}

// velocity_stepping / ccd_velocity_stepping (integrators.rs:42-175): one step.
// Returns 0, 1 (singular mass matrix), 3 (no finite impulse) or 4 (the
// continuous-collision loop did not finish); *rounds = the loop's rounds.
int oracle_vs_step(const gp_model_desc* d, int ccd, const double* q, const double* v, const double* tau, double dt,
                   double* q_out, double* v_out, int* rounds) {
  Model<double> M = load_model<double>(*d);
  std::vector<double> tt(M.nv, 0.0);
  if (tau) for (int j = 0; j < M.nv; ++j) tt[j] = tau[j];
  return vs_step(M, ccd != 0, q, v, tt.data(), dt, q_out, v_out, rounds);
}

// The contact impulse program alone (dynamics.rs:668-704): G (3nc x 3nc,
// row-major), g -> x; *iters = interior-point iterations.
int oracle_cone_program(int nc, const double* G, const double* g, double* x, int* iters) {
  const int m = 3 * nc;
  std::vector<double> Gv(G, G + size_t(m) * m), gv(g, g + m), xv;
  int st = cone_program(Gv, gv, nc, xv, iters);
  for (int k = 0; k < m; ++k) x[k] = xv[k];
  return st;
}

// mass matrix (mechanism.rs:637) and bias c(q,v) - tau_contact (dynamics.rs:233)
int oracle_mass_matrix(const gp_model_desc* d, const double* q, const double* v, double* Mout, double* cout) {
  Model<double> M = load_model<double>(*d);
  std::vector<double> tau(M.nv, 0.0), vd(M.nv);
  return crba_dynamics(M, q, v, tau.data(), vd.data(), Mout, cout);
}

// one step() (simulate.rs:20) of the given integrator
int oracle_step(const gp_model_desc* d, int integrator, int method, const double* q, const double* v,
                const double* tau, double dt, double* q_out, double* v_out, double* vdot_out) {
  if (method >= M_CRBA_LD) return step_t<long double>(d, integrator, method, q, v, tau, dt, q_out, v_out, vdot_out);
  return step_t<double>(d, integrator, method, q, v, tau, dt, q_out, v_out, vdot_out);
}

// In-place rollout of n_inst instances (instance i uses descs[i]) over n_steps
// steps with held torques, on n_threads host threads (one instance per thread
// at a time). State arrays are flattened instance-major; vdot_last (may be
// null) receives the accelerations of the last step. Returns the number of
// instances whose solve failed.
int oracle_rollout(const gp_model_desc* const* descs, int64_t n_inst, double* q, double* v,
                   const double* tau, double dt, int integrator, int method, int64_t n_steps,
                   int n_threads, double* vdot_last) {
  struct Prep { Model<double> M; AbaModel<double> A; };
  std::map<const gp_model_desc*, int> idx;
  std::vector<Prep> preps;
  std::vector<int> which(n_inst);
  std::vector<int64_t> qo(n_inst + 1, 0), vo(n_inst + 1, 0);
  for (int64_t i = 0; i < n_inst; ++i) {
    auto it = idx.find(descs[i]);
    int k;
    if (it == idx.end()) {
      Prep p; p.M = load_model<double>(*descs[i]); p.A = make_aba_model(p.M);
      k = int(preps.size()); preps.push_back(std::move(p)); idx[descs[i]] = k;
    } else k = it->second;
    which[i] = k;
    qo[i + 1] = qo[i] + preps[k].M.nq;
    vo[i + 1] = vo[i] + preps[k].M.nv;
  }
  std::atomic<int64_t> next(0);
  std::atomic<int> failures(0);
  auto worker = [&]() {
    std::vector<double> qn, vn, zero;
    for (;;) {
      int64_t i = next.fetch_add(1);
      if (i >= n_inst) break;
      const Prep& P = preps[which[i]];
      double* qi = q + qo[i];
      double* vi = v + vo[i];
      zero.assign(P.M.nv, 0.0);
      const double* ti = tau ? tau + vo[i] : zero.data();
      qn.resize(P.M.nq); vn.resize(P.M.nv);
      auto dyn = [&](const Model<double>& MM, const double* qq, const double* vv, const double* ta, double* out) {
        return dyn_any(MM, &P.A, method, qq, vv, ta, out);
      };
      for (int64_t s = 0; s < n_steps; ++s) {
        double* ad = (vdot_last && s == n_steps - 1) ? vdot_last + vo[i] : nullptr;
        const int st = integrator >= GP_VELOCITY_STEPPING
                           ? vs_step(P.M, integrator == GP_CCD_VELOCITY_STEPPING, qi, vi, ti, dt, qn.data(), vn.data())
                           : step_integrate(P.M, integrator, qi, vi, ti, dt, qn.data(), vn.data(), ad, dyn);
        if (st) {
          failures.fetch_add(1);
          break;
        }
        std::copy(qn.begin(), qn.end(), qi);
        std::copy(vn.begin(), vn.end(), vi);
      }
    }
  };
  n_threads = std::max(1, n_threads);
  std::vector<std::thread> pool;
  for (int t = 1; t < n_threads; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  return failures.load();
}

// simulate() (simulate.rs:87-112): float-accumulated loop `while t < final_time`,
// recording every state. Constant torque (NULL = zeros).
int oracle_simulate(const gp_model_desc* d, double final_time, double dt, int integrator, int method,
                    const double* q0, const double* v0, const double* tau, double* qs, double* vs,
                    int64_t max_records, int64_t* n_records) {
  Model<double> M = load_model<double>(*d);
  AbaModel<double> A = make_aba_model(M);
  std::vector<double> q(q0, q0 + M.nq), v(v0, v0 + M.nv), qn(M.nq), vn(M.nv), t0(M.nv, 0.0);
  const double* ta = tau ? tau : t0.data();
  auto dyn = [&](const Model<double>& MM, const double* qq, const double* vv, const double* tt, double* out) {
    return dyn_any(MM, &A, method, qq, vv, tt, out);
  };
  int64_t rec = 0;
  auto record = [&]() {
    if (rec < max_records) {
      if (qs) std::copy(q.begin(), q.end(), qs + rec * M.nq);
      if (vs) std::copy(v.begin(), v.end(), vs + rec * M.nv);
    }
    ++rec;
  };
  record();
  double t = 0.0;
  while (t < final_time) {
    int st = step_integrate(M, integrator, q.data(), v.data(), ta, dt, qn.data(), vn.data(), (double*)nullptr, dyn);
    if (st) { *n_records = rec; return st; }
    q = qn; v = vn;
    record();
    t += dt;
  }
  *n_records = rec;
  return 0;
}

int oracle_energy(const gp_model_desc* d, const double* q, const double* v, double* ke, double* pe) {
  Model<double> M = load_model<double>(*d);
  energy(M, q, v, ke, pe);
  return 0;
}

// MechanismState::gravitational_energy (mechanism.rs:352-362): m g z of each
// body frame's origin (not of the centre of mass).
int oracle_gravitational_energy(const gp_model_desc* d, const double* q, double* pe) {
  Model<double> M = load_model<double>(*d);
  std::vector<Iso<double>> b2r = bodies_to_root(M, q);
  double P = 0.0;
  for (int i = 0; i < M.n; ++i) P += M.I[i].m * M.g * b2r[i].t.z;
  *pe = P;
  return 0;
}

// signed distance (p - h.point) . n of every (contact point, half-space) pair,
// contact-major in the order of desc->contacts (halfspace.rs:46)
int oracle_contact_distances(const gp_model_desc* d, const double* q, double* dist) {
  Model<double> M = load_model<double>(*d);
  std::vector<Iso<double>> b2r(M.n);
  for (int i = 0; i < M.n; ++i) {
    Iso<double> XJ = joint_transform(M, i, q);
    b2r[i] = M.parent[i] < 0 ? XJ : imul(b2r[M.parent[i]], XJ);
  }
  for (int c = 0; c < d->n_contacts; ++c) {
    const gp_contact_desc& cd = d->contacts[c];
    M3<double> R = qmat(b2r[cd.body].r);
    V3<double> pw = R * V3<double>(cd.p[0], cd.p[1], cd.p[2]) + b2r[cd.body].t;
    for (int h = 0; h < d->n_halfspaces; ++h) {
      const HalfSpace<double>& hs = M.halfspaces[h];
      dist[c * d->n_halfspaces + h] = dot(pw - hs.point, hs.normal);
    }
  }
  return 0;
}

// Operation counts of one step of the ABA formulation with the given
// integrator at state (q, v): [add, mul, div, sqrt, trig, pow] of the
// operations run in FP32 by the mixed-precision kernels, then the same six of
// those run in FP64 (F64Scope: the floating pose, contact distances, u, the
// joint-space solves, and the integrator's update of q and v).
int oracle_opcount(const gp_model_desc* d, const double* q, const double* v, const double* tau,
                   int integrator, uint64_t* counts) {
  OpCounter::loading = true;  // model and state values are data, never structural zeros
  Model<Cnt> M = load_model<Cnt>(*d);
  std::vector<Cnt> qt(M.nq), vt(M.nv), tt(M.nv), qo(M.nq), vo(M.nv), ao(M.nv);
  for (int k = 0; k < M.nq; ++k) qt[k] = Cnt(q[k]);
  for (int k = 0; k < M.nv; ++k) { vt[k] = Cnt(v[k]); tt[k] = Cnt(tau ? tau[k] : 0.0); }
  const Cnt dt(0.001);
  OpCounter::loading = false;
  AbaModel<Cnt> A = make_aba_model(M);
  auto dyn = [&](const Model<Cnt>&, const Cnt* qq, const Cnt* vv, const Cnt* ta, Cnt* out) {
    --OpCounter::f64;  // the dynamics marks its own FP64 parts
    const int r = aba_dynamics(A, qq, vv, ta, out);
    ++OpCounter::f64;
    return r;
  };
  OpCounter::reset();
  int st;
  {
    F64Scope f64;  // the integrator's update of q and v
    st = step_integrate(M, integrator, qt.data(), vt.data(), tt.data(), dt, qo.data(), vo.data(), ao.data(), dyn);
  }
  if (M.n == 1 && A.kind[0] == K_FLOAT) {
    // A lone free body's D is its rigid-body inertia, the same at every step:
    // its factorisation is loop-invariant work, not part of a step.
    S6<Cnt> I6;
    M3<Cnt> cx = skew(A.c[0]);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        I6.a[a][b] = A.J[0].m[a][b];
        I6.a[a][b + 3] = cx.m[a][b];
        I6.a[a + 3][b] = cx.m[b][a];
        I6.a[a + 3][b + 3] = a == b ? A.m[0] : Cnt(0.0);
      }
    const uint64_t a0 = OpCounter::add64, m0 = OpCounter::mul64, d0 = OpCounter::div64, s0 = OpCounter::sqrt64;
    Cnt L[6][6];
    {
      F64Scope f64;
      chol6_factor(I6, L);
    }
    const uint64_t da = OpCounter::add64 - a0, dm = OpCounter::mul64 - m0, dd = OpCounter::div64 - d0,
                   ds = OpCounter::sqrt64 - s0;
    OpCounter::add64 = a0 - da; OpCounter::mul64 = m0 - dm; OpCounter::div64 = d0 - dd; OpCounter::sqrt64 = s0 - ds;
  }
  counts[0] = OpCounter::add; counts[1] = OpCounter::mul; counts[2] = OpCounter::div;
  counts[3] = OpCounter::sqrt_; counts[4] = OpCounter::trig; counts[5] = OpCounter::pow_;
  counts[6] = OpCounter::add64; counts[7] = OpCounter::mul64; counts[8] = OpCounter::div64;
  counts[9] = OpCounter::sqrt64; counts[10] = OpCounter::trig64; counts[11] = OpCounter::pow64;
  return st;
}

}  // extern "C"
