// gp_spatial.h — spatial algebra for the per-lane articulated-body kernels.
//
// Spatial vectors are [angular; linear] (the reference's order, twist.rs:23,
// wrench.rs:15). A Plucker transform parent -> child is held as (E, r):
// E rotates parent coordinates into child coordinates and r is the child
// origin in parent coordinates (for the successor->predecessor isometry
// x_p = R x_c + p of the reference, E = R^T and r = p).
// Everything is templated on the arithmetic type so one source serves the
// FP64 kernels and the FP32 sweeps of the mixed kernels.
#pragma once

#include <hip/hip_runtime.h>

#define GP_HD __host__ __device__ __forceinline__

namespace gp {

template <class R>
struct v3 {
  R x, y, z;
};
template <class R> GP_HD v3<R> mk(R a, R b, R c) { return v3<R>{a, b, c}; }
template <class R> GP_HD v3<R> operator+(v3<R> a, v3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> GP_HD v3<R> operator-(v3<R> a, v3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> GP_HD v3<R> operator-(v3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> GP_HD v3<R> operator*(R s, v3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> GP_HD R dot(v3<R> a, v3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> GP_HD v3<R> cross(v3<R> a, v3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class D, class S> GP_HD v3<D> cvt(v3<S> a) { return {D(a.x), D(a.y), D(a.z)}; }

// 3x3, row-major a[3*row + col]
template <class R>
struct m3 {
  R a[9];
};
template <class R> GP_HD v3<R> mul(const m3<R>& M, v3<R> v) {
  return {M.a[0] * v.x + M.a[1] * v.y + M.a[2] * v.z, M.a[3] * v.x + M.a[4] * v.y + M.a[5] * v.z,
          M.a[6] * v.x + M.a[7] * v.y + M.a[8] * v.z};
}
template <class R> GP_HD v3<R> mulT(const m3<R>& M, v3<R> v) {  // M^T v
  return {M.a[0] * v.x + M.a[3] * v.y + M.a[6] * v.z, M.a[1] * v.x + M.a[4] * v.y + M.a[7] * v.z,
          M.a[2] * v.x + M.a[5] * v.y + M.a[8] * v.z};
}
template <class R> GP_HD m3<R> mul(const m3<R>& A, const m3<R>& B) {
  m3<R> C;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C.a[3 * i + j] = A.a[3 * i] * B.a[j] + A.a[3 * i + 1] * B.a[3 + j] + A.a[3 * i + 2] * B.a[6 + j];
  return C;
}
template <class D, class S> GP_HD m3<D> cvt(const m3<S>& M) {
  m3<D> r;
#pragma unroll
  for (int k = 0; k < 9; ++k) r.a[k] = D(M.a[k]);
  return r;
}

// rotation matrix of a unit quaternion (x, y, z, w)
template <class R> GP_HD m3<R> quat_mat(R i, R j, R k, R w) {
  R ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  R ij = R(2) * i * j, wk = R(2) * w * k, wj = R(2) * w * j, ik = R(2) * i * k, jk = R(2) * j * k, wi = R(2) * w * i;
  return m3<R>{{ww + ii - jj - kk, ij - wk, wj + ik, wk + ij, ww - ii + jj - kk, jk - wi, ik - wj, wi + jk,
                ww - ii - jj + kk}};
}
// rotate v by unit quaternion (x, y, z, w): v + w t + u x t, t = 2 u x v
template <class R> GP_HD v3<R> quat_rot(R x, R y, R z, R w, v3<R> v) {
  v3<R> u{x, y, z};
  v3<R> t = R(2) * cross(u, v);
  return v + w * t + cross(u, t);
}

// spatial vector [w; v]
template <class R>
struct sv {
  v3<R> w, v;
};
template <class R> GP_HD sv<R> operator+(const sv<R>& a, const sv<R>& b) { return {a.w + b.w, a.v + b.v}; }
template <class R> GP_HD sv<R> operator-(const sv<R>& a, const sv<R>& b) { return {a.w - b.w, a.v - b.v}; }
template <class R> GP_HD R comp(const sv<R>& a, int k) {
  return k == 0 ? a.w.x : k == 1 ? a.w.y : k == 2 ? a.w.z : k == 3 ? a.v.x : k == 4 ? a.v.y : a.v.z;
}
template <class R> GP_HD R dot6(const sv<R>& a, const sv<R>& b) { return dot(a.w, b.w) + dot(a.v, b.v); }

// Plucker motion transform parent -> child
template <class R> GP_HD sv<R> xmotion(const m3<R>& E, v3<R> r, const sv<R>& m) {
  return {mul(E, m.w), mul(E, m.v - cross(r, m.w))};
}
// X^T f: force child -> parent
template <class R> GP_HD sv<R> xtforce(const m3<R>& E, v3<R> r, const sv<R>& f) {
  v3<R> fl = mulT(E, f.v);
  return {mulT(E, f.w) + cross(r, fl), fl};
}
// motion cross product m x s  (se3_commutator, util.rs:44)
template <class R> GP_HD sv<R> crm(const sv<R>& m, const sv<R>& s) {
  return {cross(m.w, s.w), cross(m.w, s.v) + cross(m.v, s.w)};
}
// force cross product m x* f  (dynamics.rs:90-91)
template <class R> GP_HD sv<R> crf(const sv<R>& m, const sv<R>& f) {
  return {cross(m.w, f.w) + cross(m.v, f.v), cross(m.w, f.v)};
}

// Symmetric 3x3 (xx, yy, zz, xy, xz, yz)
template <class R>
struct s3 {
  R xx, yy, zz, xy, xz, yz;
};
template <class R> GP_HD v3<R> mul(const s3<R>& S, v3<R> v) {
  return {S.xx * v.x + S.xy * v.y + S.xz * v.z, S.xy * v.x + S.yy * v.y + S.yz * v.z,
          S.xz * v.x + S.yz * v.y + S.zz * v.z};
}

// Articulated (or rigid) inertia [[A, B], [B^T, C]] with A, C symmetric.
template <class R>
struct ai6 {
  s3<R> A;
  m3<R> B;
  s3<R> C;
};
template <class R> GP_HD sv<R> mul(const ai6<R>& I, const sv<R>& m) {
  return {mul(I.A, m.w) + mul(I.B, m.v), mulT(I.B, m.w) + mul(I.C, m.v)};
}
template <class R> GP_HD void add_to(ai6<R>& a, const ai6<R>& b) {
  a.A.xx += b.A.xx; a.A.yy += b.A.yy; a.A.zz += b.A.zz; a.A.xy += b.A.xy; a.A.xz += b.A.xz; a.A.yz += b.A.yz;
#pragma unroll
  for (int k = 0; k < 9; ++k) a.B.a[k] += b.B.a[k];
  a.C.xx += b.C.xx; a.C.yy += b.C.yy; a.C.zz += b.C.zz; a.C.xy += b.C.xy; a.C.xz += b.C.xz; a.C.yz += b.C.yz;
}
// rigid-body inertia from (J about origin, c = m com, m): [[J, cx], [cx^T, m 1]]
template <class R> GP_HD ai6<R> rigid(const s3<R>& J, v3<R> c, R m) {
  ai6<R> I;
  I.A = J;
  I.B = m3<R>{{R(0), -c.z, c.y, c.z, R(0), -c.x, -c.y, c.x, R(0)}};
  I.C = s3<R>{m, m, m, R(0), R(0), R(0)};
  return I;
}
// column k of the 6x6 (as a spatial vector)
template <class R> GP_HD sv<R> col(const ai6<R>& I, int k) {
  if (k < 3) {
    v3<R> a = k == 0 ? v3<R>{I.A.xx, I.A.xy, I.A.xz} : k == 1 ? v3<R>{I.A.xy, I.A.yy, I.A.yz} : v3<R>{I.A.xz, I.A.yz, I.A.zz};
    v3<R> b{I.B.a[3 * k], I.B.a[3 * k + 1], I.B.a[3 * k + 2]};  // column k of B^T = row k of B
    return {a, b};
  }
  int j = k - 3;
  v3<R> a{I.B.a[j], I.B.a[3 + j], I.B.a[6 + j]};
  v3<R> c = j == 0 ? v3<R>{I.C.xx, I.C.xy, I.C.xz} : j == 1 ? v3<R>{I.C.xy, I.C.yy, I.C.yz} : v3<R>{I.C.xz, I.C.yz, I.C.zz};
  return {a, c};
}
// I - u u^T / D  given uD = u / D
template <class R> GP_HD ai6<R> rank1_down(const ai6<R>& I, const sv<R>& u, const sv<R>& uD) {
  ai6<R> o;
  o.A.xx = I.A.xx - uD.w.x * u.w.x; o.A.yy = I.A.yy - uD.w.y * u.w.y; o.A.zz = I.A.zz - uD.w.z * u.w.z;
  o.A.xy = I.A.xy - uD.w.x * u.w.y; o.A.xz = I.A.xz - uD.w.x * u.w.z; o.A.yz = I.A.yz - uD.w.y * u.w.z;
  const R uw[3] = {uD.w.x, uD.w.y, uD.w.z};
  const R uv[3] = {u.v.x, u.v.y, u.v.z};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) o.B.a[3 * i + j] = I.B.a[3 * i + j] - uw[i] * uv[j];
  o.C.xx = I.C.xx - uD.v.x * u.v.x; o.C.yy = I.C.yy - uD.v.y * u.v.y; o.C.zz = I.C.zz - uD.v.z * u.v.z;
  o.C.xy = I.C.xy - uD.v.x * u.v.y; o.C.xz = I.C.xz - uD.v.x * u.v.z; o.C.yz = I.C.yz - uD.v.y * u.v.z;
  return o;
}

// E^T S E for symmetric S
template <class R> GP_HD s3<R> congr(const m3<R>& E, const s3<R>& S) {
  // T = S E (3x3), then (E^T T) upper triangle
  m3<R> T;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    v3<R> c = mul(S, v3<R>{E.a[j], E.a[3 + j], E.a[6 + j]});
    T.a[j] = c.x; T.a[3 + j] = c.y; T.a[6 + j] = c.z;
  }
  auto e = [&](int i, int j) {  // (E^T T)_ij = sum_k E_ki T_kj
    return E.a[i] * T.a[j] + E.a[3 + i] * T.a[3 + j] + E.a[6 + i] * T.a[6 + j];
  };
  return s3<R>{e(0, 0), e(1, 1), e(2, 2), e(0, 1), e(0, 2), e(1, 2)};
}
// E^T B E for general B
template <class R> GP_HD m3<R> congr(const m3<R>& E, const m3<R>& B) {
  m3<R> T = mul(B, E);
  m3<R> O;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) O.a[3 * i + j] = E.a[i] * T.a[j] + E.a[3 + i] * T.a[3 + j] + E.a[6 + i] * T.a[6 + j];
  return O;
}
// X^T I X: child articulated inertia -> parent coordinates. With A' = E^T A E,
// B' = E^T B E, C' = E^T C E and rx = skew(r):
//   B_p = B' + rx C',  A_p = A' + rx B'^T - B_p rx,  C_p = C'.
template <class R> GP_HD ai6<R> xtix(const m3<R>& E, v3<R> r, const ai6<R>& I) {
  s3<R> A1 = congr(E, I.A);
  m3<R> B1 = congr(E, I.B);
  s3<R> C1 = congr(E, I.C);
  ai6<R> o;
  // rx C' columns: r x C'_col
  v3<R> c0 = cross(r, v3<R>{C1.xx, C1.xy, C1.xz});
  v3<R> c1 = cross(r, v3<R>{C1.xy, C1.yy, C1.yz});
  v3<R> c2 = cross(r, v3<R>{C1.xz, C1.yz, C1.zz});
  m3<R> Bp{{B1.a[0] + c0.x, B1.a[1] + c1.x, B1.a[2] + c2.x, B1.a[3] + c0.y, B1.a[4] + c1.y, B1.a[5] + c2.y,
            B1.a[6] + c0.z, B1.a[7] + c1.z, B1.a[8] + c2.z}};
  // rx B'^T: column j = r x (row j of B')
  v3<R> d0 = cross(r, v3<R>{B1.a[0], B1.a[1], B1.a[2]});
  v3<R> d1 = cross(r, v3<R>{B1.a[3], B1.a[4], B1.a[5]});
  v3<R> d2 = cross(r, v3<R>{B1.a[6], B1.a[7], B1.a[8]});
  // B_p rx: row i = (row i of B_p) x r
  v3<R> e0 = cross(v3<R>{Bp.a[0], Bp.a[1], Bp.a[2]}, r);
  v3<R> e1 = cross(v3<R>{Bp.a[3], Bp.a[4], Bp.a[5]}, r);
  v3<R> e2 = cross(v3<R>{Bp.a[6], Bp.a[7], Bp.a[8]}, r);
  // A_p(i,j) = A1(i,j) + d_j[i] - e_i[j]; symmetric, take upper triangle
  o.A.xx = A1.xx + d0.x - e0.x;
  o.A.yy = A1.yy + d1.y - e1.y;
  o.A.zz = A1.zz + d2.z - e2.z;
  o.A.xy = A1.xy + d1.x - e0.y;
  o.A.xz = A1.xz + d2.x - e0.z;
  o.A.yz = A1.yz + d2.y - e1.z;
  o.B = Bp;
  o.C = C1;
  return o;
}

}  // namespace gp
