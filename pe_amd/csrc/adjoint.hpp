// Adjoint of the time step (spheres and planes; relaxation models 0, 1 and 3): the per-contact map
// whose Jacobian the reverse pass needs, written once for any scalar type so that dual numbers
// give its exact directional derivatives (internal to the library).
//
// At a converged solve the impulses p and the post-solve body velocities U satisfy
//   E1:  U - u_pred - sum_c contrib_c(x, p_c) = 0                   (velocity update, HCSITS.h:2532-2535)
//   E2:  p_c - phi_c(gdot_c(x, U, p_c), p_c, geometry_c(x), mu) = 0 (one contact's NBGS update)
// where gdot_c is the contact's relative velocity with its own impulse removed
// (HCSITS.h:2343-2360).  contact_map() evaluates, for one contact, E2_c (3 values) and
// -contrib_c (12 values: dv1, dw1, dv2, dw2) from z = (x1, x2, U1, U2, p_c, mu).
#pragma once

#include "dmath.hpp"

namespace pe_amd {
namespace ad {

// Value and one directional derivative.
struct Dual {
  double v, d;
};
PE_HD Dual mkd(double v, double d = 0.0) { Dual r; r.v = v; r.d = d; return r; }
PE_HD Dual operator+(Dual a, Dual b) { return mkd(a.v + b.v, a.d + b.d); }
PE_HD Dual operator-(Dual a, Dual b) { return mkd(a.v - b.v, a.d - b.d); }
PE_HD Dual operator-(Dual a) { return mkd(-a.v, -a.d); }
PE_HD Dual operator*(Dual a, Dual b) { return mkd(a.v * b.v, a.d * b.v + a.v * b.d); }
PE_HD Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return mkd(q, (a.d - q * b.d) / b.v);
}
PE_HD Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return mkd(s, s > 0.0 ? 0.5 * a.d / s : 0.0);
}
PE_HD Dual cos(Dual a) { return mkd(::cos(a.v), -::sin(a.v) * a.d); }
PE_HD Dual sin(Dual a) { return mkd(::sin(a.v), ::cos(a.v) * a.d); }
PE_HD double val(double a) { return a; }
PE_HD double val(Dual a) { return a.v; }
PE_HD double der(double) { return 0.0; }
PE_HD double der(Dual a) { return a.d; }
// a value of type T with the given derivative (dropped for plain doubles)
template <class T>
PE_HD T with_d(double v, double d);
template <>
PE_HD double with_d<double>(double v, double) { return v; }
template <>
PE_HD Dual with_d<Dual>(double v, double d) { return mkd(v, d); }
// the value as a dual number with zero derivative
PE_HD Dual zd(double a) { return mkd(a); }
PE_HD Dual zd(Dual a) { return mkd(a.v); }
template <class T>
PE_HD T cst(double v) { return T(v); }
template <>
PE_HD Dual cst<Dual>(double v) { return mkd(v); }

template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
PE_HD V3<T> v3(T a, T b, T c) { V3<T> r; r.x = a; r.y = b; r.z = c; return r; }
template <class T>
PE_HD V3<T> operator+(V3<T> a, V3<T> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T>
PE_HD V3<T> operator-(V3<T> a, V3<T> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T>
PE_HD V3<T> scale(V3<T> a, T s) { return v3(a.x * s, a.y * s, a.z * s); }
template <class T>
PE_HD V3<T> scale(T s, V3<T> a) { return v3(s * a.x, s * a.y, s * a.z); }
template <class T>
PE_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
PE_HD V3<T> cross(V3<T> a, V3<T> b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T>
PE_HD V3<T> normalize(V3<T> a) {
  const T len = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  const T il = cst<T>(1.0) / len;
  return v3(a.x * il, a.y * il, a.z * il);
}

template <class T>
struct M3 {
  T a[9];
};
template <class T>
PE_HD V3<T> mv(const M3<T>& l, V3<T> r) {
  return v3(l.a[0] * r.x + l.a[1] * r.y + l.a[2] * r.z, l.a[3] * r.x + l.a[4] * r.y + l.a[5] * r.z,
            l.a[6] * r.x + l.a[7] * r.y + l.a[8] * r.z);
}
template <class T>
PE_HD M3<T> mm(const M3<T>& l, const M3<T>& r) {
  M3<T> m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      m.a[3 * i + j] = l.a[3 * i] * r.a[j] + l.a[3 * i + 1] * r.a[3 + j] + l.a[3 * i + 2] * r.a[6 + j];
  return m;
}
template <class T>
PE_HD M3<T> tr(const M3<T>& m) {
  M3<T> r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[3 * i + j] = m.a[3 * j + i];
  return r;
}
template <class T>
PE_HD M3<T> cols(V3<T> c1, V3<T> c2, V3<T> c3) {
  M3<T> m;
  m.a[0] = c1.x; m.a[1] = c2.x; m.a[2] = c3.x;
  m.a[3] = c1.y; m.a[4] = c2.y; m.a[5] = c3.y;
  m.a[6] = c1.z; m.a[7] = c2.z; m.a[8] = c3.z;
  return m;
}
template <class T>
PE_HD M3<T> skew_m(V3<T> v, const M3<T>& m) {  // [v]x * M
  M3<T> r;
  r.a[0] = v.y * m.a[6] - v.z * m.a[3]; r.a[1] = v.y * m.a[7] - v.z * m.a[4]; r.a[2] = v.y * m.a[8] - v.z * m.a[5];
  r.a[3] = v.z * m.a[0] - v.x * m.a[6]; r.a[4] = v.z * m.a[1] - v.x * m.a[7]; r.a[5] = v.z * m.a[2] - v.x * m.a[8];
  r.a[6] = v.x * m.a[3] - v.y * m.a[0]; r.a[7] = v.x * m.a[4] - v.y * m.a[1]; r.a[8] = v.x * m.a[5] - v.y * m.a[2];
  return r;
}
template <class T>
PE_HD M3<T> m_skew(const M3<T>& m, V3<T> v) {  // M * [v]x
  M3<T> r;
  r.a[0] = m.a[1] * v.z - m.a[2] * v.y; r.a[1] = m.a[2] * v.x - m.a[0] * v.z; r.a[2] = m.a[0] * v.y - m.a[1] * v.x;
  r.a[3] = m.a[4] * v.z - m.a[5] * v.y; r.a[4] = m.a[5] * v.x - m.a[3] * v.z; r.a[5] = m.a[3] * v.y - m.a[4] * v.x;
  r.a[6] = m.a[7] * v.z - m.a[8] * v.y; r.a[7] = m.a[8] * v.x - m.a[6] * v.z; r.a[8] = m.a[6] * v.y - m.a[7] * v.x;
  return r;
}
template <class T>
PE_HD M3<T> inv(const M3<T>& m) {
  const T* v = m.a;
  T det = v[0] * ((v[4] * v[8]) - (v[7] * v[5])) + v[1] * ((v[5] * v[6]) - (v[8] * v[3])) +
          v[2] * ((v[3] * v[7]) - (v[4] * v[6]));
  det = cst<T>(1.0) / det;
  M3<T> r;
  r.a[0] = det * ((v[4] * v[8]) - (v[5] * v[7]));
  r.a[1] = det * ((v[7] * v[2]) - (v[8] * v[1]));
  r.a[2] = det * ((v[1] * v[5]) - (v[2] * v[4]));
  r.a[3] = det * ((v[5] * v[6]) - (v[3] * v[8]));
  r.a[4] = det * ((v[8] * v[0]) - (v[6] * v[2]));
  r.a[5] = det * ((v[2] * v[3]) - (v[0] * v[5]));
  r.a[6] = det * ((v[3] * v[7]) - (v[4] * v[6]));
  r.a[7] = det * ((v[6] * v[1]) - (v[7] * v[0]));
  r.a[8] = det * ((v[0] * v[4]) - (v[1] * v[3]));
  return r;
}

// The quantities that pe's decoupled passes of InelasticCoulombContactByDecoupling read
// (HCSITS.h:2395-2520): the contact frame, lever arms, world inverse inertias, relative velocity
// with the contact's own impulse removed, gap rate, 1 / D_nn, the tangential block of D and mu.
template <class S>
struct SlideQ {
  V3<S> n, t, o, r1, r2, gdot;
  M3<S> I1, I2;
  S ddt, dni, A00, A01, A10, A11, mu;
};

template <class S>
PE_HD SlideQ<Dual> strip(const SlideQ<S>& Q) {
  SlideQ<Dual> R;
  R.n = v3(zd(Q.n.x), zd(Q.n.y), zd(Q.n.z)); R.t = v3(zd(Q.t.x), zd(Q.t.y), zd(Q.t.z));
  R.o = v3(zd(Q.o.x), zd(Q.o.y), zd(Q.o.z)); R.r1 = v3(zd(Q.r1.x), zd(Q.r1.y), zd(Q.r1.z));
  R.r2 = v3(zd(Q.r2.x), zd(Q.r2.y), zd(Q.r2.z)); R.gdot = v3(zd(Q.gdot.x), zd(Q.gdot.y), zd(Q.gdot.z));
  for (int k = 0; k < 9; ++k) { R.I1.a[k] = zd(Q.I1.a[k]); R.I2.a[k] = zd(Q.I2.a[k]); }
  R.ddt = zd(Q.ddt); R.dni = zd(Q.dni);
  R.A00 = zd(Q.A00); R.A01 = zd(Q.A01); R.A10 = zd(Q.A10); R.A11 = zd(Q.A11);
  R.mu = zd(Q.mu);
  return R;
}

// Residuals of the passes' fixed point with the normal impulse c > 0 and the tangential impulse
// on the cone's circle at angle a, x = mu c (cos a, sin a):
//   F1 = c + dni (n . (gdot + coupling(t x0 + o x1)) + ddt)    (the normal pass, HCSITS.h:2406-2414)
//   F2 = x'(a) . (As x + g_to(c)) / (mu c)                     (x minimises 0.5 x.A x + x.g_to on the
//                                                              circle, the golden-section search's
//                                                              target, HCSITS.h:2416-2509)
// where coupling(p) = (I1 (r1 x p)) x r1 + (I2 (r2 x p)) x r2, g_to = (t, o) . (gdot + coupling(n c))
// and As = (A + A^T) / 2, the part of A the objective sees (A is symmetric up to rounding).
template <class S>
PE_HD void slide_residual(const SlideQ<S>& Q, S c, S a, S& F1, S& F2) {
  const S ca = cos(a), sa = sin(a);
  const S x0 = Q.mu * c * ca, x1 = Q.mu * c * sa;
  const V3<S> pt = scale(Q.t, x0) + scale(Q.o, x1);
  const V3<S> gc = Q.gdot + cross(mv(Q.I1, cross(Q.r1, pt)), Q.r1) + cross(mv(Q.I2, cross(Q.r2, pt)), Q.r2);
  F1 = c + Q.dni * (dot(Q.n, gc) + Q.ddt);
  const V3<S> pn = scale(Q.n, c);
  const V3<S> g2 = Q.gdot + cross(mv(Q.I1, cross(Q.r1, pn)), Q.r1) + cross(mv(Q.I2, cross(Q.r2, pn)), Q.r2);
  const S as = cst<S>(0.5) * (Q.A01 + Q.A10);
  const S y0 = Q.A00 * x0 + as * x1 + dot(Q.t, g2);
  const S y1 = as * x0 + Q.A11 * x1 + dot(Q.o, g2);
  F2 = -sa * y0 + ca * y1;
}

// Jacobian of (F1, F2) with respect to (c, a) at values, with the residual.
PE_HD void slide_jacobian(const SlideQ<Dual>& Q0, double c, double a, double J[4], double F[2]) {
  Dual F1, F2;
  slide_residual(Q0, mkd(c, 1.0), mkd(a), F1, F2);
  J[0] = F1.d; J[2] = F2.d;
  F[0] = F1.v; F[1] = F2.v;
  slide_residual(Q0, mkd(c), mkd(a, 1.0), F1, F2);
  J[1] = F1.d; J[3] = F2.d;
}

// The sliding contact's impulse (c, x0, x1) in the contact frame as a function of the contact's
// inputs: the fixed point (c, a) of the decoupled passes with the exact minimiser on the circle,
// found by Newton's method from the forward run's converged impulse (c0, a0), and differentiated
// by the implicit function theorem: d(c, a) = -J^-1 dF.  The golden-section search of the forward
// run approximates this minimiser; with enough steps it meets it to rounding.
template <class T>
PE_HD void slide_fixed_point(const SlideQ<T>& Q, double c0, double a0, T& pc0, T& pc1, T& pc2) {
  const SlideQ<Dual> Q0 = strip(Q);
  double c = c0, a = a0, J[4], F[2];
  slide_jacobian(Q0, c, a, J, F);
  for (int it = 0; it < 4; ++it) {
    const double det = J[0] * J[3] - J[1] * J[2];
    if (!(det != 0.0) || !isfinite(det)) break;
    const double dc = -(J[3] * F[0] - J[1] * F[1]) / det;
    const double da = -(-J[2] * F[0] + J[0] * F[1]) / det;
    double Jn[4], Fn[2];
    slide_jacobian(Q0, c + dc, a + da, Jn, Fn);
    if (!(fabs(Fn[0]) + fabs(Fn[1]) < fabs(F[0]) + fabs(F[1]))) break;   // no further progress
    c += dc; a += da;
    for (int k = 0; k < 4; ++k) J[k] = Jn[k];
    F[0] = Fn[0]; F[1] = Fn[1];
  }
  // derivative of F with respect to the inputs at fixed (c, a)
  T F1, F2;
  slide_residual(Q, cst<T>(c), cst<T>(a), F1, F2);
  const double det = J[0] * J[3] - J[1] * J[2];
  double dc = 0.0, da = 0.0;
  if (det != 0.0 && isfinite(det)) {
    dc = -(J[3] * der(F1) - J[1] * der(F2)) / det;
    da = -(-J[2] * der(F1) + J[0] * der(F2)) / det;
  }
  const T cT = with_d<T>(c, dc), aT = with_d<T>(a, da);
  pc0 = cT;
  pc1 = Q.mu * cT * cos(aT);
  pc2 = Q.mu * cT * sin(aT);
}

// Constants of one contact (at the evaluation point).
struct ContactConst {
  double rad1, rad2;         // radii (rad2 unused when body 2 is a plane)
  double im1, im2;           // inverse masses (0 for fixed bodies)
  double I1[9], I2[9];       // world-frame inverse inertias as the forward step used them
  double pn[3], pd;          // plane normal and displacement (plane contacts)
  double xp[3];              // plane position (plane contacts)
  int plane;                 // body 2 is a plane
  double erp, dtinv;
};

constexpr int kAdjIn = 22;   // x1(3) x2(3) U1(6) U2(6) p(3) mu
constexpr int kAdjOut = 15;  // E2 (3), -contrib (12)

// One contact's part of E1/E2 (see the top of this file) for relaxation model MODEL (0, 1 or 3).
// pe's branch structure is followed on the values, so the derivative is the one of the branch
// (active mode) the contact is in.  Model 3's sliding branch is the fixed point of its decoupled
// passes with the exact minimiser on the friction circle (slide_fixed_point), located from the
// converged impulse p; a sliding contact whose normal impulse is zero has zero impulse.
template <class T, int MODEL>
PE_HD void contact_map(const ContactConst& K, const T* z, T* out) {
  const V3<T> x1 = v3(z[0], z[1], z[2]);
  const V3<T> x2 = K.plane ? v3(cst<T>(K.xp[0]), cst<T>(K.xp[1]), cst<T>(K.xp[2])) : v3(z[3], z[4], z[5]);
  const V3<T> v1 = v3(z[6], z[7], z[8]), w1 = v3(z[9], z[10], z[11]);
  const V3<T> v2 = v3(z[12], z[13], z[14]), w2 = v3(z[15], z[16], z[17]);
  const V3<T> p = v3(z[18], z[19], z[20]);
  const T mu = z[21];
  // geometry: collideSphereSphere / collideSpherePlane (MaxContacts.h:547-568, 777-792)
  V3<T> n, gpos;
  T dist;
  if (K.plane) {
    n = v3(cst<T>(K.pn[0]), cst<T>(K.pn[1]), cst<T>(K.pn[2]));
    dist = dot(n, x1) - cst<T>(K.rad1) - cst<T>(K.pd);
    gpos = x1 - scale(cst<T>(K.rad1) + dist, n);
  } else {
    const V3<T> d = x1 - x2;
    dist = sqrt(dot(d, d)) - cst<T>(K.rad1) - cst<T>(K.rad2);
    n = normalize(d);
    gpos = x2 + scale(n, cst<T>(K.rad2) + cst<T>(0.5) * dist);
  }
  // contact caching, HCSITS.h:1741-1818
  const V3<T> r1 = gpos - x1, r2 = gpos - x2;
  V3<T> t;
  if (fabs(val(n.x)) < fabs(val(n.y))) {
    if (fabs(val(n.x)) < fabs(val(n.z))) t = v3(cst<T>(0.0), n.z, -n.y);
    else t = v3(n.y, -n.x, cst<T>(0.0));
  } else {
    if (fabs(val(n.y)) < fabs(val(n.z))) t = v3(-n.z, cst<T>(0.0), n.x);
    else t = v3(n.y, -n.x, cst<T>(0.0));
  }
  t = normalize(t);
  const V3<T> o = cross(n, t);
  const M3<T> C = cols(n, t, o);
  if (val(dist) < 0.0) dist = dist * cst<T>(K.erp);
  const T ddt = dist * cst<T>(K.dtinv);
  M3<T> I1, I2;
  for (int k = 0; k < 9; ++k) { I1.a[k] = cst<T>(K.I1[k]); I2.a[k] = cst<T>(K.I2[k]); }
  const T im1 = cst<T>(K.im1), im2 = cst<T>(K.im2);
  // relative velocity with the own impulse removed (HCSITS.h:2343-2360)
  const V3<T> dv1 = v1 - scale(im1, p), dw1 = w1 - mv(I1, cross(r1, p));
  const V3<T> dv2 = v2 + scale(im2, p), dw2 = w2 + mv(I2, cross(r2, p));
  const V3<T> gdot = (dv1 - dv2) + cross(dw1, r1) - cross(dw2, r2);
  V3<T> g = v3(dot(n, gdot), dot(t, gdot), dot(o, gdot));
  g.x = g.x + ddt;
  V3<T> pn;  // new impulse, world frame
  if (val(g.x) >= 0.0) {
    pn = v3(cst<T>(0.0), cst<T>(0.0), cst<T>(0.0));
  } else {
    M3<T> Km;
    {
      const M3<T> A = m_skew(skew_m(r1, I1), r1), B = m_skew(skew_m(r2, I2), r2);
      for (int k = 0; k < 9; ++k) Km.a[k] = -(A.a[k] + B.a[k]);
      Km.a[0] = Km.a[0] + (im1 + im2);
      Km.a[4] = Km.a[4] + (im1 + im2);
      Km.a[8] = Km.a[8] + (im1 + im2);
    }
    const M3<T> D = mm(mm(tr(C), Km), C);
    const T dni = cst<T>(1.0) / D.a[0];
    if (MODEL == 0) {
      pn = scale(n, -(dni * g.x));
    } else if (MODEL == 1) {
      const M3<T> Di = inv(D);
      const V3<T> s = mv(Di, g);
      T pc0 = -s.x, pc1 = -s.y, pc2 = -s.z;
      T flimit = mu * pc0;
      T fsq = pc1 * pc1 + pc2 * pc2;
      if (val(fsq) > val(flimit) * val(flimit) || val(pc0) < 0.0) {
        // ApproximateInelasticCoulombContactByDecoupling, HCSITS.h:2186-2226
        T det = D.a[4] * D.a[8] - D.a[5] * D.a[7];
        det = cst<T>(1.0) / det;
        const T T00 = det * D.a[8], T01 = -det * D.a[5], T10 = -det * D.a[7], T11 = det * D.a[4];
        V3<T> ptmp = scale(dot(t, p), t) + scale(dot(o, p), o);
        const T gn = g.x + dot(n, cross(mv(I1, cross(r1, ptmp)), r1) + cross(mv(I2, cross(r2, ptmp)), r2));
        pc0 = -(dni * gn);
        if (val(pc0) < 0.0) pc0 = cst<T>(0.0);
        ptmp = scale(n, pc0);
        const V3<T> g2 = gdot + cross(mv(I1, cross(r1, ptmp)), r1) + cross(mv(I2, cross(r2, ptmp)), r2);
        const T gt = dot(t, g2), go = dot(o, g2);
        pc1 = -(T00 * gt + T01 * go);
        pc2 = -(T10 * gt + T11 * go);
        flimit = mu * pc0;
        fsq = pc1 * pc1 + pc2 * pc2;
        if (val(fsq) > val(flimit) * val(flimit)) {
          const T fr = flimit / sqrt(fsq);
          pc1 = pc1 * fr;
          pc2 = pc2 * fr;
        }
      }
      pn = mv(C, v3(pc0, pc1, pc2));
    } else if (MODEL == 3) {
      const M3<T> Di = inv(D);
      const V3<T> s = mv(Di, g);
      T pc0 = -s.x, pc1 = -s.y, pc2 = -s.z;
      const T flimit = mu * pc0;
      const T fsq = pc1 * pc1 + pc2 * pc2;
      if (val(fsq) > val(flimit) * val(flimit) || val(pc0) < 0.0) {
        // dynamic (HCSITS.h:2395-2520): located from the converged impulse's contact-frame parts
        const double c0 = val(n.x) * val(p.x) + val(n.y) * val(p.y) + val(n.z) * val(p.z);
        const double q1 = val(t.x) * val(p.x) + val(t.y) * val(p.y) + val(t.z) * val(p.z);
        const double q2 = val(o.x) * val(p.x) + val(o.y) * val(p.y) + val(o.z) * val(p.z);
        if (!(c0 > 0.0)) {
          pc0 = pc1 = pc2 = cst<T>(0.0);
        } else if (val(mu) == 0.0) {
          // no friction: the tangential impulse is zero and the normal pass alone remains
          pc0 = -(dni * g.x);
          pc1 = pc2 = cst<T>(0.0);
        } else {
          SlideQ<T> Q;
          Q.n = n; Q.t = t; Q.o = o; Q.r1 = r1; Q.r2 = r2; Q.gdot = gdot; Q.I1 = I1; Q.I2 = I2;
          Q.ddt = ddt; Q.dni = dni; Q.mu = mu;
          Q.A00 = D.a[4]; Q.A01 = D.a[5]; Q.A10 = D.a[7]; Q.A11 = D.a[8];
          slide_fixed_point(Q, c0, atan2(q2, q1), pc0, pc1, pc2);
        }
      }
      pn = mv(C, v3(pc0, pc1, pc2));
    }
  }
  out[0] = p.x - pn.x;
  out[1] = p.y - pn.y;
  out[2] = p.z - pn.z;
  // -(velocity corrections of p), HCSITS.h:2532-2535
  const V3<T> a1 = scale(im1, p), b1 = mv(I1, cross(r1, p));
  const V3<T> a2 = scale(im2, p), b2 = mv(I2, cross(r2, p));
  out[3] = -a1.x; out[4] = -a1.y; out[5] = -a1.z;
  out[6] = -b1.x; out[7] = -b1.y; out[8] = -b1.z;
  out[9] = a2.x; out[10] = a2.y; out[11] = a2.z;
  out[12] = b2.x; out[13] = b2.y; out[14] = b2.z;
}

}  // namespace ad

// One recorded step of a differentiable run (device pointers).
struct AdjTape {
  uint32_t nc;
  const uint32_t* b1;   // sweep order, with kFixedBit
  const uint32_t* b2;
  const double* p;      // 3 per contact: converged impulse
  const double* mu;     // per contact: friction
  const double* pos;    // 3 per body, at the start of the step
  const double* sb;     // kSB per body (inverse mass, world inverse inertia)
  const double* vel;    // 6 per body, after the step (v, w)
};

// Body constants the per-contact map needs.
struct AdjConsts {
  const uint8_t* shape;
  const double* param;
  const double* pnormal;
  const double* pd;
  const double* ppos;   // body positions (planes do not move)
  double erp, dtinv;
};

void adj_jacobian(int model, const AdjTape& T, const AdjConsts& A, double* jac, hipStream_t s);
void adj_prepare(uint32_t nc, const double* jac, double* minv, hipStream_t s);
void adj_init(uint32_t n, const uint8_t* fixed, const double* xbar, const double* vbar, const double* wbar, double dt,
              double* alpha, hipStream_t s);
void adj_sweep(const AdjTape& T, uint32_t begin, uint32_t end, const double* jac, const double* minv, double* alpha,
               double* lambda, unsigned long long* delta, hipStream_t s);
void adj_assemble(const AdjTape& T, const double* jac, const double* alpha, const double* lambda, double* xbar,
                  double* mubar, hipStream_t s);
void adj_finish(uint32_t n, const double* alpha, double* vbar, double* wbar, hipStream_t s);
void adj_restore_R(uint32_t n, const double* q, double* R, hipStream_t s);
void adj_pack_vel(uint32_t n, const double* v, const double* w, double* vel, hipStream_t s);
void adj_pack_p(uint32_t nc, const double* f, size_t cap, double* p, double* mu, hipStream_t s);

}  // namespace pe_amd
