// adjoint_math.hpp -- reverse-mode derivatives (vector-Jacobian products) of
// the per-particle FP64 arithmetic in pic_math.hpp: the field interpolation,
// the Boris and relativistic (Vay) pushes, the particle BCs and the current
// deposit.  Each *_vjp takes the forward inputs, recomputes what it needs and
// adds the input adjoints to the caller's accumulators.  Divisions by a
// recomputed quantity are multiplications by its reciprocal here (the
// forward's bit-exact operation order matters only in pic_math.hpp).
// Compiles as device code and as plain host C++.
#pragma once

#include <cmath>

#include "pic_math.hpp"

namespace emo {

EMO_HD inline void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// y = a x b with adjoint yb: ab += b x yb, bb += yb x a
EMO_HD inline void cross3_vjp(const double a[3], const double b[3], const double yb[3], double ab[3],
                              double bb[3]) {
  double t[3];
  cross3(b, yb, t);
  for (int q = 0; q < 3; q++) ab[q] += t[q];
  cross3(yb, a, t);
  for (int q = 0; q < 3; q++) bb[q] += t[q];
}

// interp_fields: fld = I(pos, cv).  Adds posb (3) and cvb[3..20] (18).
EMO_HD inline void interp_fields_vjp(const Geo &g, const double cv[21], const double pos[3], const double fb[6],
                                     double posb[3], double cvb[21]) {
  const double *L = g.d;
  const double iL[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
  double ds[3], dsb[3] = {0, 0, 0};
  for (int q = 0; q < 3; q++) ds[q] = pos[q] - cv[q];
  for (int q = 0; q < 3; q++) {
    int j = (q + 1) % 3, k = (q + 2) % 3;
    const double eb = fb[q] * (iL[j] * iL[k]);
    const double c0 = cv[3 + 4 * q], c1 = cv[3 + 4 * q + 1], c2 = cv[3 + 4 * q + 2], c3 = cv[3 + 4 * q + 3];
    cvb[3 + 4 * q] += eb * (L[j] - ds[j]) * (L[k] - ds[k]);
    cvb[3 + 4 * q + 1] += eb * ds[j] * (L[k] - ds[k]);
    cvb[3 + 4 * q + 2] += eb * ds[j] * ds[k];
    cvb[3 + 4 * q + 3] += eb * (L[j] - ds[j]) * ds[k];
    dsb[j] += eb * (-(L[k] - ds[k]) * c0 + (L[k] - ds[k]) * c1 + ds[k] * c2 - ds[k] * c3);
    dsb[k] += eb * (-(L[j] - ds[j]) * c0 - ds[j] * c1 + ds[j] * c2 + (L[j] - ds[j]) * c3);
    const double bb = fb[3 + q] * iL[q];
    const double d0 = cv[15 + 2 * q], d1 = cv[15 + 2 * q + 1];
    cvb[15 + 2 * q] += bb * (L[q] - ds[q]);
    cvb[15 + 2 * q + 1] += bb * ds[q];
    dsb[q] += bb * (d1 - d0);
  }
  for (int q = 0; q < 3; q++) posb[q] += dsb[q];
}

// boris_step: (x, v, f) -> (x', v').  Given xb', vb' adds xb, vb, fb.
EMO_HD inline void boris_step_vjp(const double v[3], const double f[6], double q, double m, double dt,
                                  const double xb1[3], const double vb1[3], double xb[3], double vb[3],
                                  double fb[6]) {
  double q_p = (0.5 * dt * q) / m;
  q_p *= kUnitAcc;
  const double E[3] = {f[0], f[1], f[2]}, B[3] = {f[3], f[4], f[5]};
  double t[3] = {q_p * B[0], q_p * B[1], q_p * B[2]};
  double bsqr = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
  double tsqr = 2.0 / (1.0 + q_p * q_p * bsqr);
  double s[3] = {tsqr * t[0], tsqr * t[1], tsqr * t[2]};
  double vm[3] = {v[0] + q_p * E[0], v[1] + q_p * E[1], v[2] + q_p * E[2]};
  double vmxt[3];
  cross3(vm, t, vmxt);
  double vi[3] = {vm[0] + vmxt[0], vm[1] + vmxt[1], vm[2] + vmxt[2]};
  // reverse
  double vnb[3], Eb[3] = {0, 0, 0}, Bb[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) {
    xb[d] += xb1[d];
    vnb[d] = vb1[d] + dt * xb1[d];
    Eb[d] += q_p * vnb[d];  // vn = vp + qp E
  }
  double vmb[3] = {vnb[0], vnb[1], vnb[2]};  // vp = vm + vi x s
  double vib[3] = {0, 0, 0}, sb[3] = {0, 0, 0};
  cross3_vjp(vi, s, vnb, vib, sb);
  double tb[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) vmb[d] += vib[d];  // vi = vm + vm x t
  cross3_vjp(vm, t, vib, vmb, tb);
  double tsqrb = 0.0;
  for (int d = 0; d < 3; d++) {
    tb[d] += tsqr * sb[d];
    tsqrb += sb[d] * t[d];
  }
  const double bsqrb = tsqrb * (-tsqr * tsqr * q_p * q_p * 0.5);
  for (int d = 0; d < 3; d++) {
    Bb[d] += 2.0 * B[d] * bsqrb + q_p * tb[d];
    vb[d] += vmb[d];
    Eb[d] += q_p * vmb[d];
  }
  for (int d = 0; d < 3; d++) {
    fb[d] += Eb[d];
    fb[3 + d] += Bb[d];
  }
}

// vay_step: (x, v, gamma, f) -> (x', v', gamma').  Given xb', vb', gb' adds
// xb, vb, gb, fb.
EMO_HD inline void vay_step_vjp(const double v[3], double gamma, const double f[6], double q, double m, double dt,
                                const double xb1[3], const double vb1[3], double gb1, double xb[3], double vb[3],
                                double &gb, double fb[6]) {
  double q_p = (0.5 * dt * q) / m;
  q_p *= kUnitAcc;
  const double E[3] = {f[0], f[1], f[2]}, B[3] = {f[3], f[4], f[5]};
  double vXb[3];
  cross3(v, B, vXb);
  double uh[3], up[3];
  for (int d = 0; d < 3; d++) {
    uh[d] = gamma * v[d] + q_p * (E[d] + vXb[d]);
    up[d] = uh[d] + q_p * E[d];
  }
  double gp2 = 1.0 + up[0] * up[0] + up[1] * up[1] + up[2] * up[2];
  double tau[3] = {q_p * B[0], q_p * B[1], q_p * B[2]};
  double tau2 = tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2];
  double sigma = gp2 - tau2;
  double ustar = up[0] * tau[0] + up[1] * tau[1] + up[2] * tau[2];
  double r = sqrt(sigma * sigma + 4 * (tau2 + ustar * ustar));
  double gi = sqrt(0.5 * (sigma + r));
  const double ig = 1.0 / gi, ir = 1.0 / r;
  double tt[3] = {tau[0] * ig, tau[1] * ig, tau[2] * ig};
  double s = 1.0 / (1.0 + tt[0] * tt[0] + tt[1] * tt[1] + tt[2] * tt[2]);
  double ud = ustar * ig;
  double upxt[3];
  cross3(up, tt, upxt);
  double w[3], ui[3], vn[3];
  for (int d = 0; d < 3; d++) {
    w[d] = up[d] + ud * tt[d] + upxt[d];
    ui[d] = s * w[d];
    vn[d] = ui[d] * ig;
  }
  // reverse
  double vnb[3];
  double gib = gb1;
  for (int d = 0; d < 3; d++) {
    xb[d] += xb1[d];
    vnb[d] = vb1[d] + dt * xb1[d];
  }
  double uib[3];
  double vnb_vn = 0.0;
  for (int d = 0; d < 3; d++) {
    uib[d] = vnb[d] * ig;
    vnb_vn += vnb[d] * vn[d];
  }
  gib += -vnb_vn * ig;
  double sb = 0.0, wb[3];
  for (int d = 0; d < 3; d++) {
    sb += uib[d] * w[d];
    wb[d] = s * uib[d];
  }
  double upb[3] = {wb[0], wb[1], wb[2]}, ttb[3] = {0, 0, 0};
  cross3_vjp(up, tt, wb, upb, ttb);
  double udb = 0.0;
  for (int d = 0; d < 3; d++) {
    ttb[d] += ud * wb[d];
    udb += wb[d] * tt[d];
  }
  double ustarb = udb * ig;
  gib += -udb * ustar * (ig * ig);
  for (int d = 0; d < 3; d++) ttb[d] += sb * (-s * s) * 2.0 * tt[d];
  double taub[3];
  double ttb_tau = 0.0;
  for (int d = 0; d < 3; d++) {
    taub[d] = ttb[d] * ig;
    ttb_tau += ttb[d] * tau[d];
  }
  gib += -ttb_tau * (ig * ig);
  double sigmab = gib * 0.25 * ig;
  const double rb = gib * 0.25 * ig;
  sigmab += rb * sigma * ir;
  double tau2b = rb * 2.0 * ir;
  ustarb += rb * 4.0 * ustar * ir;
  for (int d = 0; d < 3; d++) {
    upb[d] += ustarb * tau[d];
    taub[d] += ustarb * up[d];
  }
  const double gp2b = sigmab;
  tau2b += -sigmab;
  for (int d = 0; d < 3; d++) {
    taub[d] += 2.0 * tau[d] * tau2b;
    upb[d] += 2.0 * up[d] * gp2b;
  }
  double Eb[3], Bb[3], uhb[3];
  for (int d = 0; d < 3; d++) {
    Bb[d] = q_p * taub[d];
    uhb[d] = upb[d];
    Eb[d] = q_p * upb[d];
  }
  double qu[3];
  double gsum = 0.0;
  for (int d = 0; d < 3; d++) {
    gsum += uhb[d] * v[d];
    vb[d] += gamma * uhb[d];
    Eb[d] += q_p * uhb[d];
    qu[d] = q_p * uhb[d];
  }
  gb += gsum;
  cross3_vjp(v, B, qu, vb, Bb);
  for (int d = 0; d < 3; d++) {
    fb[d] += Eb[d];
    fb[3 + d] += Bb[d];
  }
}

// particle_bcs: reflections flip the sign of both adjoints in their
// dimension; periodic shifts leave them unchanged.
EMO_HD inline void particle_bcs_vjp(const int flips[3], double xb[3], double vb[3]) {
  for (int d = 0; d < 3; d++)
    if (flips[d] & 1) {
      xb[d] = -xb[d];
      vb[d] = -vb[d];
    }
}

// deposit_row (current part): r[0..11] = J(pos, vel).  Given rb[0..11] adds
// posb, velb.
EMO_HD inline void deposit_row_vjp(const Geo &g, const double cellverts[3], const double pos[3],
                                   const double vel[3], double pcharge, const double rb[12], double posb[3],
                                   double velb[3]) {
  const double *L = g.d;
  const double iL[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
  const double qv = (kUnitRhoJ * pcharge) * (iL[0] * iL[1] * iL[2]);  // charge / cell volume
  double rp[3] = {pos[0] - cellverts[0], pos[1] - cellverts[1], pos[2] - cellverts[2]};
  double rpb[3] = {0, 0, 0};
  for (int q = 0; q < 3; q++) {
    int j = (q + 1) % 3, k = (q + 2) % 3;
    const double dcdv = qv * (iL[j] * iL[k]);
    const double c = dcdv * vel[q];
    const double w0 = (L[j] - rp[j]) * (L[k] - rp[k]), w1 = rp[j] * (L[k] - rp[k]), w2 = rp[j] * rp[k],
                 w3 = (L[j] - rp[j]) * rp[k];
    const double b0 = rb[4 * q], b1 = rb[4 * q + 1], b2 = rb[4 * q + 2], b3 = rb[4 * q + 3];
    velb[q] += dcdv * (b0 * w0 + b1 * w1 + b2 * w2 + b3 * w3);
    rpb[j] += c * (-b0 * (L[k] - rp[k]) + b1 * (L[k] - rp[k]) + b2 * rp[k] - b3 * rp[k]);
    rpb[k] += c * (-b0 * (L[j] - rp[j]) - b1 * rp[j] + b2 * rp[j] + b3 * (L[j] - rp[j]));
  }
  for (int d = 0; d < 3; d++) posb[d] += rpb[d];
}

// Padded-array offsets that the 12 current entries of deposit_row reach in
// the fixed-order grid sum (Grid::addJ, src/grid/oGrid.cpp:91-111): entry 3
// repeats entry 1's Jx[ix][iy+1][iz] (the reference's duplicate target).
EMO_HD inline void deposit_target_index(const Geo &g, int ix, int iy, int iz, long long idx[12]) {
  idx[0] = gidx(g, ix, iy, iz);
  idx[1] = gidx(g, ix, iy + 1, iz);
  idx[2] = gidx(g, ix, iy + 1, iz + 1);
  idx[3] = gidx(g, ix, iy + 1, iz);
  idx[4] = gidx(g, ix, iy, iz);
  idx[5] = gidx(g, ix, iy, iz + 1);
  idx[6] = gidx(g, ix + 1, iy, iz + 1);
  idx[7] = gidx(g, ix + 1, iy, iz);
  idx[8] = gidx(g, ix, iy, iz);
  idx[9] = gidx(g, ix + 1, iy, iz);
  idx[10] = gidx(g, ix + 1, iy + 1, iz);
  idx[11] = gidx(g, ix, iy + 1, iz);
}

}  // namespace emo
