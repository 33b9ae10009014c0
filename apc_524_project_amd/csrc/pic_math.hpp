// pic_math.hpp -- the FP64 per-particle arithmetic of the plain form, written
// once and used by the forward kernels, the adjoint's recomputation and the
// host-side checks of the adjoint.  Every function follows the reference's
// operation order; translation units that include it are compiled with
// -ffp-contract=off so that the GPU results equal the serial CPU ones bit
// for bit.
#pragma once

#include "geo.hpp"

namespace emo {

// Grid::getCellID (src/grid/oGrid.cpp:208-222); -1 outside the padded grid
// (where the reference asserts).
EMO_HD inline int cell_of(const Geo &g, double x, double y, double z, int *ix, int *iy, int *iz) {
  int a = (int)((x - g.x0[0]) * g.id[0]);
  a += 1;
  int b = (int)((y - g.x0[1]) * g.id[1]);
  b += 1;
  int c = (int)((z - g.x0[2]) * g.id[2]);
  c += 1;
  *ix = a;
  *iy = b;
  *iz = c;
  if (a < 0 || a >= g.nc[0] || b < 0 || b >= g.nc[1] || c < 0 || c >= g.nc[2]) return -1;
  return (g.nc[1] * g.nc[2]) * a + g.nc[2] * b + c;
}

// Grid::getFieldInterpolatorVec (src/grid/oGrid.cpp:153-197): the cell's
// least vertex (Q22: x0 + dx * stored index), 4 E edges per component and
// the 2 (B + B_tm1)/2 faces per component.
EMO_HD inline void cell_vector(const Geo &g, const double *const E[3], const double *const B[3],
                               const double *const Bt[3], int ix, int iy, int iz, double cv[21]) {
  cv[0] = g.x0[0] + g.d[0] * ix;
  cv[1] = g.x0[1] + g.d[1] * iy;
  cv[2] = g.x0[2] + g.d[2] * iz;
  const double *Ex = E[0], *Ey = E[1], *Ez = E[2];
  cv[3] = Ex[gidx(g, ix, iy, iz)];
  cv[4] = Ex[gidx(g, ix, iy + 1, iz)];
  cv[5] = Ex[gidx(g, ix, iy + 1, iz + 1)];
  cv[6] = Ex[gidx(g, ix, iy, iz + 1)];
  cv[7] = Ey[gidx(g, ix, iy, iz)];
  cv[8] = Ey[gidx(g, ix, iy, iz + 1)];
  cv[9] = Ey[gidx(g, ix + 1, iy, iz + 1)];
  cv[10] = Ey[gidx(g, ix + 1, iy, iz)];
  cv[11] = Ez[gidx(g, ix, iy, iz)];
  cv[12] = Ez[gidx(g, ix + 1, iy, iz)];
  cv[13] = Ez[gidx(g, ix + 1, iy + 1, iz)];
  cv[14] = Ez[gidx(g, ix, iy + 1, iz)];
  cv[15] = (Bt[0][gidx(g, ix, iy, iz)] + B[0][gidx(g, ix, iy, iz)]) / 2;
  cv[16] = (Bt[0][gidx(g, ix + 1, iy, iz)] + B[0][gidx(g, ix + 1, iy, iz)]) / 2;
  cv[17] = (Bt[1][gidx(g, ix, iy, iz)] + B[1][gidx(g, ix, iy, iz)]) / 2;
  cv[18] = (Bt[1][gidx(g, ix, iy + 1, iz)] + B[1][gidx(g, ix, iy + 1, iz)]) / 2;
  cv[19] = (Bt[2][gidx(g, ix, iy, iz)] + B[2][gidx(g, ix, iy, iz)]) / 2;
  cv[20] = (Bt[2][gidx(g, ix, iy, iz + 1)] + B[2][gidx(g, ix, iy, iz + 1)]) / 2;
}

// The padded-array offsets of the 12 E and 6 B entries of cell_vector,
// in the same order (cv[3 + q]).
EMO_HD inline void cell_vector_index(const Geo &g, int ix, int iy, int iz, long long idx[18]) {
  idx[0] = gidx(g, ix, iy, iz);
  idx[1] = gidx(g, ix, iy + 1, iz);
  idx[2] = gidx(g, ix, iy + 1, iz + 1);
  idx[3] = gidx(g, ix, iy, iz + 1);
  idx[4] = gidx(g, ix, iy, iz);
  idx[5] = gidx(g, ix, iy, iz + 1);
  idx[6] = gidx(g, ix + 1, iy, iz + 1);
  idx[7] = gidx(g, ix + 1, iy, iz);
  idx[8] = gidx(g, ix, iy, iz);
  idx[9] = gidx(g, ix + 1, iy, iz);
  idx[10] = gidx(g, ix + 1, iy + 1, iz);
  idx[11] = gidx(g, ix, iy + 1, iz);
  idx[12] = gidx(g, ix, iy, iz);
  idx[13] = gidx(g, ix + 1, iy, iz);
  idx[14] = gidx(g, ix, iy, iz);
  idx[15] = gidx(g, ix, iy + 1, iz);
  idx[16] = gidx(g, ix, iy, iz);
  idx[17] = gidx(g, ix, iy, iz + 1);
}

// Interpolator::interpolate_fields (src/particles/interpolate.cpp:11-47)
EMO_HD inline void interp_fields(const Geo &g, const double cv[21], const double pos[3], double fld[6]) {
  const double *lcell = g.d;
  double ds[3];
  for (int q = 0; q < 3; q++) ds[q] = pos[q] - cv[q];
  for (int q = 0; q < 3; q++) {
    int j = (q + 1) % 3, k = (q + 2) % 3;
    double eField = 0;
    eField += (lcell[j] - ds[j]) * (lcell[k] - ds[k]) * cv[3 + 4 * q];
    eField += ds[j] * (lcell[k] - ds[k]) * cv[3 + 4 * q + 1];
    eField += ds[j] * ds[k] * cv[3 + 4 * q + 2];
    eField += (lcell[j] - ds[j]) * ds[k] * cv[3 + 4 * q + 3];
    eField /= lcell[j] * lcell[k];
    double bField = 0;
    bField += (lcell[q] - ds[q]) * cv[3 + 12 + 2 * q];
    bField += ds[q] * cv[3 + 12 + 2 * q + 1];
    bField /= lcell[q];
    fld[q] = eField;
    fld[3 + q] = bField;
  }
}

// Boris::Step (src/pusher/boris.cpp:12-75), operation for operation
EMO_HD inline void boris_step(double xp[3], double vp[3], const double f[6], double q, double m, double dt) {
  double q_p = (0.5 * dt * q) / m;
  q_p *= kUnitAcc;
  double ex = f[0], ey = f[1], ez = f[2];
  double bx = f[3], by = f[4], bz = f[5];
  double tx = q_p * bx, ty = q_p * by, tz = q_p * bz;
  double bsqr = bx * bx + by * by + bz * bz;
  double tsqr = 2.0 / (1.0 + q_p * q_p * bsqr);
  double sx = tsqr * tx, sy = tsqr * ty, sz = tsqr * tz;
  double v_mx = vp[0] + q_p * ex;
  double v_my = vp[1] + q_p * ey;
  double v_mz = vp[2] + q_p * ez;
  double v_ix = v_mx + v_my * tz - v_mz * ty;
  double v_iy = v_my + v_mz * tx - v_mx * tz;
  double v_iz = v_mz + v_mx * ty - v_my * tx;
  double v_px = v_mx + v_iy * sz - v_iz * sy;
  double v_py = v_my + v_iz * sx - v_ix * sz;
  double v_pz = v_mz + v_ix * sy - v_iy * sx;
  double vx_new = v_px + q_p * ex;
  double vy_new = v_py + q_p * ey;
  double vz_new = v_pz + q_p * ez;
  xp[0] = xp[0] + dt * vx_new;
  xp[1] = xp[1] + dt * vy_new;
  xp[2] = xp[2] + dt * vz_new;
  vp[0] = vx_new;
  vp[1] = vy_new;
  vp[2] = vz_new;
}

// Relativistic_Boris::Step (src/pusher/relativisticBoris.cpp:28-140)
EMO_HD inline void vay_step(double xp[3], double vp[3], double &gam, const double f[6], double q, double m,
                            double dt) {
  double vx = vp[0], vy = vp[1], vz = vp[2];
  double gamma = gam;
  double q_p = (0.5 * dt * q) / m;
  q_p *= kUnitAcc;
  double ex = f[0], ey = f[1], ez = f[2];
  double bx = f[3], by = f[4], bz = f[5];
  double vXbx = vy * bz - vz * by;
  double vXby = vz * bx - vx * bz;
  double vXbz = vx * by - vy * bx;
  double uhx = gamma * vx + q_p * (ex + vXbx);
  double uhy = gamma * vy + q_p * (ey + vXby);
  double uhz = gamma * vz + q_p * (ez + vXbz);
  double upx = uhx + q_p * ex;
  double upy = uhy + q_p * ey;
  double upz = uhz + q_p * ez;
  double gammap_sq = 1.0 + upx * upx + upy * upy + upz * upz;
  double taux = q_p * bx, tauy = q_p * by, tauz = q_p * bz;
  double tau_sq = taux * taux + tauy * tauy + tauz * tauz;
  double sigma = gammap_sq - tau_sq;
  double ustar = upx * taux + upy * tauy + upz * tauz;
  double gamma_int = sqrt(sigma * sigma + 4 * (tau_sq + ustar * ustar));
  gamma_int = sqrt(0.5 * (sigma + gamma_int));
  double tx = taux / gamma_int, ty = tauy / gamma_int, tz = tauz / gamma_int;
  double s = 1.0 / (1.0 + tx * tx + ty * ty + tz * tz);
  double udott = ustar / gamma_int;
  double upXt_x = upy * tz - upz * ty;
  double upXt_y = upz * tx - upx * tz;
  double upXt_z = upx * ty - upy * tx;
  double uint_x = s * (upx + udott * tx + upXt_x);
  double uint_y = s * (upy + udott * ty + upXt_y);
  double uint_z = s * (upz + udott * tz + upXt_z);
  double vx_new = uint_x / gamma_int;
  double vy_new = uint_y / gamma_int;
  double vz_new = uint_z / gamma_int;
  xp[0] = xp[0] + dt * vx_new;
  xp[1] = xp[1] + dt * vy_new;
  xp[2] = xp[2] + dt * vz_new;
  vp[0] = vx_new;
  vp[1] = vy_new;
  vp[2] = vz_new;
  gam = gamma_int;
}

// The six particle BCs in the order x-,x+,y-,y+,z-,z+
// (src/particles/particle_handler.cpp:287-295): periodic shifts by
// (xMax - xMin) on strict compares (bc_p_periodic.cpp:44-87), reflecting
// mirrors and flips v (bc_p_reflecting.cpp:43-53).  flips[d] counts the
// reflections (for the adjoint), sh[d] the periodic shifts.
EMO_HD inline void particle_bcs(double xp[3], double vp[3], const double xmin[3], const double xmax[3],
                                const int pbc[6], int flips[3]) {
  for (int d = 0; d < 3; d++) flips[d] = 0;
  for (int side = 0; side < 6; side++) {
    int d = side >> 1, right = side & 1;
    double xMin = xmin[d], xMax = xmax[d];
    double x = xp[d];
    if (pbc[side] == 0) {
      if (!right && x < xMin) xp[d] = x + (xMax - xMin);
      if (right && x > xMax) xp[d] = x - (xMax - xMin);
    } else {
      if (right && x > xMax) {
        xp[d] = 2.0 * xMax - x;
        vp[d] = -vp[d];
        flips[d]++;
      }
      if (!right && x < xMin) {
        xp[d] = 2.0 * xMin - x;
        vp[d] = -vp[d];
        flips[d]++;
      }
    }
  }
}

// Depositor::deposit_particle_J and _Rho (src/particles/deposit.cpp:12-62)
// for one particle: the 12 edge currents (and 8 vertex charges) relative to
// the cell's least vertex (Grid::getCellVertex, src/grid/oGrid.cpp:233-244).
EMO_HD inline void deposit_row(const Geo &g, const double cellverts[3], const double pos[3], const double vel[3],
                               double pcharge, bool rho, double r[20]) {
  const double *lcell = g.d;
  const double cellVolume = lcell[0] * lcell[1] * lcell[2];
  double relpos[3] = {pos[0] - cellverts[0], pos[1] - cellverts[1], pos[2] - cellverts[2]};
  for (int q = 0; q < 3; q++) {
    int j = (q + 1) % 3, k = (q + 2) % 3;
    double cellArea = lcell[j] * lcell[k];
    double curIn1D = kUnitRhoJ * pcharge * vel[q];
    double cur2AreaRatio = (curIn1D / cellVolume) / cellArea;
    r[4 * q] = cur2AreaRatio * (lcell[j] - relpos[j]) * (lcell[k] - relpos[k]);
    r[4 * q + 1] = cur2AreaRatio * relpos[j] * (lcell[k] - relpos[k]);
    r[4 * q + 2] = cur2AreaRatio * relpos[j] * relpos[k];
    r[4 * q + 3] = cur2AreaRatio * (lcell[j] - relpos[j]) * relpos[k];
  }
  if (rho) {
    double cc = kUnitRhoJ * (pcharge / cellVolume) / cellVolume;
    r[12] = cc * (lcell[0] - relpos[0]) * (lcell[1] - relpos[1]) * (lcell[2] - relpos[2]);
    r[13] = cc * relpos[0] * (lcell[1] - relpos[1]) * (lcell[2] - relpos[2]);
    r[14] = cc * relpos[0] * relpos[1] * (lcell[2] - relpos[2]);
    r[15] = cc * (lcell[0] - relpos[0]) * relpos[1] * (lcell[2] - relpos[2]);
    r[16] = cc * (lcell[0] - relpos[0]) * relpos[1] * relpos[2];
    r[17] = cc * relpos[0] * relpos[1] * relpos[2];
    r[18] = cc * relpos[0] * (lcell[1] - relpos[1]) * relpos[2];
    r[19] = cc * (lcell[0] - relpos[0]) * (lcell[1] - relpos[1]) * relpos[2];
  }
}

}  // namespace emo
