// Host build of the per-particle forward arithmetic and its reverse-mode
// derivatives (apc_524_project_amd/csrc/pic_math.hpp, adjoint_math.hpp), for
// finite-difference checks from tests/test_adjoint_math.py.  Test code only.
#include "../../apc_524_project_amd/csrc/adjoint_math.hpp"

using namespace emo;

static Geo make_geo(const double d[3]) {
  Geo g{};
  for (int q = 0; q < 3; q++) {
    g.n[q] = 8;
    g.nt[q] = 11;
    g.nc[q] = 10;
    g.x0[q] = 0.0;
    g.d[q] = d[q];
    g.id[q] = 1.0 / d[q];
  }
  g.sx = 11 * 11;
  g.sy = 11;
  g.ntot = 11 * 11 * 11;
  g.ncell = 1000;
  return g;
}

extern "C" {

// out[6] = interp(cv, pos)
void am_interp(const double d[3], const double cv[21], const double pos[3], double out[6]) {
  Geo g = make_geo(d);
  interp_fields(g, cv, pos, out);
}
void am_interp_vjp(const double d[3], const double cv[21], const double pos[3], const double fb[6],
                   double posb[3], double cvb[21]) {
  Geo g = make_geo(d);
  for (int q = 0; q < 3; q++) posb[q] = 0;
  for (int q = 0; q < 21; q++) cvb[q] = 0;
  interp_fields_vjp(g, cv, pos, fb, posb, cvb);
}

// state[7] = x3, v3, gamma; f[6]; rel selects the pusher
void am_push(int rel, double state[7], const double f[6], double q, double m, double dt) {
  if (rel) vay_step(state, state + 3, state[6], f, q, m, dt);
  else boris_step(state, state + 3, f, q, m, dt);
}
void am_push_vjp(int rel, const double state[7], const double f[6], double q, double m, double dt,
                 const double sb1[7], double sb[7], double fb[6]) {
  for (int k = 0; k < 7; k++) sb[k] = 0;
  for (int k = 0; k < 6; k++) fb[k] = 0;
  if (rel) vay_step_vjp(state + 3, state[6], f, q, m, dt, sb1, sb1 + 3, sb1[6], sb, sb + 3, sb[6], fb);
  else boris_step_vjp(state + 3, f, q, m, dt, sb1, sb1 + 3, sb, sb + 3, fb);
}

// r[12] = deposit_row(pos, vel) currents for the cell whose least vertex is cv3
void am_deposit(const double d[3], const double cv3[3], const double pos[3], const double vel[3], double q,
                double r[12]) {
  Geo g = make_geo(d);
  double r20[20];
  deposit_row(g, cv3, pos, vel, q, false, r20);
  for (int k = 0; k < 12; k++) r[k] = r20[k];
}
void am_deposit_vjp(const double d[3], const double cv3[3], const double pos[3], const double vel[3], double q,
                    const double rb[12], double posb[3], double velb[3]) {
  Geo g = make_geo(d);
  for (int k = 0; k < 3; k++) posb[k] = velb[k] = 0;
  deposit_row_vjp(g, cv3, pos, vel, q, rb, posb, velb);
}
}
