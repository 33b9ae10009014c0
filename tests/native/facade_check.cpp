// facade_check.cpp -- C entry points over the facade's single-particle
// classes (Interpolator, Depositor, the pushers), so that the tests can drive
// them from Python.  Built with the facade sources and linked against
// libemopic.so; none of these calls touches the device.
#include "../../apc_524_project_amd/host/emopic.hpp"

using namespace emopic;

extern "C" {

// Interpolator::interpolate_fields: out6 = {e1, e2, e3, b1, b2, b3}
void fc_interp(double *pos, double *lcell, double *cellvars, double *out6) {
  Field_part f;
  Interpolator in;
  in.interpolate_fields(pos, lcell, cellvars, &f);
  out6[0] = f.e1;
  out6[1] = f.e2;
  out6[2] = f.e3;
  out6[3] = f.b1;
  out6[4] = f.b2;
  out6[5] = f.b3;
}

// Depositor::deposit_particle_J / _Rho, accumulating into obj
void fc_deposit(int rho, const double *x, const double *v, double q, double *lcell, double *cellverts, double *obj) {
  Particle p;
  for (int d = 0; d < 3; d++) {
    p.x[d] = x[d];
    p.v[d] = v[d];
  }
  p.q = q;
  Depositor dep;
  if (rho) dep.deposit_particle_Rho(&p, lcell, cellverts, obj);
  else dep.deposit_particle_J(&p, lcell, cellverts, obj);
}

}  // extern "C"
