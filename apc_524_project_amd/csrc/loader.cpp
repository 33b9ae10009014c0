// loader.cpp -- host particle loader: Input::ProcessInfo's super-particle
// scaling (src/IO/check.cpp:53-111), Particle_Handler::Load
// (src/particles/particle_handler.cpp:34-85) with the reference's RNG, and
// zeroAvgChargeAndVelocity_ (:379-444).
//
// The RNG is Numerical Recipes' ran2 (Press et al., Numerical Recipes in C,
// 2nd ed., s7.1: two Schrage-method LCGs combined through a 32-entry
// Bays-Durham shuffle) with Box-Muller pairs, in the seeding and call order
// of src/utils/RNG.cpp:23-141.  ran2's shuffle table has no skip-ahead, so
// the stream is generated serially on the host.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "emo_internal.hpp"

namespace emo {

class Random_Number_Generator {
 public:
  explicit Random_Number_Generator(long seed) {
    long idum = seed;
    if (idum == 0) idum = 1;
    else if (idum < 0) idum = -idum;
    idum2_ = idum;
    for (int j = kNtab + 7; j >= 0; j--) {
      long k = idum / kIq1;
      idum = kIa1 * (idum - k * kIq1) - k * kIr1;
      if (idum < 0) idum += kIm1;
      if (j < kNtab) iv_[j] = idum;
    }
    idum_ = idum;
    iy_ = iv_[0];
  }

  double getUniform() {
    long k = idum_ / kIq1;
    idum_ = kIa1 * (idum_ - k * kIq1) - k * kIr1;
    if (idum_ < 0) idum_ += kIm1;
    k = idum2_ / kIq2;
    idum2_ = kIa2 * (idum2_ - k * kIq2) - k * kIr2;
    if (idum2_ < 0) idum2_ += kIm2;
    int j = (int)(iy_ / kNdiv);
    iy_ = iv_[j] - idum2_;
    iv_[j] = idum_;
    if (iy_ < 1) iy_ += kIm1 - 1;
    double temp = (1.0 / kIm1) * iy_;
    const double rnmx = 1.0 - DBL_EPSILON;
    return temp > rnmx ? rnmx : temp;
  }

  double getGaussian(double mu, double sigma) {
    if (!generate_) {
      generate_ = true;
      return z1_ * sigma + mu;
    }
    double u1, u2;
    do {
      u1 = getUniform();
      u2 = getUniform();
    } while (u1 <= DBL_MIN);
    const double two_pi = 2.0 * 3.14159265358979323846;
    // The reference's g++ -O3 build evaluates cos and sin of the same
    // argument with one glibc sincos call; sincos can differ from separate
    // sin/cos calls in the last bit, so it is called explicitly here.
    double sn, cs;
    ::sincos(two_pi * u2, &sn, &cs);
    z0_ = std::sqrt(-2.0 * std::log(u1)) * cs;
    z1_ = std::sqrt(-2.0 * std::log(u1)) * sn;
    generate_ = false;
    return z0_ * sigma + mu;
  }

 private:
  static constexpr long kIm1 = 2147483563L, kIm2 = 2147483399L;
  static constexpr long kIa1 = 40014L, kIa2 = 40692L;
  static constexpr long kIq1 = 53668L, kIq2 = 52774L;
  static constexpr long kIr1 = 12211L, kIr2 = 3791L;
  static constexpr int kNtab = 32;
  static constexpr long kNdiv = 1 + (kIm1 - 1) / kNtab;
  long idum_ = 0, idum2_ = 123456789, iy_ = 0, iv_[kNtab] = {};
  double z0_ = 0.0, z1_ = 0.0;
  bool generate_ = true;
};

// Input::ProcessInfo: super_ratio = dens_phys / (Np / V); q, m, T scaled;
// dens_frac normalised.
int process_info(const emo_loader *ld, emo_config *cfg, double *dens_out, double *temp_out) {
  if (ld->nspecies < 1 || ld->nspecies > EMO_NSPEC) return EMO_ERR_CONFIG;
  double total_volume = 1.0;
  for (int i = 0; i < 3; i++) total_volume *= cfg->Lxyz[i];
  if (!(total_volume > 0)) return EMO_ERR_CONFIG;
  double n0 = ld->nparticles_tot / total_volume;
  double dens_phys = ld->dens_phys;
  if (dens_phys < 0) dens_phys = -dens_phys;
  else if (dens_phys == 0) dens_phys = n0;
  double super_ratio = n0 > 0 ? dens_phys / n0 : -1.0;
  double cden = 0.0;
  for (int i = 0; i < ld->nspecies; i++) {
    if (ld->dens_frac[i] < 0) return EMO_ERR_CONFIG;
    // the displacement loader divides by the wavenumber
    if (ld->pert_eps[i] != 0.0 && ld->pert_k[i] == 0.0) return EMO_ERR_CONFIG;
    cden += ld->dens_frac[i];
  }
  for (int i = 0; i < ld->nspecies; i++) {
    double mass = ld->mass_ratio[i] < 0 ? -ld->mass_ratio[i] : ld->mass_ratio[i];
    double temp = ld->temp[i] < 0 ? -ld->temp[i] : ld->temp[i];
    if (mass == 0) return EMO_ERR_CONFIG;
    double charge = ld->charge_ratio[i];
    double dens = ld->dens_frac[i];
    if (super_ratio > 0) {
      dens /= cden;
      mass *= super_ratio;
      charge *= super_ratio;
      temp *= super_ratio;
    }
    cfg->mass[i] = mass;
    cfg->charge[i] = charge;
    cfg->isTestParticle[i] = ld->isTestParticle[i];
    if (dens_out) dens_out[i] = dens;
    if (temp_out) temp_out[i] = temp;
  }
  cfg->nspecies = ld->nspecies;
  return EMO_OK;
}

// Generates the initial particle state on the host in the reference's order.
int load_maxwellian_host(const emo_config *cfg, const emo_loader *ld, std::vector<double> &x,
                         std::vector<double> &v, std::vector<double> &gam,
                         std::vector<int32_t> &sp, std::vector<int64_t> &id, std::vector<double> *dx_deps) {
  emo_config c = *cfg;
  double dens[EMO_NSPEC], temp[EMO_NSPEC];
  int rc = process_info(ld, &c, dens, temp);
  if (rc) return rc;
  int64_t npart = ld->nparticles_tot;
  x.resize(3 * npart);
  if (dx_deps) dx_deps->assign(npart, 0.0);
  v.resize(3 * npart);
  gam.resize(npart);
  sp.resize(npart);
  id.resize(npart);
  Random_Number_Generator rng((long)ld->seed);
  int ispec = 0;
  double cd = dens[0];
  for (int64_t ip = 0; ip < npart; ip++) {
    if (ip >= cd * npart) {
      ispec += 1;
      cd += dens[ispec];
    }
    if (ispec >= ld->nspecies) return EMO_ERR_CONFIG;
    double m = c.mass[ispec];
    double vth = kUnitVth * std::sqrt(temp[ispec] / m);
    double *xp = &x[3 * ip], *vp = &v[3 * ip];
    xp[0] = rng.getUniform() * cfg->Lxyz[0] + cfg->xyz0[0];
    xp[1] = rng.getUniform() * cfg->Lxyz[1] + cfg->xyz0[1];
    xp[2] = rng.getUniform() * cfg->Lxyz[2] + cfg->xyz0[2];
    vp[0] = rng.getGaussian(0.0, vth);
    vp[1] = rng.getGaussian(0.0, vth);
    vp[2] = rng.getGaussian(0.0, vth);
    // extensions (zero by default): displacement loader along x and drift.
    // The sensitivity dx/deps = -sin(k x_u)/k exists wherever a wavenumber
    // is set, also at eps = 0 (the linear response about the uniform load);
    // the displacement itself is applied only for eps != 0.
    if (dx_deps && ld->pert_k[ispec] != 0.0) {
      double k = ld->pert_k[ispec];
      (*dx_deps)[ip] = -std::sin(k * (xp[0] - cfg->xyz0[0])) / k;
    }
    if (ld->pert_eps[ispec] != 0.0) {
      double k = ld->pert_k[ispec];
      double xr = xp[0] - cfg->xyz0[0];
      xr = xr - (ld->pert_eps[ispec] / k) * std::sin(k * xr);
      if (xr < 0) xr += cfg->Lxyz[0];
      if (xr >= cfg->Lxyz[0]) xr -= cfg->Lxyz[0];
      xp[0] = xr + cfg->xyz0[0];
    }
    vp[0] += ld->drift[ispec][0];
    vp[1] += ld->drift[ispec][1];
    vp[2] += ld->drift[ispec][2];
    // gamma = sqrt(1 - v^2) as stored by the reference (particle_handler.cpp:77)
    gam[ip] = std::sqrt(1 - vp[0] * vp[0] - vp[1] * vp[1] - vp[2] * vp[2]);
    sp[ip] = ispec;
    id[ip] = ip;
  }
  bool all_periodic = true;
  for (int i = 0; i < 6; i++)
    if (cfg->fields_bound[i] != EMO_BC_PERIODIC) all_periodic = false;
  int za = ld->zero_avg < 0 ? (all_periodic && ld->poisson_init) : ld->zero_avg;
  if (!za) return EMO_OK;
  // zeroAvgChargeAndVelocity_: subtract the mean velocity of each group of
  // equal-charge particles, in index order; fail on a net charge.
  double qTot = 0.0;
  double min_q = (double)FLT_MAX;
  std::vector<unsigned char> counted(npart, 0);
  int64_t qCount;
  do {
    int64_t qRefIdx = 0;
    qCount = 0;
    double vx = 0.0, vy = 0.0, vz = 0.0;
    for (int64_t i = 0; i < npart; i++) {
      if (counted[i] == 0) {
        if (qRefIdx == 0) qRefIdx = i;
        if (c.charge[sp[i]] == c.charge[sp[qRefIdx]]) {
          counted[i] = 1;
          qCount += 1;
          vx += v[3 * i];
          vy += v[3 * i + 1];
          vz += v[3 * i + 2];
        }
      }
    }
    if (qCount > 0) {
      vx /= qCount;
      vy /= qCount;
      vz /= qCount;
      for (int64_t i = 0; i < npart; i++) {
        if (counted[i] == 1) {
          counted[i] = 2;
          v[3 * i] -= vx;
          v[3 * i + 1] -= vy;
          v[3 * i + 2] -= vz;
          qTot += c.charge[sp[i]];
          double aq = std::fabs(c.charge[sp[i]]);
          min_q = aq < min_q ? aq : min_q;
        }
      }
    }
  } while (qCount > 0);
  if (std::fabs(qTot) >= min_q / npart) return EMO_ERR_INVARIANT;
  return EMO_OK;
}

}  // namespace emo

namespace emo {
// d x0 / d pert_eps of every particle (load order) for the displacement
// loader above: -sin(k x_u) / k along x, from the same random stream
int loader_dx_deps(const emo_config *cfg, const emo_loader *ld, std::vector<double> &out) {
  std::vector<double> x, v, gam;
  std::vector<int32_t> sp;
  std::vector<int64_t> id;
  return load_maxwellian_host(cfg, ld, x, v, gam, sp, id, &out);
}
}  // namespace emo
