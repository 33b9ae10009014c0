// boundary.cpp -- host side of the field boundary conditions: the order of
// Grid::executeBC (src/grid/grid.cpp:577-585), the periodic ghost copies of
// BC_F_Periodic (src/boundaries/b_fields/bc_f_periodic.cpp:47-52), LightBC's
// pulse scalars (src/grid/lightBC.cpp:10-198, src/utils/evaluateGaussianPulse.cpp:7-30)
// and PoissonBC's constant planes (src/poisson/poissonBC.cpp:67-80).  The
// scalars are evaluated here in FP64 with the host libm; the device kernels
// only fill and copy planes.
#include <cmath>

#include <cmath>
#include <vector>

#include "emo_internal.hpp"

namespace emo {
namespace {

double gaussian_pulses(int nw, const double *amp, const double *omg, const double *pha,
                       const double *del, const double *dwt, double t, double dt) {
  double val = 0.0;
  for (int i = 0; i < nw; i++) {
    double tmp = amp[i] * std::cos(omg[i] * (t + dt) + pha[i] * kPi / 180.0);
    val += tmp * std::exp(-std::pow((t - del[i]) * dwt[i], 2));
  }
  return val;
}

// plane index of Grid::setFieldAlongEdge (src/grid/oGrid.cpp:291-431):
// fid 0..5 = Ex..Bz, upper side uses the last physical or the ghost plane
int edge_plane_index(const Geo &g, int fid, int dim, bool upper) {
  if (!upper) return 1;
  int n_ = g.nc[dim];
  int comp = fid % 3;
  bool is_e = fid < 3;
  if (is_e) return comp == dim ? n_ - 1 - 1 : n_ - 1;
  return comp == dim ? n_ - 1 : n_ - 1 - 1;
}

template <typename T>
void light_bc(emo_ctx *c, FieldSet<T> &fs, int side, double t, double dt) {
  const emo_config &cf = c->cfg;
  int sign = (side & 1) ? 1 : -1;
  int dim = side / 2;
  int sidev = sign * (dim + 1);
  // waves registered on this side, per polarisation (LightBC ctor :27-68)
  int nw[3] = {0, 0, 0};
  double amp[3][EMO_NWAVE], omg[3][EMO_NWAVE], pha[3][EMO_NWAVE], del[3][EMO_NWAVE], dwt[3][EMO_NWAVE];
  for (int i = 0; i < cf.nwaves; i++) {
    if (cf.inSide[i] == sidev) {
      int pid = cf.inPolE[i] - 1;
      int k = nw[pid]++;
      amp[pid][k] = cf.peakamps[i];
      omg[pid][k] = cf.omegas[i];
      pha[pid][k] = cf.phases[i];
      del[pid][k] = cf.delays[i];
      dwt[pid][k] = cf.invWidths[i];
    }
  }
  // the two transverse (E, B) pairs then the normal pair (:80-136)
  static const int EpolT[3][3] = {{1, 2, 0}, {2, 0, 1}, {0, 1, 2}};
  static const int BpolT[3][3] = {{2, 1, 0}, {0, 2, 1}, {1, 0, 2}};
  double dx = c->geo.d[dim];
  double bPhase = (dx + sign * dt) / 2;
  PlaneFill fills[6];
  int nf = 0;
  for (int i = 0; i < 3; i++) {
    int eid = EpolT[dim][i], bid = BpolT[dim][i];
    double ewave = 0.0, bwave = 0.0;
    bool has = nw[eid] > 0 && eid != dim;
    if (has) {
      ewave = gaussian_pulses(nw[eid], amp[eid], omg[eid], pha[eid], del[eid], dwt[eid], t, 0.0);
      int bsign = sign * (int)std::pow(-1, i + 1);
      bwave = bsign * gaussian_pulses(nw[eid], amp[eid], omg[eid], pha[eid], del[eid], dwt[eid], t, bPhase);
    }
    for (int pass = 0; pass < 2; pass++) {
      PlaneFill pf;
      int fid = pass == 0 ? eid : 3 + bid;
      pf.f = pass == 0 ? (void *)fs.E[eid] : (void *)fs.B[bid];  // member arrays
      pf.dim = dim;
      pf.idx = edge_plane_index(c->geo, fid, dim, sign > 0);
      pf.base = pass == 0 ? cf.E0[eid] : cf.B0[bid];
      pf.wave = has ? (pass == 0 ? ewave : bwave) : 0.0;
      pf.prof = has ? c->wprof[dim][fid] : nullptr;
      fills[nf++] = pf;
    }
  }
  fk_plane_fill_multi<T>(c, fills, nf);  // six different arrays
}

}  // namespace

template <typename T>
void refresh_fields(emo_ctx *c, FieldSet<T> &fs, double t, double dt) {
  // Q23: B through the fixed fieldPtr_ allocations
  T *f[6] = {fs.E[0], fs.E[1], fs.E[2], fs.Bfix[0], fs.Bfix[1], fs.Bfix[2]};
  // runs of consecutive periodic sides go in one launch; a LightBC side
  // between them keeps its place in the side order (its fills cover the
  // whole plane, ghosts included, and the later copies read them)
  int ns = 0, ds[6], srcs[6], dsts[6];
  for (int side = 0; side < 6; side++) {
    int d = side / 2, right = side & 1;
    if (c->cfg.fields_bound[side] == EMO_BC_PERIODIC) {
      ds[ns] = d;
      srcs[ns] = right ? 1 : c->geo.n[d];
      dsts[ns] = right ? c->geo.n[d] + 1 : 0;
      ns++;
    } else {
      fk_plane_copy_sides<T>(c, f, 6, ns, ds, srcs, dsts);
      ns = 0;
      light_bc<T>(c, fs, side, t, dt);
    }
  }
  fk_plane_copy_sides<T>(c, f, 6, ns, ds, srcs, dsts);
}

template <typename T>
void potential_bc(emo_ctx *c, T *u, const double *bound) {
  // runs of consecutive periodic sides in one launch (their sources are
  // physical points, their destinations ghost points); an external side's
  // fill keeps its place in the side order (it writes physical plane 1 on
  // the left, which later copies read)
  T *f[1] = {u};
  int ns = 0, ds[6], srcs[6], dsts[6];
  for (int side = 0; side < 6; side++) {
    int d = side / 2, right = side & 1;
    if (c->cfg.fields_bound[side] == EMO_BC_PERIODIC) {
      ds[ns] = d;
      srcs[ns] = right ? 1 : c->geo.n[d];
      dsts[ns] = right ? c->geo.n[d] + 1 : 0;
      ns++;
    } else {
      fk_plane_copy_sides<T>(c, f, 1, ns, ds, srcs, dsts);
      ns = 0;
      PlaneFill pf;
      pf.f = u;
      pf.dim = d;
      pf.idx = right ? 1 + c->geo.n[d] : 1;  // PoissonBC offset_ (poissonBC.cpp:15-20)
      pf.base = bound[side];
      pf.wave = 0.0;
      pf.prof = nullptr;
      fk_plane_fill<T>(c, pf);
    }
  }
  fk_plane_copy_sides<T>(c, f, 1, ns, ds, srcs, dsts);
}

// The transverse profile exp(-(ya^2 + yb^2) / waist^2) of the wave part on
// the planes normal to dim, for field fid (0..5 = Ex..Bz): ya, yb the
// distances of the field's staggered points (Grid::getGridPhys,
// src/grid/grid.cpp:415-440: E_i offset dx_i/2 along i, B_i offset d/2 along
// the two other axes) from the box centre, over the full padded plane.
// Computed on the host with the host's libm, once per context, so that both
// precisions' plane fills and the adjoint's use the same values.
int build_waist_profiles(emo_ctx *c) {
  const emo_config &cf = c->cfg;
  if (!(cf.waist > 0) || cf.nwaves <= 0) return EMO_OK;
  const Geo &g = c->geo;
  for (int dim = 0; dim < 3; dim++) {
    if (cf.fields_bound[2 * dim] == EMO_BC_PERIODIC && cf.fields_bound[2 * dim + 1] == EMO_BC_PERIODIC) continue;
    const int u = dim == 0 ? 1 : 0, w = dim == 2 ? 1 : 2;
    const size_t np = (size_t)g.nt[u] * g.nt[w];
    std::vector<double> h(np);
    for (int fid = 0; fid < 6; fid++) {
      const int comp = fid % 3;
      double off[3];
      for (int q = 0; q < 3; q++) {
        if (fid < 3) off[q] = q == comp ? 0.5 * g.d[q] : 0.0;
        else off[q] = q == comp ? 0.0 : 0.5 * g.d[q];
      }
      const double cu = cf.xyz0[u] + 0.5 * cf.Lxyz[u], cw = cf.xyz0[w] + 0.5 * cf.Lxyz[w];
      for (int pu = 0; pu < g.nt[u]; pu++)
        for (int pw = 0; pw < g.nt[w]; pw++) {
          const double ya = g.x0[u] + g.d[u] * (pu - 1) + off[u] - cu;
          const double yb = g.x0[w] + g.d[w] * (pw - 1) + off[w] - cw;
          h[(size_t)pu * g.nt[w] + pw] = std::exp(-(ya * ya + yb * yb) / (cf.waist * cf.waist));
        }
      if (hipMalloc((void **)&c->wprof[dim][fid], np * sizeof(double)) != hipSuccess) {
        c->wprof[dim][fid] = nullptr;
        c->err = "hipMalloc failed for the pulse profile";
        return EMO_ERR_ALLOC;
      }
      if (ctx_h2d(c, c->wprof[dim][fid], h.data(), np * sizeof(double)) != hipSuccess) {
        c->err = "hipMemcpy failed for the pulse profile";
        return EMO_ERR_HIP;
      }
    }
  }
  return EMO_OK;
}

void free_waist_profiles(emo_ctx *c) {
  for (auto &row : c->wprof)
    for (double *&p : row) {
      if (p) (void)hipFree(p);
      p = nullptr;
    }
}

template void refresh_fields<double>(emo_ctx *, FieldSet<double> &, double, double);
template void refresh_fields<float>(emo_ctx *, FieldSet<float> &, double, double);
template void potential_bc<double>(emo_ctx *, double *, const double *);
template void potential_bc<float>(emo_ctx *, float *, const double *);

}  // namespace emo
