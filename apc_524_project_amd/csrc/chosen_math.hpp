// chosen_math.hpp -- the per-particle FP32 arithmetic of the chosen form:
// gather from a cell's gather vector, relativistic (Vay) or plain Boris push
// on u = gamma v, position update in cell units with the particle BCs, move
// code.  Used by the forward kernels (chosen.hip) and, recomputed, by the
// FP32 adjoint (adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "emo_internal.hpp"
#include "movecode.hpp"

namespace emo {

constexpr int kSpeciesShift = 26;
constexpr int32_t kCellMask = (1 << kSpeciesShift) - 1;


struct SpeciesF32 {
  float qp[EMO_NSPEC];  // ((0.5 dt q)/m) UNIT_ACC
  float jc[EMO_NSPEC];  // UNIT_RHOJ q / V: current per unit velocity per particle
  int test[EMO_NSPEC];
};

inline SpeciesF32 species_f32(const emo_ctx *c, double dt) {
  SpeciesF32 s;
  double V = c->geo.d[0] * c->geo.d[1] * c->geo.d[2];
  for (int k = 0; k < EMO_NSPEC; k++) {
    double q = c->species.q[k], m = c->species.m[k];
    double qp = (0.5 * dt * q) / m;
    qp *= kUnitAcc;
    s.qp[k] = (float)qp;
    s.jc[k] = (float)(kUnitRhoJ * q / V);
    s.test[k] = c->species.test[k];
  }
  return s;
}


__device__ inline float as_f(int v) { return __int_as_float(v); }
__device__ inline int as_i(float v) { return __float_as_int(v); }

// ---------------------------------------------------------------------------
// One particle's gather + push + particle BCs.  The file is compiled without
// floating-point contraction (every fused multiply-add is an explicit fmaf),
// so the result does not depend on where the function is inlined.
// Branch-free: a wavefront's lanes take different sides of the boundary
// tests.
// ---------------------------------------------------------------------------
struct PushK {
  Geo g;
  const float4 *cv;  // per cell, the 18 gather values (k_cellvec_f32)
  SpeciesF32 s;
  float dtid[3];  // dt / dx per dim
  int lper[3], rper[3];  // periodic particle BC on the lower / upper side
  int per[3];     // periodic particle BCs per dim
  FastDiv div1, div2;  // division by nc[1], nc[2]
};

// the push constants of the context for step dt (chosen.hip)
PushK fp32_push_consts(emo_ctx *c, double dt);

struct Pushed {
  int cell;     // new cell id
  int ijk[3];   // its padded indices
  bool flip[3];  // reflected (u flipped) in dim d
  float f[3];   // new offset
  float u[3];   // new u (after the BCs)
  float v[3];   // new velocity (deposit): u[d] * ig
  float ig;     // 1 / gamma of the new u (1 non-relativistic)
  bool outside;  // the new position left the padded grid (then clamped into it)
};


constexpr int kCellVec = 5;


// padded indices of a cell id (division by multiply-high)
__device__ inline void decode_fast(const PushK &a, int cell, int &ix, int &iy, int &iz) {
  const uint32_t t = fastdiv((uint32_t)cell, a.div2);
  iz = cell - (int)(t * (uint32_t)a.g.nc[2]);
  const uint32_t x = fastdiv(t, a.div1);
  iy = (int)(t - x * (uint32_t)a.g.nc[1]);
  ix = (int)x;
}

// One dimension of the move in cell units: split the new position pf of a
// particle of padded cell ix into cell and offset, then the particle BCs of
// the two sides in the reference's order (src/particles/particle_handler.cpp:
// 287-295).  p = ix-1 is the physical cell, P = p + f the position / dx.
// Periodic (bc_p_periodic.cpp:44-87): P < 0 -> P + n; P > n -> P - n.
// Reflecting (bc_p_reflecting.cpp:43-53): mirror at 0 or n and flip u.
__device__ inline void move_dim(float pf, int ix, int n, bool lper, bool rper, int &ix_out, float &f_out,
                                float &u, bool &flipped) {
  const float fl = floorf(pf);
  float fr = pf - fl;
  int p = ix - 1 + (int)fl;
  const bool round_up = fr >= 1.f;  // rounding of a tiny negative offset
  fr = round_up ? 0.f : fr;
  p = round_up ? p + 1 : p;
  {
    const bool L = p < 0;
    const bool fz = fr == 0.f;
    const int pl = lper ? p + n : (fz ? -p : -p - 1);
    const float fl2 = (lper || fz) ? fr : 1.f - fr;
    p = L ? pl : p;
    fr = L ? fl2 : fr;
    u = (L && !lper) ? -u : u;
    flipped = L && !lper;
  }
  {
    const bool R = p > n || (p == n && fr > 0.f);
    const bool fz = fr == 0.f;
    const int pr = rper ? p - n : (fz ? 2 * n - p : 2 * n - p - 1);
    const float fr2 = (rper || fz) ? fr : 1.f - fr;
    p = R ? pr : p;
    fr = R ? fr2 : fr;
    u = (R && !rper) ? -u : u;
    flipped = flipped != (R && !rper);
  }
  ix_out = p + 1;
  f_out = fr;
}

// move_dim when both sides of the dimension are periodic
__device__ inline void move_dim_periodic(float pf, int ix, int n, int &ix_out, float &f_out) {
  const float fl = floorf(pf);
  float fr = pf - fl;
  int p = ix - 1 + (int)fl;
  const bool round_up = fr >= 1.f;
  fr = round_up ? 0.f : fr;
  p = round_up ? p + 1 : p;
  p = p < 0 ? p + n : p;
  p = (p > n || (p == n && fr > 0.f)) ? p - n : p;
  ix_out = p + 1;
  f_out = fr;
}

// The cell's gather vector (five 16-byte loads).
struct CellVec {
  float4 Ex, Ey, Ez, B, Bz;
};
__device__ inline CellVec load_cellvec(const PushK &a, int w) {
  const float4 *cv = a.cv + (long long)(w & kCellMask) * kCellVec;
  return CellVec{cv[0], cv[1], cv[2], cv[3], cv[4]};
}

// The gather (Interpolator::interpolate_fields, src/particles/interpolate.cpp:
// 11-47) from a cell's gather vector at offsets f: E bilinear over the 4
// edges, B linear over 2 faces, with the weights 2 - f and f - 1 (Q22).
__device__ __forceinline__ void gather_cell(const CellVec &V, float fx, float fy, float fz, float e[3], float b[3]) {
  const float ax = 2.f - fx, bx = fx - 1.f;
  const float ay = 2.f - fy, by = fy - 1.f;
  const float az = 2.f - fz, bz = fz - 1.f;
  const float4 vEx = V.Ex, vEy = V.Ey, vEz = V.Ez, vB = V.B, vBz = V.Bz;
  e[0] = fmaf(az, fmaf(ay, vEx.x, by * vEx.y), bz * fmaf(by, vEx.z, ay * vEx.w));
  e[1] = fmaf(ax, fmaf(az, vEy.x, bz * vEy.y), bx * fmaf(bz, vEy.z, az * vEy.w));
  e[2] = fmaf(ay, fmaf(ax, vEz.x, bx * vEz.y), by * fmaf(bx, vEz.z, ax * vEz.w));
  b[0] = fmaf(ax, vB.x, bx * vB.y);
  b[1] = fmaf(ay, vB.z, by * vB.w);
  b[2] = fmaf(az, vBz.x, bz * vBz.y);
}

// (ix, iy, iz): the padded indices of the particle's cell (every particle
// of a cell shares them, so the caller decodes the cell id once); q_p:
// ((0.5 dt q)/m) UNIT_ACC of the particle's species; g0: the stored gamma
// for the first relativistic push (Q3), if use_g0
template <bool REL, bool PER>
__device__ __forceinline__ void push_particle(const PushK &a, const CellVec &V, int ix, int iy, int iz, float fx,
                                              float fy, float fz, float ux, float uy, float uz, float q_p,
                                              bool use_g0, float g0, Pushed &o, unsigned long long *flag) {
  const Geo &g = a.g;
  float eg[3], bg[3];
  gather_cell(V, fx, fy, fz, eg, bg);
  const float ex = eg[0], ey = eg[1], ez = eg[2];
  const float bxf = bg[0], byf = bg[1], bzf = bg[2];
  float ig_new = 1.f;
  if (REL) {
    // Relativistic_Boris::Step (src/pusher/relativisticBoris.cpp:28-140) on
    // u = gamma v.  Reciprocals and square roots are the hardware's
    // single-instruction forms (about 1 ulp).
    const float ig = use_g0 ? __builtin_amdgcn_rcpf(g0)
                            : __builtin_amdgcn_rsqf(fmaf(ux, ux, fmaf(uy, uy, fmaf(uz, uz, 1.f))));
    const float vx0 = ux * ig, vy0 = uy * ig, vz0 = uz * ig;
    const float upx = fmaf(q_p, ex, fmaf(q_p, ex + fmaf(vy0, bzf, -(vz0 * byf)), ux));
    const float upy = fmaf(q_p, ey, fmaf(q_p, ey + fmaf(vz0, bxf, -(vx0 * bzf)), uy));
    const float upz = fmaf(q_p, ez, fmaf(q_p, ez + fmaf(vx0, byf, -(vy0 * bxf)), uz));
    const float taux = q_p * bxf, tauy = q_p * byf, tauz = q_p * bzf;
    const float tau_sq = fmaf(taux, taux, fmaf(tauy, tauy, tauz * tauz));
    const float sigma = fmaf(upx, upx, fmaf(upy, upy, fmaf(upz, upz, 1.f))) - tau_sq;
    const float ustar = fmaf(upx, taux, fmaf(upy, tauy, upz * tauz));
    const float root = __builtin_amdgcn_sqrtf(fmaf(sigma, sigma, 4.f * fmaf(ustar, ustar, tau_sq)));
    const float igi = __builtin_amdgcn_rsqf(0.5f * (sigma + root));
    const float tx = taux * igi, ty = tauy * igi, tz = tauz * igi;
    const float s = __builtin_amdgcn_rcpf(fmaf(tx, tx, fmaf(ty, ty, fmaf(tz, tz, 1.f))));
    const float udott = ustar * igi;
    ux = s * fmaf(udott, tx, upx + fmaf(upy, tz, -(upz * ty)));
    uy = s * fmaf(udott, ty, upy + fmaf(upz, tx, -(upx * tz)));
    uz = s * fmaf(udott, tz, upz + fmaf(upx, ty, -(upy * tx)));
    ig_new = igi;
  } else {
    // Boris::Step (src/pusher/boris.cpp:12-75)
    const float tx = q_p * bxf, ty = q_p * byf, tz = q_p * bzf;
    const float bsqr = fmaf(bxf, bxf, fmaf(byf, byf, bzf * bzf));
    const float tsqr = 2.f * __builtin_amdgcn_rcpf(fmaf(q_p * q_p, bsqr, 1.f));
    const float sx_ = tsqr * tx, sy_ = tsqr * ty, sz_ = tsqr * tz;
    const float v_mx = fmaf(q_p, ex, ux), v_my = fmaf(q_p, ey, uy), v_mz = fmaf(q_p, ez, uz);
    const float v_ix = v_mx + fmaf(v_my, tz, -(v_mz * ty));
    const float v_iy = v_my + fmaf(v_mz, tx, -(v_mx * tz));
    const float v_iz = v_mz + fmaf(v_mx, ty, -(v_my * tx));
    ux = fmaf(q_p, ex, v_mx + fmaf(v_iy, sz_, -(v_iz * sy_)));
    uy = fmaf(q_p, ey, v_my + fmaf(v_iz, sx_, -(v_ix * sz_)));
    uz = fmaf(q_p, ez, v_mz + fmaf(v_ix, sy_, -(v_iy * sx_)));
  }
  // x += dt v, in cell units, split into cell index and offset; BCs
  int jx, jy, jz;
  if (PER) {  // every particle BC periodic
    o.flip[0] = o.flip[1] = o.flip[2] = false;
    move_dim_periodic(fmaf(a.dtid[0], ux * ig_new, fx), ix, g.n[0], jx, o.f[0]);
    move_dim_periodic(fmaf(a.dtid[1], uy * ig_new, fy), iy, g.n[1], jy, o.f[1]);
    move_dim_periodic(fmaf(a.dtid[2], uz * ig_new, fz), iz, g.n[2], jz, o.f[2]);
  } else {
    move_dim(fmaf(a.dtid[0], ux * ig_new, fx), ix, g.n[0], a.lper[0], a.rper[0], jx, o.f[0], ux, o.flip[0]);
    move_dim(fmaf(a.dtid[1], uy * ig_new, fy), iy, g.n[1], a.lper[1], a.rper[1], jy, o.f[1], uy, o.flip[1]);
    move_dim(fmaf(a.dtid[2], uz * ig_new, fz), iz, g.n[2], a.lper[2], a.rper[2], jz, o.f[2], uz, o.flip[2]);
  }
  o.outside = jx < 0 || jx >= g.nc[0] || jy < 0 || jy >= g.nc[1] || jz < 0 || jz >= g.nc[2];
  if (o.outside) {
    if (flag) atomicOr(flag, kFlagOutOfGrid);
    jx = min(max(jx, 1), g.n[0]);
    jy = min(max(jy, 1), g.n[1]);
    jz = min(max(jz, 1), g.n[2]);
  }
  o.cell = (jx * g.nc[1] + jy) * g.nc[2] + jz;
  o.ijk[0] = jx;
  o.ijk[1] = jy;
  o.ijk[2] = jz;
  o.u[0] = ux;
  o.u[1] = uy;
  o.u[2] = uz;
  o.v[0] = ux * ig_new;
  o.v[1] = uy * ig_new;
  o.v[2] = uz * ig_new;
  o.ig = ig_new;
}


// ---------------------------------------------------------------------------
// Reverse mode of push_particle followed by the particle's deposit
// (Depositor::deposit_particle_J with the weights 2 - f and f - 1): given
// the adjoints of the particle's outputs -- the new offset f' and u' after
// the BCs -- and the current adjoint Jb at the grid points its deposit adds
// to, the adjoints of its inputs f, u (and of the stored gamma g0 of the
// first relativistic push) and of the 18 gather-vector entries it read
// (cvb, the order of CellVec).  The forward is recomputed from the inputs
// with the same expressions and the same single-instruction reciprocals and
// square roots, with each expression's multiply-adds fused (which changes
// only rounding, and the same way wherever the function is inlined).  Cell changes are piecewise constant: this is the derivative
// away from cell faces.  jc = 0 for particles that deposit nothing (test
// particles).
// ---------------------------------------------------------------------------
__device__ inline void cross3(const float a[3], const float b[3], float c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <bool REL, bool PER>
__device__ inline void push_deposit_vjp(const PushK &a, const CellVec &V, int w, const float fin[3],
                                        const float uin[3], float q_p, bool use_g0, float g0, float jc,
                                        const float *const Jb[3], const float fbo[3], const float ubo[3],
                                        float fb[3], float ub[3], float &g0b, float cvb[18]) {
#pragma clang fp contract(on)
  const Geo &g = a.g;
  int ic[3];
  decode_fast(a, w & kCellMask, ic[0], ic[1], ic[2]);
  // ---- forward ----
  const float A[3] = {2.f - fin[0], 2.f - fin[1], 2.f - fin[2]};
  const float Bw[3] = {fin[0] - 1.f, fin[1] - 1.f, fin[2] - 1.f};
  const float Ex[4] = {V.Ex.x, V.Ex.y, V.Ex.z, V.Ex.w}, Ey[4] = {V.Ey.x, V.Ey.y, V.Ey.z, V.Ey.w};
  const float Ez[4] = {V.Ez.x, V.Ez.y, V.Ez.z, V.Ez.w};
  const float Bf[6] = {V.B.x, V.B.y, V.B.z, V.B.w, V.Bz.x, V.Bz.y};
  float e[3], b[3];
  e[0] = A[2] * (A[1] * Ex[0] + Bw[1] * Ex[1]) + Bw[2] * (Bw[1] * Ex[2] + A[1] * Ex[3]);
  e[1] = A[0] * (A[2] * Ey[0] + Bw[2] * Ey[1]) + Bw[0] * (Bw[2] * Ey[2] + A[2] * Ey[3]);
  e[2] = A[1] * (A[0] * Ez[0] + Bw[0] * Ez[1]) + Bw[1] * (Bw[0] * Ez[2] + A[0] * Ez[3]);
  for (int d = 0; d < 3; d++) b[d] = A[d] * Bf[2 * d] + Bw[d] * Bf[2 * d + 1];
  const float u[3] = {uin[0], uin[1], uin[2]};
  float ig = 1.f, v0[3], cr[3], up[3], tau[3], tau2 = 0.f, sigma = 0.f, ustar = 0.f, root = 1.f, igi = 1.f;
  float t[3], s = 1.f, udott = 0.f, wv[3], un[3], sv[3], vm[3], vi[3], D = 1.f, tsqr = 0.f;
  if (REL) {
    ig = use_g0 ? __builtin_amdgcn_rcpf(g0) : __builtin_amdgcn_rsqf(1.f + dot3(u, u));
    for (int d = 0; d < 3; d++) v0[d] = u[d] * ig;
    cross3(v0, b, cr);
    for (int d = 0; d < 3; d++) {
      up[d] = u[d] + q_p * (e[d] + cr[d]) + q_p * e[d];
      tau[d] = q_p * b[d];
    }
    tau2 = dot3(tau, tau);
    sigma = 1.f + dot3(up, up) - tau2;
    ustar = dot3(up, tau);
    root = __builtin_amdgcn_sqrtf(sigma * sigma + 4.f * (ustar * ustar + tau2));
    igi = __builtin_amdgcn_rsqf(0.5f * (sigma + root));
    for (int d = 0; d < 3; d++) t[d] = tau[d] * igi;
    D = 1.f + dot3(t, t);
    s = __builtin_amdgcn_rcpf(D);
    udott = ustar * igi;
    float uxt[3];
    cross3(up, t, uxt);
    for (int d = 0; d < 3; d++) {
      wv[d] = up[d] + udott * t[d] + uxt[d];
      un[d] = s * wv[d];
    }
  } else {
    for (int d = 0; d < 3; d++) t[d] = q_p * b[d];
    D = 1.f + q_p * q_p * dot3(b, b);
    tsqr = 2.f * __builtin_amdgcn_rcpf(D);
    for (int d = 0; d < 3; d++) {
      sv[d] = tsqr * t[d];
      vm[d] = u[d] + q_p * e[d];
    }
    float c1[3];
    cross3(vm, t, c1);
    for (int d = 0; d < 3; d++) vi[d] = vm[d] + c1[d];
    float c2[3];
    cross3(vi, sv, c2);
    for (int d = 0; d < 3; d++) un[d] = q_p * e[d] + vm[d] + c2[d];
  }
  // move: new offsets and the BCs' reflections
  float fo[3], uo[3];
  int jc3[3];
  bool flip[3] = {false, false, false};
  for (int d = 0; d < 3; d++) {
    uo[d] = un[d];
    const float pf = fin[d] + a.dtid[d] * (un[d] * igi);
    if (PER)
      move_dim_periodic(pf, ic[d], g.n[d], jc3[d], fo[d]);
    else
      move_dim(pf, ic[d], g.n[d], a.lper[d], a.rper[d], jc3[d], fo[d], uo[d], flip[d]);
  }
  // ---- deposit^T at the new cell ----
  float fob[3] = {fbo[0], fbo[1], fbo[2]};
  float vob[3] = {0.f, 0.f, 0.f};
  const bool in_grid = jc3[0] >= 0 && jc3[0] < g.nc[0] && jc3[1] >= 0 && jc3[1] < g.nc[1] && jc3[2] >= 0 &&
                       jc3[2] < g.nc[2];
  if (jc != 0.f && in_grid) {
    // 32-bit point indices (a padded grid holds fewer than 2^31 points)
    int idx[12];
    idx[0] = gidx(g, jc3[0], jc3[1], jc3[2]);
    idx[1] = gidx(g, jc3[0], jc3[1] + 1, jc3[2]);
    idx[2] = gidx(g, jc3[0], jc3[1] + 1, jc3[2] + 1);
    idx[3] = gidx(g, jc3[0], jc3[1] + 1, jc3[2]);
    idx[4] = gidx(g, jc3[0], jc3[1], jc3[2]);
    idx[5] = gidx(g, jc3[0], jc3[1], jc3[2] + 1);
    idx[6] = gidx(g, jc3[0] + 1, jc3[1], jc3[2] + 1);
    idx[7] = gidx(g, jc3[0] + 1, jc3[1], jc3[2]);
    idx[8] = gidx(g, jc3[0], jc3[1], jc3[2]);
    idx[9] = gidx(g, jc3[0] + 1, jc3[1], jc3[2]);
    idx[10] = gidx(g, jc3[0] + 1, jc3[1] + 1, jc3[2]);
    idx[11] = gidx(g, jc3[0], jc3[1] + 1, jc3[2]);
    const float Ao[3] = {2.f - fo[0], 2.f - fo[1], 2.f - fo[2]};
    const float Bo[3] = {fo[0] - 1.f, fo[1] - 1.f, fo[2] - 1.f};
    for (int q = 0; q < 3; q++) {
      const int j = (q + 1) % 3, k = (q + 2) % 3;
      const float r0 = Jb[q][idx[4 * q]], r1 = Jb[q][idx[4 * q + 1]];
      const float r2 = Jb[q][idx[4 * q + 2]], r3 = Jb[q][idx[4 * q + 3]];
      const float vq = uo[q] * igi;
      // w = {A_j A_k, B_j A_k, B_j B_k, A_j B_k}
      vob[q] += jc * (Ao[j] * Ao[k] * r0 + Bo[j] * Ao[k] * r1 + Bo[j] * Bo[k] * r2 + Ao[j] * Bo[k] * r3);
      fob[j] += jc * vq * (-Ao[k] * r0 + Ao[k] * r1 + Bo[k] * r2 - Bo[k] * r3);
      fob[k] += jc * vq * (-Ao[j] * r0 - Bo[j] * r1 + Bo[j] * r2 + Ao[j] * r3);
    }
  }
  // ---- reverse ----
  // BCs: a reflection maps f -> 1 - f and u -> -u (so v -> -v)
  float pfb[3], unb[3], vnb[3];
  for (int d = 0; d < 3; d++) {
    const float sg = flip[d] ? -1.f : 1.f;
    pfb[d] = sg * fob[d];
    unb[d] = sg * ubo[d];
    vnb[d] = sg * vob[d] + a.dtid[d] * pfb[d];  // the deposit's v and the position update
    fb[d] = pfb[d];
  }
  float igib = 0.f;
  for (int d = 0; d < 3; d++) {
    unb[d] += vnb[d] * igi;
    igib += vnb[d] * un[d];
  }
  float eb[3] = {0.f, 0.f, 0.f}, bb[3] = {0.f, 0.f, 0.f};
  g0b = 0.f;
  if (REL) {
    // un = s w,  w = up + udott t + up x t
    const float sb = dot3(unb, wv);
    float wb[3];
    for (int d = 0; d < 3; d++) wb[d] = s * unb[d];
    float upb[3], tb[3], c1[3], c2[3];
    cross3(t, wb, c1);   // d(up x t)/dup^T wb = t x wb
    cross3(wb, up, c2);  // d(up x t)/dt^T wb = wb x up
    const float udottb = dot3(wb, t);
    for (int d = 0; d < 3; d++) {
      upb[d] = wb[d] + c1[d];
      tb[d] = udott * wb[d] + c2[d];
    }
    // s = 1 / D,  D = 1 + t.t
    const float Db = -s * s * sb;
    for (int d = 0; d < 3; d++) tb[d] += 2.f * t[d] * Db;
    // udott = ustar igi
    float ustarb = udottb * igi;
    igib += udottb * ustar;
    // t = tau igi
    float taub[3];
    for (int d = 0; d < 3; d++) {
      taub[d] = tb[d] * igi;
      igib += tb[d] * tau[d];
    }
    // igi = h^(-1/2),  h = (sigma + root) / 2
    const float hb = igib * (-0.5f) * igi * igi * igi;
    float sigmab = 0.5f * hb;
    const float rootb = 0.5f * hb;
    // root = sqrt(R),  R = sigma^2 + 4 (ustar^2 + tau2)
    const float Rb = rootb * 0.5f * __builtin_amdgcn_rcpf(root);
    sigmab += 2.f * sigma * Rb;
    ustarb += 8.f * ustar * Rb;
    float tau2b = 4.f * Rb;
    // ustar = up . tau;  sigma = 1 + up.up - tau2;  tau2 = tau.tau
    for (int d = 0; d < 3; d++) {
      upb[d] += ustarb * tau[d] + 2.f * up[d] * sigmab;
      taub[d] += ustarb * up[d];
    }
    tau2b -= sigmab;
    for (int d = 0; d < 3; d++) taub[d] += 2.f * tau[d] * tau2b;
    for (int d = 0; d < 3; d++) bb[d] += q_p * taub[d];
    // up = u + q_p (e + v0 x b) + q_p e
    float crb[3], v0b[3], c3[3], c4[3];
    for (int d = 0; d < 3; d++) {
      ub[d] = upb[d];
      eb[d] += 2.f * q_p * upb[d];
      crb[d] = q_p * upb[d];
    }
    cross3(b, crb, c3);  // v0b = b x crb
    cross3(crb, v0, c4); // bb += crb x v0
    for (int d = 0; d < 3; d++) {
      v0b[d] = c3[d];
      bb[d] += c4[d];
    }
    // v0 = u ig
    float igb = 0.f;
    for (int d = 0; d < 3; d++) {
      ub[d] += v0b[d] * ig;
      igb += v0b[d] * u[d];
    }
    if (use_g0) {
      g0b = -igb * ig * ig;
    } else {
      const float k3 = -igb * ig * ig * ig;
      for (int d = 0; d < 3; d++) ub[d] += k3 * u[d];
    }
  } else {
    // un = q_p e + vm + vi x sv
    float vmb[3], vib[3], svb[3], tb[3], c1[3], c2[3];
    cross3(sv, unb, c1);  // vib = sv x unb
    cross3(unb, vi, c2);  // svb = unb x vi
    for (int d = 0; d < 3; d++) {
      eb[d] += q_p * unb[d];
      vmb[d] = unb[d];
      vib[d] = c1[d];
      svb[d] = c2[d];
    }
    // vi = vm + vm x t
    float c5[3], c6[3];
    cross3(t, vib, c5);
    cross3(vib, vm, c6);
    for (int d = 0; d < 3; d++) {
      vmb[d] += vib[d] + c5[d];
      tb[d] = c6[d];
    }
    // vm = u + q_p e
    for (int d = 0; d < 3; d++) {
      ub[d] = vmb[d];
      eb[d] += q_p * vmb[d];
    }
    // sv = tsqr t,  tsqr = 2 / D,  D = 1 + q_p^2 b.b,  t = q_p b
    const float tsqrb = dot3(svb, t);
    for (int d = 0; d < 3; d++) tb[d] += tsqr * svb[d];
    const float Db = -tsqr * __builtin_amdgcn_rcpf(D) * tsqrb;
    for (int d = 0; d < 3; d++) bb[d] += 2.f * q_p * q_p * b[d] * Db + q_p * tb[d];
  }
  // ---- gather^T ----
  float Ab[3] = {0.f, 0.f, 0.f}, Bwb[3] = {0.f, 0.f, 0.f};
  // e[0] = A2 (A1 Ex0 + B1 Ex1) + B2 (B1 Ex2 + A1 Ex3)
  cvb[0] = eb[0] * A[2] * A[1];
  cvb[1] = eb[0] * A[2] * Bw[1];
  cvb[2] = eb[0] * Bw[2] * Bw[1];
  cvb[3] = eb[0] * Bw[2] * A[1];
  Ab[2] += eb[0] * (A[1] * Ex[0] + Bw[1] * Ex[1]);
  Bwb[2] += eb[0] * (Bw[1] * Ex[2] + A[1] * Ex[3]);
  Ab[1] += eb[0] * (A[2] * Ex[0] + Bw[2] * Ex[3]);
  Bwb[1] += eb[0] * (A[2] * Ex[1] + Bw[2] * Ex[2]);
  // e[1] = A0 (A2 Ey0 + B2 Ey1) + B0 (B2 Ey2 + A2 Ey3)
  cvb[4] = eb[1] * A[0] * A[2];
  cvb[5] = eb[1] * A[0] * Bw[2];
  cvb[6] = eb[1] * Bw[0] * Bw[2];
  cvb[7] = eb[1] * Bw[0] * A[2];
  Ab[0] += eb[1] * (A[2] * Ey[0] + Bw[2] * Ey[1]);
  Bwb[0] += eb[1] * (Bw[2] * Ey[2] + A[2] * Ey[3]);
  Ab[2] += eb[1] * (A[0] * Ey[0] + Bw[0] * Ey[3]);
  Bwb[2] += eb[1] * (A[0] * Ey[1] + Bw[0] * Ey[2]);
  // e[2] = A1 (A0 Ez0 + B0 Ez1) + B1 (B0 Ez2 + A0 Ez3)
  cvb[8] = eb[2] * A[1] * A[0];
  cvb[9] = eb[2] * A[1] * Bw[0];
  cvb[10] = eb[2] * Bw[1] * Bw[0];
  cvb[11] = eb[2] * Bw[1] * A[0];
  Ab[1] += eb[2] * (A[0] * Ez[0] + Bw[0] * Ez[1]);
  Bwb[1] += eb[2] * (Bw[0] * Ez[2] + A[0] * Ez[3]);
  Ab[0] += eb[2] * (A[1] * Ez[0] + Bw[1] * Ez[3]);
  Bwb[0] += eb[2] * (A[1] * Ez[1] + Bw[1] * Ez[2]);
  // b[d] = A_d Bf[2d] + B_d Bf[2d+1]
  for (int d = 0; d < 3; d++) {
    cvb[12 + 2 * d] = bb[d] * A[d];
    cvb[13 + 2 * d] = bb[d] * Bw[d];
    Ab[d] += bb[d] * Bf[2 * d];
    Bwb[d] += bb[d] * Bf[2 * d + 1];
  }
  for (int d = 0; d < 3; d++) fb[d] += Bwb[d] - Ab[d];
}

}  // namespace emo
