// ============================================================================
// axi_kernels.hpp — device code of the axisymmetric solver (face and cell
// kernels, dt minimum); included by axi.hip. Scheme: CPUSolver,
// src/cfd_cpu.js:7-210, in FP32.
// ============================================================================
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace axik {

// No multiply-add contraction in any axisymmetric kernel: each result is the
// rounded product / sum the source states, whatever the code's layout (the
// HIP default fuses across statements, and which pairs fuse changes with how
// the code is arranged). Against the oracle this keeps TJ, HBTF and LBTF
// within 7.7e-6 per variable after 100 iterations.
#pragma clang fp contract(off)

constexpr int kB = 256;
constexpr float kR = 287.05f;

struct AxiConsts {
  float comps[32];
  float fs_r, fs_u, fs_p;
  float cfl, hf;
  int order, visc;
  float capf;        // dt cap factor of the global (1) / dtRatio caps, read at every replay
};

struct AxiView {
  int W, H, N;
  const float *rc, *dx, *dr, *rf;
  const uint8_t *solid, *comp;
  const float *phi, *ae, *an, *wgt, *lmix, *tgt, *rate, *yrate, *mfd;
};

__device__ __forceinline__ float clampf(float v, float a, float b) { return v < a ? a : v > b ? b : v; }
__device__ __forceinline__ float minmod(float a, float b) { return a * b <= 0.f ? 0.f : (fabsf(a) < fabsf(b) ? a : b); }

struct Prim { float r, u, v, p, Y, g; };

// CPUSolver.prim for one cell (src/cfd_cpu.js:67-74)
__device__ __forceinline__ Prim prim(const float *__restrict__ Q, int N, int k) {
  Prim P;
  P.r = fmaxf(Q[k], 1e-4f);
  P.u = Q[N + k] / P.r; P.v = Q[2 * N + k] / P.r;
  P.Y = clampf(Q[4 * N + k] / P.r, 0.f, 1.f);
  P.g = 1.4f - 0.07f * P.Y;
  P.p = fmaxf(1.f, (P.g - 1.f) * (Q[3 * N + k] - 0.5f * P.r * (P.u * P.u + P.v * P.v)));
  return P;
}

// hllc (src/cfd_cpu.js:7-31); states (rho, un, ut, p, Y) with gamma each side
__device__ void hllc(float rL, float uL, float vL, float pL, float YL, float gL, float rR, float uR, float vR, float pR, float YR,
                     float gR, float *F) {
  const float cL = sqrtf(gL * pL / rL), cR = sqrtf(gR * pR / rR);
  const float SL = fminf(uL - cL, uR - cR), SR = fmaxf(uL + cL, uR + cR);
  const float EL = pL / (gL - 1.f) + 0.5f * rL * (uL * uL + vL * vL), ER = pR / (gR - 1.f) + 0.5f * rR * (uR * uR + vR * vR);
  if (SL >= 0.f) { F[0] = rL * uL; F[1] = rL * uL * uL + pL; F[2] = rL * uL * vL; F[3] = uL * (EL + pL); F[4] = rL * uL * YL; return; }
  if (SR <= 0.f) { F[0] = rR * uR; F[1] = rR * uR * uR + pR; F[2] = rR * uR * vR; F[3] = uR * (ER + pR); F[4] = rR * uR * YR; return; }
  const float SM = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR));
  const bool left = SM >= 0.f;
  const float rK = left ? rL : rR, uK = left ? uL : uR, vK = left ? vL : vR, pK = left ? pL : pR, YK = left ? YL : YR;
  const float EK = left ? EL : ER, S = left ? SL : SR;
  const float f = rK * (S - uK) / (S - SM);
  const float Es = f * (EK / rK + (SM - uK) * (SM + pK / (rK * (S - uK))));
  F[0] = rK * uK + S * (f - rK);
  F[1] = rK * uK * uK + pK + S * (f * SM - rK * uK);
  F[2] = rK * uK * vK + S * (f * vK - rK * vK);
  F[3] = uK * (EK + pK) + S * (Es - EK);
  F[4] = rK * uK * YK + S * (f * YK - rK * YK);
}

// farfield ghost (src/cfd_cpu.js:35-48), outward normal (nx, nr); G = rho, u, v, p, Y
__device__ void farfield(const AxiConsts &c, const Prim &P, float nx, float nr, float *G) {
  const float un = P.u * nx + P.v * nr, cs = sqrtf(P.g * P.p / P.r);
  const float g0 = 1.4f, c0 = sqrtf(g0 * c.fs_p / c.fs_r), un0 = c.fs_u * nx;
  if (un >= cs) { G[0] = P.r; G[1] = P.u; G[2] = P.v; G[3] = P.p; G[4] = P.Y; return; }
  if (un0 <= -c0) { G[0] = c.fs_r; G[1] = c.fs_u; G[2] = 0.f; G[3] = c.fs_p; G[4] = 0.f; return; }
  const float Rp = un + 2.f * cs / (P.g - 1.f), Rm = un0 - 2.f * c0 / (g0 - 1.f);
  const float unb = 0.5f * (Rp + Rm);
  float gb, s, ut, vt, Yb;
  if (unb > 0.f) { gb = P.g; s = P.p / powf(P.r, P.g); ut = P.u - un * nx; vt = P.v - un * nr; Yb = P.Y; }
  else { gb = g0; s = c.fs_p / powf(c.fs_r, g0); ut = c.fs_u - un0 * nx; vt = -un0 * nr; Yb = 0.f; }
  const float cb = fmaxf(1.f, 0.25f * (gb - 1.f) * (Rp - Rm));
  const float rb = powf(cb * cb / (gb * s), 1.f / (gb - 1.f));
  G[0] = rb; G[1] = ut + unb * nx; G[2] = vt + unb * nr; G[3] = rb * cb * cb / gb; G[4] = Yb;
}

__device__ __forceinline__ void sel(const Prim &P, int dir, float *o) {
  o[0] = P.r; o[1] = dir ? P.v : P.u; o[2] = dir ? P.u : P.v; o[3] = P.p; o[4] = P.Y;
}

// MUSCL face states (recon, src/cfd_cpu.js:86-94). hm / hp: the outer cell on
// that side takes part (both full cells and the face inside the grid); the
// outer primitives are always passed in, and every choice is a value select,
// so the caller's loads of them are issued with all the others.
__device__ __forceinline__ void recon(const Prim &A, const Prim &Am, bool hm, const Prim &B, const Prim &Bp, bool hp, int dir,
                                      bool second, float *L, float *R) {
  float A5[5], B5[5], M5[5], P5[5];
  sel(A, dir, A5); sel(B, dir, B5); sel(Am, dir, M5); sel(Bp, dir, P5);
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const float l = A5[q] + 0.5f * minmod(A5[q] - M5[q], B5[q] - A5[q]);
    const float r = B5[q] - 0.5f * minmod(B5[q] - A5[q], P5[q] - B5[q]);
    L[q] = second && hm ? l : A5[q];
    R[q] = second && hp ? r : B5[q];
  }
  if (!second) return;
  const bool badL = L[0] <= 0.f || L[3] <= 0.f, badR = R[0] <= 0.f || R[3] <= 0.f;
#pragma unroll
  for (int q = 0; q < 5; q++) { L[q] = badL ? A5[q] : L[q]; R[q] = badR ? B5[q] : R[q]; }
}

// x face between cells a = f - 1 and b = f of one row (src/cfd_cpu.js:98-113):
// the aperture-weighted flux (0 where the face is closed). Pm / Pp are the
// outer cells a - 1 / b + 1; every cell index is clamped into the row by the
// caller, and the tests here are value selects on already-loaded data.
__device__ __forceinline__ void x_face(const AxiConsts &c, int W, int f, const Prim &Pm, const Prim &Pa, const Prim &Pb, const Prim &Pp,
                                       float pm, float pa, float pb, float pp, bool sa, bool sb, float aea, float *out) {
  const int a = f - 1, b = f;
  const bool second = c.order == 2;
  float F[5], G[5], L[5], R[5];
  const float al = a < 0 ? (sb ? 0.f : 1.f) : b >= W ? (sa ? 0.f : 1.f) : aea;
  if (a < 0 || b >= W) {                 // farfield ghost at x = lo / x = hi (two faces per row)
    const Prim &P = a < 0 ? Pb : Pa;
    farfield(c, P, a < 0 ? -1.f : 1.f, 0.f, G);
    if (a < 0) hllc(G[0], G[1], G[2], G[3], G[4], 1.4f - 0.07f * G[4], Pb.r, Pb.u, Pb.v, Pb.p, Pb.Y, Pb.g, F);
    else hllc(Pa.r, Pa.u, Pa.v, Pa.p, Pa.Y, Pa.g, G[0], G[1], G[2], G[3], G[4], 1.4f - 0.07f * G[4], F);
  } else {
    const bool fa = pa > 0.999f && pb > 0.999f;
    const bool hm = fa && a - 1 >= 0 && pm > 0.999f, hp = fa && b + 1 < W && pp > 0.999f;
    recon(Pa, Pm, hm, Pb, Pp, hp, 0, second, L, R);
    hllc(L[0], L[1], L[2], L[3], L[4], 1.4f - 0.07f * L[4], R[0], R[1], R[2], R[3], R[4], 1.4f - 0.07f * R[4], F);
  }
  const bool open = al > 0.f;
#pragma unroll
  for (int q = 0; q < 5; q++) out[q] = open ? al * F[q] : 0.f;
}

// r face between rows a = f - 1 and b = f of one column (src/cfd_cpu.js:115-143):
// out = the stored flux in (rho, rho u, rho v, E, rho Y) order with the
// mixing-length shear, heat and species fluxes, *nut its eddy viscosity.
__device__ __forceinline__ void r_face(const AxiConsts &c, int H, int f, const Prim &Pm, const Prim &Pa, const Prim &Pb, const Prim &Pp,
                                       float pm, float pa, float pb, float pp, bool sa, float ana, float la, float lb, float rca,
                                       float rcb, float *out, float *nutp) {
  const int a = f - 1, b = f;
  const bool second = c.order == 2;
  float F[5], G[5], L[5], R[5];
  float G2[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  float nut = 0.f;
  const float al = a < 0 ? 0.f : b >= H ? (sa ? 0.f : 1.f) : ana;
  if (b >= H) {                            // farfield ghost at r = top
    farfield(c, Pa, 0.f, 1.f, G);
    hllc(Pa.r, Pa.v, Pa.u, Pa.p, Pa.Y, Pa.g, G[0], G[2], G[1], G[3], G[4], 1.4f - 0.07f * G[4], F);
  } else {
    const bool fa = pa > 0.999f && pb > 0.999f;
    const bool hm = fa && a - 1 >= 0 && pm > 0.999f, hp = fa && b + 1 < H && pp > 0.999f;
    recon(Pa, Pm, hm, Pb, Pp, hp, 1, second, L, R);
    hllc(L[0], L[1], L[2], L[3], L[4], 1.4f - 0.07f * L[4], R[0], R[1], R[2], R[3], R[4], 1.4f - 0.07f * R[4], F);
    if (c.visc) {            // mixing-length shear, heat and species fluxes (src/cfd_cpu.js:129-140)
      const float drc = rcb - rca;
      const float dudr = (Pb.u - Pa.u) / drc;
      const float l = 0.5f * (la + lb);
      nut = l * l * fabsf(dudr);
      const float rho = 0.5f * (Pa.r + Pb.r), mu = rho * nut;
      const float Ta = Pa.p / (Pa.r * kR), Tb = Pb.p / (Pb.r * kR);
      const float ga = 0.5f * (Pa.g + Pb.g), cp = ga * kR / (ga - 1.f);
      const float tau = mu * dudr;
      G2[1] = tau;
      G2[3] = 0.5f * (Pa.u + Pb.u) * tau + mu * cp / 0.9f * (Tb - Ta) / drc;
      G2[4] = mu / 0.9f * (Pb.Y - Pa.Y) / drc;
    }
  }
  const bool open = al > 0.f;
  out[0] = open ? al * F[0] : 0.f;
  out[1] = open ? al * (F[2] - G2[1]) : 0.f;
  out[2] = open ? al * F[1] : 0.f;
  out[3] = open ? al * (F[3] - G2[3]) : 0.f;
  out[4] = open ? al * (F[4] - G2[4]) : 0.f;
  *nutp = open ? nut : 0.f;
}

// One thread per face (src/cfd_cpu.js:97-143). Every load of a face (the four
// cells' states and volume fractions, the face's aperture, the solid flags) is
// issued before any of them is tested, at indices clamped into the grid: one
// memory round trip per face instead of one per decision. Closed faces compute
// a flux too and store 0 through a select.
__global__ void __launch_bounds__(kB) face_kernel(AxiView v, const AxiConsts *__restrict__ cp, const float *__restrict__ Q, float *__restrict__ Fx,
                                                  float *__restrict__ Fr, float *__restrict__ nutF) {
  const AxiConsts &c = *cp;
  const int W = v.W, H = v.H, N = v.N;
  const int NXF = (W + 1) * H, NRF = W * (H + 1);
  const int t = blockIdx.x * kB + threadIdx.x;
  float F[5];
  if (t < NXF) {
    const int j = t / (W + 1), f = t % (W + 1), a = f - 1, b = f;
    // cells a-1, a, b, b+1 of row j, clamped into the row
    const int ia = a < 0 ? 0 : a, ib = b >= W ? W - 1 : b, im = ia > 0 ? ia - 1 : 0, ip = ib + 1 < W ? ib + 1 : W - 1;
    const int ka = j * W + ia, kb = j * W + ib, km = j * W + im, kp = j * W + ip;
    const Prim Pa = prim(Q, N, ka), Pb = prim(Q, N, kb), Pm = prim(Q, N, km), Pp = prim(Q, N, kp);
    x_face(c, W, f, Pm, Pa, Pb, Pp, v.phi[km], v.phi[ka], v.phi[kb], v.phi[kp], v.solid[ka] != 0, v.solid[kb] != 0, v.ae[ka], F);
#pragma unroll
    for (int q = 0; q < 5; q++) Fx[q * NXF + t] = F[q];
    return;
  }
  const int u = t - NXF;
  if (u >= NRF) return;
  const int f = u / W, i = u % W, a = f - 1, b = f;
  // cells a-1, a, b, b+1 of column i, clamped into the column
  const int ja = a < 0 ? 0 : a, jb = b >= H ? H - 1 : b, jm = ja > 0 ? ja - 1 : 0, jp = jb + 1 < H ? jb + 1 : H - 1;
  const int ka = ja * W + i, kb = jb * W + i, km = jm * W + i, kp = jp * W + i;
  const Prim Pa = prim(Q, N, ka), Pb = prim(Q, N, kb), Pm = prim(Q, N, km), Pp = prim(Q, N, kp);
  float nut;
  r_face(c, H, f, Pm, Pa, Pb, Pp, v.phi[km], v.phi[ka], v.phi[kb], v.phi[kp], v.solid[ka] != 0, v.an[ka], v.lmix[ka], v.lmix[kb],
         v.rc[ja], v.rc[jb], F, &nut);
#pragma unroll
  for (int q = 0; q < 5; q++) Fr[q * NRF + u] = F[q];
  nutF[u] = nut;
}

// Everything a cell update reads besides its face fluxes.
struct CellIn {
  float q0[5], qb[5];
  float dx, dr, ax0, ax1, rcj, ph, aek, aem, ank, anm, mfd, wgt, tgtk, rate, yrate, dts;
  int id;
  bool sol;
};
__device__ __forceinline__ void load_cell(const AxiView &v, const float *__restrict__ Qin, const float *Qb, const float *__restrict__ dt,
                                          int mode, int k, int i, int j, CellIn &x) {
  const int N = v.N, W = v.W;
  const float *Qbb = mode == 1 ? Qb : Qin;
#pragma unroll
  for (int q = 0; q < 5; q++) { x.q0[q] = Qin[q * N + k]; x.qb[q] = Qbb[q * N + k]; }
  x.sol = v.solid[k] != 0;
  x.dx = v.dx[i]; x.dr = v.dr[j]; x.ax0 = v.rf[j]; x.ax1 = v.rf[j + 1]; x.rcj = v.rc[j]; x.ph = v.phi[k];
  x.aek = v.ae[k]; x.aem = v.ae[i > 0 ? k - 1 : k]; x.ank = v.an[k]; x.anm = v.an[j > 0 ? k - W : k];
  x.id = v.comp[k];
  x.mfd = v.mfd[k]; x.wgt = v.wgt[k]; x.tgtk = v.tgt[k]; x.rate = v.rate[k]; x.yrate = v.yrate[k];
  x.dts = dt[k];
}

// Cell update (src/cfd_cpu.js:145-192). mode 0: stage 1 with dt, 1: stage 2,
// 2: dt only, 3: stage 1 with dt capped at *cap. fxa / fxb: x faces i and
// i + 1, fra / frb: r faces j and j + 1, nu0 / nu1 their eddy viscosities.
// Solid cells keep their state (modes 0 and 3 copy it to the output) and
// write nothing else.
__device__ __forceinline__ void cell_update(const AxiView &v, const AxiConsts &c, int k, int i, int j, const CellIn &x, const float *fxa,
                                            const float *fxb, const float *fra, const float *frb, float nu0, float nu1, int mode,
                                            const float *__restrict__ cap, float *Qout, float *__restrict__ dt, float *__restrict__ src,
                                            float *__restrict__ mfxo) {
  const int N = v.N;
  const float *q0 = x.q0, *qb = x.qb;
  if (x.sol) {
    if (mode == 0 || mode == 3)
#pragma unroll
      for (int q = 0; q < 5; q++) Qout[q * N + k] = q0[q];
    return;
  }
  const float dx = x.dx, dr = x.dr, ax0 = x.ax0, ax1 = x.ax1, ph = x.ph;
  Prim P;
  P.r = fmaxf(q0[0], 1e-4f);
  P.u = q0[1] / P.r; P.v = q0[2] / P.r;
  P.Y = clampf(q0[4] / P.r, 0.f, 1.f);
  P.g = 1.4f - 0.07f * P.Y;
  P.p = fmaxf(1.f, (P.g - 1.f) * (q0[3] - 0.5f * P.r * (P.u * P.u + P.v * P.v)));
  float d = x.dts;
  if (mode == 0 || mode == 2) {
    const float nu = fmaxf(nu0, nu1);
    const float cs = sqrtf(P.g * P.p / P.r);
    d = c.cfl * fmaxf(ph, 0.25f) / ((fabsf(P.u) + cs) / dx + (fabsf(P.v) + cs) / dr + 2.2f * nu / (dr * dr));
    dt[k] = d;
    if (mode == 2) return;
  }
  const float vol = x.rcj * dr;
  float R[5];
#pragma unroll
  for (int q = 0; q < 5; q++) R[q] = (-(fxb[q] - fxa[q]) / dx - (frb[q] * ax1 - fra[q] * ax0) / vol) / ph;
  const float aw = i > 0 ? x.aem : 1.f, as = j > 0 ? x.anm : 0.f;
  R[1] += P.p * (x.aek - aw) / (dx * ph);
  R[2] += P.p * (x.ank * ax1 - as * ax0) / (vol * ph);
  const float mfx = 0.5f * (fxa[0] + fxb[0]);
  float sk = 0.f;
  const int id = x.id;
  if (id) {
    const float rho = q0[0], g = P.g, cp = g * kR / (g - 1.f);
    const float T = P.p / (rho * kR), Tt = T + 0.5f * (P.u * P.u + P.v * P.v) / cp;
    const float ty = c.comps[id * 4 + 2], mf = mfx / ph;
    if (ty == 1.f || ty == 2.f) {     // blade row: tau^(wgt dx) on the upstream face enthalpy
      const float e = c.comps[id * 4 + 1], fac = ty == 1.f ? 1.f : clampf(mf / (0.3f * x.mfd), 0.f, 1.f);
      const float mw = fxa[0], HU = mw > 1e-9f ? fxa[3] / mw : cp * Tt;
      const float w = (HU > 0.f ? HU : cp * Tt) * expm1f(c.comps[id * 4] * x.wgt * fac * dx) / dx;
      R[1] += ty == 1.f ? e * rho * w : rho * w / e;
      R[3] += mf * w;
      sk = mf * w;
    } else if (x.rate > 0.f) {         // heat source relaxed to the target total temperature
      const float K = c.comps[id * 4 + 1] != 0.f ? c.comps[id * 4 + 1] : 1.f;
      const float tgt = c.comps[id * 4] + (x.tgtk - c.comps[id * 4]) * K;
      const float m0 = fmaxf(mf, 0.f);
      const float q = m0 * cp * fmaxf(0.f, tgt - Tt) * x.rate;
      R[3] += q;
      const float mfuel = q / c.hf;
      R[0] += mfuel;
      sk = mfuel;
      R[4] += m0 * (1.f - P.Y) * x.yrate + mfuel;
    }
  }
  src[k] = sk; mfxo[k] = mfx;
  if (mode == 3) { d = fminf(d, *cap); dt[k] = d; }
  float n[5];
  if (mode == 1) for (int q = 0; q < 5; q++) n[q] = 0.5f * qb[q] + 0.5f * (q0[q] + d * R[q]);
  else for (int q = 0; q < 5; q++) n[q] = q0[q] + d * R[q];
  // floors (CPUSolver.fix, src/cfd_cpu.js:187-192)
  const float r = fmaxf(n[0], 1e-3f);
  n[0] = r;
  const float Y = clampf(n[4] / r, 0.f, 1.f);
  n[4] = Y * r;
  const float g = 1.4f - 0.07f * Y, ke = 0.5f * (n[1] * n[1] + n[2] * n[2]) / r;
  if ((g - 1.f) * (n[3] - ke) < 50.f) n[3] = 50.f / (g - 1.f) + ke;
  for (int q = 0; q < 5; q++) Qout[q * N + k] = n[q];
}

// One thread per cell (src/cfd_cpu.js:145-192) over the stored face fluxes.
// The cell's loads (state, face fluxes, geometry, zone data, the stored dt)
// are issued before the solid flag is tested.
__global__ void __launch_bounds__(kB) cell_kernel(AxiView v, const AxiConsts *__restrict__ cp, const float *__restrict__ Qin, const float *Qb,
                                                  float *Qout, const float *__restrict__ Fx, const float *__restrict__ Fr,
                                                  const float *__restrict__ nutF, float *__restrict__ dt, float *__restrict__ src,
                                                  float *__restrict__ mfxo, int mode, const float *__restrict__ cap) {
  const AxiConsts &c = *cp;
  const int W = v.W, N = v.N;
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= N) return;
  const int NXF = (W + 1) * v.H, NRF = W * (v.H + 1);
  const int i = k % W, j = k / W;
  const int fx0 = j * (W + 1) + i, fx1 = fx0 + 1, fr0 = j * W + i, fr1 = fr0 + W;
  CellIn x;
  load_cell(v, Qin, Qb, dt, mode, k, i, j, x);
  float fxa[5], fxb[5], fra[5], frb[5];
#pragma unroll
  for (int q = 0; q < 5; q++) { fxa[q] = Fx[q * NXF + fx0]; fxb[q] = Fx[q * NXF + fx1]; fra[q] = Fr[q * NRF + fr0]; frb[q] = Fr[q * NRF + fr1]; }
  cell_update(v, c, k, i, j, x, fxa, fxb, fra, frb, nutF[fr0], nutF[fr1], mode, cap, Qout, dt, src, mfxo);
}

// ---------------------------------------------------------------------------
// Fused stage: one block of kFT threads per tile of kTW x kTH cells does the
// faces and the cell updates of its tile. Phase 1: every cell of the tile and
// its 2-cell ring (indices clamped into the grid, as the face kernel clamps)
// is loaded once, all of a thread's loads issued before any is used, and its
// primitives, volume fraction, solid flag, apertures and mixing length go to
// LDS. Phase 2: the tile's x and r faces from LDS into LDS (a face on a tile's
// edge is evaluated by both tiles, the same bits). Phase 3: every cell of the
// tile updates from its four faces; its own loads were issued at the start.
// Same arithmetic as face_kernel + cell_kernel (the same functions, no
// multiply-add contraction), without the face fluxes' round trip through
// memory and with one launch per stage instead of two. kMode 0: stage 1 with
// dt, 1: stage 2.
// ---------------------------------------------------------------------------
// A tile row has kTW + 1 = 64 x faces and a face row kTW r faces: each is one
// wavefront's work, so every wave evaluates faces of one kind at a time.
// Tile height TH and waves per block NW are template parameters.
constexpr int kTW = 63, kPW = kTW + 4;
template <int TH>
struct AxiTile {
  static constexpr int PH = TH + 4;
  float P[6][PH][kPW];       // r u v p Y g of cells (j0 - 2 .. j0 + TH + 1) x (i0 - 2 .. i0 + kTW + 1), indices clamped into the grid
  float phi[PH][kPW], ae[PH][kPW], an[PH][kPW], lmix[PH][kPW];
  uint8_t sol[PH][kPW];
  float rc[PH];
  float Fx[5][TH][kTW + 1];  // x faces i0 .. i0 + kTW of the tile's rows
  float Fr[5][TH + 1][kTW];  // r faces j0 .. j0 + TH of the tile's columns
  float nut[TH + 1][kTW];
};
template <int TH>
__device__ __forceinline__ Prim lds_prim(const AxiTile<TH> &t, int r, int c) {
  Prim P;
  P.r = t.P[0][r][c]; P.u = t.P[1][r][c]; P.v = t.P[2][r][c]; P.p = t.P[3][r][c]; P.Y = t.P[4][r][c]; P.g = t.P[5][r][c];
  return P;
}
// prim() on already-loaded conserved values
__device__ __forceinline__ Prim prim_of(const float *q) {
  Prim P;
  P.r = fmaxf(q[0], 1e-4f);
  P.u = q[1] / P.r; P.v = q[2] / P.r;
  P.Y = clampf(q[4] / P.r, 0.f, 1.f);
  P.g = 1.4f - 0.07f * P.Y;
  P.p = fmaxf(1.f, (P.g - 1.f) * (q[3] - 0.5f * P.r * (P.u * P.u + P.v * P.v)));
  return P;
}
// true in every lane when no lane of the wave has an open face (the wave skips the flux)
__device__ __forceinline__ bool wave_all_closed(bool open) {
#ifdef __AMDGCN__
  return __builtin_amdgcn_ballot_w64(open) == 0;
#else
  return !open;
#endif
}

// The block's work on tile t (LDS, passed in so a CPU build can share it between threads).
template <int kMode, int TH, int NW>
__device__ __forceinline__ void fused_tile(const AxiView &v, const AxiConsts *__restrict__ cp, const float *__restrict__ Qin,
                                           const float *Qb, float *Qout, float *__restrict__ dt, float *__restrict__ src,
                                           float *__restrict__ mfxo, int ntx, AxiTile<TH> &t) {
  constexpr int FT = 64 * NW, PH = TH + 4, NP = kPW * PH, NU = 2 * TH + 1;
  constexpr int PR = (NP + FT - 1) / FT, FR = (NU + NW - 1) / NW, CR = (TH + NW - 1) / NW;
  const AxiConsts &c = *cp;
  const int W = v.W, H = v.H, N = v.N;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int i0 = (blockIdx.x % ntx) * kTW, j0 = (blockIdx.x / ntx) * TH;
  // phase 1: every load of this thread's ring cells, then their LDS entries
  float q[PR][5], ph[PR], ae[PR], an[PR], lm[PR];
  uint8_t so[PR];
#pragma unroll
  for (int m = 0; m < PR; m++) {
    const int e = tid + m * FT, ee = e < NP ? e : 0;
    const int r = ee / kPW, cc = ee % kPW;
    const int gi = min(max(i0 - 2 + cc, 0), W - 1), gj = min(max(j0 - 2 + r, 0), H - 1), kk = gj * W + gi;
#pragma unroll
    for (int z = 0; z < 5; z++) q[m][z] = Qin[z * N + kk];
    ph[m] = v.phi[kk]; ae[m] = v.ae[kk]; an[m] = v.an[kk]; lm[m] = v.lmix[kk]; so[m] = v.solid[kk];
  }
  const float rcr = tid < PH ? v.rc[min(max(j0 - 2 + tid, 0), H - 1)] : 0.f;
  // this thread's own cells (lane < kTW, rows wave, wave + NW, ...): their loads go out now, the latency overlaps the face work
  CellIn x[CR];
#pragma unroll
  for (int m = 0; m < CR; m++) {
    const int ci = i0 + lane, cj = j0 + wave + m * NW;
    const bool own = lane < kTW && wave + m * NW < TH && ci < W && cj < H;
    load_cell(v, Qin, Qb, dt, kMode, own ? cj * W + ci : 0, own ? ci : 0, own ? cj : 0, x[m]);
  }
#pragma unroll
  for (int m = 0; m < PR; m++) {
    const int e = tid + m * FT;
    if (e >= NP) continue;
    const int r = e / kPW, cc = e % kPW;
    const Prim P = prim_of(q[m]);
    t.P[0][r][cc] = P.r; t.P[1][r][cc] = P.u; t.P[2][r][cc] = P.v; t.P[3][r][cc] = P.p; t.P[4][r][cc] = P.Y; t.P[5][r][cc] = P.g;
    t.phi[r][cc] = ph[m]; t.ae[r][cc] = ae[m]; t.an[r][cc] = an[m]; t.lmix[r][cc] = lm[m]; t.sol[r][cc] = so[m];
  }
  if (tid < PH) t.rc[tid] = rcr;
  __syncthreads();
  // phase 2: face units, one per wave at a time: x faces of tile row u (u < TH), r faces of face row u - TH.
  // A wave whose faces are all closed stores zeros without evaluating a flux (the face functions give 0 there too).
#pragma unroll
  for (int m = 0; m < FR; m++) {
    const int u0 = wave + m * NW;
    float F[5];
    if (u0 < TH) {
      const int tr = u0, fl = lane, f = i0 + fl, j = j0 + tr;
      const int r = tr + 2, ca = fl + 1;                 // LDS column of cell a = f - 1
      const bool in = j < H && f <= W;
      const float al = f - 1 < 0 ? (t.sol[r][ca + 1] ? 0.f : 1.f) : f >= W ? (t.sol[r][ca] ? 0.f : 1.f) : t.ae[r][ca];
      if (wave_all_closed(in && al > 0.f)) {
        if (in)
#pragma unroll
          for (int z = 0; z < 5; z++) t.Fx[z][tr][fl] = 0.f;
      } else if (in) {
        x_face(c, W, f, lds_prim(t, r, ca - 1), lds_prim(t, r, ca), lds_prim(t, r, ca + 1), lds_prim(t, r, ca + 2), t.phi[r][ca - 1],
               t.phi[r][ca], t.phi[r][ca + 1], t.phi[r][ca + 2], t.sol[r][ca] != 0, t.sol[r][ca + 1] != 0, t.ae[r][ca], F);
#pragma unroll
        for (int z = 0; z < 5; z++) t.Fx[z][tr][fl] = F[z];
      }
    } else if (u0 < NU) {
      const int tr = u0 - TH, cl = lane, f = j0 + tr, i = i0 + cl;
      const int col = cl < kTW ? cl + 2 : 2, ra = tr + 1;   // LDS row of cell a = f - 1
      const bool in = cl < kTW && i < W && f <= H;
      const float al = f - 1 < 0 ? 0.f : f >= H ? (t.sol[ra][col] ? 0.f : 1.f) : t.an[ra][col];
      if (wave_all_closed(in && al > 0.f)) {
        if (in) {
#pragma unroll
          for (int z = 0; z < 5; z++) t.Fr[z][tr][cl] = 0.f;
          t.nut[tr][cl] = 0.f;
        }
      } else if (in) {
        float nut;
        r_face(c, H, f, lds_prim(t, ra - 1, col), lds_prim(t, ra, col), lds_prim(t, ra + 1, col), lds_prim(t, ra + 2, col),
               t.phi[ra - 1][col], t.phi[ra][col], t.phi[ra + 1][col], t.phi[ra + 2][col], t.sol[ra][col] != 0, t.an[ra][col],
               t.lmix[ra][col], t.lmix[ra + 1][col], t.rc[ra], t.rc[ra + 1], F, &nut);
#pragma unroll
        for (int z = 0; z < 5; z++) t.Fr[z][tr][cl] = F[z];
        t.nut[tr][cl] = nut;
      }
    }
  }
  __syncthreads();
  // phase 3: the cell updates
#pragma unroll
  for (int m = 0; m < CR; m++) {
    const int tx = lane, ty = wave + m * NW, ci = i0 + tx, cj = j0 + ty;
    if (!(tx < kTW && ty < TH && ci < W && cj < H)) continue;
    float fxa[5], fxb[5], fra[5], frb[5];
#pragma unroll
    for (int z = 0; z < 5; z++) { fxa[z] = t.Fx[z][ty][tx]; fxb[z] = t.Fx[z][ty][tx + 1]; fra[z] = t.Fr[z][ty][tx]; frb[z] = t.Fr[z][ty + 1][tx]; }
    cell_update(v, c, cj * W + ci, ci, cj, x[m], fxa, fxb, fra, frb, t.nut[ty][tx], t.nut[ty + 1][tx], kMode, nullptr, Qout, dt, src, mfxo);
  }
}
template <int kMode, int TH, int NW>
__global__ void __launch_bounds__(64 * NW) fused_kernel(AxiView v, const AxiConsts *__restrict__ cp, const float *__restrict__ Qin,
                                                        const float *Qb, float *Qout, float *__restrict__ dt, float *__restrict__ src,
                                                        float *__restrict__ mfxo, int ntx) {
  __shared__ AxiTile<TH> t;
  fused_tile<kMode, TH, NW>(v, cp, Qin, Qb, Qout, dt, src, mfxo, ntx, t);
}

// Minimum dt over fluid cells (CPUSolver.step global / dtRatio cap, src/cfd_cpu.js:198-202).
__global__ void __launch_bounds__(kB) dtmin_kernel(const float *__restrict__ dt, const uint8_t *__restrict__ solid, int N,
                                                   float *__restrict__ partial) {
  __shared__ float lds[kB / 64];
  float m = INFINITY;
  for (int k = blockIdx.x * kB + threadIdx.x; k < N; k += gridDim.x * kB)
    if (!solid[k]) m = fminf(m, dt[k]);
  for (int off = 32; off > 0; off >>= 1) m = fminf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = lds[0];
    for (int w = 1; w < kB / 64; w++) a = fminf(a, lds[w]);
    partial[blockIdx.x] = a;
  }
}
// The factor comes from the device constants, so a captured graph replays the
// current global / dtRatio choice.
__global__ void dtcap_kernel(const float *__restrict__ partial, int nb, const AxiConsts *__restrict__ cp, float *__restrict__ cap) {
  if (threadIdx.x != 0) return;
  float a = INFINITY;
  for (int b = 0; b < nb; b++) a = fminf(a, partial[b]);
  *cap = a * cp->capf;
}


// ---------------------------------------------------------------------------
// Readback monitors in FP64: what the application's convergence loop reads
// back after each chunk of iterations (src/app_cfd.js:361-404).
//   monitor_kernel  thrust parts of cfdForces with the solver's own sources
//                   (src/cfd_post.js:44-75), the density residual of the
//                   readback (src/app_cfd.js:368-371) and the count of
//                   non-physical cells (src/app_cfd.js:377): block partial sums.
//   section_kernel  the sums of cfdSection (src/cfd_post.js:83-102) at
//                   sections (the control stations, station averages), one
//                   thread per section, rows in the host's order.
// Per-cell terms are the host's double operations in the host's order
// (contraction off, NaN-propagating max / min / clamp as in JavaScript); only
// the order of the grid-wide sums differs from the host's sequential loop.
// ---------------------------------------------------------------------------
constexpr int kMonK = 14;         // Fb, Fw, mfuel, pw[0..7], sum of d^2, count, non-physical cells
constexpr int kMonMaxSec = 4;
constexpr int kMonBlocks = 512;
constexpr double kPi = 3.141592653589793;
constexpr double kRGas = 287.05;

struct MonIn {
  double comps[32];
  double P0;
  int prev_valid;                 // the density snapshot holds the previous readback
};

__device__ __forceinline__ double jmax(double a, double b) { return (a != a || b != b) ? (a != a ? a : b) : (a > b ? a : b); }
__device__ __forceinline__ double jmin(double a, double b) { return (a != a || b != b) ? (a != a ? a : b) : (a < b ? a : b); }
__device__ __forceinline__ double jclamp(double v, double a, double b) { return v < a ? a : v > b ? b : v; }

struct PrimD { double r, u, v, p, Y, g; };
// cfdPrim (src/cfd_post.js:4-8) on the FP32 state, in doubles
__device__ __forceinline__ PrimD prim_d(const float *__restrict__ Q, int N, int k) {
#pragma clang fp contract(off)
  PrimD q;
  q.r = jmax((double)Q[k], 1e-4);
  q.u = (double)Q[N + k] / q.r;
  q.v = (double)Q[2 * N + k] / q.r;
  q.Y = jclamp((double)Q[4 * N + k] / q.r, 0.0, 1.0);
  q.g = 1.4 - 0.07 * q.Y;
  q.p = jmax(1.0, (q.g - 1.0) * ((double)Q[3 * N + k] - 0.5 * q.r * (q.u * q.u + q.v * q.v)));
  return q;
}

__global__ void __launch_bounds__(kB) monitor_kernel(AxiView v, const double *__restrict__ rcD, const double *__restrict__ drD,
                                                     const double *__restrict__ dxD, const float *__restrict__ Q,
                                                     const float *__restrict__ src, const float *__restrict__ mfx,
                                                     const float *__restrict__ prev, const MonIn *__restrict__ in,
                                                     double *__restrict__ partials) {
#pragma clang fp contract(off)
  __shared__ double lds[kB / 64][kMonK];
  double acc[kMonK];
#pragma unroll
  for (int c = 0; c < kMonK; c++) acc[c] = 0.0;
  const int N = v.N, W = v.W;
  const bool pv = in->prev_valid != 0;
  const double P0 = in->P0;
  for (int k = blockIdx.x * kB + threadIdx.x; k < N; k += gridDim.x * kB) {
    // divergence check of the readback: every seventh cell, solid or not, with
    // energy or density not positive (src/app_cfd.js:377)
    if (k % 7 == 0 && (!(Q[3 * N + k] > 0.f) || !(Q[k] > 0.f))) acc[13] += 1.0;
    if (v.solid[k]) continue;
    const int i = k % W, j = k / W;
    const double A = 2.0 * kPi * rcD[j] * drD[j], V = A * dxD[i] * (double)v.phi[k];
    const int id = v.comp[k];
    if (id) {
      const double ty = in->comps[id * 4 + 2];
      if (ty == 1.0 || ty == 2.0) {
        const PrimD q = prim_d(Q, N, k);
        const double e = in->comps[id * 4 + 1], mf = (double)mfx[k] / (double)v.phi[k];
        const double w = fabs(mf) > 1e-9 ? (double)src[k] / mf : 0.0;
        acc[0] += (ty == 1.0 ? e * q.r * w : q.r * w / e) * V;
        acc[3 + id] += (double)src[k] * V;
      } else if (v.rate[k] > 0.f) {
        acc[2] += (double)src[k] * V;
      }
    }
    const double aw = i > 0 ? (double)v.ae[k - 1] : 1.0, dA = (double)v.ae[k] - aw;
    if (dA != 0.0) acc[1] += (prim_d(Q, N, k).p - P0) * dA * A;
    if (pv && (k & 1) == 0) {
      const double d = ((double)Q[k] - (double)prev[k]) / (double)prev[k];
      if (d == d) { acc[11] += d * d; acc[12] += 1.0; }
    }
  }
  // block sum: lanes, then the four waves in order
#pragma unroll
  for (int c = 0; c < kMonK; c++)
    for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_down(acc[c], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < kMonK; c++) lds[wave][c] = acc[c];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int c = 0; c < kMonK; c++) {
      double a = 0.0;
      for (int w = 0; w < kB / 64; w++) a += lds[w][c];
      partials[blockIdx.x * kMonK + c] = a;
    }
}

// partials of nb blocks -> out[0..kMonK): one wave per component, lane l
// sums blocks l, l + 64, ..., then a fixed shuffle tree (the same order at
// every call)
constexpr int kMonFinishThreads = 64 * kMonK;
__global__ void __launch_bounds__(kMonFinishThreads) monitor_finish(const double *__restrict__ partials, int nb, double *__restrict__ out) {
  const int c = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double a = 0.0;
  for (int b = lane; b < nb; b += 64) a += partials[b * kMonK + c];
  for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
  if (lane == 0) out[c] = a;
}

// cfdSection's sums (src/cfd_post.js:83-102) for section s = blockIdx.x at
// column col[s] over rows[s H + j] != 0, with the solver's mean axial face
// mass flux as the cell mass flow; written to out[kSecK s + c]:
//   0 mass flow, 1 sum of max(mf, 0) (the weight w), 2 area A, 3 sum p dA,
//   4..10 w-weighted sums of Tt, Pt, T, p, |V|, M, Y
// The rows' terms are computed in parallel, kSecBlock rows at a time, into
// LDS; thread c then adds term c over the rows in row order, as the host's
// sequential loop does, so the sums are the host's to the bit.
constexpr int kSecK = 11;
constexpr int kSecBlock = 256;    // threads per block of section_kernel (one block per section)
__global__ void __launch_bounds__(kSecBlock) section_kernel(AxiView v, const double *__restrict__ rcD, const double *__restrict__ drD,
                                                            const float *__restrict__ Q, const float *__restrict__ mfx, int nsec,
                                                            const int *__restrict__ col, const uint8_t *__restrict__ rows,
                                                            double *__restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double term[kSecK][kSecBlock];
  const int s = blockIdx.x, t = threadIdx.x;
  if (s >= nsec) return;
  const int W = v.W, H = v.H, N = v.N, i = col[s];
  double acc = 0.0;                 // thread t < kSecK: the running sum of term t
  for (int j0 = 0; j0 < H; j0 += kSecBlock) {
    const int j = j0 + t;
    const bool on = j < H && rows[(size_t)s * H + j] != 0;
    if (!on) {
      // a row that does not count adds +0: the sums start at +0 and never become -0
      // (x + (-x) rounds to +0), so acc + 0 == acc to the bit
      for (int c = 0; c < kSecK; c++) term[c][t] = 0.0;
    } else {
      const int k = j * W + i;
      const PrimD q = prim_d(Q, N, k);
      const double al = jmin(1.0, 0.5 * ((i > 0 ? (double)v.ae[k - 1] : 1.0) + (double)v.ae[k]));
      const double dA = 2.0 * kPi * rcD[j] * drD[j] * al;
      const double T = q.p / (q.r * kRGas), cp = q.g * kRGas / (q.g - 1.0), V2 = q.u * q.u + q.v * q.v, Tt = T + V2 / (2.0 * cp);
      const double Pt = q.p * pow(Tt / T, q.g / (q.g - 1.0));
      const double mf = (double)mfx[k] * 2.0 * kPi * rcD[j] * drD[j], wk = jmax(mf, 0.0);
      term[0][t] = mf; term[1][t] = wk; term[2][t] = dA; term[3][t] = q.p * dA;
      term[4][t] = wk * Tt; term[5][t] = wk * Pt; term[6][t] = wk * T; term[7][t] = wk * q.p; term[8][t] = wk * sqrt(V2);
      term[9][t] = wk * sqrt(V2 / (q.g * q.p / q.r)); term[10][t] = wk * q.Y;
    }
    __syncthreads();
    if (t < kSecK) {
      const int n = min(kSecBlock, H - j0);
#pragma unroll 8
      for (int r = 0; r < n; r++) acc += term[t][r];
    }
    __syncthreads();
  }
  if (t < kSecK) out[(size_t)kSecK * s + t] = acc;
}

}  // namespace axik
