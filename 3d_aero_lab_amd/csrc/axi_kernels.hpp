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
  const bool second = c.order == 2;
  float F[5], G[5], L[5], R[5];
  if (t < NXF) {
    const int j = t / (W + 1), f = t % (W + 1), a = f - 1, b = f;
    // cells a-1, a, b, b+1 of row j, clamped into the row
    const int ia = a < 0 ? 0 : a, ib = b >= W ? W - 1 : b, im = ia > 0 ? ia - 1 : 0, ip = ib + 1 < W ? ib + 1 : W - 1;
    const int ka = j * W + ia, kb = j * W + ib, km = j * W + im, kp = j * W + ip;
    const uint8_t sa = v.solid[ka], sb = v.solid[kb];
    const float aea = v.ae[ka], pa = v.phi[ka], pb = v.phi[kb], pm = v.phi[km], pp = v.phi[kp];
    const Prim Pa = prim(Q, N, ka), Pb = prim(Q, N, kb), Pm = prim(Q, N, km), Pp = prim(Q, N, kp);
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
    for (int q = 0; q < 5; q++) Fx[q * NXF + t] = open ? al * F[q] : 0.f;
    return;
  }
  const int u = t - NXF;
  if (u >= NRF) return;
  const int f = u / W, i = u % W, a = f - 1, b = f;
  // cells a-1, a, b, b+1 of column i, clamped into the column
  const int ja = a < 0 ? 0 : a, jb = b >= H ? H - 1 : b, jm = ja > 0 ? ja - 1 : 0, jp = jb + 1 < H ? jb + 1 : H - 1;
  const int ka = ja * W + i, kb = jb * W + i, km = jm * W + i, kp = jp * W + i;
  const uint8_t sa = v.solid[ka];
  const float ana = v.an[ka], pa = v.phi[ka], pb = v.phi[kb], pm = v.phi[km], pp = v.phi[kp];
  const float la = v.lmix[ka], lb = v.lmix[kb], rca = v.rc[ja], rcb = v.rc[jb];
  const Prim Pa = prim(Q, N, ka), Pb = prim(Q, N, kb), Pm = prim(Q, N, km), Pp = prim(Q, N, kp);
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
  Fr[u] = open ? al * F[0] : 0.f;
  Fr[NRF + u] = open ? al * (F[2] - G2[1]) : 0.f;
  Fr[2 * NRF + u] = open ? al * F[1] : 0.f;
  Fr[3 * NRF + u] = open ? al * (F[3] - G2[3]) : 0.f;
  Fr[4 * NRF + u] = open ? al * (F[4] - G2[4]) : 0.f;
  nutF[u] = open ? nut : 0.f;
}

// One thread per cell (src/cfd_cpu.js:145-192). mode 0: stage 1 with dt,
// 1: stage 2, 2: dt only, 3: stage 1 with dt capped at *cap. The cell's loads
// (state, face fluxes, geometry, zone data, the stored dt) are issued before
// the solid flag is tested; solid cells keep their state (modes 0 and 3 copy
// it to the output) and write nothing else.
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
  // every load of the cell
  const bool sol = v.solid[k] != 0;
  const float *Qbb = mode == 1 ? Qb : Qin;
  float q0[5], qb[5], fxa[5], fxb[5], fra[5], frb[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    q0[q] = Qin[q * N + k]; qb[q] = Qbb[q * N + k];
    fxa[q] = Fx[q * NXF + fx0]; fxb[q] = Fx[q * NXF + fx1]; fra[q] = Fr[q * NRF + fr0]; frb[q] = Fr[q * NRF + fr1];
  }
  const float dx = v.dx[i], dr = v.dr[j], ax0 = v.rf[j], ax1 = v.rf[j + 1], rcj = v.rc[j], ph = v.phi[k];
  const float nu0 = nutF[j * W + i], nu1 = nutF[(j + 1) * W + i];
  const float aek = v.ae[k], aem = v.ae[i > 0 ? k - 1 : k], ank = v.an[k], anm = v.an[j > 0 ? k - W : k];
  const int id = v.comp[k];
  const float mfd = v.mfd[k], wgt = v.wgt[k], tgtk = v.tgt[k], rate = v.rate[k], yrate = v.yrate[k];
  const float dts = dt[k];
  if (sol) {
    if (mode == 0 || mode == 3)
#pragma unroll
      for (int q = 0; q < 5; q++) Qout[q * N + k] = q0[q];
    return;
  }
  Prim P;
  P.r = fmaxf(q0[0], 1e-4f);
  P.u = q0[1] / P.r; P.v = q0[2] / P.r;
  P.Y = clampf(q0[4] / P.r, 0.f, 1.f);
  P.g = 1.4f - 0.07f * P.Y;
  P.p = fmaxf(1.f, (P.g - 1.f) * (q0[3] - 0.5f * P.r * (P.u * P.u + P.v * P.v)));
  float d = dts;
  if (mode == 0 || mode == 2) {
    const float nu = fmaxf(nu0, nu1);
    const float cs = sqrtf(P.g * P.p / P.r);
    d = c.cfl * fmaxf(ph, 0.25f) / ((fabsf(P.u) + cs) / dx + (fabsf(P.v) + cs) / dr + 2.2f * nu / (dr * dr));
    dt[k] = d;
    if (mode == 2) return;
  }
  const float vol = rcj * dr;
  float R[5];
#pragma unroll
  for (int q = 0; q < 5; q++) R[q] = (-(fxb[q] - fxa[q]) / dx - (frb[q] * ax1 - fra[q] * ax0) / vol) / ph;
  const float aw = i > 0 ? aem : 1.f, as = j > 0 ? anm : 0.f;
  R[1] += P.p * (aek - aw) / (dx * ph);
  R[2] += P.p * (ank * ax1 - as * ax0) / (vol * ph);
  const float mfx = 0.5f * (fxa[0] + fxb[0]);
  float sk = 0.f;
  if (id) {
    const float rho = q0[0], g = P.g, cp = g * kR / (g - 1.f);
    const float T = P.p / (rho * kR), Tt = T + 0.5f * (P.u * P.u + P.v * P.v) / cp;
    const float ty = c.comps[id * 4 + 2], mf = mfx / ph;
    if (ty == 1.f || ty == 2.f) {     // blade row: tau^(wgt dx) on the upstream face enthalpy
      const float e = c.comps[id * 4 + 1], fac = ty == 1.f ? 1.f : clampf(mf / (0.3f * mfd), 0.f, 1.f);
      const float mw = fxa[0], HU = mw > 1e-9f ? fxa[3] / mw : cp * Tt;
      const float w = (HU > 0.f ? HU : cp * Tt) * expm1f(c.comps[id * 4] * wgt * fac * dx) / dx;
      R[1] += ty == 1.f ? e * rho * w : rho * w / e;
      R[3] += mf * w;
      sk = mf * w;
    } else if (rate > 0.f) {           // heat source relaxed to the target total temperature
      const float K = c.comps[id * 4 + 1] != 0.f ? c.comps[id * 4 + 1] : 1.f;
      const float tgt = c.comps[id * 4] + (tgtk - c.comps[id * 4]) * K;
      const float m0 = fmaxf(mf, 0.f);
      const float q = m0 * cp * fmaxf(0.f, tgt - Tt) * rate;
      R[3] += q;
      const float mfuel = q / c.hf;
      R[0] += mfuel;
      sk = mfuel;
      R[4] += m0 * (1.f - P.Y) * yrate + mfuel;
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
//   section_kernel  mass flow and mass-weighted Tt sums of cfdSection
//                   (src/cfd_post.js:83-102) at the control stations, one
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

// partials of nb blocks -> out[0..kMonK), block order
__global__ void monitor_finish(const double *__restrict__ partials, int nb, double *__restrict__ out) {
  const int c = threadIdx.x;
  if (c >= kMonK) return;
  double a = 0.0;
  for (int b = 0; b < nb; b++) a += partials[b * kMonK + c];
  out[c] = a;
}

// out[kMonK + 3 s + {0, 1, 2}] = mass flow, sum of max(mf, 0), sum of max(mf, 0) Tt at column col[s] over rows[s H + j] != 0
__global__ void section_kernel(AxiView v, const double *__restrict__ rcD, const double *__restrict__ drD, const float *__restrict__ Q,
                               const float *__restrict__ mfx, int nsec, const int *__restrict__ col, const uint8_t *__restrict__ rows,
                               double *__restrict__ out) {
#pragma clang fp contract(off)
  const int s = threadIdx.x;
  if (s >= nsec) return;
  const int W = v.W, H = v.H, N = v.N, i = col[s];
  double m = 0.0, w = 0.0, wTt = 0.0;
  for (int j = 0; j < H; j++) {
    if (!rows[s * H + j]) continue;
    const int k = j * W + i;
    const PrimD q = prim_d(Q, N, k);
    const double T = q.p / (q.r * kRGas), cp = q.g * kRGas / (q.g - 1.0), V2 = q.u * q.u + q.v * q.v, Tt = T + V2 / (2.0 * cp);
    const double mf = (double)mfx[k] * 2.0 * kPi * rcD[j] * drD[j], wk = jmax(mf, 0.0);
    m += mf; w += wk; wTt += wk * Tt;
  }
  out[kMonK + 3 * s] = m;
  out[kMonK + 3 * s + 1] = w;
  out[kMonK + 3 * s + 2] = wTt;
}

}  // namespace axik
