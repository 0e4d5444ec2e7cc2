// ============================================================================
// euler3d_tile.hpp — tiled, face-once SSP-RK2 stage of the 3D Euler cut-cell
// scheme for gfx950 (a wavefront = one 64-cell x row of a tile, or two 32-cell rows).
//
// Same arithmetic as the reference's per-cell residual (CfdCPU.cellResidual,
// src/air_cfd.js:198-236) but every face flux is evaluated once and shared by
// the two cells of the face: the face state L = P_i + s_i/2, R = P_{i+1} -
// s_{i+1}/2 uses slopes that depend only on the cell they belong to (three
// cells of that cell's own stencil and its two faces), so the flux of face
// i+1/2 seen from cell i (sg = +1) and from cell i+1 (sg = -1) is the same
// number (src/air_cfd.js:213-223).
//
// Work decomposition of one block (TX x TY threads, TX = 64 or 32; a
// wavefront covers 64 / TX rows of the tile):
//   x-y tile of TX x TY columns, marching planes k0..k1-1 of a z chunk.
//   Tile-edge work: the y faces below row 0 on threads 0..TX-1 (wave 0), the x
//   faces left of column 0 on threads 64..64+TY-1 (wave 1), the tile-edge
//   slopes on threads 128.. (waves 2, 3, ...).
//   LDS per plane: primitives and flags of the tile plus a 2-cell halo ring,
//   x / y apertures, x / y slopes, x / y face fluxes.
//   Registers per column: primitives of planes k-1..k+2, z slopes of k and
//   k+1, the z flux of face k-1/2 carried from the previous plane.
// Per plane: load (own plane k+2 into the queue, halo of plane k into LDS
// from registers loaded one plane earlier), slopes, fluxes, cell update;
// three barriers. In the balanced launch each wave's issue priority falls as
// it advances through its chunk (progress_prio).
// Divisions use the hardware reciprocal, square root and reciprocal square
// root (1 ulp).
// ============================================================================
#pragma once
#include "euler3d.hpp"

namespace aero3d {

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Keeps a per-lane value in a vector register, opaque to the optimiser from
// this point on. Tests of a lane's position (tile-edge lanes, halo-ring cells)
// do not change from plane to plane; hoisted out of the plane loop they become
// 64-bit lane masks in scalar registers, which do not fit and are spilled to
// vector-register lanes and reloaded on every plane. Made opaque at the top of
// each plane, they are recomputed from a few vector registers instead. Only the
// device pass needs this (elsewhere, e.g. a CPU build of this header, it is a
// no-op).
template <class T>
__device__ __forceinline__ void keep_vgpr(T &x) {
#ifdef __AMDGCN__
  asm volatile("" : "+v"(x));
#else
  (void)x;
#endif
}

// Issue priority of this wave from the share of its chunk already done: 3 in
// the first quarter down to 0 in the last. Two blocks share each CU, and at
// equal priority the SIMDs issue the older block's waves first, so the older
// block finishes some 15 % of the launch ahead and the younger then runs alone
// at half the occupancy. With the priority falling as a block advances, the
// block that is behind is issued first and both finish together.
__device__ __forceinline__ void progress_prio(int done, int len) {
  // quarter q = floor(4 done / len), tested as 4 done < (q + 1) len (no division)
  const int d4 = 4 * done;
  if (d4 < len) __builtin_amdgcn_s_setprio(3);
  else if (d4 < 2 * len) __builtin_amdgcn_s_setprio(2);
  else if (d4 < 3 * len) __builtin_amdgcn_s_setprio(1);
  else __builtin_amdgcn_s_setprio(0);
}

// Global access by a 32-bit element index from a wave-uniform base: the byte
// offset is formed in 32 bits, so the access takes the scalar-base + vector-
// offset form with no 64-bit address arithmetic per access (planes of up to
// 2^30 cells).
template <class T>
__device__ __forceinline__ T gld(const T *__restrict__ base, unsigned i) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void gst(T *__restrict__ base, unsigned i, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + i * (unsigned)sizeof(T)) = v;
}

// minmod(a, b) = 0 when a b <= 0, else the one of smaller magnitude
// (src/air_cfd.js:196): the median of a, b and 0, one instruction.
__device__ __forceinline__ float mmod3(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.f); }

// Primitives (rho, u, v, w, p) of conserved a[] (cfdPrim5, src/air_cfd.js:169).
__device__ __forceinline__ void prim_fast(float a0, float a1, float a2, float a3, float a4, float *o) {
  const float r = fmaxf(a0, 1e-6f), ir = frcp(r), u = a1 * ir, v = a2 * ir, w = a3 * ir;
  o[0] = r; o[1] = u; o[2] = v; o[3] = w;
  o[4] = fmaxf(kGm1 * (a4 - 0.5f * r * (u * u + v * v + w * w)), 1e-7f);
}

// Total enthalpy per volume E + p of a side of a face from its primitives
// (rho, p, velocity): 3.5 p + r |v|^2 / 2 for gamma = 1.4.
__device__ __forceinline__ float hllc_enthalpy(float r, float p, float v0, float v1, float v2) {
  return p * 3.5f + 0.5f * r * (v0 * v0 + v1 * v1 + v2 * v2);
}

// HLLC flux along axis d (src/air_cfd.js:179-195) with four hardware
// transcendentals per face: two reciprocal square roots for the sound speeds,
// the reciprocals of the contact-speed denominator and of S - SM. The
// reference's F + S (U* - U) is regrouped through one common factor (below),
// so that no 1/r or 1/(S - u) is needed; the tangential momentum fluxes are
// the mass flux times the upwind side's velocity components.
template <int d>
__device__ __forceinline__ void hllc_fast(const float *L, const float *R, float *out) {
  const float rL = L[0], rR = R[0], pL = L[4], pR = R[4], uL = L[1 + d], uR = R[1 + d];
  const float gpL = kGamma * pL, gpR = kGamma * pR;   // a = sqrt(g p / r) = g p / sqrt(g p r)
  const float aL = gpL * frsq(gpL * rL), aR = gpR * frsq(gpR * rR);
  const float SL = fminf(uL - aL, uR - aR), SR = fmaxf(uL + aL, uR + aR);
  // each branch forms the total enthalpy of the chosen side only (hllc_enthalpy)
  if (SL >= 0.f || SR <= 0.f) {              // supersonic: the upwind side's flux
    const bool left = SL >= 0.f;
    const float r = left ? rL : rR, u = left ? uL : uR, p = left ? pL : pR;
    const float v0 = left ? L[1] : R[1], v1 = left ? L[2] : R[2], v2 = left ? L[3] : R[3];
    const float H = hllc_enthalpy(r, p, v0, v1, v2);
    const float m = r * u;
    // each component stored once (no update of an element picked by d)
    out[0] = m;
    out[1] = d == 0 ? m * v0 + p : m * v0;
    out[2] = d == 1 ? m * v1 + p : m * v1;
    out[3] = d == 2 ? m * v2 + p : m * v2;
    out[4] = u * H;
    return;
  }
  const float AL = rL * (SL - uL), AR = rR * (SR - uR);   // contact speed from r (S - u) of both sides
  const float SM = (pR - pL + uL * AL - uR * AR) * frcp(AL - AR);
  const bool left = SM >= 0.f;
  const float S = left ? SL : SR, uK = left ? uL : uR, rK = left ? rL : rR, pK = left ? pL : pR;
  const float v0 = left ? L[1] : R[1], v1 = left ? L[2] : R[2], v2 = left ? L[3] : R[3];
  const float HK = hllc_enthalpy(rK, pK, v0, v1, v2);
  // F + S (U* - U) with the star state U* = f (1, SM, v_t, E / r + (SM - u)(SM + p / (r (S - u)))),
  // f = r (S - u) / (S - SM): every component of S (U* - U) reduces through
  // g = S (SM - u) / (S - SM), giving the mass flux m = r (u + g), the normal
  // momentum flux r u^2 + p + S r g = m u + p + t and the energy flux
  // (u + g)(E + p) + t SM, with t = r g (S - u)
  const float SmU = S - uK, g = S * (SM - uK) * frcp(S - SM);
  const float ug = uK + g, t = rK * g * SmU;
  const float m = rK * ug;                   // mass flux
  const float mn = fmaf(m, uK, pK) + t;      // normal momentum
  out[0] = m;
  out[1] = d == 0 ? mn : m * v0;
  out[2] = d == 1 ? mn : m * v1;
  out[3] = d == 2 ? mn : m * v2;
  out[4] = fmaf(ug, HK, t * SM);
}

// Face flux of axis d between left cell (PL, slope sL) and right cell (PR,
// slope sR) with aperture a; zero when the face is closed. A face between two
// cells that are not fluid is computed like any other (no cell update uses
// it): the reference closes every face that touches a solid cell, and the
// tile gives the columns and rows outside the box aperture 0, so such faces
// are fixed- or ghost-cell faces.
// kPin (operands read from LDS): every operand is an input of one empty asm
// statement before the face is tested, so all of its LDS reads are issued
// together and waited for once; otherwise the reads are sunk below the test,
// a second LDS round trip. Input operands only: read-write ones made the
// compiler copy the values out of the paired LDS reads' registers.
template <int d, bool kPin = false>
__device__ __forceinline__ void face_flux(float *PL, float *sL, float *PR, float *sR, float a, float *F) {
  if constexpr (kPin) {
#ifdef __AMDGCN__
    // one statement, so the reads are waited for once (one wait per pinned value otherwise)
    asm volatile("" ::"v"(PL[0]), "v"(PL[1]), "v"(PL[2]), "v"(PL[3]), "v"(PL[4]), "v"(sL[0]), "v"(sL[1]), "v"(sL[2]),
                 "v"(sL[3]), "v"(sL[4]), "v"(PR[0]), "v"(PR[1]), "v"(PR[2]), "v"(PR[3]), "v"(PR[4]), "v"(sR[0]),
                 "v"(sR[1]), "v"(sR[2]), "v"(sR[3]), "v"(sR[4]), "v"(a));
#endif
  }
  if (!(a > 0.f)) {
#pragma unroll
    for (int v = 0; v < 5; v++) F[v] = 0.f;
    return;
  }
  float L[5], R[5];
#pragma unroll
  for (int v = 0; v < 5; v++) { L[v] = PL[v] + 0.5f * sL[v]; R[v] = PR[v] - 0.5f * sR[v]; }
  // first-order fallback where a reconstructed density or pressure is not
  // positive; rare, so the selections sit behind a wave-uniform test
  const bool bad = L[0] <= 0.f || L[4] <= 0.f || R[0] <= 0.f || R[4] <= 0.f;
  if (__builtin_amdgcn_ballot_w64(bad)) {
#pragma unroll
    for (int v = 0; v < 5; v++) { L[v] = bad ? PL[v] : L[v]; R[v] = bad ? PR[v] : R[v]; }
  }
  hllc_fast<d>(L, R, F);
}

// minmod slope of cell c from its neighbours (src/air_cfd.js:213-216):
// zero unless the three cells are fluid and both faces of c fully open.
__device__ __forceinline__ void slope5(const float *Pm, const float *P0, const float *Pp, bool ok, float *s) {
#pragma unroll
  for (int v = 0; v < 5; v++) s[v] = ok ? mmod3(P0[v] - Pm[v], Pp[v] - P0[v]) : 0.f;
}

// The slope test of a cell from LDS: flags of the cell and its two neighbours
// and the apertures of its two faces. All five are read before any is tested
// (a short-circuit test would read them one LDS round trip at a time).
__device__ __forceinline__ bool slope_ok(bool o2, int tm, int t0, int tp, float am, float ap) {
#ifdef __AMDGCN__
  asm volatile("" ::"v"(tm), "v"(t0), "v"(tp), "v"(am), "v"(ap));
#endif
  return o2 & (tm == 0) & (t0 == 0) & (tp == 0) & (am == 1.f) & (ap == 1.f);
}

template <int TX, int TY>
struct TileSmem {
  static constexpr int PW = TX + 4, PH = TY + 4;
  // components innermost: neighbouring cells' values lie within one paired
  // read's offset range of one base address (fewer address adds and LDS
  // instructions than one plane per component); cells are 20 bytes apart,
  // so 32 consecutive lanes reading one component hit 32 different banks
  float P[PH][PW][5];        // primitives, tile + 2-cell halo (corners unused)
  int8_t T[PH][PW];          // flags: 0..5 cells, -1 freestream ghost, -2 mirror ghost
  float A0[TY][TX + 3];      // ap0 of cells x = -2..TX (x faces)
  float A1[TY + 3][TX];      // ap1 of rows y = -2..TY (y faces)
  float Sx[TY][TX + 2][5];   // x slopes of cells -1..TX
  float Sy[TY + 2][TX][5];   // y slopes of rows -1..TY
  float Fx[TY][TX + 1][5];   // x fluxes of faces -1/2..TX-1/2
  float Fy[TY + 1][TX][5];   // y fluxes of faces -1/2..TY-1/2
};

// Classify and load the primitives of cell (i, j, k) of one case: freestream
// ghost outside the box, mirror below y = 0, else the primitives of the stored
// state (src/air_cfd.js:205-210). Fixed cells (flags 3..5) hold their fixed
// conserved state in every state buffer (the solver writes it at reset, write
// and set_fix, and every stage writes it), so their primitives come from the
// stored state like any other cell's: no per-lane lookup and no branch on the
// flag, and the flag and the five values are loaded in one memory round trip
// (cell 0 stands in for cells outside the box). P is written once per
// component from register values.
__device__ __forceinline__ int load_prim(const CaseView &cv, const StageConsts &kc, const float *__restrict__ qin, size_t ns,
                                         int i, int j, int k, float *P) {
  const bool inbox = !(i < 0 || i >= cv.NX || j >= cv.NY || k < 0 || k >= cv.NZ);
  const bool mirror = j < 0;
  const int jj = mirror ? -1 - j : j;
  const unsigned q = inbox ? (unsigned)((k * cv.NY + jj) * cv.NX + i) : 0u;
  const int tq = gld(cv.t, q);
  float w[5];
  prim_fast(gld(qin, q), gld(qin + ns, q), gld(qin + 2 * ns, q), gld(qin + 3 * ns, q), gld(qin + 4 * ns, q), w);
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const float x = !inbox ? kc.Fw[v] : w[v];
    P[v] = (v == 2 && inbox && mirror) ? -x : x;
  }
  return !inbox ? -1 : (mirror ? -2 : tq);
}

// One stage over the z chunk [k0, k1) of tile (i0, j0) of case cv.
// Writes out[] for every cell of the chunk; dtout (if non-null) gets dt.
// kPrio: set the wave priority from the chunk's progress (for launches whose
// blocks are all resident at once; where blocks are dispatched over many
// rounds, oldest-first issue finishes blocks sooner and frees their slots).
// kEarlyY: the y face's LDS operands are read before the z face is computed
// (hides their latency; costs the registers of 15 values across the z face).
// kMode: 0 decides at run time (stage 1 when qin == bse, dt written when
// dtout is non-null); 1 stage 1 (bse == qin), 2 stage 2 without dt, 3 stage 2
// with dt. The fixed modes hold fewer base pointers in scalar registers
// (fewer of them spilled to vector-register lanes and reloaded per plane).
template <int TX, int TY, bool kPrio = false, int kMode = 0, bool kEarlyY = false>
__device__ __forceinline__ void tile_stage(const CaseView &cv, const StageConsts &kc, const float *__restrict__ qin,
                                           const float *__restrict__ bse, float *__restrict__ out, float *__restrict__ dtout,
                                           size_t ns, float wa, float wb, int i0, int j0, int k0, int k1,
                                           TileSmem<TX, TY> &sm) {
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int NX = cv.NX, NY = cv.NY, NZ = cv.NZ, NXY = NX * NY;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i < NX && j < NY;
  const unsigned qc = own ? (unsigned)(j * NX + i) : 0u;   // column base index (plane 0)
  const bool o2 = kc.order == 2;
  // RK weights out = wA base + wB (in + dt R): stage 1 (0, 1), stage 2 (1/2, 1/2); run-time mode takes the arguments
  const float wA = kMode == 1 ? 0.f : (kMode >= 2 ? 0.5f : wa), wB = kMode == 1 ? 1.f : (kMode >= 2 ? 0.5f : wb);
  const float ihx = own ? frcp(cv.dx[i]) : 0.f, ihy = own ? frcp(cv.dy[j]) : 0.f;
  const float *__restrict__ ap0 = cv.ap, *__restrict__ ap1 = cv.ap + cv.N, *__restrict__ ap2 = cv.ap + 2 * (size_t)cv.N,
                            *__restrict__ ap3 = cv.ap + 3 * (size_t)cv.N;
  // z-face aperture of the own column; 1 past the box's bottom and top (open
  // boundary faces), 0 for columns outside the box (their faces are never used)
  auto apz = [&](int k) -> float { return !own ? 0.f : (k < 0 || k >= NZ) ? 1.f : gld(ap2, (unsigned)(k * NXY) + qc); };

  // halo ring (no corners: 4 rows of TX, 4 columns of TY), NHP cells per
  // thread; each cell's raw state is loaded one plane ahead (ha / ht) so the
  // loads of plane k + 1 are in flight while plane k is computed
  constexpr int NT = TX * TY, NH = 4 * TX + 4 * TY, NHP = (NH + NT - 1) / NT, PW = TileSmem<TX, TY>::PW;
  int hoff[NHP];             // cell index (hy + 2) * PW + hx + 2 in sm.P / sm.T, -1: no cell
  unsigned hq[NHP];          // plane-0 index of the cell (mirrored row below y = 0)
  bool hin[NHP], hmir[NHP];  // inside the box in x / y; mirror ghost
  float ha[NHP][5];
  int ht[NHP];
#pragma unroll
  for (int p = 0; p < NHP; p++) {
    const int h = tid + p * NT;
    int hx = 0, hy = 0;
    if (h < 4 * TX) { const int r = h / TX; hx = h % TX; hy = r < 2 ? r - 2 : TY + r - 2; }
    else { const int c = (h - 4 * TX) / TY; hy = (h - 4 * TX) % TY; hx = c < 2 ? c - 2 : TX + c - 2; }
    const int gi = i0 + hx, gj = j0 + hy;
    hoff[p] = h < NH ? (hy + 2) * PW + hx + 2 : -1;
    hin[p] = h < NH && gi >= 0 && gi < NX && gj < NY;
    hmir[p] = gj < 0;
    hq[p] = hin[p] ? (unsigned)((hmir[p] ? -1 - gj : gj) * NX + gi) : 0u;
  }
  // copies of the lane's thread index and halo-ring data made opaque per plane (keep_vgpr)
  int tidv = tid;
  int hoffv[NHP], hinv[NHP], hmirv[NHP];
#pragma unroll
  for (int p = 0; p < NHP; p++) { hoffv[p] = hoff[p]; hinv[p] = hin[p]; hmirv[p] = hmir[p]; }
  // raw state of plane kk (cell 0 stands in outside the box)
  auto halo_load = [&](int kk) {
#pragma unroll
    for (int p = 0; p < NHP; p++) {
      const unsigned q = hinv[p] ? (unsigned)(kk * NXY) + hq[p] : 0u;
      ht[p] = gld(cv.t, q);
#pragma unroll
      for (int v = 0; v < 5; v++) ha[p][v] = gld(qin + v * ns, q);
    }
  };
  halo_load(k0);
  // apertures written to LDS each plane (own x / y faces, the tile-edge face
  // of lanes 0..2TY-1 and 64..64+2TX-1) and the far-face aperture of this
  // lane's tile-edge slope, loaded one plane ahead
  const int et = (tid + TX * TY - 128) % (TX * TY);     // edge slope e on thread 128 + e (waves 2, 3, ...)
  int etv = et;
  float aX = 1.f, aY = 1.f, aE = 1.f, aS = 1.f;
  auto ap_load = [&](int kk) {
    const unsigned qq = (unsigned)(kk * NXY) + qc;
    aX = own ? gld(ap0, qq) : 0.f;   // columns / rows outside the box: faces never used
    aY = own ? gld(ap1, qq) : 0.f;
    if (tidv < 2 * TY) {
      const int r = tidv >> 1, ii = (tidv & 1) - 2, gi = i0 + ii, gj = j0 + r;
      aE = (gi < 0 || gi >= NX || gj >= NY) ? 1.f : gld(ap0, (unsigned)((kk * NY + gj) * NX + gi));
    } else if (tidv >= 64 && tidv < 64 + 2 * TX) {
      const int u = tidv - 64, c = u >> 1, jj = (u & 1) - 2, gi = i0 + c, gj = j0 + jj;
      aE = (gj < 0 || gj >= NY || gi >= NX) ? 1.f : gld(ap1, (unsigned)((kk * NY + gj) * NX + gi));
    }
    if (etv < 2 * TY) {
      if (etv & 1) {
        const int gi = i0 + TX, gj = j0 + (etv >> 1);
        aS = (gi < 0 || gi >= NX || gj >= NY) ? 1.f : gld(ap0, (unsigned)((kk * NY + gj) * NX + gi));
      }
    } else if (etv < 2 * TY + 2 * TX) {
      const int u = etv - 2 * TY;
      if (u & 1) {
        const int gi = i0 + (u >> 1), gj = j0 + TY;
        aS = (gj < 0 || gj >= NY || gi >= NX) ? 1.f : gld(ap1, (unsigned)((kk * NY + gj) * NX + gi));
      }
    }
  };
  ap_load(k0);
  // operands of the cell update (conserved input and base state, phi, baffle
  // slope, dz), loaded one plane ahead so that their memory round trip
  // overlaps a whole plane instead of one barrier
  constexpr bool kS1 = kMode == 1;   // stage 1: the base state is the input state (one read)
  float uQ[5] = {}, uB[5] = {}, uPh = 1.f, uSu = 0.f, uDz = 1.f;
  auto upd_load = [&](int kk) {
    if (own) {
      const unsigned qq = (unsigned)(kk * NXY) + qc;
#pragma unroll
      for (int c = 0; c < 5; c++) uQ[c] = gld(qin + c * ns, qq);
      if (!kS1 && !(kMode == 0 && qin == bse)) {
#pragma unroll
        for (int c = 0; c < 5; c++) uB[c] = gld(bse + c * ns, qq);
      }
      uPh = gld(cv.phi, qq); uSu = gld(ap3, qq); uDz = cv.dz[kk];
    }
  };
  upd_load(k0);

  // ---- z queue preamble: planes k0-2 .. k0+1 (after the first plane's halo,
  // aperture and update loads are issued, so all go out in one round trip)
  float Pz[4][5];          // primitives of planes k-1, k, k+1, k+2 (own column)
  int Tz[4];
  float Pm2[5];
  // columns outside the box (partial tiles) carry freestream ghosts
  const int Tm2 = load_prim(cv, kc, qin, ns, own ? i : -1, j, k0 - 2, Pm2);
#pragma unroll
  for (int m = 0; m < 3; m++) Tz[m] = load_prim(cv, kc, qin, ns, own ? i : -1, j, k0 - 1 + m, Pz[m]);
  float Am1 = apz(k0 - 2), A0z = apz(k0 - 1), A1z = apz(k0);   // faces k0-3/2, k0-1/2, k0+1/2
  float sz0[5], sz1[5], Fzm[5];
  slope5(Pm2, Pz[0], Pz[1], o2 && Tm2 == 0 && Tz[0] == 0 && Tz[1] == 0 && Am1 == 1.f && A0z == 1.f, sz0);   // s(k0-1)
  slope5(Pz[0], Pz[1], Pz[2], o2 && Tz[0] == 0 && Tz[1] == 0 && Tz[2] == 0 && A0z == 1.f && A1z == 1.f, sz1); // s(k0)
  face_flux<2>(Pz[0], sz0, Pz[1], sz1, A0z, Fzm);                                             // F(k0-1/2)
#pragma unroll
  for (int v = 0; v < 5; v++) sz0[v] = sz1[v];                                                              // sz0 = s(k)
  // queue: Pz[0] = k-1, Pz[1] = k, Pz[2] = k+1; Pz[3] from the raw state of
  // plane k+2 (zr / zt), loaded at the end of the plane before
  float sUpPrev = (own && k0 > 0) ? gld(ap3, (unsigned)((k0 - 1) * NXY) + qc) : 0.f;
  float zr[5];
  int zt;
  bool zin;
  auto zraw_load = [&](int kk) {
    zin = own && kk < NZ;
    const unsigned q = zin ? (unsigned)(kk * NXY) + qc : 0u;
    zt = gld(cv.t, q);
#pragma unroll
    for (int v = 0; v < 5; v++) zr[v] = gld(qin + v * ns, q);
  };
  zraw_load(k0 + 2);

  for (int k = k0; k < k1; k++) {
    if constexpr (kPrio) progress_prio(k - k0, k1 - k0);
    keep_vgpr(tidv); keep_vgpr(etv);
#pragma unroll
    for (int p = 0; p < NHP; p++) { keep_vgpr(hoffv[p]); keep_vgpr(hinv[p]); keep_vgpr(hmirv[p]); }
    // ---------------- load phase
    {   // plane k+2 of the own column (never a mirror row): load_prim on the raw state loaded a plane earlier
      float w[5];
      prim_fast(zr[0], zr[1], zr[2], zr[3], zr[4], w);
#pragma unroll
      for (int v = 0; v < 5; v++) Pz[3][v] = zin ? w[v] : kc.Fw[v];
      Tz[3] = zin ? zt : -1;
    }
    const float A2z = apz(k + 1);                   // face k+3/2
    // tile interior of plane k from the queue
#pragma unroll
    for (int v = 0; v < 5; v++) sm.P[ty + 2][tx + 2][v] = Pz[1][v];
    sm.T[ty + 2][tx + 2] = (int8_t)Tz[1];
    // halo ring of plane k from the prefetched raw state (primitives as
    // load_prim), then the prefetch of plane k + 1
#pragma unroll
    for (int p = 0; p < NHP; p++) {
      if (hoffv[p] < 0) continue;
      float w[5];
      prim_fast(ha[p][0], ha[p][1], ha[p][2], ha[p][3], ha[p][4], w);
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const float x = !hinv[p] ? kc.Fw[v] : w[v];
        (&sm.P[0][0][0])[hoffv[p] * 5 + v] = (v == 2 && hinv[p] && hmirv[p]) ? -x : x;
      }
      (&sm.T[0][0])[hoffv[p]] = (int8_t)(!hinv[p] ? -1 : (hmirv[p] ? -2 : ht[p]));
    }
    if (k + 1 < k1) halo_load(k + 1);
    // apertures: own x / y faces, plus the two faces left of / below the tile
    const unsigned qk = (unsigned)(k * NXY) + qc;
    sm.A0[ty][tx + 2] = aX;
    sm.A1[ty + 2][tx] = aY;
    if (tidv < 2 * TY) sm.A0[tidv >> 1][(tidv & 1)] = aE;
    else if (tidv >= 64 && tidv < 64 + 2 * TX) sm.A1[(tidv - 64) & 1][(tidv - 64) >> 1] = aE;
    const float aSk = aS;
    if (k + 1 < k1) ap_load(k + 1);
    __syncthreads();

    // ---------------- slopes: own x and y, then the tile-edge extras
    {
      const int xi = tx + 2, yj = ty + 2;
      float P0[5], Pm[5], Pp[5], s[5];
#pragma unroll
      for (int v = 0; v < 5; v++) { P0[v] = sm.P[yj][xi][v]; Pm[v] = sm.P[yj][xi - 1][v]; Pp[v] = sm.P[yj][xi + 1][v]; }
      const int t0 = sm.T[yj][xi];
      bool ok = slope_ok(o2, sm.T[yj][xi - 1], t0, sm.T[yj][xi + 1], sm.A0[ty][tx + 1], sm.A0[ty][tx + 2]);
      slope5(Pm, P0, Pp, ok, s);
#pragma unroll
      for (int v = 0; v < 5; v++) sm.Sx[ty][tx + 1][v] = s[v];
#pragma unroll
      for (int v = 0; v < 5; v++) { Pm[v] = sm.P[yj - 1][xi][v]; Pp[v] = sm.P[yj + 1][xi][v]; }
      ok = slope_ok(o2, sm.T[yj - 1][xi], t0, sm.T[yj + 1][xi], sm.A1[ty + 1][tx], sm.A1[ty + 2][tx]);
      slope5(Pm, P0, Pp, ok, s);
#pragma unroll
      for (int v = 0; v < 5; v++) sm.Sy[ty + 1][tx][v] = s[v];
    }
    // the tile-edge slopes go to waves 2.. (waves 0 and 1 take the tile-edge fluxes)
    for (int e = etv; e < 2 * TY + 2 * TX; e += TX * TY) {
      float P0[5], Pm[5], Pp[5], s[5];
      if (e < 2 * TY) {                                 // x slope of cell -1 or TX in row r
        const int r = e >> 1, cx = (e & 1) ? TX : -1, xi = cx + 2, yj = r + 2;
#pragma unroll
        for (int v = 0; v < 5; v++) { P0[v] = sm.P[yj][xi][v]; Pm[v] = sm.P[yj][xi - 1][v]; Pp[v] = sm.P[yj][xi + 1][v]; }
        // faces of cell cx: ap0 of cells cx-1 and cx; cell TX's + face lies outside the tile
        const int gi = i0 + cx, gj = j0 + r;
        const float am = sm.A0[r][cx + 1];
        const float apf = (cx == -1) ? sm.A0[r][cx + 2] : aSk;
        (void)gi; (void)gj;
        const bool ok = slope_ok(o2, sm.T[yj][xi - 1], sm.T[yj][xi], sm.T[yj][xi + 1], am, apf);
        slope5(Pm, P0, Pp, ok, s);
#pragma unroll
        for (int v = 0; v < 5; v++) sm.Sx[r][cx + 1][v] = s[v];
      } else {                                          // y slope of row -1 or TY in column c
        const int u = e - 2 * TY, c = u >> 1, ry = (u & 1) ? TY : -1, xi = c + 2, yj = ry + 2;
#pragma unroll
        for (int v = 0; v < 5; v++) { P0[v] = sm.P[yj][xi][v]; Pm[v] = sm.P[yj - 1][xi][v]; Pp[v] = sm.P[yj + 1][xi][v]; }
        const int gi = i0 + c, gj = j0 + ry;
        const float am = sm.A1[ry + 1][c];
        const float apf = (ry == -1) ? sm.A1[ry + 2][c] : aSk;
        (void)gi; (void)gj;
        const bool ok = slope_ok(o2, sm.T[yj - 1][xi], sm.T[yj][xi], sm.T[yj + 1][xi], am, apf);
        slope5(Pm, P0, Pp, ok, s);
#pragma unroll
        for (int v = 0; v < 5; v++) sm.Sy[ry + 1][c][v] = s[v];
      }
    }
    // z slope of plane k+1 (own column): cells k, k+1, k+2, faces k+1/2, k+3/2
    float sz1n[5];
    slope5(Pz[1], Pz[2], Pz[3], o2 && Tz[1] == 0 && Tz[2] == 0 && Tz[3] == 0 && A1z == 1.f && A2z == 1.f, sz1n);
    __syncthreads();

    // ---------------- fluxes: own + x, + z, + y faces, tile-edge - faces
    float Fzp[5];
    {
      const int xi = tx + 2, yj = ty + 2;
      float PL[5], PR[5], sL[5], sR[5], F[5];
#pragma unroll
      for (int v = 0; v < 5; v++) { PL[v] = sm.P[yj][xi][v]; PR[v] = sm.P[yj][xi + 1][v]; sL[v] = sm.Sx[ty][tx + 1][v]; sR[v] = sm.Sx[ty][tx + 2][v]; }
      face_flux<0, true>(PL, sL, PR, sR, sm.A0[ty][tx + 2], F);
#pragma unroll
      for (int v = 0; v < 5; v++) sm.Fx[ty][tx + 1][v] = F[v];
      if constexpr (kEarlyY) {   // the y face's LDS reads go out before the z face (registers only) is computed
        float PRy[5], sLy[5], sRy[5];
#pragma unroll
        for (int v = 0; v < 5; v++) { PRy[v] = sm.P[yj + 1][xi][v]; sLy[v] = sm.Sy[ty + 1][tx][v]; sRy[v] = sm.Sy[ty + 2][tx][v]; }
        const float Ay = sm.A1[ty + 2][tx];
        face_flux<2>(Pz[1], sz0, Pz[2], sz1n, A1z, Fzp);
        face_flux<1, true>(PL, sLy, PRy, sRy, Ay, F);
      } else {
#pragma unroll
        for (int v = 0; v < 5; v++) { PR[v] = sm.P[yj + 1][xi][v]; sL[v] = sm.Sy[ty + 1][tx][v]; sR[v] = sm.Sy[ty + 2][tx][v]; }
        face_flux<1, true>(PL, sL, PR, sR, sm.A1[ty + 2][tx], F);
      }
#pragma unroll
      for (int v = 0; v < 5; v++) sm.Fy[ty + 1][tx][v] = F[v];
    }
    if (tidv < TX) {                                    // y face below row 0, column tid
      const int c = tidv, xi = c + 2;
      float PL[5], PR[5], sL[5], sR[5], F[5];
#pragma unroll
      for (int v = 0; v < 5; v++) { PL[v] = sm.P[1][xi][v]; PR[v] = sm.P[2][xi][v]; sL[v] = sm.Sy[0][c][v]; sR[v] = sm.Sy[1][c][v]; }
      face_flux<1, true>(PL, sL, PR, sR, sm.A1[1][c], F);
#pragma unroll
      for (int v = 0; v < 5; v++) sm.Fy[0][c][v] = F[v];
    } else if (tidv >= 64 && tidv < 64 + TY) {           // x face left of column 0, row r (wave 1)
      const int r = tidv - 64, yj = r + 2;
      float PL[5], PR[5], sL[5], sR[5], F[5];
#pragma unroll
      for (int v = 0; v < 5; v++) { PL[v] = sm.P[yj][1][v]; PR[v] = sm.P[yj][2][v]; sL[v] = sm.Sx[r][0][v]; sR[v] = sm.Sx[r][1][v]; }
      face_flux<0, true>(PL, sL, PR, sR, sm.A0[r][1], F);
#pragma unroll
      for (int v = 0; v < 5; v++) sm.Fx[r][0][v] = F[v];
    }
    if constexpr (!kEarlyY) face_flux<2>(Pz[1], sz0, Pz[2], sz1n, A1z, Fzp);
    // own apertures into registers (LDS apertures are rewritten by the next plane's load)
    const float ax_m = sm.A0[ty][tx + 1], ax_p = sm.A0[ty][tx + 2], ay_m = sm.A1[ty + 1][tx], ay_p = sm.A1[ty + 2][tx];
    // this plane's update operands (loaded one plane earlier), then the next plane's loads
    float Qc[5], Bq[5];
    const float phq = uPh, sUpq = uSu, dzq = uDz;
#pragma unroll
    for (int c = 0; c < 5; c++) { Qc[c] = uQ[c]; Bq[c] = (kS1 || (kMode == 0 && qin == bse)) ? uQ[c] : uB[c]; }
    if (k + 1 < k1) upd_load(k + 1);
    __syncthreads();

    // ---------------- cell update
    if (own) {
      const unsigned q = qk;
      const int tq = Tz[1];
      if (tq != 0) {               // solid and fixed cells keep their state (a fixed cell holds its fixed state)
#pragma unroll
        for (int c = 0; c < 5; c++) gst(out + c * ns, q, Bq[c]);
        if (kMode == 3 || (kMode == 0 && dtout)) gst(dtout, q, 0.f);
      } else {
        const float ihz = frcp(dzq);
        const float *C = Pz[1];
        const float pC = C[4];
        float R[5];
        // each face's flux weighted by aperture / width (weights formed once per cell), + faces
        // subtracted, - faces added; then the wall pressure p_C (A+ - A-) / h on each axis
        const float wxp = ax_p * ihx, wxm = ax_m * ihx, wyp = ay_p * ihy, wym = ay_m * ihy, wzp = A1z * ihz, wzm = A0z * ihz;
#pragma unroll
        for (int c = 0; c < 5; c++) {
          float r = -(sm.Fx[ty][tx + 1][c] * wxp);
          r = fmaf(sm.Fx[ty][tx][c], wxm, r);
          r = fmaf(-sm.Fy[ty + 1][tx][c], wyp, r);
          r = fmaf(sm.Fy[ty][tx][c], wym, r);
          r = fmaf(-Fzp[c], wzp, r);
          R[c] = fmaf(Fzm[c], wzm, r);
        }
        R[1] = fmaf(pC, wxp - wxm, R[1]);
        R[2] = fmaf(pC, wyp - wym, R[2]);
        R[3] = fmaf(pC, wzp - wzm, R[3]);
        // baffle transpiration (src/air_cfd.js:226-233)
        const float sUp = sUpq, sDn = sUpPrev;
        if (sUp != 0.f || sDn != 0.f) {
          const float r = C[0], u = C[1], v = C[2], w = C[3];
          const float Hh = (pC * 2.5f + 0.5f * r * (u * u + v * v + w * w) + pC) * frcp(r);
          const float m = (sDn != 0.f ? (1.f - A0z) * r * u * sDn : 0.f) - (sUp != 0.f ? (1.f - A1z) * r * u * sUp : 0.f);
          const float mh = m * ihz;
          R[0] += mh; R[1] += mh * u; R[2] += mh * v; R[3] += mh * w; R[4] += mh * Hh;
        }
        const float ph = phq;
        float Bw[5];
        if constexpr (kS1) {      // stage 1: the base state is the input state, whose primitives are C
#pragma unroll
          for (int c = 0; c < 5; c++) Bw[c] = C[c];
        } else {
          prim_fast(Bq[0], Bq[1], Bq[2], Bq[3], Bq[4], Bw);
        }
        const float gpB = kGamma * Bw[4];
        const float a = gpB * frsq(gpB * Bw[0]);      // sqrt(g p / r) with one transcendental, as in the HLLC
        const float den = (fabsf(Bw[1]) + a) * ihx + (fabsf(Bw[2]) + a) * ihy + (fabsf(Bw[3]) + a) * ihz;
        const float cph = kc.cfl * fmaxf(ph, 0.25f);
        // dt = cfl max(phi, 1/4) / den; the update takes dt / phi = cfl max(phi, 1/4) / (den phi) in every stage
        // mode (the same bits whether or not dt is written out)
        constexpr bool kNoDt = kMode == 1 || kMode == 2;
        const float dt = kNoDt ? 0.f : cph * frcp(den);
        const float dtp = cph * frcp(den * ph);
        float n[5];
#pragma unroll
        for (int c = 0; c < 5; c++) {
          // stage 1 is Q + dt R itself (a product 0 * base would not be dropped by the compiler)
          if constexpr (kS1) n[c] = Qc[c] + dtp * R[c];
          else n[c] = wA * Bq[c] + wB * (Qc[c] + dtp * R[c]);
        }
        n[0] = fmaxf(n[0], 1e-4f);
        const float ke = 0.5f * (n[1] * n[1] + n[2] * n[2] + n[3] * n[3]) * frcp(n[0]);
        if (kGm1 * (n[4] - ke) < 1e-5f) n[4] = 2.5e-5f + ke;
#pragma unroll
        for (int c = 0; c < 5; c++) gst(out + c * ns, q, n[c]);
        if (kMode == 3 || (kMode == 0 && dtout)) gst(dtout, q, dt);
      }
      sUpPrev = sUpq;
    }
    // ---------------- shift the z queue
#pragma unroll
    for (int v = 0; v < 5; v++) {
      Pz[0][v] = Pz[1][v]; Pz[1][v] = Pz[2][v]; Pz[2][v] = Pz[3][v];
      sz0[v] = sz1n[v]; Fzm[v] = Fzp[v];
    }
    Tz[0] = Tz[1]; Tz[1] = Tz[2]; Tz[2] = Tz[3];
    A0z = A1z; A1z = A2z;
    if (k + 1 < k1) zraw_load(k + 3);
    (void)Am1;
  }
}

}  // namespace aero3d
