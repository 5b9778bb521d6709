// iris_yee2.h -- two Yee steps per pass over the bulk of the grid.
//
// k_yee2 advances E and H from step n to n+2 in one pass over z-chunks whose
// planes z0-1 .. z1+2 all lie between the UPML slabs and away from the TF/SF
// planes (regions I and III and the structure).  Per cell it performs the
// operations of two single steps (reference code/fdtd.f90:57-248 with the
// periodic wraps of code/boundary.f90:163-348); without a Bloch phase
// (exp(+-i ak) = 1, the x/y wraps are plain copies) its results are exactly
// those of two launches of k_yee_fused, with one they agree to rounding (see
// PH below).  What it saves is HBM traffic: step n is read and step n+2
// written once.
//
// Work decomposition.  A workgroup is kT2X lanes (x) by kT2R waves (y) and
// covers the footprint x0-1 .. x0+kT2X-2, y0-1 .. y0+kT2R-2 of an output tile
// of kT2OX x kT2OY cells.  The E update reads H at c-x, c-y and the H update
// reads E at c+x, c+y, so the footprint reaches one cell back and two ahead.
// Lanes outside 1..n take their values from the periodic image (the wrap
// copies the reference writes).  The workgroup marches z with a pipeline:
// at plane z it forms E1(z), H1(z-1), E2(z-1), H2(z-2) (1: step n+1,
// 2: step n+2); the z-1 and z-2 values of each lane stay in registers, and the
// lateral neighbours come through LDS planes: E1, H1 and E2 double-buffered
// by plane parity with one barrier per plane (FP32 fields), or E1 double and
// H1, E2 single with two barriers per plane (FP64 fields).
//
// Besides E2/H2 (and the H min-side ghost planes) of its chunk it writes:
//   - E1 and H1 (with H's min-side ghosts) of its first plane when the chunk
//     below is not a two-step chunk, and H1 of its last plane when the chunk
//     above is not: the single-step passes of the neighbouring chunks read
//     these at step n+1;
//   - the DFT face samples of step n+1 of its own cells (the samples of step
//     n+2 are gathered afterwards from the output buffer).
#pragma once
#include "iris_kernels.h"

namespace iris {

constexpr int kT2X = 64;  // lanes per wave (the workgroup's x dimension)
// Waves per workgroup (footprint rows) and the waves per SIMD the register
// allocation must allow: FP32 fields 16 rows, FP64 fields 12 rows, one
// workgroup per CU (144 KB of LDS either way).
// The UPML variant carries D and B through the pipeline as well (twelve
// more complex registers per lane) and uses 12 / 8 rows.
// W = footprint cells per tile row: 64 (one row per wave) or 32 (two rows
// per wave).  The 32-wide plain tile with FP32 fields has 8 waves (16 rows,
// 72 KB of LDS), so two workgroups share a CU and one computes while the
// other waits at its barrier; the other 32-wide tiles keep the wave count of
// the 64-wide ones (twice the rows: fewer halo cells per output).
template <typename T, bool PML = false, int W = 64> struct Yee2Shape {
  static constexpr bool HALF = sizeof(T) == 4 && !PML && W == 32;
  static constexpr int R = sizeof(T) == 4 ? (PML ? 12 : (HALF ? 8 : 16)) : (PML ? 8 : 12);  // waves
  static constexpr int RY = R * kT2X / W;  // footprint rows
  static constexpr int OX = W - 3;         // output cells per tile row
  static constexpr int OY = RY - 3;        // output rows per tile
  static constexpr int WG_PER_CU = HALF ? 2 : 1;
  static constexpr int WAVES_PER_EU = (R * WG_PER_CU + 3) / 4;
};

// Face membership of cell (x, y) in box g, 5 bits: on the bxi / bxf x face
// (1, 2) or the byi / byf y face (4, 8) -- each within the box's range in the
// other lateral coordinate -- and inside the box's x and y ranges (16: its z
// faces).  Computed once per lane, so that the z march tests a plane with a
// few instructions: edge lanes alone carry lateral bits, and the z faces are
// two planes of the march.
__device__ inline unsigned face_bits(const BoxGeom& g, int x, int y) {
  const bool inx = x >= g.bxi && x <= g.bxf, iny = y >= g.byi && y <= g.byf;
  unsigned f = 0;
  if (g.npts[0] && iny) f |= (x == g.bxi ? 1u : 0u) | (x == g.bxf ? 2u : 0u);
  if (g.npts[1] && inx) f |= (y == g.byi ? 4u : 0u) | (y == g.byf ? 8u : 0u);
  if (g.npts[2] && inx && iny) f |= 16u;
  return f;
}
// the x-face bits of both boxes in a face_bits2 set
constexpr unsigned kXFaceBits = 3u | (3u << 5);
// the bits of both boxes of a sampling set (5 a box)
template <typename T> struct Yee2Samples;
template <typename T>
__device__ inline unsigned face_bits2(const Yee2Samples<T>& smp, int x, int y) {
  unsigned f = 0;
  for (int b = 0; b < 2; ++b)
    if (b < smp.nbox) f |= face_bits(smp.g[b], x, y) << (5 * b);
  return f;
}

// Index in the concatenated face-point list of box g (inverse of box_point,
// reference code/dataprocess.f90:813-901): store E (base 0) or H (base 3) of
// cell (x,y,z) into every face point the cell is; f = the cell's face bits
// for g.
template <typename C>
__device__ inline void sample_cell(const BoxGeom& g, C* __restrict__ samp, unsigned f, int x, int y, int z, int base,
                                   const C v[3]) {
  if ((f & 15u) && z >= g.bzi && z <= g.bzf) {
    if (f & 3u) {
      const int64_t q = g.start[0] + 6 * ((int64_t)(y - g.byi) + (int64_t)g.nyb * (z - g.bzi)) + base;
      if (f & 1u) { samp[q] = v[0]; samp[q + 1] = v[1]; samp[q + 2] = v[2]; }
      if (f & 2u) {
        const int64_t r = q + g.npts[0] / 2;
        samp[r] = v[0]; samp[r + 1] = v[1]; samp[r + 2] = v[2];
      }
    }
    if (f & 12u) {
      const int64_t q = g.start[1] + 6 * ((int64_t)(x - g.bxi) + (int64_t)g.nxb * (z - g.bzi)) + base;
      if (f & 4u) { samp[q] = v[0]; samp[q + 1] = v[1]; samp[q + 2] = v[2]; }
      if (f & 8u) {
        const int64_t r = q + g.npts[1] / 2;
        samp[r] = v[0]; samp[r + 1] = v[1]; samp[r + 2] = v[2];
      }
    }
  }
  if ((f & 16u) && (z == g.bzi || z == g.bzf)) {
    const int64_t q = g.start[2] + 6 * ((int64_t)(x - g.bxi) + (int64_t)g.nxb * (y - g.byi)) + base;
    if (z == g.bzi) { samp[q] = v[0]; samp[q + 1] = v[1]; samp[q + 2] = v[2]; }
    if (z == g.bzf) {
      const int64_t r = q + g.npts[2] / 2;
      samp[r] = v[0]; samp[r + 1] = v[1]; samp[r + 2] = v[2];
    }
  }
}

// The x faces alone (f: the cell's face bits for g): the points whose
// gather from the fields costs a cache line each (consecutive points of an x
// face are rows apart).  The two-step kernels store their step n+2 values
// there themselves; the gather of that step skips them on their planes.
template <typename C>
__device__ inline void sample_xface(const BoxGeom& g, C* __restrict__ samp, unsigned f, int y, int z, int base,
                                    const C v[3]) {
  if ((f & 3u) && z >= g.bzi && z <= g.bzf) {
    const int64_t q = g.start[0] + 6 * ((int64_t)(y - g.byi) + (int64_t)g.nyb * (z - g.bzi)) + base;
    if (f & 1u) { samp[q] = v[0]; samp[q + 1] = v[1]; samp[q + 2] = v[2]; }
    if (f & 2u) {
      const int64_t r = q + g.npts[0] / 2;
      samp[r] = v[0]; samp[r + 1] = v[1]; samp[r + 2] = v[2];
    }
  }
}

// E update of one cell from its curl outside the UPML: region I or III
// (diagonal epsa_inv) or the structure (per-cell eps_inv, complex only when
// its imaginary part is stored), as e_cell does (code/fdtd.f90:86-147).
template <typename T>
__device__ inline void e_update_plain(const StepArgs<T>& a, int64_t col, int z, const cplx<T> e[3],
                                      const cplx<T> c[3], cplx<T> n[3]) {
  if (z < a.hi || z > a.hf) {
    const T* ac = (z < a.hi) ? a.a1 : a.a3;
    n[0] = cmadd<T>(e[0], ac[0], c[0]);
    n[1] = cmadd<T>(e[1], ac[1], c[1]);
    n[2] = cmadd<T>(e[2], ac[2], c[2]);
  } else {
    const int64_t eidx = col + a.sxy * (z - a.hi);
    const T er = a.eps_re[eidx];
    if (a.eps_im) {
      const cplx<T> ei = mk<T>(er, a.eps_im[eidx]);
      n[0] = cadd(e[0], cmul(ei, c[0]));
      n[1] = cadd(e[1], cmul(ei, c[1]));
      n[2] = cadd(e[2], cmul(ei, c[2]));
    } else {
      n[0] = cmadd<T>(e[0], er, c[0]);
      n[1] = cmadd<T>(e[1], er, c[1]);
      n[2] = cmadd<T>(e[2], er, c[2]);
    }
  }
}

// z-UPML updates of one cell with the operations of e_cell / h_cell
// (reference code/boundary_II.f90:84-99, :170-183): from the old aux value
// and the curl, the new aux value and the new field.
template <typename T>
__device__ inline void e_upml3(const PmlPlane<T>& P, T e02, const cplx<T> e[3], const cplx<T> c[3],
                               const cplx<T> dold[3], cplx<T> dnew[3], cplx<T> n[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const cplx<T> cc = cscale(e02, c[k]);
    dnew[k] = cadd(cscale(P.cd[k], dold[k]), cscale(P.gd[k], cc));
    n[k] = cadd(cscale(P.ce[k], e[k]), cscale(P.ge[k], csub(cscale(P.fp[k], dnew[k]), cscale(P.fm[k], dold[k]))));
  }
}
template <typename T>
__device__ inline void h_upml3(const PmlPlane<T>& P, const cplx<T> h[3], const cplx<T> c[3], const cplx<T> bold[3],
                               cplx<T> bnew[3], cplx<T> n[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    bnew[k] = csub(cscale(P.cd[k], bold[k]), cscale(P.gb[k], c[k]));
    n[k] = cadd(cscale(P.ce[k], h[k]), cscale(P.gh[k], csub(cscale(P.fp[k], bnew[k]), cscale(P.fm[k], bold[k]))));
  }
}

// Streaming store of a complex value (the L2 does not keep the line).
template <typename C> __device__ inline void st_nt(C* p, C v) {
  using E = decltype(v.x);
  typedef E V2 __attribute__((ext_vector_type(2)));
  V2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<V2*>(p));
}
template <typename C> __device__ inline void store3_nt(C* f, int64_t i, int64_t cs, const C v[3]) {
  st_nt(f + i, v[0]);
  st_nt(f + i + cs, v[1]);
  st_nt(f + i + 2 * cs, v[2]);
}

// Sampling by the two-step kernels: up to two boxes; samp[b] is the box's
// ring (slot 0), null when the box has no points.  The kernels take the ring
// slot of step n+1 (every face of their own cells; < 0: not sampled) and of
// step n+2 (x faces only; < 0: left to the gather).
template <typename T> struct Yee2Samples {
  BoxGeom g[2];
  cplx<T>* samp[2];
  int nbox;
};

// PH = true: grids with a Bloch phase in x or y.  A footprint lane outside
// 1..n then works in the frame of the cell it stands for: every value it
// loads is multiplied by that cell's Bloch factor (exp(-i ak) below 1,
// exp(+i ak) above n, the product at corners), so what it computes is the
// periodic image with its phase and its LDS neighbours need no correction.
// The results equal those of single steps up to rounding in the cells next
// to the period boundary (the single-step pass applies the factor after
// forming an image instead of before).
// PML = true: chunks that reach into the z-UPML slabs (planes 2 .. nz-2; the
// wrap in z is left to single steps).  On UPML planes the lane also reads
// D (of step n) with E and B with H, keeps D and B of step n+1 beside E1 and
// H1, and stores D and B of step n+2 into a.dn / a.bn (a third aux buffer:
// other workgroups still read the step-n values for their halos).  With
// the export below, D of step n+1 on plane z0 also goes to db: the
// single-step pass under the chunk recomputes E of step n+2 there.
// The work of one workgroup: tile (bx, by) of chunk bz.
template <typename T, bool PH, bool PML, int kT2W>
__device__ __forceinline__ void yee2_tile(const StepArgs<T>& a, cplx<T>* __restrict__ eb, cplx<T>* __restrict__ hb,
                                          cplx<T>* __restrict__ db, const int* __restrict__ chunks, int bx, int by,
                                          int bz, const Yee2Samples<T>& smp, int slot, int slot2, int xbase) {
  using C = cplx<T>;
  constexpr int kT2R = Yee2Shape<T, PML, kT2W>::RY, kT2OX = Yee2Shape<T, PML, kT2W>::OX,
                kT2OY = Yee2Shape<T, PML, kT2W>::OY;
  // LDS planes by the parity of z.  With FP32 fields H1 and E2 are double
  // buffered as E1 is, and one barrier a plane orders every reuse: a slot
  // written at plane z is read after that plane's barrier and written again
  // at plane z+2, after plane z+1's barrier, which follows all of plane z's
  // reads.  FP64 fields (twice the bytes) keep single H1 and E2 slots and a
  // second barrier before E2 is rewritten.
  constexpr int NB = sizeof(T) == 4 ? 2 : 1;
  __shared__ C s1[2][3][kT2R][kT2W];   // E1 of planes z (z & 1) and z-1
  __shared__ C sh[NB][3][kT2R][kT2W];  // H1 of plane z-1
  __shared__ C s2[NB][3][kT2R][kT2W];  // E2 of plane z-1
  // footprint cell (cx, cy) of this lane: one row per wave, or two
  const int cx = threadIdx.x % kT2W, cy = threadIdx.y * (kT2X / kT2W) + threadIdx.x / kT2W;
  const int nx = a.nx, ny = a.ny;
  const int64_t cs = a.cs, sxy = a.sxy;
  const int craw = chunks[bz];
  const int z0 = craw & kChunkZMask;
  const int z1 = z0 + ((craw >> kChunkLenShift) & kChunkLenMask);
  const bool exp_below = (craw & kChunkExportBelow) != 0, exp_above = (craw & kChunkExportAbove) != 0;

  // footprint cell of this lane and its periodic source
  const int xg = xbase + bx * kT2OX + cx, yg = by * kT2OY + cy;  // x0-1+cx, y0-1+cy
  const int xs = xg < 1 ? xg + nx : (xg > nx ? xg - nx : xg);
  const int ys = yg < 1 ? yg + ny : (yg > ny ? yg - ny : yg);
  const bool own = cx >= 1 && cx <= kT2OX && cy >= 1 && cy <= kT2OY && xg <= nx && yg <= ny;
  const int64_t col = (int64_t)(xs + a.xoff) + a.sx * ys;
  const int64_t colxm = col - 1;     // x-1 (the ghost x = 0 when xs = 1)
  const int64_t colym = col - a.sx;  // y-1 (the ghost y = 0 when ys = 1)
  // LDS neighbours, clamped into the tile (edge lanes compute values nobody uses)
  const int lp = min(cx + 1, kT2W - 1), lm = max(cx - 1, 0);
  const int wp = min(cy + 1, kT2R - 1), wm = max(cy - 1, 0);
  const C* __restrict__ e0 = a.e;
  const C* __restrict__ h0 = a.h;
  // Bloch factor of the cell this lane stands for (1 inside); every lane
  // multiplies, so the loads carry no per-lane branch (a factor of exactly 1
  // leaves a value unchanged)
  C phi = mk<T>((T)1, (T)0);
  if constexpr (PH) {
    if (xg < 1) phi = a.fxm; else if (xg > nx) phi = a.fxp;
    if (yg < 1) phi = cmul(phi, a.fym); else if (yg > ny) phi = cmul(phi, a.fyp);
  }
  auto ld = [&](const C* __restrict__ f, int64_t i) -> C {
    if constexpr (PH) return cmul(phi, f[i]);
    return f[i];
  };

  auto upml = [&](int z) { return PML && (z <= a.pmlzi || z >= a.pmlzf); };
  const unsigned fbits = face_bits2(smp, xs, ys);  // this lane's faces of the DFT boxes
  auto auxi = [&](int z, int k) -> int64_t {  // aux index of component k of this lane's cell on UPML plane z
    const int top = z >= a.pmlzf;
    return a.aux(top, k, col + sxy * (top ? (z - a.pmlzf) : (z - 1)));
  };

  // pipeline registers of this lane: H0(z-1), E1(z-1), H1(z-2), E2(z-2), and
  // on UPML planes D1(z-1), B1(z-2)
  C hz1[3], e1p[3] = {}, h1pp[3] = {}, e2p[3] = {};
  C d1p[3] = {}, b1pp[3] = {};
  {
    const int64_t i = col + sxy * (z0 - 2);
    hz1[0] = ld(h0, i);
    hz1[1] = ld(h0, i + cs);
    hz1[2] = ld(h0, i + 2 * cs);
  }

  for (int z = z0 - 1; z <= z1 + 2; ++z) {
    const int64_t iz = col + sxy * z;
    // ---- E1(z) from H0(z), H0(z-1), E0(z) (code/fdtd.f90:78-96, :137-147)
    C h0z[3], e1[3], d1n[3] = {}, b1n[3] = {};
    {
      h0z[0] = ld(h0, iz);
      h0z[1] = ld(h0, iz + cs);
      h0z[2] = ld(h0, iz + 2 * cs);
      const C h3ym = ld(h0, colym + sxy * z + 2 * cs), h1ym = ld(h0, colym + sxy * z);
      const C h3xm = ld(h0, colxm + sxy * z + 2 * cs), h2xm = ld(h0, colxm + sxy * z + cs);
      const C c[3] = {cadd(csub(csub(h0z[2], h3ym), h0z[1]), hz1[1]),
                      cadd(csub(csub(h0z[0], hz1[0]), h0z[2]), h3xm),
                      cadd(csub(csub(h0z[1], h2xm), h0z[0]), h1ym)};
      const C eo[3] = {ld(e0, iz), ld(e0, iz + cs), ld(e0, iz + 2 * cs)};
      if (upml(z)) {
        const C dold[3] = {ld(a.d, auxi(z, 0)), ld(a.d, auxi(z, 1)), ld(a.d, auxi(z, 2))};
        e_upml3<T>(a.pml[z], a.e02, eo, c, dold, d1n, e1);
      } else {
        e_update_plain<T>(a, col, z, eo, c, e1);
      }
    }
    // ---- H1(z-1) from H0(z-1), E1(z-1) (own and LDS +x/+y), E1(z)
    // (code/fdtd.f90:182-195)
    C h1n[3] = {};
    const int pz = z & 1, pm = pz ^ 1;  // LDS slot of E1(z) and of E1(z-1)
    if (z >= z0) {
      const C e3yp = s1[pm][2][wp][cx], e1yp = s1[pm][0][wp][cx];
      const C e3xp = s1[pm][2][cy][lp], e2xp = s1[pm][1][cy][lp];
      const C c0 = cadd(csub(csub(e3yp, e1p[2]), e1[1]), e1p[1]);
      const C c1 = cadd(csub(csub(e1[0], e1p[0]), e3xp), e1p[2]);
      const C c2 = cadd(csub(csub(e2xp, e1p[1]), e1yp), e1p[0]);
      if (upml(z - 1)) {
        const C bold[3] = {ld(a.b, auxi(z - 1, 0)), ld(a.b, auxi(z - 1, 1)), ld(a.b, auxi(z - 1, 2))};
        const C cc[3] = {c0, c1, c2};
        h_upml3<T>(a.pml[z - 1], hz1, cc, bold, b1n, h1n);
      } else {
        h1n[0] = cmsub<T>(hz1[0], a.dtcq2, c0);
        h1n[1] = cmsub<T>(hz1[1], a.dtcq2, c1);
        h1n[2] = cmsub<T>(hz1[2], a.dtcq2, c2);
      }
    }
    // slot pz last held E1(z-2), read before the previous plane's second
    // barrier; sh was last read before the previous plane's second barrier
    s1[pz][0][cy][cx] = e1[0];
    s1[pz][1][cy][cx] = e1[1];
    s1[pz][2][cy][cx] = e1[2];
    const int bz = NB == 2 ? pz : 0, bm = NB == 2 ? pm : 0;  // H1 / E2 slots of this plane and the last
    sh[bz][0][cy][cx] = h1n[0];
    sh[bz][1][cy][cx] = h1n[1];
    sh[bz][2][cy][cx] = h1n[2];
    __syncthreads();
    // ---- E2(z-1) from E1(z-1), H1(z-1) (own and LDS -x/-y), H1(z-2)
    C e2n[3] = {}, d2n[3] = {};
    if (z >= z0 + 1) {
      const C h3ym = sh[bz][2][wm][cx], h1ym = sh[bz][0][wm][cx];
      const C h3xm = sh[bz][2][cy][lm], h2xm = sh[bz][1][cy][lm];
      const C c[3] = {cadd(csub(csub(h1n[2], h3ym), h1n[1]), h1pp[1]),
                      cadd(csub(csub(h1n[0], h1pp[0]), h1n[2]), h3xm),
                      cadd(csub(csub(h1n[1], h2xm), h1n[0]), h1ym)};
      if (upml(z - 1)) {
        e_upml3<T>(a.pml[z - 1], a.e02, e1p, c, d1p, d2n, e2n);
      } else {
        e_update_plain<T>(a, col, z - 1, e1p, c, e2n);
      }
    }
    // ---- H2(z-2) from H1(z-2), E2(z-2) (own and LDS +x/+y), E2(z-1)
    C h2n[3] = {}, b2n[3] = {};
    if (z >= z0 + 2) {
      const C e3yp = s2[bm][2][wp][cx], e1yp = s2[bm][0][wp][cx];
      const C e3xp = s2[bm][2][cy][lp], e2xp = s2[bm][1][cy][lp];
      const C c0 = cadd(csub(csub(e3yp, e2p[2]), e2n[1]), e2p[1]);
      const C c1 = cadd(csub(csub(e2n[0], e2p[0]), e3xp), e2p[2]);
      const C c2 = cadd(csub(csub(e2xp, e2p[1]), e1yp), e2p[0]);
      if (upml(z - 2)) {
        const C cc[3] = {c0, c1, c2};
        h_upml3<T>(a.pml[z - 2], h1pp, cc, b1pp, b2n, h2n);
      } else {
        h2n[0] = cmsub<T>(h1pp[0], a.dtcq2, c0);
        h2n[1] = cmsub<T>(h1pp[1], a.dtcq2, c1);
        h2n[2] = cmsub<T>(h1pp[2], a.dtcq2, c2);
      }
    }
    if constexpr (NB == 1) __syncthreads();  // the E2(z-2) and H1(z-1) neighbours have been read
    s2[bz][0][cy][cx] = e2n[0];
    s2[bz][1][cy][cx] = e2n[1];
    s2[bz][2][cy][cx] = e2n[2];

    if (own) {
      // step n+2 of the chunk's planes, with H's min-side ghost planes
      if (z >= z0 + 1 && z - 1 <= z1) {
        if (a.nt_out) store3_nt(a.en, col + sxy * (z - 1), cs, e2n);
        else store3(a.en, col + sxy * (z - 1), cs, e2n);
        if (upml(z - 1)) {
          a.dn[auxi(z - 1, 0)] = d2n[0];
          a.dn[auxi(z - 1, 1)] = d2n[1];
          a.dn[auxi(z - 1, 2)] = d2n[2];
        }
        if (slot2 >= 0 && (fbits & kXFaceBits))
          for (int b = 0; b < smp.nbox; ++b)
            if (smp.samp[b])
              sample_xface(smp.g[b], smp.samp[b] + (int64_t)slot2 * smp.g[b].ntot, (fbits >> (5 * b)) & 31u, ys,
                           z - 1, 0, e2n);
      }
      if (z >= z0 + 2) {
        const int64_t ih = col + sxy * (z - 2);
        if (a.nt_out) store3_nt(a.hn, ih, cs, h2n);
        else store3(a.hn, ih, cs, h2n);
        if (upml(z - 2)) {
          a.bn[auxi(z - 2, 0)] = b2n[0];
          a.bn[auxi(z - 2, 1)] = b2n[1];
          a.bn[auxi(z - 2, 2)] = b2n[2];
        }
        if (xs == nx) store3_phase(a.hn, ih - nx, cs, a.fxm, h2n);
        if (ys == ny) store3_phase(a.hn, ih - a.sx * ny, cs, a.fym, h2n);
        if (slot2 >= 0 && (fbits & kXFaceBits))
          for (int b = 0; b < smp.nbox; ++b)
            if (smp.samp[b])
              sample_xface(smp.g[b], smp.samp[b] + (int64_t)slot2 * smp.g[b].ntot, (fbits >> (5 * b)) & 31u, ys,
                           z - 2, 3, h2n);
      }
      // step n+1 planes read by the single-step passes of the neighbouring chunks
      if (exp_below && z == z0) {
        store3(eb, iz, cs, e1);
        if (upml(z)) {
          db[auxi(z, 0)] = d1n[0];
          db[auxi(z, 1)] = d1n[1];
          db[auxi(z, 2)] = d1n[2];
        }
      }
      if (exp_below && z - 1 == z0) {
        const int64_t ih = col + sxy * z0;
        store3(hb, ih, cs, h1n);
        if (xs == nx) store3_phase(hb, ih - nx, cs, a.fxm, h1n);
        if (ys == ny) store3_phase(hb, ih - a.sx * ny, cs, a.fym, h1n);
      }
      if (exp_above && z - 1 == z1) store3(hb, col + sxy * z1, cs, h1n);
      // DFT samples of step n+1
      for (int b = 0; slot >= 0 && b < smp.nbox; ++b) {
        if (!smp.samp[b]) continue;
        C* const sp = smp.samp[b] + (int64_t)slot * smp.g[b].ntot;
        const unsigned f = (fbits >> (5 * b)) & 31u;
        if (z <= z1 && z >= z0) sample_cell(smp.g[b], sp, f, xs, ys, z, 0, e1);
        if (z - 1 >= z0 && z - 1 <= z1) sample_cell(smp.g[b], sp, f, xs, ys, z - 1, 3, h1n);
      }
    }
    // rotate the pipeline
    for (int k = 0; k < 3; ++k) {
      hz1[k] = h0z[k];
      e1p[k] = e1[k];
      if (z >= z0) h1pp[k] = h1n[k];
      e2p[k] = e2n[k];
      if constexpr (PML) {
        d1p[k] = d1n[k];
        if (z >= z0) b1pp[k] = b1n[k];
      }
    }
  }
}

template <typename T, bool PH, bool PML = false, int W = 64>
__global__ __launch_bounds__((kT2X * Yee2Shape<T, PML, W>::R), (Yee2Shape<T, PML, W>::WAVES_PER_EU)) void k_yee2(
    StepArgs<T> a, cplx<T>* __restrict__ eb, cplx<T>* __restrict__ hb, cplx<T>* __restrict__ db,
    const int* __restrict__ chunks, TileMap tm, Yee2Samples<T> smp, int slot, int slot2) {
  int bx, by, bz;
  tile_of(tm, blockIdx.x, &bx, &by, &bz);
  yee2_tile<T, PH, PML, W>(a, eb, hb, db, chunks, bx, by, bz, smp, slot, slot2, tm.x0);
}

// Batched two-step pass: the chunks of many grids in one launch.  Entry i of
// a table holds one grid's arguments for the pass from one of its two
// buffer sets (the pair of steps alternates between them); workgroups
// wg_start[i] .. wg_start[i+1]-1 are its tiles, a multiple of the XCD count
// so that the tile order's XCD dealing holds within every grid (the padding
// workgroups return at once).  The DFT samples of step n+1 go to ring slot
// `slot` of the grid's boxes (no sample when slot < 0), the x faces of step
// n+2 to slot2 (none when slot2 < 0).
template <typename T> struct Batch2Inst {
  StepArgs<T> a;
  cplx<T>* eb;
  cplx<T>* hb;
  cplx<T>* db;
  const int* chunks;
  TileMap tm;
  Yee2Samples<T> smp;  // ring bases (slot 0)
  int ox = 0, sl = 0, ns = 0;  // tile shape of the two-cells-a-lane kernel (k_yee2p_batch)
};

template <typename T, bool PH, bool PML, int W>
__global__ __launch_bounds__((kT2X * Yee2Shape<T, PML, W>::R), (Yee2Shape<T, PML, W>::WAVES_PER_EU)) void k_yee2_batch(
    const Batch2Inst<T>* __restrict__ inst, const int* __restrict__ wg_start, int ninst, int slot, int slot2) {
  const int blk = blockIdx.x;
  const int i = batch_instance(wg_start, ninst, blk);
  const Batch2Inst<T>& I = inst[i];
  const int L = blk - wg_start[i];
  if (L >= I.tm.nblk) return;  // padding (no barrier has been reached)
  int bx, by, bz;
  tile_of(I.tm, L, &bx, &by, &bz);
  yee2_tile<T, PH, PML, W>(I.a, I.eb, I.hb, I.db, I.chunks, bx, by, bz, I.smp, slot, slot2, I.tm.x0);
}

// DFT samples of one step from a buffer that holds that step only on the
// planes outside the two-step chunks (zmask[z] = 0); the two-step kernel
// stores the samples of the others itself.
template <typename T>
__global__ __launch_bounds__(256) void k_dft_sample_masked(StepArgs<T> a, BoxGeom g, cplx<T>* samp, int slot,
                                                           const unsigned char* __restrict__ zmask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.ntot) return;
  int c, x, y, z;
  box_point(g, i, &c, &x, &y, &z);
  if (zmask[z]) return;
  const int64_t idx = (int64_t)(x + a.xoff) + a.sx * y + a.sxy * z;
  samp[(int64_t)slot * g.ntot + i] = (c < 3) ? a.e[idx + c * a.cs] : a.h[idx + (c - 3) * a.cs];
}

}  // namespace iris
