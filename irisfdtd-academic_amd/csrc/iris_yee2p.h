// iris_yee2p.h -- the two-step pass with FP32 fields, two x-cells per lane.
//
// The operations per cell are those of k_yee2 (two Yee steps per pass,
// reference code/fdtd.f90:57-248 with the wraps of code/boundary.f90:163-348)
// in the same order, so the results are bitwise those of k_yee2 at any tile
// width.  What changes is the shape of the work:
//   - a lane holds two adjacent x-cells (cell a at odd x, cell b = a + 1) of
//     every value, so each load and store moves 16 bytes, as with FP64
//     fields, and a wave covers a row of 128 footprint cells (124 of them
//     output cells: x0-2 .. x0+125 for x0 .. x0+123);
//   - x-neighbours come from the lane next door by DPP wave shifts (cell a's
//     -x neighbour is the previous lane's cell b, cell b's +x neighbour the
//     next lane's cell a; the other two are the lane's own), so no LDS and
//     no second global load serve the x direction;
//   - LDS carries only the y-neighbours, and of those only the components a
//     neighbour reads (comps 3 and 1: the +y/-y terms of the curls), for E1,
//     H1 and E2 double-buffered by plane parity: 144 KB, one barrier a plane.
// Narrow tiles.  A tile row of ox output cells takes sl = ox/2 + 2 lanes;
// a wave holds ns = 64/sl such rows side by side (segments), so the
// workgroup covers 12 ns footprint rows and outputs ox x (12 ns - 3) cells.
// The full tile is ox = 124, sl = 64, ns = 1.  Narrow tiles take the
// columns beside the full tiles when nx is not a multiple of 124 (16 of 512)
// and whole grids narrower than 124, in the same launch.  The DPP shifts
// cross from one segment into the next; as at the wave's ends, what they
// bring is read only by the halo cells at the segment's ends.  Lanes beyond
// the last segment stand in for the first lane of the tile's last row (in
// bounds) and write nothing.
// Pairs never straddle the period boundary when nx is even (cell a is odd),
// so both cells of a lane are inside, or both are images on the same side
// with one Bloch factor.  The shift beyond the wave's ends (lane 0's -x,
// lane 63's +x) reads 0; those values feed only footprint cells no output
// depends on (x0-2 and x0+124, x0+125: the halo is one cell behind and two
// ahead of the output).
#pragma once
#include "iris_yee2.h"

namespace iris {

// one component of a lane's two cells
struct P2 {
  float2 a, b;
};

__device__ __forceinline__ P2 padd(P2 u, P2 v) { return P2{cadd(u.a, v.a), cadd(u.b, v.b)}; }
__device__ __forceinline__ P2 psub(P2 u, P2 v) { return P2{csub(u.a, v.a), csub(u.b, v.b)}; }
__device__ __forceinline__ P2 pscale(float s, P2 u) { return P2{cscale(s, u.a), cscale(s, u.b)}; }
__device__ __forceinline__ P2 pmul(float2 f, P2 u) { return P2{cmul(f, u.a), cmul(f, u.b)}; }
// e + s*c and h - s*c as fused multiply-adds (cmadd / cmsub per cell)
__device__ __forceinline__ P2 pmadd(P2 e, float s, P2 c) { return P2{cmadd<float>(e.a, s, c.a), cmadd<float>(e.b, s, c.b)}; }
__device__ __forceinline__ P2 pmsub(P2 h, float s, P2 c) { return P2{cmsub<float>(h.a, s, c.a), cmsub<float>(h.b, s, c.b)}; }

__device__ __forceinline__ P2 pld(const float2* __restrict__ f, int64_t i) {
  const float4 v = *reinterpret_cast<const float4*>(f + i);
  return P2{make_float2(v.x, v.y), make_float2(v.z, v.w)};
}
__device__ __forceinline__ void pst(float2* f, int64_t i, P2 v) {
  *reinterpret_cast<float4*>(f + i) = make_float4(v.a.x, v.a.y, v.b.x, v.b.y);
}
__device__ __forceinline__ void pst_nt(float2* f, int64_t i, P2 v) {
  typedef float V4 __attribute__((ext_vector_type(4)));
  V4 w;
  w.x = v.a.x;
  w.y = v.a.y;
  w.z = v.b.x;
  w.w = v.b.y;
  __builtin_nontemporal_store(w, reinterpret_cast<V4*>(f + i));
}

// DPP wave shifts: the value of the next lane (wave_shl:1) and of the
// previous lane (wave_shr:1); 0 beyond the wave's ends.  Every lane of the
// wave must be active where they are used.
// bound_ctrl set: an invalid source lane reads 0 with no old value to set up
// (measured 5-6% faster per launch at 512 x 512 x 1024 than keeping an old
// value of 0 in a register)
__device__ __forceinline__ float from_next(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float from_prev(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
// the +x neighbours of a lane's cells: its own cell b, the next lane's cell a
__device__ __forceinline__ P2 xplus(P2 v) { return P2{v.b, make_float2(from_next(v.a.x), from_next(v.a.y))}; }
// the -x neighbours: the previous lane's cell b, its own cell a
__device__ __forceinline__ P2 xminus(P2 v) { return P2{make_float2(from_prev(v.b.x), from_prev(v.b.y)), v.a}; }

constexpr int kP2R = 12;            // waves per workgroup (footprint rows of a full tile)
constexpr int kP2W = 128;           // footprint cells per row of a full tile
constexpr int kP2OX = kP2W - 4;     // its output cells per row (x0 .. x0+123)
constexpr int kP2OY = kP2R - 3;     // its output rows

// One shape of tiles in a launch: the tile map (tm.x0: x of the first
// output cell of tile column 0, less 1), ox output cells per tile row, sl
// lanes per row, ns rows per wave.
struct P2Part {
  TileMap tm;
  int ox, sl, ns;
};

// E update of a lane's two cells outside the UPML (e_update_plain per cell)
__device__ __forceinline__ void p_e_update(const StepArgs<float>& a, int64_t col, int z, const P2 e[3], const P2 c[3],
                                           P2 n[3]) {
  if (z < a.hi || z > a.hf) {
    const float* ac = (z < a.hi) ? a.a1 : a.a3;
#pragma unroll
    for (int k = 0; k < 3; ++k) n[k] = pmadd(e[k], ac[k], c[k]);
  } else {
    const int64_t eidx = col + a.sxy * (z - a.hi);
    const float2 er = *reinterpret_cast<const float2*>(a.eps_re + eidx);  // cells a, b
    if (a.eps_im) {
      const float2 eim = *reinterpret_cast<const float2*>(a.eps_im + eidx);
      const float2 ea = make_float2(er.x, eim.x), eb = make_float2(er.y, eim.y);
#pragma unroll
      for (int k = 0; k < 3; ++k) n[k] = P2{cadd(e[k].a, cmul(ea, c[k].a)), cadd(e[k].b, cmul(eb, c[k].b))};
    } else {
#pragma unroll
      for (int k = 0; k < 3; ++k) n[k] = P2{cmadd<float>(e[k].a, er.x, c[k].a), cmadd<float>(e[k].b, er.y, c[k].b)};
    }
  }
}

template <bool PH>
__device__ __forceinline__ void yee2p_tile(const StepArgs<float>& a, float2* __restrict__ eb, float2* __restrict__ hb,
                                           const int* __restrict__ chunks, int bx, int by, int bz,
                                           const Yee2Samples<float>& smp, int slot, int slot2, int x0, int ox, int sl,
                                           int ns) {
  using C = float2;
  constexpr int NL = kP2R * 64;  // LDS entries per plane and component: rows x lanes of a row
  __shared__ P2 s1[2][2][NL];    // E1 comps 3, 1 of planes z (z & 1) and z-1
  __shared__ P2 sh[2][2][NL];    // H1 comps 3, 1
  __shared__ P2 s2[2][2][NL];    // E2 comps 3, 1
  const int lane = threadIdx.x, seg = lane / sl;
  const bool idle = seg >= ns;
  const int ry = kP2R * ns, oy = ry - 3;  // footprint and output rows of the tile
  const int cx = idle ? 0 : lane - seg * sl, cy = idle ? ry - 1 : threadIdx.y * ns + seg;
  const int nx = a.nx, ny = a.ny;
  const int64_t cs = a.cs, sxy = a.sxy;
  const int craw = chunks[bz];
  const int z0 = craw & kChunkZMask;
  const int z1 = z0 + ((craw >> kChunkLenShift) & kChunkLenMask);
  const bool exp_below = (craw & kChunkExportBelow) != 0, exp_above = (craw & kChunkExportAbove) != 0;

  // cell a of this lane (odd x) and its periodic source; cell b is the next
  const int xg = x0 + bx * ox - 1 + 2 * cx, yg = by * oy + cy;
  const int xs = xg < 1 ? xg + nx : (xg > nx ? xg - nx : xg);
  const int ys = yg < 1 ? yg + ny : (yg > ny ? yg - ny : yg);
  const bool own = !idle && cx >= 1 && cx <= ox / 2 && cy >= 1 && cy <= oy && xg <= nx && yg <= ny;
  const int64_t col = (int64_t)(xs + a.xoff) + a.sx * ys;
  const int64_t colym = col - a.sx;  // y-1 (the ghost y = 0 when ys = 1)
  // LDS entries of this lane's row and of the rows above and below (clamped
  // into the tile: edge rows compute values nobody uses)
  const int li = cy * sl + cx, lp = min(cy + 1, ry - 1) * sl + cx, lm = max(cy - 1, 0) * sl + cx;
  const C* __restrict__ e0 = a.e;
  const C* __restrict__ h0 = a.h;
  C phi = mk<float>(1.0f, 0.0f);
  if constexpr (PH) {
    if (xg < 1) phi = a.fxm; else if (xg > nx) phi = a.fxp;
    if (yg < 1) phi = cmul(phi, a.fym); else if (yg > ny) phi = cmul(phi, a.fyp);
  }
  auto ld = [&](const C* __restrict__ f, int64_t i) -> P2 {
    if constexpr (PH) return pmul(phi, pld(f, i));
    return pld(f, i);
  };
  auto store = [&](float2* f, int64_t i, P2 v) {
    if (a.nt_out) pst_nt(f, i, v);
    else pst(f, i, v);
  };
  // the DFT-box faces of this lane's cells a (bits 0..9) and b (10..19)
  const unsigned fbits = face_bits2(smp, xs, ys) | (face_bits2(smp, xs + 1, ys) << 10);

  // pipeline registers: H0(z-1), E1(z-1), H1(z-2), E2(z-2)
  P2 hz1[3], e1p[3] = {}, h1pp[3] = {}, e2p[3] = {};
  {
    const int64_t i = col + sxy * (z0 - 2);
    hz1[0] = ld(h0, i);
    hz1[1] = ld(h0, i + cs);
    hz1[2] = ld(h0, i + 2 * cs);
  }

  for (int z = z0 - 1; z <= z1 + 2; ++z) {
    const int64_t iz = col + sxy * z;
    const int pz = z & 1, pm = pz ^ 1;
    // ---- E1(z) from H0(z), H0(z-1), E0(z) (code/fdtd.f90:78-96, :137-147)
    P2 h0z[3], e1[3];
    {
      h0z[0] = ld(h0, iz);
      h0z[1] = ld(h0, iz + cs);
      h0z[2] = ld(h0, iz + 2 * cs);
      const P2 h3ym = ld(h0, colym + sxy * z + 2 * cs), h1ym = ld(h0, colym + sxy * z);
      const P2 h3xm = xminus(h0z[2]), h2xm = xminus(h0z[1]);
      const P2 c[3] = {padd(psub(psub(h0z[2], h3ym), h0z[1]), hz1[1]),
                       padd(psub(psub(h0z[0], hz1[0]), h0z[2]), h3xm),
                       padd(psub(psub(h0z[1], h2xm), h0z[0]), h1ym)};
      const P2 eo[3] = {ld(e0, iz), ld(e0, iz + cs), ld(e0, iz + 2 * cs)};
      p_e_update(a, col, z, eo, c, e1);
    }
    // ---- H1(z-1) from H0(z-1), E1(z-1) (own, +x by DPP, +y from LDS), E1(z)
    // (code/fdtd.f90:182-195)
    P2 h1n[3] = {};
    {
      const P2 e3xp = xplus(e1p[2]), e2xp = xplus(e1p[1]);
      if (z >= z0) {
        const P2 e3yp = s1[pm][0][lp], e1yp = s1[pm][1][lp];
        const P2 c0 = padd(psub(psub(e3yp, e1p[2]), e1[1]), e1p[1]);
        const P2 c1 = padd(psub(psub(e1[0], e1p[0]), e3xp), e1p[2]);
        const P2 c2 = padd(psub(psub(e2xp, e1p[1]), e1yp), e1p[0]);
        h1n[0] = pmsub(hz1[0], a.dtcq2, c0);
        h1n[1] = pmsub(hz1[1], a.dtcq2, c1);
        h1n[2] = pmsub(hz1[2], a.dtcq2, c2);
      }
    }
    if (!idle) {
      s1[pz][0][li] = e1[2];
      s1[pz][1][li] = e1[0];
      sh[pz][0][li] = h1n[2];
      sh[pz][1][li] = h1n[0];
    }
    __syncthreads();
    // ---- E2(z-1) from E1(z-1), H1(z-1) (own, -x by DPP, -y from LDS), H1(z-2)
    P2 e2n[3] = {};
    {
      const P2 h3xm = xminus(h1n[2]), h2xm = xminus(h1n[1]);
      if (z >= z0 + 1) {
        const P2 h3ym = sh[pz][0][lm], h1ym = sh[pz][1][lm];
        const P2 c[3] = {padd(psub(psub(h1n[2], h3ym), h1n[1]), h1pp[1]),
                         padd(psub(psub(h1n[0], h1pp[0]), h1n[2]), h3xm),
                         padd(psub(psub(h1n[1], h2xm), h1n[0]), h1ym)};
        p_e_update(a, col, z - 1, e1p, c, e2n);
      }
    }
    // ---- H2(z-2) from H1(z-2), E2(z-2) (own, +x by DPP, +y from LDS), E2(z-1)
    P2 h2n[3] = {};
    {
      const P2 e3xp = xplus(e2p[2]), e2xp = xplus(e2p[1]);
      if (z >= z0 + 2) {
        const P2 e3yp = s2[pm][0][lp], e1yp = s2[pm][1][lp];
        const P2 c0 = padd(psub(psub(e3yp, e2p[2]), e2n[1]), e2p[1]);
        const P2 c1 = padd(psub(psub(e2n[0], e2p[0]), e3xp), e2p[2]);
        const P2 c2 = padd(psub(psub(e2xp, e2p[1]), e1yp), e2p[0]);
        h2n[0] = pmsub(h1pp[0], a.dtcq2, c0);
        h2n[1] = pmsub(h1pp[1], a.dtcq2, c1);
        h2n[2] = pmsub(h1pp[2], a.dtcq2, c2);
      }
    }
    if (!idle) {
      s2[pz][0][li] = e2n[2];
      s2[pz][1][li] = e2n[0];
    }

    if (own) {
      // step n+2 of the chunk's planes, with H's min-side ghost planes
      if (z >= z0 + 1 && z - 1 <= z1) {
        const int64_t ie = col + sxy * (z - 1);
        store(a.en, ie, e2n[0]);
        store(a.en, ie + cs, e2n[1]);
        store(a.en, ie + 2 * cs, e2n[2]);
        if (slot2 >= 0 && (fbits & (kXFaceBits | (kXFaceBits << 10))))
          for (int b = 0; b < smp.nbox; ++b) {
            if (!smp.samp[b]) continue;
            C* const sp = smp.samp[b] + (int64_t)slot2 * smp.g[b].ntot;
            const C va[3] = {e2n[0].a, e2n[1].a, e2n[2].a}, vb[3] = {e2n[0].b, e2n[1].b, e2n[2].b};
            sample_xface(smp.g[b], sp, (fbits >> (5 * b)) & 31u, ys, z - 1, 0, va);
            sample_xface(smp.g[b], sp, (fbits >> (10 + 5 * b)) & 31u, ys, z - 1, 0, vb);
          }
      }
      if (z >= z0 + 2) {
        const int64_t ih = col + sxy * (z - 2);
        store(a.hn, ih, h2n[0]);
        store(a.hn, ih + cs, h2n[1]);
        store(a.hn, ih + 2 * cs, h2n[2]);
        if (xs + 1 == nx) {
          const C g[3] = {h2n[0].b, h2n[1].b, h2n[2].b};
          store3_phase(a.hn, ih + 1 - nx, cs, a.fxm, g);
        }
        if (ys == ny) {
          const int64_t ig = ih - a.sx * ny;
          pst(a.hn, ig, pmul(a.fym, h2n[0]));
          pst(a.hn, ig + cs, pmul(a.fym, h2n[1]));
          pst(a.hn, ig + 2 * cs, pmul(a.fym, h2n[2]));
        }
        if (slot2 >= 0 && (fbits & (kXFaceBits | (kXFaceBits << 10))))
          for (int b = 0; b < smp.nbox; ++b) {
            if (!smp.samp[b]) continue;
            C* const sp = smp.samp[b] + (int64_t)slot2 * smp.g[b].ntot;
            const C va[3] = {h2n[0].a, h2n[1].a, h2n[2].a}, vb[3] = {h2n[0].b, h2n[1].b, h2n[2].b};
            sample_xface(smp.g[b], sp, (fbits >> (5 * b)) & 31u, ys, z - 2, 3, va);
            sample_xface(smp.g[b], sp, (fbits >> (10 + 5 * b)) & 31u, ys, z - 2, 3, vb);
          }
      }
      // step n+1 planes read by the single-step passes of the neighbouring chunks
      if (exp_below && z == z0) {
        pst(eb, iz, e1[0]);
        pst(eb, iz + cs, e1[1]);
        pst(eb, iz + 2 * cs, e1[2]);
      }
      // H of step n+1 for the single steps of the chunk below, and on the
      // record planes of a gradient run's recording
      if ((exp_below && z - 1 == z0) ||
          (z - 1 >= z0 && z - 1 <= z1 && z - 1 >= a.rec_z0 && z - 1 <= a.rec_z1)) {
        const int64_t ih = col + sxy * (z - 1);
        pst(hb, ih, h1n[0]);
        pst(hb, ih + cs, h1n[1]);
        pst(hb, ih + 2 * cs, h1n[2]);
        if (xs + 1 == nx) {
          const C g[3] = {h1n[0].b, h1n[1].b, h1n[2].b};
          store3_phase(hb, ih + 1 - nx, cs, a.fxm, g);
        }
        if (ys == ny) {
          const int64_t ig = ih - a.sx * ny;
          pst(hb, ig, pmul(a.fym, h1n[0]));
          pst(hb, ig + cs, pmul(a.fym, h1n[1]));
          pst(hb, ig + 2 * cs, pmul(a.fym, h1n[2]));
        }
      }
      if (exp_above && z - 1 == z1) {
        const int64_t ih = col + sxy * z1;
        pst(hb, ih, h1n[0]);
        pst(hb, ih + cs, h1n[1]);
        pst(hb, ih + 2 * cs, h1n[2]);
      }
      // DFT samples of step n+1
      for (int b = 0; slot >= 0 && b < smp.nbox; ++b) {
        if (!smp.samp[b]) continue;
        C* const sp = smp.samp[b] + (int64_t)slot * smp.g[b].ntot;
        const unsigned fa = (fbits >> (5 * b)) & 31u, fb = (fbits >> (10 + 5 * b)) & 31u;
        if (z <= z1 && z >= z0) {
          const C va[3] = {e1[0].a, e1[1].a, e1[2].a}, vb[3] = {e1[0].b, e1[1].b, e1[2].b};
          sample_cell(smp.g[b], sp, fa, xs, ys, z, 0, va);
          sample_cell(smp.g[b], sp, fb, xs + 1, ys, z, 0, vb);
        }
        if (z - 1 >= z0 && z - 1 <= z1) {
          const C va[3] = {h1n[0].a, h1n[1].a, h1n[2].a}, vb[3] = {h1n[0].b, h1n[1].b, h1n[2].b};
          sample_cell(smp.g[b], sp, fa, xs, ys, z - 1, 3, va);
          sample_cell(smp.g[b], sp, fb, xs + 1, ys, z - 1, 3, vb);
        }
      }
    }
    // rotate the pipeline
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      hz1[k] = h0z[k];
      e1p[k] = e1[k];
      if (z >= z0) h1pp[k] = h1n[k];
      e2p[k] = e2n[k];
    }
  }
}

// Workgroups 0 .. p0.tm.nblk-1 are tiles of p0, the next npad0 - p0.tm.nblk
// return at once (the count of p0 rounded up to the XCD count, so that the
// tile order's XCD dealing holds in both parts), the rest tiles of p1.
template <bool PH>
__global__ __launch_bounds__(64 * kP2R, 3) void k_yee2p(StepArgs<float> a, float2* __restrict__ eb,
                                                        float2* __restrict__ hb, const int* __restrict__ chunks,
                                                        P2Part p0, P2Part p1, int npad0, Yee2Samples<float> smp,
                                                        int slot, int slot2) {
  int L = blockIdx.x;
  const bool second = L >= npad0;
  if (!second && L >= p0.tm.nblk) return;  // padding (no barrier has been reached)
  if (second) L -= npad0;
  const P2Part& P = second ? p1 : p0;
  int bx, by, bz;
  tile_of(P.tm, L, &bx, &by, &bz);
  yee2p_tile<PH>(a, eb, hb, chunks, bx, by, bz, smp, slot, slot2, P.tm.x0, P.ox, P.sl, P.ns);
}

// the batched form (Batch2Inst tables, as k_yee2_batch; one entry per part)
template <bool PH>
__global__ __launch_bounds__(64 * kP2R, 3) void k_yee2p_batch(const Batch2Inst<float>* __restrict__ inst,
                                                              const int* __restrict__ wg_start, int ninst, int slot,
                                                              int slot2) {
  const int blk = blockIdx.x;
  const int i = batch_instance(wg_start, ninst, blk);
  const Batch2Inst<float>& I = inst[i];
  const int L = blk - wg_start[i];
  if (L >= I.tm.nblk) return;  // padding (no barrier has been reached)
  int bx, by, bz;
  tile_of(I.tm, L, &bx, &by, &bz);
  yee2p_tile<PH>(I.a, I.eb, I.hb, I.chunks, bx, by, bz, I.smp, slot, slot2, I.tm.x0, I.ox, I.sl, I.ns);
}

}  // namespace iris
