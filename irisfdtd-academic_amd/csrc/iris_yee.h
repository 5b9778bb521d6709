// iris_yee.h -- gfx950 Yee-step kernels of the MI355X FDTD engine.
//
// One time step (reference code/driver.f90:355-450) is, by default, one
// launch of k_yee_fused over double-buffered fields; the in-place two-pass
// form (k_epass, k_hpass, k_tfsf) performs the same operations and is kept
// for comparison (it is selected with IRIS_FUSED=0).
//   E update: bulk Yee (code/fdtd.f90:57-153) and z-UPML
//     (code/boundary_II.f90:60-138), max-side Bloch wraps
//     (code/boundary.f90:193-348);
//   H update: bulk (code/fdtd.f90:157-248) and z-UPML
//     (code/boundary_II.f90:142-220), min-side Bloch wraps;
//   TF/SF incident field (code/tfsf.f90:231-355) on the two planes the
//     corrections read, and the corrections (code/tfsf.f90:76-119 for 3D
//     grids, :163-227 for 1D/2D).
// The reference double-buffers e/h and swaps pointers; every value it reads
// is an interior value of the previous step or a ghost refreshed before it is
// read.  Ghost cells it writes but never reads (E min-side, H max-side,
// corners, TF/SF corrections of E ghosts) are not written.
#pragma once
#include "iris_device.h"

namespace iris {

// ---------------------------------------------------------------------------
// Per-cell updates shared by the in-place passes and the fused pass.
// ---------------------------------------------------------------------------

// New E at one cell from its curl H (already formed in the reference's
// order, code/fdtd.f90:78-83).  Reads E_old at idx and, in the UPML slabs,
// D_old; writes D_new when store_aux.
template <typename T>
__device__ inline void e_cell(const StepArgs<T>& a, int64_t col, int iz, int64_t idx, cplx<T> c0, cplx<T> c1,
                              cplx<T> c2, bool store_aux, cplx<T> n[3]) {
  using C = cplx<T>;
  const C e0 = a.e[idx], e1 = a.e[idx + a.cs], e2 = a.e[idx + 2 * a.cs];
  if (iz <= a.pmlzi || iz >= a.pmlzf) {
    // z-UPML, reference code/boundary_II.f90:84-99
    const int k = iz >= a.pmlzf;
    const int64_t pidx = col + a.sxy * (k ? (iz - a.pmlzf) : (iz - 1));
    const PmlPlane<T>& P = a.pml[iz];
    c0 = cscale(a.e02, c0);
    c1 = cscale(a.e02, c1);
    c2 = cscale(a.e02, c2);
    const int64_t i0 = a.aux(k, 0, pidx), i1 = a.aux(k, 1, pidx), i2 = a.aux(k, 2, pidx);
    const C d0o = a.d[i0], d1o = a.d[i1], d2o = a.d[i2];
    const C d0 = cadd(cscale(P.cd[0], d0o), cscale(P.gd[0], c0));
    const C d1 = cadd(cscale(P.cd[1], d1o), cscale(P.gd[1], c1));
    const C d2 = cadd(cscale(P.cd[2], d2o), cscale(P.gd[2], c2));
    if (store_aux) {
      a.dn[i0] = d0;
      a.dn[i1] = d1;
      a.dn[i2] = d2;
    }
    n[0] = cadd(cscale(P.ce[0], e0), cscale(P.ge[0], csub(cscale(P.fp[0], d0), cscale(P.fm[0], d0o))));
    n[1] = cadd(cscale(P.ce[1], e1), cscale(P.ge[1], csub(cscale(P.fp[1], d1), cscale(P.fm[1], d1o))));
    n[2] = cadd(cscale(P.ce[2], e2), cscale(P.ge[2], csub(cscale(P.fp[2], d2), cscale(P.fm[2], d2o))));
  } else if (iz < a.hi) {
    // region I, reference code/fdtd.f90:86-96 (off-diagonal terms are 0)
    n[0] = cmadd<T>(e0, a.a1[0], c0);
    n[1] = cmadd<T>(e1, a.a1[1], c1);
    n[2] = cmadd<T>(e2, a.a1[2], c2);
  } else if (iz <= a.hf) {
    // structure, reference code/fdtd.f90:115-117
    const int64_t eidx = col + a.sxy * (iz - a.hi);
    const T er = a.eps_re[eidx];
    if (a.eps_im) {
      const C ei = mk<T>(er, a.eps_im[eidx]);
      n[0] = cadd(e0, cmul(ei, c0));
      n[1] = cadd(e1, cmul(ei, c1));
      n[2] = cadd(e2, cmul(ei, c2));
    } else {
      n[0] = cmadd<T>(e0, er, c0);
      n[1] = cmadd<T>(e1, er, c1);
      n[2] = cmadd<T>(e2, er, c2);
    }
  } else {
    // region III, reference code/fdtd.f90:137-147
    n[0] = cmadd<T>(e0, a.a3[0], c0);
    n[1] = cmadd<T>(e1, a.a3[1], c1);
    n[2] = cmadd<T>(e2, a.a3[2], c2);
  }
}

// New H at one cell from its curl E (code/fdtd.f90:182-187) and H_old;
// B_old/B_new in the UPML slabs.
template <typename T>
__device__ inline void h_cell(const StepArgs<T>& a, int64_t col, int iz, cplx<T> h0, cplx<T> h1, cplx<T> h2,
                              cplx<T> c0, cplx<T> c1, cplx<T> c2, cplx<T> n[3]) {
  using C = cplx<T>;
  if (iz <= a.pmlzi || iz >= a.pmlzf) {
    // z-UPML, reference code/boundary_II.f90:170-183
    const int k = iz >= a.pmlzf;
    const int64_t pidx = col + a.sxy * (k ? (iz - a.pmlzf) : (iz - 1));
    const PmlPlane<T>& P = a.pml[iz];
    const int64_t i0 = a.aux(k, 0, pidx), i1 = a.aux(k, 1, pidx), i2 = a.aux(k, 2, pidx);
    const C b0o = a.b[i0], b1o = a.b[i1], b2o = a.b[i2];
    const C b0 = csub(cscale(P.cd[0], b0o), cscale(P.gb[0], c0));
    const C b1 = csub(cscale(P.cd[1], b1o), cscale(P.gb[1], c1));
    const C b2 = csub(cscale(P.cd[2], b2o), cscale(P.gb[2], c2));
    a.bn[i0] = b0;
    a.bn[i1] = b1;
    a.bn[i2] = b2;
    n[0] = cadd(cscale(P.ce[0], h0), cscale(P.gh[0], csub(cscale(P.fp[0], b0), cscale(P.fm[0], b0o))));
    n[1] = cadd(cscale(P.ce[1], h1), cscale(P.gh[1], csub(cscale(P.fp[1], b1), cscale(P.fm[1], b1o))));
    n[2] = cadd(cscale(P.ce[2], h2), cscale(P.gh[2], csub(cscale(P.fp[2], b2), cscale(P.fm[2], b2o))));
  } else {
    // bulk, reference code/fdtd.f90:191-195 (mu_inv_bbox = 1)
    n[0] = cmsub<T>(h0, a.dtcq2, c0);
    n[1] = cmsub<T>(h1, a.dtcq2, c1);
    n[2] = cmsub<T>(h2, a.dtcq2, c2);
  }
}

template <typename C> __device__ inline void store3(C* f, int64_t i, int64_t cs, const C v[3]) {
  f[i] = v[0];
  f[i + cs] = v[1];
  f[i + 2 * cs] = v[2];
}
template <typename C> __device__ inline void store3_phase(C* f, int64_t i, int64_t cs, C ph, const C v[3]) {
  f[i] = cmul(ph, v[0]);
  f[i + cs] = cmul(ph, v[1]);
  f[i + 2 * cs] = cmul(ph, v[2]);
}

// ---------------------------------------------------------------------------
// In-place E pass.  Thread = one (ix,iy) column, marching zchunk planes
// upward; the z-1 values of H1/H2 stay in registers between planes.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_epass(StepArgs<T> a) {
  using C = cplx<T>;
  const int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (ix > a.nx || iy > a.ny) return;
  const int z0 = blockIdx.z * a.zchunk + 1;
  const int z1 = min(z0 + a.zchunk - 1, a.nz);
  const int64_t col = (int64_t)(ix + a.xoff) + a.sx * iy;
  const int64_t cs = a.cs;
  int64_t idx = col + a.sxy * z0;
  C h1m = a.h[idx - a.sxy];
  C h2m = a.h[idx - a.sxy + cs];
  for (int iz = z0; iz <= z1; ++iz, idx += a.sxy) {
    const C h1 = a.h[idx], h2 = a.h[idx + cs], h3 = a.h[idx + 2 * cs];
    const C h3ym = a.h[idx - a.sx + 2 * cs], h1ym = a.h[idx - a.sx];
    const C h3xm = a.h[idx - 1 + 2 * cs], h2xm = a.h[idx - 1 + cs];
    // curl H, reference code/fdtd.f90:78-83
    const C c0 = cadd(csub(csub(h3, h3ym), h2), h2m);
    const C c1 = cadd(csub(csub(h1, h1m), h3), h3xm);
    const C c2 = cadd(csub(csub(h2, h2xm), h1), h1ym);
    h1m = h1;
    h2m = h2;
    C n[3];
    e_cell<T>(a, col, iz, idx, c0, c1, c2, true, n);
    store3(a.e, idx, cs, n);
    // max-side Bloch wraps, reference code/boundary.f90:205-212,269-278,333-342
    if (ix == 1) store3_phase(a.e, idx + a.nx, cs, a.fxp, n);
    if (iy == 1) store3_phase(a.e, idx + a.sx * a.ny, cs, a.fyp, n);
    if (iz == 1) store3_phase(a.e, idx + a.sxy * a.nz, cs, a.fzp, n);
  }
}

// ---------------------------------------------------------------------------
// In-place H pass.  Same column marching; E1/E2 at z+1 become the next
// plane's own values.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_hpass(StepArgs<T> a) {
  using C = cplx<T>;
  const int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (ix > a.nx || iy > a.ny) return;
  const int z0 = blockIdx.z * a.zchunk + 1;
  const int z1 = min(z0 + a.zchunk - 1, a.nz);
  const int64_t col = (int64_t)(ix + a.xoff) + a.sx * iy;
  const int64_t cs = a.cs;
  int64_t idx = col + a.sxy * z0;
  C e1 = a.e[idx];
  C e2 = a.e[idx + cs];
  for (int iz = z0; iz <= z1; ++iz, idx += a.sxy) {
    const C e1p = a.e[idx + a.sxy], e2p = a.e[idx + a.sxy + cs];
    const C e3 = a.e[idx + 2 * cs];
    const C e3yp = a.e[idx + a.sx + 2 * cs], e1yp = a.e[idx + a.sx];
    const C e3xp = a.e[idx + 1 + 2 * cs], e2xp = a.e[idx + 1 + cs];
    // curl E, reference code/fdtd.f90:182-187
    const C c0 = cadd(csub(csub(e3yp, e3), e2p), e2);
    const C c1 = cadd(csub(csub(e1p, e1), e3xp), e3);
    const C c2 = cadd(csub(csub(e2xp, e2), e1yp), e1);
    e1 = e1p;
    e2 = e2p;
    C n[3];
    h_cell<T>(a, col, iz, a.h[idx], a.h[idx + cs], a.h[idx + 2 * cs], c0, c1, c2, n);
    store3(a.h, idx, cs, n);
    // min-side Bloch wraps, reference code/boundary.f90:177-183,241-247,305-311
    if (ix == a.nx) store3_phase(a.h, idx - a.nx, cs, a.fxm, n);
    if (iy == a.ny) store3_phase(a.h, idx - a.sx * a.ny, cs, a.fym, n);
    if (iz == a.nz) store3_phase(a.h, idx - a.sxy * a.nz, cs, a.fzm, n);
  }
}

// ---------------------------------------------------------------------------
// TF/SF source.  The Gaussian-modulated plane wave depends on (ix, iz, it)
// only (u%y = 0), evaluated in FP64 exactly as reference code/tfsf.f90:327-346.
// ---------------------------------------------------------------------------

// incident values of comps {1,3,5} (i_pol=2) or {2,4,6} (i_pol=1) at (ix, iz)
__device__ inline void tfsf_incident(const TfsfConst& t, int it, int ix, int iz, double2* inc) {
  if (t.tab) {
    // matched oblique field: the table's value at ix = 0 times e^{i phix ix}
    const double2* v = t.tab + ((int64_t)(it + t.shift) * 2 + (iz - (t.iz_min - 1))) * 3;
    double sn, cs;
    sincos(t.phix * (double)ix, &sn, &cs);
    const double2 ph = make_double2(cs, sn);
    for (int k = 0; k < 3; ++k) inc[k] = cmul(v[k], ph);
    return;
  }
  const double c0dt_it = t.c0dt * (double)(it + t.shift);
  const double A = (t.q * (double)(ix - t.ixo)) * t.ux + (t.q * ((double)iz - t.zopump)) * t.uz;
  const double kzz = t.ko * ((t.sq * t.sign_kz) * A - c0dt_it);
  const double tt = (A - c0dt_it / t.sq) / t.dk;
  const double env = exp(-(tt * tt));
  double sn, cs;
  sincos(kzz, &sn, &cs);
  if (t.i_pol == 2) {
    // ex = s*exp(i kzz)*env*uz ; ez = -exp(i kzz)*env*ux ; hy = uz*s*ex - ux*ez
    double2 ex = make_double2(((t.sign_kz * cs) * env) * t.uz, ((t.sign_kz * sn) * env) * t.uz);
    double2 ez = make_double2(-((cs * env) * t.ux), -((sn * env) * t.ux));
    const double us = t.uz * t.sign_kz;
    double2 hy = make_double2(us * ex.x - t.ux * ez.x, us * ex.y - t.ux * ez.y);
    if (t.cosine == 1) { ex.y = 0.0; ez.y = 0.0; hy.y = 0.0; }
    inc[0] = make_double2(t.win * (t.amp * ex.x), t.win * (t.amp * ex.y));     // comp 1
    inc[1] = make_double2(t.win * (t.amp * ez.x), t.win * (t.amp * ez.y));     // comp 3
    inc[2] = make_double2(t.win * (t.hamp * hy.x), t.win * (t.hamp * hy.y));   // comp 5
  } else {
    // hx = -s*exp(i kzz)*env*uz ; hz = exp(i kzz)*env*ux ; ey = ux*hz - uz*s*hx
    double2 hx = make_double2(-(((t.sign_kz * cs) * env) * t.uz), -(((t.sign_kz * sn) * env) * t.uz));
    double2 hz = make_double2((cs * env) * t.ux, (sn * env) * t.ux);
    const double us = t.uz * t.sign_kz;
    double2 ey = make_double2(t.ux * hz.x - us * hx.x, t.ux * hz.y - us * hx.y);
    if (t.cosine == 1) { ey.y = 0.0; hz.y = 0.0; hx.y = 0.0; }
    inc[0] = make_double2(t.win * (t.amp * ey.x), t.win * (t.amp * ey.y));     // comp 2
    inc[1] = make_double2(t.win * (t.hamp * hx.x), t.win * (t.hamp * hx.y));   // comp 4
    inc[2] = make_double2(t.win * (t.hamp * hz.x), t.win * (t.hamp * hz.y));   // comp 6
  }
}

template <typename T> __device__ inline cplx<T> to_c(double2 v) { return mk<T>((T)v.x, (T)v.y); }

// field value +/- an FP64 increment, rounded to the field precision first
template <typename T>
__device__ inline cplx<T> add_inc(cplx<T> v, double2 inc) {
  if constexpr (sizeof(T) == 8) return cadd(v, inc);
  else return cadd(v, to_c<T>(inc));
}
template <typename T>
__device__ inline cplx<T> sub_inc(cplx<T> v, double2 inc) {
  if constexpr (sizeof(T) == 8) return csub(v, inc);
  else return csub(v, to_c<T>(inc));
}
inline __device__ double2 dscale(double s, double2 v) { return make_double2(s * v.x, s * v.y); }

// Correction added to the E value stored at (ix, iz_min): component comp
// (0 or 1) gets +inc (sign +1) or -inc (sign -1).
__device__ inline double2 tfsf_e_increment(const TfsfConst& t, int it, int ix, int* comp, int* sign) {
  double2 inc[3];
  if (t.i_pol == 2) {
    *comp = 0;
    *sign = 1;
    if (!t.mode2d) {
      // Ex += eps_in(1)*inc5(iz_min-1), reference code/tfsf.f90:95-98
      tfsf_incident(t, it, ix, t.iz_min - 1, inc);
      return dscale(t.ein, inc[2]);
    }
    // 1D/2D: Ex += inc5(iz_min)/eps, reference code/tfsf.f90:178
    tfsf_incident(t, it, ix, t.iz_min, inc);
    return cdiv_scaled(inc[2], make_double2(t.eps_ms, 0.0));
  }
  // Ey -= eps_in(1)*inc4(iz_min-1), reference code/tfsf.f90:101-105
  *comp = 1;
  *sign = -1;
  tfsf_incident(t, it, ix, t.iz_min - 1, inc);
  return dscale(t.ein, inc[1]);
}

// Plane of the H correction and its increment at ix (reference
// code/tfsf.f90:108-118 for 3D, :189-190 for 1D/2D).
__device__ inline int tfsf_h_plane(const TfsfConst& t) {
  return (t.i_pol == 2 && t.mode2d) ? t.iz_min : t.iz_min - 1;
}
__device__ inline double2 tfsf_h_increment(const TfsfConst& t, int it, int ix, int* comp, int* sign) {
  double2 inc[3];
  tfsf_incident(t, it, ix, t.iz_min, inc);
  *comp = (t.i_pol == 2) ? 1 : 0;   // Hy += dtcq2*inc1 ; Hx -= dtcq2*inc2
  *sign = (t.i_pol == 2) ? 1 : -1;
  return dscale(t.dtcq2, inc[0]);
}

template <typename T>
__device__ inline void apply_inc(cplx<T>* v, int comp, int sign, double2 d) {
  v[comp] = sign > 0 ? add_inc<T>(v[comp], d) : sub_inc<T>(v[comp], d);
}

// In-place TF/SF pass.  Thread = (ix, iy) with ix in [0,nx], iy in [0,ny].
// Only cells that a later step reads are corrected: interior E, and interior
// H plus the one ghost line the next E pass reads (ix=0 for Hy, iy=0 for Hx).
template <typename T>
__global__ __launch_bounds__(256) void k_tfsf(StepArgs<T> a, const TfsfConst* tp, int it) {
  using C = cplx<T>;
  const TfsfConst& t = *tp;
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  if (ix > a.nx || iy > a.ny) return;
  const int64_t col = (int64_t)(ix + a.xoff) + a.sx * iy;
  int comp, sign;
  if (ix >= 1 && iy >= 1) {
    const double2 d = tfsf_e_increment(t, it, ix, &comp, &sign);
    const int64_t i = col + a.sxy * t.iz_min + comp * a.cs;
    C v[3];
    v[comp] = a.e[i];
    apply_inc<T>(v, comp, sign, d);
    a.e[i] = v[comp];
  }
  const double2 d = tfsf_h_increment(t, it, ix, &comp, &sign);
  const bool live = (t.i_pol == 2) ? iy >= 1 : ix >= 1;
  if (live) {
    const int64_t i = col + a.sxy * tfsf_h_plane(t) + comp * a.cs;
    C v[3];
    v[comp] = a.h[i];
    apply_inc<T>(v, comp, sign, d);
    a.h[i] = v[comp];
  }
}

// ---------------------------------------------------------------------------
// Fused E+H pass.  One sweep per time step over double-buffered fields:
// E_new(z) is computed from E_old/H_old, and one plane behind it H_new(z-1)
// from H_old and E_new, so E and H are each read once and written once
// (96 B/cell in FP32 instead of 144 for the two in-place passes).
//
// Block = a kFTX x kFTY (x,y) tile marching a z-chunk: kFTY waves for the
// tile's own cells plus one "halo" wave.  H_new at the tile's last
// column/row needs E_new at x+1 / y+1, which belong to the neighbouring
// tiles, so the halo wave recomputes E_new on the +y row and (lanes
// 0..kFTY-1) on the +x column from the old fields; the values are identical
// to what the neighbour stores.  Cells at index nx+1, ny+1 or nz+1 are the
// Bloch images of index 1 (reference code/boundary.f90:193-348), computed
// from the index-1 inputs and multiplied by exp(+i ak).  E_new of the current
// plane travels to the H update of the next iteration through LDS (two
// barriers per plane).  The TF/SF corrections are applied to the values
// stored, after the H update has used the uncorrected E, as in the
// reference; the H min-side ghost planes (read by the next step's E) are
// stored with their own corrections.  E's max-side ghosts are not stored:
// in this scheme nothing reads them.
// ---------------------------------------------------------------------------
#ifndef IRIS_FUSED_TY
#define IRIS_FUSED_TY 8
#endif
#ifndef IRIS_FUSED_MINBLOCKS
#define IRIS_FUSED_MINBLOCKS 1
#endif
constexpr int kFTX = 64, kFTY = IRIS_FUSED_TY;

// E_new at source cell (xs,ys,zs) (all in 1..n) from the old fields; the
// z-1 values of H1/H2 come in h1m/h2m, H_old of the cell is returned in hown.
template <typename T, bool GEN>
__device__ inline void e_item(const StepArgs<T>& a, int xs, int ys, int zs, cplx<T> h1m, cplx<T> h2m,
                              bool store_aux, cplx<T> n[3], cplx<T> hown[3]) {
  using C = cplx<T>;
  const int64_t cs = a.cs;
  const int64_t col = (int64_t)(xs + a.xoff) + a.sx * ys;
  const int64_t idx = col + a.sxy * zs;
  const C h1 = a.h[idx], h2 = a.h[idx + cs], h3 = a.h[idx + 2 * cs];
  const C h3ym = a.h[idx - a.sx + 2 * cs], h1ym = a.h[idx - a.sx];
  const C h3xm = a.h[idx - 1 + 2 * cs], h2xm = a.h[idx - 1 + cs];
  const C c0 = cadd(csub(csub(h3, h3ym), h2), h2m);
  const C c1 = cadd(csub(csub(h1, h1m), h3), h3xm);
  const C c2 = cadd(csub(csub(h2, h2xm), h1), h1ym);
  if constexpr (GEN) {
    e_cell<T>(a, col, zs, idx, c0, c1, c2, store_aux, n);
  } else {
    // bulk planes: region I or III only (reference code/fdtd.f90:86-96, :137-147)
    const T* ac = (zs < a.hi) ? a.a1 : a.a3;
    n[0] = cmadd<T>(a.e[idx], ac[0], c0);
    n[1] = cmadd<T>(a.e[idx + cs], ac[1], c1);
    n[2] = cmadd<T>(a.e[idx + 2 * cs], ac[2], c2);
  }
  hown[0] = h1;
  hown[1] = h2;
  hown[2] = h3;
}

// Tile order of a fused launch.  Workgroups are dealt to the 8 XCDs
// round-robin by workgroup id, and each XCD has its own L2.  With run > 0
// the tiles, in (x, y, chunk) order, are cut into runs of `run` tiles and the
// runs are dealt to the XCDs round-robin: the workgroups of one XCD process
// whole runs, so x- and y-neighbours inside a run, whose halo rows and
// columns overlap, share an L2, while every chunk is still spread over all
// XCDs.  Tiles past the last whole round of runs, and run = 0, keep the
// identity order t = L.
// chunks[] word of the two-step pass: first plane (bits 0-19), planes - 1
// (bits 20-27), export flags (bits 28-29).  A single-step chunk word is its
// first plane (bits 0-19) and its planes (bits 20-27; 0: a.zchunk planes),
// up to nz.
constexpr int kChunkZMask = (1 << 20) - 1;
constexpr int kChunkLenShift = 20, kChunkLenMask = 0xff;
constexpr int kChunkExportBelow = 1 << 28;
constexpr int kChunkExportAbove = 1 << 29;

struct TileMap {
  int tx, ty;  // tiles along x and y
  int nblk;    // workgroups in the launch (tx * ty * chunks)
  int run;     // tiles per run (0: identity order)
  int x0 = 0;  // two-step pass: x of the footprint's first cell in tile column 0, less its usual place
               // (a strip of columns beside tiles of another kernel)
};
constexpr int kNumXcd = 8;

__device__ inline void tile_of(const TileMap& m, int L, int* bx, int* by, int* bz) {
  int t = L;
  if (m.run > 0) {
    const int round = kNumXcd * m.run;
    if (L < (m.nblk / round) * round) {
      const int x = L % kNumXcd, loc = L / kNumXcd;
      t = ((loc / m.run) * kNumXcd + x) * m.run + loc % m.run;
    }
  }
  *bx = t % m.tx;
  *by = (t / m.tx) % m.ty;
  *bz = t / (m.tx * m.ty);
}

template <typename C> __device__ inline void phase3(C ph, C v[3]) {
  v[0] = cmul(ph, v[0]);
  v[1] = cmul(ph, v[1]);
  v[2] = cmul(ph, v[2]);
}

// GEN = false: the lean variant for z-chunks whose planes z0-1..z1+1 are all
// bulk (between the UPML slabs, outside the structure and the TF/SF planes,
// z1 < nz); GEN = true handles every other chunk.  chunks[] lists the first
// plane of each chunk this launch processes; the grid is one-dimensional,
// tm.nblk workgroups, and tm gives each its tile.
// The work of workgroup `blk` of a fused launch (tile and chunk from tm).
template <typename T, bool GEN>
__device__ __forceinline__ void yee_fused_tile(const StepArgs<T>& a, const TfsfConst* tp, int it, const int* chunks,
                                               const TileMap& tm, int blk) {
  using C = cplx<T>;
  __shared__ C se[3][kFTY + 1][kFTX + 1];
  const int lane = threadIdx.x, wy = threadIdx.y;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t cs = a.cs;
  int tbx, tby, tbz;
  tile_of(tm, blk, &tbx, &tby, &tbz);
  const int x0 = tbx * kFTX + 1, y0 = tby * kFTY + 1;
  const int cw = chunks[tbz];
  const int z0 = cw & kChunkZMask, zl = (cw >> kChunkLenShift) & kChunkLenMask;
  const int z1 = min(z0 + (zl ? zl : a.zchunk) - 1, nz);
  const bool halo = wy == kFTY;
  // item A: own cell, or a Bloch image inside a partial tile, or (halo wave)
  // the +y neighbour row; item B (halo wave lanes < kFTY): +x neighbour column
  const int xa = x0 + lane, ya = y0 + wy;
  bool va = xa <= nx + 1 && ya <= ny + 1 && !(xa == nx + 1 && ya == ny + 1);
  if (halo) va = va && xa <= nx;
  const bool owna = !halo && xa <= nx && ya <= ny;
  const int xsa = (xa == nx + 1) ? 1 : xa, ysa = (ya == ny + 1) ? 1 : ya;
  const int xb = x0 + kFTX, yb = y0 + lane;
  const bool vb = halo && lane < kFTY && xb <= nx + 1 && yb <= ny;
  const int xsb = (xb == nx + 1) ? 1 : xb;
  const int64_t cola = (int64_t)(xa + a.xoff) + a.sx * ya;  // own store column
  const bool tf_on = GEN && tp->on != 0;
  const int tf_ez = GEN ? tp->iz_min : 0, tf_hz = GEN ? tfsf_h_plane(*tp) : 0;

  C a1m{}, a2m{}, a3m{}, b1m{}, b2m{}, ep[3] = {};
  if (va) {
    const int64_t i = (int64_t)(xsa + a.xoff) + a.sx * ysa + a.sxy * (z0 - 1);
    a1m = a.h[i];
    a2m = a.h[i + cs];
    a3m = a.h[i + 2 * cs];
  }
  if (vb) {
    const int64_t i = (int64_t)(xsb + a.xoff) + a.sx * yb + a.sxy * (z0 - 1);
    b1m = a.h[i];
    b2m = a.h[i + cs];
  }

  for (int z = z0; z <= z1 + 1; ++z) {
    const bool real_plane = z <= z1;
    const bool image_z = GEN && z == nz + 1;
    const int zs = image_z ? 1 : z;
    C na[3], nb[3], ha[3], hb[3];
    const bool doa = va && (real_plane || owna);
    if (doa) {
      // z-1 values for this E; on the image plane they come from the ghost
      // plane 0, while a1m/a2m keep H_old(nz) for the H update below
      C z1m = a1m, z2m = a2m;
      if (image_z) {
        const int64_t i = (int64_t)(xsa + a.xoff) + a.sx * ysa;
        z1m = a.h[i];
        z2m = a.h[i + cs];
      }
      e_item<T, GEN>(a, xsa, ysa, zs, z1m, z2m, owna && real_plane, na, ha);
      if (xa == nx + 1) phase3(a.fxp, na);
      if (ya == ny + 1) phase3(a.fyp, na);
      if (image_z) phase3(a.fzp, na);
    }
    const bool dob = vb && real_plane;
    if (dob) {
      e_item<T, GEN>(a, xsb, yb, zs, b1m, b2m, false, nb, hb);
      if (xb == nx + 1) phase3(a.fxp, nb);
    }
    // H update of plane z-1: H_old(z-1) is (a1m,a2m,a3m), E_new(z-1) own is
    // ep, its +x/+y neighbours are in LDS, E_new(z) own is na
    if (owna && z > z0) {
      const int zh = z - 1;
      const C e3yp = se[2][wy + 1][lane], e1yp = se[0][wy + 1][lane];
      const C e3xp = se[2][wy][lane + 1], e2xp = se[1][wy][lane + 1];
      const C c0 = cadd(csub(csub(e3yp, ep[2]), na[1]), ep[1]);
      const C c1 = cadd(csub(csub(na[0], ep[0]), e3xp), ep[2]);
      const C c2 = cadd(csub(csub(e2xp, ep[1]), e1yp), ep[0]);
      C hn[3];
      if constexpr (GEN) {
        h_cell<T>(a, cola, zh, a1m, a2m, a3m, c0, c1, c2, hn);
      } else {
        hn[0] = cmsub<T>(a1m, a.dtcq2, c0);
        hn[1] = cmsub<T>(a2m, a.dtcq2, c1);
        hn[2] = cmsub<T>(a3m, a.dtcq2, c2);
      }
      C gx[3] = {cmul(a.fxm, hn[0]), cmul(a.fxm, hn[1]), cmul(a.fxm, hn[2])};
      C gy[3] = {cmul(a.fym, hn[0]), cmul(a.fym, hn[1]), cmul(a.fym, hn[2])};
      if (tf_on && zh == tf_hz) {
        int comp, sign;
        const double2 d = tfsf_h_increment(*tp, it, xa, &comp, &sign);
        apply_inc<T>(hn, comp, sign, d);
        if (tp->i_pol == 2) {
          // ghost x=0 carries the incident field at ix=0 (reference range ix_min=0)
          if (xa == nx) apply_inc<T>(gx, comp, sign, tfsf_h_increment(*tp, it, 0, &comp, &sign));
        } else {
          apply_inc<T>(gy, comp, sign, d);  // ghost y=0, same ix
        }
      }
      const int64_t ih = cola + a.sxy * zh;
      store3(a.hn, ih, cs, hn);
      if (xa == nx) store3(a.hn, ih - nx, cs, gx);
      if (ya == ny) store3(a.hn, ih - a.sx * ny, cs, gy);
      if (GEN && zh == nz) store3_phase(a.hn, ih - a.sxy * nz, cs, a.fzm, hn);
    }
    if (doa) { a1m = ha[0]; a2m = ha[1]; a3m = ha[2]; }
    if (dob) { b1m = hb[0]; b2m = hb[1]; }
    if (!real_plane) break;
    if (owna) {
      C st[3] = {na[0], na[1], na[2]};
      if (tf_on && z == tf_ez) {
        int comp, sign;
        const double2 d = tfsf_e_increment(*tp, it, xa, &comp, &sign);
        apply_inc<T>(st, comp, sign, d);
      }
      store3(a.en, cola + a.sxy * z, cs, st);
    }
    __syncthreads();
    if (doa) { se[0][wy][lane] = na[0]; se[1][wy][lane] = na[1]; se[2][wy][lane] = na[2]; }
    if (dob) { se[0][lane][kFTX] = nb[0]; se[1][lane][kFTX] = nb[1]; se[2][lane][kFTX] = nb[2]; }
    __syncthreads();
    if (owna) { ep[0] = na[0]; ep[1] = na[1]; ep[2] = na[2]; }
  }
}

template <typename T, bool GEN>
__global__ __launch_bounds__(kFTX * (kFTY + 1), IRIS_FUSED_MINBLOCKS) void k_yee_fused(StepArgs<T> a, const TfsfConst* tp, int it,
                                                                                      const int* chunks, TileMap tm) {
  yee_fused_tile<T, GEN>(a, tp, it, chunks, tm, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Batched fused pass: one launch advances many independent grids (a sweep
// over angles, Bloch phases or structures) by one step.  Instance i owns
// workgroups wg_start[i] .. wg_start[i+1]-1; its arguments for the two
// buffer parities sit in device memory.  Per workgroup the work is exactly
// that of the single-grid pass.
// ---------------------------------------------------------------------------
template <typename T> struct BatchInst {
  StepArgs<T> a[2];         // parity 0: current buffers -> other; parity 1: the reverse
  const TfsfConst* tp;
  const int* chunks[2];     // lean chunks, general chunks
  TileMap tm[2];
};

__device__ inline int batch_instance(const int* __restrict__ wg_start, int ninst, int blk) {
  int lo = 0, hi = ninst;  // wg_start[0] = 0 <= blk < wg_start[ninst]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (wg_start[mid] <= blk) lo = mid; else hi = mid;
  }
  return lo;
}

template <typename T, bool GEN>
__global__ __launch_bounds__(kFTX * (kFTY + 1), IRIS_FUSED_MINBLOCKS) void k_yee_fused_batch(
    const BatchInst<T>* __restrict__ inst, const int* __restrict__ wg_start, int ninst, int par, int it) {
  const int blk = blockIdx.x;
  const int i = batch_instance(wg_start, ninst, blk);
  const BatchInst<T>& I = inst[i];
  yee_fused_tile<T, GEN>(I.a[par], I.tp, it, I.chunks[GEN ? 1 : 0], I.tm[GEN ? 1 : 0], blk - wg_start[i]);
}

}  // namespace iris
