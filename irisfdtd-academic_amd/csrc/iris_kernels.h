// iris_kernels.h -- gfx950 kernels of the MI355X FDTD engine.
//
// One time step (reference code/driver.f90:355-450) is four launches on one
// stream, all updating the fields in place (the reference double-buffers
// e/h and swaps pointers; every value it reads is reproduced in place):
//   k_epass   E update: bulk Yee (code/fdtd.f90:57-153) and z-UPML
//             (code/boundary_II.f90:60-138) in one z-marching pass, with the
//             max-side Bloch wraps of E (code/boundary.f90:193-348) written by
//             the threads that own the wrapped cells;
//   k_hpass   H update: bulk (code/fdtd.f90:157-248) + z-UPML
//             (code/boundary_II.f90:142-220) + min-side Bloch wraps of H;
//   k_tfsf    TF/SF incident field (code/tfsf.f90:231-355) on the two planes
//             the corrections read, and the corrections themselves
//             (code/tfsf.f90:76-119 for 3D grids, :163-227 for 1D/2D);
//   k_dft_*   on-the-fly DFT (code/dataprocess.f90:109-209): samples of the
//             box faces are buffered and folded into the FP64 accumulators
//             every K samples, in the reference's accumulation order.
// Ghost cells the reference writes but never reads (E min-side, H max-side,
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
  const C e0 = a.e[0][idx], e1 = a.e[1][idx], e2 = a.e[2][idx];
  if (iz <= a.pmlzi || iz >= a.pmlzf) {
    // z-UPML, reference code/boundary_II.f90:84-99
    const int k = iz >= a.pmlzf;
    const int64_t pidx = col + a.sxy * (k ? (iz - a.pmlzf) : (iz - 1));
    const PmlPlane<T>& P = a.pml[iz];
    c0 = cscale(a.e02, c0);
    c1 = cscale(a.e02, c1);
    c2 = cscale(a.e02, c2);
    const C d0o = a.d[k][0][pidx], d1o = a.d[k][1][pidx], d2o = a.d[k][2][pidx];
    const C d0 = cadd(cscale(P.cd[0], d0o), cscale(P.gd[0], c0));
    const C d1 = cadd(cscale(P.cd[1], d1o), cscale(P.gd[1], c1));
    const C d2 = cadd(cscale(P.cd[2], d2o), cscale(P.gd[2], c2));
    if (store_aux) {
      a.dn[k][0][pidx] = d0;
      a.dn[k][1][pidx] = d1;
      a.dn[k][2][pidx] = d2;
    }
    n[0] = cadd(cscale(P.ce[0], e0), cscale(P.ge[0], csub(cscale(P.fp[0], d0), cscale(P.fm[0], d0o))));
    n[1] = cadd(cscale(P.ce[1], e1), cscale(P.ge[1], csub(cscale(P.fp[1], d1), cscale(P.fm[1], d1o))));
    n[2] = cadd(cscale(P.ce[2], e2), cscale(P.ge[2], csub(cscale(P.fp[2], d2), cscale(P.fm[2], d2o))));
  } else if (iz < a.hi) {
    // region I, reference code/fdtd.f90:86-96 (off-diagonal terms are 0)
    n[0] = cadd(e0, cscale(a.a1[0], c0));
    n[1] = cadd(e1, cscale(a.a1[1], c1));
    n[2] = cadd(e2, cscale(a.a1[2], c2));
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
      n[0] = cadd(e0, cscale(er, c0));
      n[1] = cadd(e1, cscale(er, c1));
      n[2] = cadd(e2, cscale(er, c2));
    }
  } else {
    // region III, reference code/fdtd.f90:137-147
    n[0] = cadd(e0, cscale(a.a3[0], c0));
    n[1] = cadd(e1, cscale(a.a3[1], c1));
    n[2] = cadd(e2, cscale(a.a3[2], c2));
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
    const C b0o = a.b[k][0][pidx], b1o = a.b[k][1][pidx], b2o = a.b[k][2][pidx];
    const C b0 = csub(cscale(P.cd[0], b0o), cscale(P.gb[0], c0));
    const C b1 = csub(cscale(P.cd[1], b1o), cscale(P.gb[1], c1));
    const C b2 = csub(cscale(P.cd[2], b2o), cscale(P.gb[2], c2));
    a.bn[k][0][pidx] = b0;
    a.bn[k][1][pidx] = b1;
    a.bn[k][2][pidx] = b2;
    n[0] = cadd(cscale(P.ce[0], h0), cscale(P.gh[0], csub(cscale(P.fp[0], b0), cscale(P.fm[0], b0o))));
    n[1] = cadd(cscale(P.ce[1], h1), cscale(P.gh[1], csub(cscale(P.fp[1], b1), cscale(P.fm[1], b1o))));
    n[2] = cadd(cscale(P.ce[2], h2), cscale(P.gh[2], csub(cscale(P.fp[2], b2), cscale(P.fm[2], b2o))));
  } else {
    // bulk, reference code/fdtd.f90:191-195 (mu_inv_bbox = 1)
    n[0] = csub(h0, cscale(a.dtcq2, c0));
    n[1] = csub(h1, cscale(a.dtcq2, c1));
    n[2] = csub(h2, cscale(a.dtcq2, c2));
  }
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
  int64_t idx = col + a.sxy * z0;
  C h1m = a.h[0][idx - a.sxy];
  C h2m = a.h[1][idx - a.sxy];
  for (int iz = z0; iz <= z1; ++iz, idx += a.sxy) {
    const C h1 = a.h[0][idx], h2 = a.h[1][idx], h3 = a.h[2][idx];
    const C h3ym = a.h[2][idx - a.sx], h1ym = a.h[0][idx - a.sx];
    const C h3xm = a.h[2][idx - 1], h2xm = a.h[1][idx - 1];
    // curl H, reference code/fdtd.f90:78-83
    const C c0 = cadd(csub(csub(h3, h3ym), h2), h2m);
    const C c1 = cadd(csub(csub(h1, h1m), h3), h3xm);
    const C c2 = cadd(csub(csub(h2, h2xm), h1), h1ym);
    h1m = h1;
    h2m = h2;
    C n[3];
    e_cell<T>(a, col, iz, idx, c0, c1, c2, true, n);
    a.e[0][idx] = n[0];
    a.e[1][idx] = n[1];
    a.e[2][idx] = n[2];
    // max-side Bloch wraps, reference code/boundary.f90:205-212,269-278,333-342
    if (ix == 1) {
      a.e[0][idx + a.nx] = cmul(a.fxp, n[0]);
      a.e[1][idx + a.nx] = cmul(a.fxp, n[1]);
      a.e[2][idx + a.nx] = cmul(a.fxp, n[2]);
    }
    if (iy == 1) {
      const int64_t g = idx + a.sx * a.ny;
      a.e[0][g] = cmul(a.fyp, n[0]);
      a.e[1][g] = cmul(a.fyp, n[1]);
      a.e[2][g] = cmul(a.fyp, n[2]);
    }
    if (iz == 1) {
      const int64_t g = idx + a.sxy * a.nz;
      a.e[0][g] = cmul(a.fzp, n[0]);
      a.e[1][g] = cmul(a.fzp, n[1]);
      a.e[2][g] = cmul(a.fzp, n[2]);
    }
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
  int64_t idx = col + a.sxy * z0;
  C e1 = a.e[0][idx];
  C e2 = a.e[1][idx];
  for (int iz = z0; iz <= z1; ++iz, idx += a.sxy) {
    const C e1p = a.e[0][idx + a.sxy], e2p = a.e[1][idx + a.sxy];
    const C e3 = a.e[2][idx];
    const C e3yp = a.e[2][idx + a.sx], e1yp = a.e[0][idx + a.sx];
    const C e3xp = a.e[2][idx + 1], e2xp = a.e[1][idx + 1];
    // curl E, reference code/fdtd.f90:182-187
    const C c0 = cadd(csub(csub(e3yp, e3), e2p), e2);
    const C c1 = cadd(csub(csub(e1p, e1), e3xp), e3);
    const C c2 = cadd(csub(csub(e2xp, e2), e1yp), e1);
    e1 = e1p;
    e2 = e2p;
    C n[3];
    h_cell<T>(a, col, iz, a.h[0][idx], a.h[1][idx], a.h[2][idx], c0, c1, c2, n);
    a.h[0][idx] = n[0];
    a.h[1][idx] = n[1];
    a.h[2][idx] = n[2];
    // min-side Bloch wraps, reference code/boundary.f90:177-183,241-247,305-311
    if (ix == a.nx) {
      a.h[0][idx - a.nx] = cmul(a.fxm, n[0]);
      a.h[1][idx - a.nx] = cmul(a.fxm, n[1]);
      a.h[2][idx - a.nx] = cmul(a.fxm, n[2]);
    }
    if (iy == a.ny) {
      const int64_t g = idx - a.sx * a.ny;
      a.h[0][g] = cmul(a.fym, n[0]);
      a.h[1][g] = cmul(a.fym, n[1]);
      a.h[2][g] = cmul(a.fym, n[2]);
    }
    if (iz == a.nz) {
      const int64_t g = idx - a.sxy * a.nz;
      a.h[0][g] = cmul(a.fzm, n[0]);
      a.h[1][g] = cmul(a.fzm, n[1]);
      a.h[2][g] = cmul(a.fzm, n[2]);
    }
  }
}

// ---------------------------------------------------------------------------
// TF/SF source.  The Gaussian-modulated plane wave depends on (ix, iz, it)
// only (u%y = 0), evaluated in FP64 exactly as reference code/tfsf.f90:327-346.
// ---------------------------------------------------------------------------
struct TfsfArgs {
  double q, ux, uz, zopump, ko, dk, sq, sign_kz, c0dt_it, amp, hamp, win;
  double ein, dtcq2, eps_ms;
  int ixo, iz_min, i_pol, cosine, mode2d;
};

// incident (E or H component, complex) of comps {1,3,5} (i_pol=2) or
// {2,4,6} (i_pol=1) at (ix, iz)
__device__ inline void tfsf_incident(const TfsfArgs& t, int ix, int iz, double2* inc) {
  const double A = (t.q * (double)(ix - t.ixo)) * t.ux + (t.q * ((double)iz - t.zopump)) * t.uz;
  const double kzz = t.ko * ((t.sq * t.sign_kz) * A - t.c0dt_it);
  const double tt = (A - t.c0dt_it / t.sq) / t.dk;
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

template <typename T>
__device__ inline void add_to(cplx<T>* f, int64_t i, double2 inc) {
  if constexpr (sizeof(T) == 8) {
    f[i] = cadd(f[i], inc);
  } else {
    f[i] = cadd(f[i], to_c<T>(inc));
  }
}
template <typename T>
__device__ inline void sub_from(cplx<T>* f, int64_t i, double2 inc) {
  if constexpr (sizeof(T) == 8) {
    f[i] = csub(f[i], inc);
  } else {
    f[i] = csub(f[i], to_c<T>(inc));
  }
}

// Thread = (ix, iy) with ix in [0,nx], iy in [0,ny].  Only cells that a
// later step reads are corrected: interior E, and interior H plus the one
// ghost line the next E pass reads (ix=0 for Hy, iy=0 for Hx).
template <typename T>
__global__ __launch_bounds__(256) void k_tfsf(StepArgs<T> a, TfsfArgs t) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y * blockDim.y + threadIdx.y;
  if (ix > a.nx || iy > a.ny) return;
  const int64_t col = (int64_t)(ix + a.xoff) + a.sx * iy;
  const bool ein_ok = ix >= 1 && iy >= 1;
  if (t.i_pol == 2 && !t.mode2d) {
    // reference code/tfsf.f90:95-98 (Ex at iz_min) and :114-118 (Hy at iz_min-1)
    double2 inc_lo[3], inc_hi[3];
    tfsf_incident(t, ix, t.iz_min - 1, inc_lo);
    tfsf_incident(t, ix, t.iz_min, inc_hi);
    if (ein_ok) {
      add_to<T>(a.e[0], col + a.sxy * t.iz_min, make_double2(t.ein * inc_lo[2].x, t.ein * inc_lo[2].y));
    }
    if (iy >= 1) {
      add_to<T>(a.h[1], col + a.sxy * (t.iz_min - 1), make_double2(t.dtcq2 * inc_hi[0].x, t.dtcq2 * inc_hi[0].y));
    }
  } else if (t.i_pol == 2) {
    // 1D/2D grids, reference code/tfsf.f90:176-193: Ex += inc5(iz_min)/eps,
    // Hy(iz_min) += dtcq2*inc1(iz_min)
    double2 inc_hi[3];
    tfsf_incident(t, ix, t.iz_min, inc_hi);
    if (ein_ok) {
      add_to<T>(a.e[0], col + a.sxy * t.iz_min, cdiv_scaled(inc_hi[2], make_double2(t.eps_ms, 0.0)));
    }
    if (iy >= 1) {
      add_to<T>(a.h[1], col + a.sxy * t.iz_min, make_double2(t.dtcq2 * inc_hi[0].x, t.dtcq2 * inc_hi[0].y));
    }
  } else {
    // i_pol = 1, reference code/tfsf.f90:101-112: Ey -= inc4(iz_min-1)/eps(1),
    // Hx(iz_min-1) -= dtcq2*inc2(iz_min)
    double2 inc_lo[3], inc_hi[3];
    tfsf_incident(t, ix, t.iz_min - 1, inc_lo);
    tfsf_incident(t, ix, t.iz_min, inc_hi);
    if (ein_ok) {
      sub_from<T>(a.e[1], col + a.sxy * t.iz_min, make_double2(t.ein * inc_lo[1].x, t.ein * inc_lo[1].y));
    }
    if (ix >= 1) {
      sub_from<T>(a.h[0], col + a.sxy * (t.iz_min - 1), make_double2(t.dtcq2 * inc_hi[0].x, t.dtcq2 * inc_hi[0].y));
    }
  }
}

// ---------------------------------------------------------------------------
// Fused E+H pass.  One sweep per time step over double-buffered fields:
// E_new(z) is computed from E_old/H_old, and one plane behind it H_new(z-1)
// from H_old and E_new, so E and H are each read once and written once
// (96 B/cell in FP32 instead of 144 for the two in-place passes).
//
// Block = a 64 x 4 (x,y) tile marching a z-chunk, 4 waves for the tile's own
// cells plus a fifth "halo" wave.  H_new at the tile's last column/row needs
// E_new at x+1 / y+1, which belong to the neighbouring tiles, so the halo
// wave recomputes E_new on the +y row and (lanes 0..3) on the +x column from
// the old fields; the values are identical to what the neighbour stores.
// Cells at index nx+1, ny+1 or nz+1 are the Bloch images of index 1
// (reference code/boundary.f90:193-348), computed from the index-1 inputs and
// multiplied by exp(+i ak).  E_new of the current plane travels to the H
// update of the next iteration through LDS (two barriers per plane).
// The TF/SF corrections (code/tfsf.f90:95-118, :176-193) are applied to the
// values stored, after the H update has used the uncorrected E, as in the
// reference; the H min-side ghost planes (read by the next step's E) are
// stored with their own corrections.  E's max-side ghosts are not stored:
// in this scheme nothing reads them.
// ---------------------------------------------------------------------------
constexpr int kFTX = 64, kFTY = 4;

// E_new at source cell (xs,ys,zs) (all in 1..n) from the old fields; the
// z-1 values of H1/H2 come in h1m/h2m, H_old of the cell is returned in hown.
template <typename T>
__device__ inline void e_item(const StepArgs<T>& a, int xs, int ys, int zs, cplx<T> h1m, cplx<T> h2m,
                              bool store_aux, cplx<T> n[3], cplx<T> hown[3]) {
  using C = cplx<T>;
  const int64_t col = (int64_t)(xs + a.xoff) + a.sx * ys;
  const int64_t idx = col + a.sxy * zs;
  const C h1 = a.h[0][idx], h2 = a.h[1][idx], h3 = a.h[2][idx];
  const C h3ym = a.h[2][idx - a.sx], h1ym = a.h[0][idx - a.sx];
  const C h3xm = a.h[2][idx - 1], h2xm = a.h[1][idx - 1];
  const C c0 = cadd(csub(csub(h3, h3ym), h2), h2m);
  const C c1 = cadd(csub(csub(h1, h1m), h3), h3xm);
  const C c2 = cadd(csub(csub(h2, h2xm), h1), h1ym);
  e_cell<T>(a, col, zs, idx, c0, c1, c2, store_aux, n);
  hown[0] = h1;
  hown[1] = h2;
  hown[2] = h3;
}

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

template <typename T>
__global__ __launch_bounds__(kFTX * (kFTY + 1)) void k_yee_fused(StepArgs<T> a, TfsfArgs t, int tfsf_on) {
  using C = cplx<T>;
  __shared__ C se[3][kFTY + 1][kFTX + 1];
  const int lane = threadIdx.x, wy = threadIdx.y;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int x0 = blockIdx.x * kFTX + 1, y0 = blockIdx.y * kFTY + 1;
  const int z0 = blockIdx.z * a.zchunk + 1;
  const int z1 = min(z0 + a.zchunk - 1, nz);
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

  C a1m{}, a2m{}, a3m{}, b1m{}, b2m{}, ep[3] = {};
  if (va) {
    const int64_t i = (int64_t)(xsa + a.xoff) + a.sx * ysa + a.sxy * (z0 - 1);
    a1m = a.h[0][i];
    a2m = a.h[1][i];
    a3m = a.h[2][i];
  }
  if (vb) {
    const int64_t i = (int64_t)(xsb + a.xoff) + a.sx * yb + a.sxy * (z0 - 1);
    b1m = a.h[0][i];
    b2m = a.h[1][i];
  }

  for (int z = z0; z <= z1 + 1; ++z) {
    const bool real_plane = z <= z1;
    const bool image_z = z == nz + 1;
    const int zs = image_z ? 1 : z;
    C na[3], nb[3], ha[3], hb[3];
    const bool doa = va && (real_plane || owna);
    if (doa) {
      // z-1 values for this E; on the image plane they come from the ghost
      // plane 0, while a1m/a2m keep H_old(nz) for the H update below
      C z1m = a1m, z2m = a2m;
      if (image_z) {
        const int64_t i = (int64_t)(xsa + a.xoff) + a.sx * ysa;
        z1m = a.h[0][i];
        z2m = a.h[1][i];
      }
      e_item<T>(a, xsa, ysa, zs, z1m, z2m, owna && real_plane, na, ha);
      if (xa == nx + 1) { na[0] = cmul(a.fxp, na[0]); na[1] = cmul(a.fxp, na[1]); na[2] = cmul(a.fxp, na[2]); }
      if (ya == ny + 1) { na[0] = cmul(a.fyp, na[0]); na[1] = cmul(a.fyp, na[1]); na[2] = cmul(a.fyp, na[2]); }
      if (image_z) { na[0] = cmul(a.fzp, na[0]); na[1] = cmul(a.fzp, na[1]); na[2] = cmul(a.fzp, na[2]); }
    }
    const bool dob = vb && real_plane;
    if (dob) {
      e_item<T>(a, xsb, yb, zs, b1m, b2m, false, nb, hb);
      if (xb == nx + 1) { nb[0] = cmul(a.fxp, nb[0]); nb[1] = cmul(a.fxp, nb[1]); nb[2] = cmul(a.fxp, nb[2]); }
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
      h_cell<T>(a, cola, zh, a1m, a2m, a3m, c0, c1, c2, hn);
      C gx[3], gy[3];
      gx[0] = cmul(a.fxm, hn[0]); gx[1] = cmul(a.fxm, hn[1]); gx[2] = cmul(a.fxm, hn[2]);
      gy[0] = cmul(a.fym, hn[0]); gy[1] = cmul(a.fym, hn[1]); gy[2] = cmul(a.fym, hn[2]);
      if (tfsf_on) {
        if (t.i_pol == 2) {
          // Hy += dtcq2*inc1 on iz_min-1 (3D) or iz_min (1D/2D), ghost x=0 with inc1(0)
          if (zh == (t.mode2d ? t.iz_min : t.iz_min - 1)) {
            double2 inc[3];
            tfsf_incident(t, xa, t.iz_min, inc);
            hn[1] = add_inc<T>(hn[1], make_double2(t.dtcq2 * inc[0].x, t.dtcq2 * inc[0].y));
            if (xa == nx) {
              tfsf_incident(t, 0, t.iz_min, inc);
              gx[1] = add_inc<T>(gx[1], make_double2(t.dtcq2 * inc[0].x, t.dtcq2 * inc[0].y));
            }
          }
        } else if (zh == t.iz_min - 1) {
          // Hx -= dtcq2*inc2 on iz_min-1, ghost y=0 too (same ix)
          double2 inc[3];
          tfsf_incident(t, xa, t.iz_min, inc);
          const double2 d = make_double2(t.dtcq2 * inc[0].x, t.dtcq2 * inc[0].y);
          hn[0] = sub_inc<T>(hn[0], d);
          gy[0] = sub_inc<T>(gy[0], d);
        }
      }
      const int64_t ih = cola + a.sxy * zh;
      a.hn[0][ih] = hn[0];
      a.hn[1][ih] = hn[1];
      a.hn[2][ih] = hn[2];
      if (xa == nx) {
        a.hn[0][ih - nx] = gx[0];
        a.hn[1][ih - nx] = gx[1];
        a.hn[2][ih - nx] = gx[2];
      }
      if (ya == ny) {
        const int64_t g = ih - a.sx * ny;
        a.hn[0][g] = gy[0];
        a.hn[1][g] = gy[1];
        a.hn[2][g] = gy[2];
      }
      if (zh == nz) {
        const int64_t g = ih - a.sxy * nz;
        a.hn[0][g] = cmul(a.fzm, hn[0]);
        a.hn[1][g] = cmul(a.fzm, hn[1]);
        a.hn[2][g] = cmul(a.fzm, hn[2]);
      }
    }
    if (doa) { a1m = ha[0]; a2m = ha[1]; a3m = ha[2]; }
    if (dob) { b1m = hb[0]; b2m = hb[1]; }
    if (!real_plane) break;
    if (owna) {
      C st[3] = {na[0], na[1], na[2]};
      if (tfsf_on && z == t.iz_min) {
        double2 inc[3];
        if (t.i_pol == 2) {
          if (!t.mode2d) {
            tfsf_incident(t, xa, t.iz_min - 1, inc);
            st[0] = add_inc<T>(st[0], make_double2(t.ein * inc[2].x, t.ein * inc[2].y));
          } else {
            tfsf_incident(t, xa, t.iz_min, inc);
            st[0] = add_inc<T>(st[0], cdiv_scaled(inc[2], make_double2(t.eps_ms, 0.0)));
          }
        } else {
          tfsf_incident(t, xa, t.iz_min - 1, inc);
          st[1] = sub_inc<T>(st[1], make_double2(t.ein * inc[1].x, t.ein * inc[1].y));
        }
      }
      const int64_t ie = cola + a.sxy * z;
      a.en[0][ie] = st[0];
      a.en[1][ie] = st[1];
      a.en[2][ie] = st[2];
    }
    __syncthreads();
    if (doa) { se[0][wy][lane] = na[0]; se[1][wy][lane] = na[1]; se[2][wy][lane] = na[2]; }
    if (dob) { se[0][lane][kFTX] = nb[0]; se[1][lane][kFTX] = nb[1]; se[2][lane][kFTX] = nb[2]; }
    __syncthreads();
    if (owna) { ep[0] = na[0]; ep[1] = na[1]; ep[2] = na[2]; }
  }
}

// ---------------------------------------------------------------------------
// On-the-fly DFT.
// ---------------------------------------------------------------------------
struct BoxGeom {
  int bxi, bxf, byi, byf, bzi, bzf;
  int nxb, nyb, nzb;
  int64_t npts[3];    // complex points per pair (0 when the pair is skipped)
  int64_t start[3];   // offset of each pair in the concatenated point index
  int64_t ntot;       // sum of active npts
};

// (component 0..5, x, y, z) of concatenated point index i, following the
// gather table of reference code/dataprocess.f90:813-901
__device__ inline void box_point(const BoxGeom& g, int64_t i, int* comp, int* x, int* y, int* z) {
  int j = 0;
  if (g.npts[1] && i >= g.start[1]) j = 1;
  if (g.npts[2] && i >= g.start[2]) j = 2;
  int64_t r = i - g.start[j];
  const int64_t half = g.npts[j] / 2;
  const bool second = r >= half;
  if (second) r -= half;
  *comp = (int)(r % 6);
  const int64_t q = r / 6;
  if (j == 0) {
    *x = second ? g.bxf : g.bxi;
    *y = (int)(q % g.nyb) + g.byi;
    *z = (int)(q / g.nyb) + g.bzi;
  } else if (j == 1) {
    *x = (int)(q % g.nxb) + g.bxi;
    *y = second ? g.byf : g.byi;
    *z = (int)(q / g.nxb) + g.bzi;
  } else {
    *x = (int)(q % g.nxb) + g.bxi;
    *y = (int)(q / g.nxb) + g.byi;
    *z = second ? g.bzf : g.bzi;
  }
}

// copy the box-face samples of this step into slot `slot` of the buffer
template <typename T>
__global__ __launch_bounds__(256) void k_dft_sample(StepArgs<T> a, BoxGeom g, cplx<T>* samp, int slot) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.ntot) return;
  int c, x, y, z;
  box_point(g, i, &c, &x, &y, &z);
  const int64_t idx = (int64_t)(x + a.xoff) + a.sx * y + a.sxy * z;
  samp[(int64_t)slot * g.ntot + i] = (c < 3) ? a.e[c][idx] : a.h[c - 3][idx];
}

// phase table ph[k][iw] = exp(i*fw(iw)*n_k), n_k = n0 + k*stride
// (reference code/dataprocess.f90:145: cdexp(ci*w*it) = (cos, sin)(w*it))
__global__ void k_dft_phase(const double* fw, int nof, int nk, int n0, int stride, double2* ph) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nof * nk) return;
  const int k = t / nof, iw = t % nof;
  const double arg = fw[iw] * (double)(n0 + k * stride);
  double s, c;
  sincos(arg, &s, &c);
  ph[t] = make_double2(c, s);
}

// emg(i,iw) += dt*F_k(i)*ph[k][iw] for the nk buffered samples, one sample at
// a time as the reference does each step; the running sum stays in a
// register across the K samples.  Thread = one point-component; its K
// samples (scaled by dt) stay in registers for the sweep over all
// wavelengths, so each sample is read once per fold and each accumulator is
// read and written once per fold.
template <typename T, int KMAX>
__global__ __launch_bounds__(256) void k_dft_accum(BoxGeom g, const cplx<T>* samp, int nk, const double2* ph,
                                                   int nof, double dt, double2* emg) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.ntot) return;
  int j = 0;
  if (g.npts[1] && i >= g.start[1]) j = 1;
  if (g.npts[2] && i >= g.start[2]) j = 2;
  const int64_t np = g.npts[j];
  const int64_t r = i - g.start[j];
  double2* acc_base = emg + g.start[j] * (int64_t)nof + r;  // emg pair block: (np, nof)
  double2 f[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    if (k < nk) {
      const cplx<T> s = samp[(int64_t)k * g.ntot + i];
      f[k] = make_double2(dt * (double)s.x, dt * (double)s.y);
    }
  }
  if (nk == KMAX) {
    // four wavelengths per iteration keep four accumulator loads in flight
    int iw = 0;
    for (; iw + 4 <= nof; iw += 4) {
      double2 acc[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = acc_base[(int64_t)(iw + u) * np];
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = cadd(acc[u], cmul(f[k], ph[k * nof + iw + u]));
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc_base[(int64_t)(iw + u) * np] = acc[u];
    }
    for (; iw < nof; ++iw) {
      double2 acc = acc_base[(int64_t)iw * np];
#pragma unroll
      for (int k = 0; k < KMAX; ++k) acc = cadd(acc, cmul(f[k], ph[k * nof + iw]));
      acc_base[(int64_t)iw * np] = acc;
    }
  } else {
    for (int iw = 0; iw < nof; ++iw) {
      double2 acc = acc_base[(int64_t)iw * np];
      const double2* p = ph + iw;
#pragma unroll
      for (int k = 0; k < KMAX; ++k)
        if (k < nk) acc = cadd(acc, cmul(f[k], p[k * nof]));
      acc_base[(int64_t)iw * np] = acc;
    }
  }
}

// Face fluxes, reference code/dataprocess.f90:241-271: block = (pair, iw,
// half); writes the unscaled sum with the reference's sign into
// out[iw*6 + face].
__global__ __launch_bounds__(256) void k_flux(BoxGeom g, const double2* emg, int nof, int active_mask,
                                              double* out) {
  const int j = blockIdx.x / 2, half = blockIdx.x % 2, iw = blockIdx.y;
  if (!((active_mask >> j) & 1)) return;
  const int64_t np = g.npts[j];
  const double2* a = emg + g.start[j] * (int64_t)nof + (int64_t)iw * np + (half ? np / 2 : 0);
  const int64_t npt = np / 12;  // points per face
  int i1, i2, i3, i4;  // (Ex,Ey,Ez,Hx,Hy,Hz) = 0..5
  if (j == 0) { i1 = 1; i2 = 5; i3 = 2; i4 = 4; }        // Ey Hz* - Ez Hy*
  else if (j == 1) { i1 = 2; i2 = 3; i3 = 0; i4 = 5; }   // Ez Hx* - Ex Hz*
  else { i1 = 0; i2 = 4; i3 = 1; i4 = 3; }               // Ex Hy* - Ey Hx*
  double s = 0.0;
  for (int64_t p = threadIdx.x; p < npt; p += blockDim.x) {
    const double2* q = a + 6 * p;
    const double2 A = q[i1], B = q[i2], Cc = q[i3], D = q[i4];
    const double v1 = A.x * B.x - A.y * (-B.y);
    const double v2 = Cc.x * D.x - Cc.y * (-D.y);
    s += v1 - v2;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[iw * 6 + 2 * j + half] = half ? red[0] : -red[0];
}

}  // namespace iris
