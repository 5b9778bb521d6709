/*
 * iris_oracle.c -- FP64 CPU restatement of IrisFDTD-Academic's time loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP engine is
 * checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the CPU baseline;
 * the product path never links or calls it.
 *
 * It follows the reference operation by operation, in the reference's
 * evaluation order and with no FMA contraction (build with
 * -ffp-contract=off), so that on the same inputs it reproduces the
 * reference binary built by oracle/patch_ref.py bit for bit:
 *
 *   loop order            code/driver.f90:338-450
 *   e_fdtd / h_fdtd       code/fdtd.f90:57-153 / :157-248
 *   UPML_Array_e / _h     code/boundary_II.f90:60-138 / :142-220
 *   Bloch wraps           code/boundary.f90:163-348
 *   tfsf_gauss            code/tfsf.f90:231-355
 *   tfsf_update_3DArray   code/tfsf.f90:76-153
 *   tfsf_update_2DArray   code/tfsf.f90:163-227
 *   scattering_fourier    code/dataprocess.f90:109-209
 *   scattering_fourier_out (numbers only) code/dataprocess.f90:213-277
 *   init_store_pts_fourier code/dataprocess.f90:793-903
 *
 * Storage is the reference's: complex AoS fields (3,0:nx+1,0:ny+1,0:nz+1),
 * double-buffered e/h and UPML d/b with pointer swaps.  The API mirrors
 * include/iris.h with an oracle_ prefix so tests drive both the same way.
 *
 * Arithmetic conventions of the reference build (LLVM Flang, x86-64, no FMA):
 *  - real*complex promotes the real to (r,0); for finite operands this is
 *    exactly (r*re, r*im), which is what cmul_r computes;
 *  - complex*complex is (ac-bd, ad+bc);
 *  - complex/complex (also complex/real, which Fortran promotes) goes through
 *    compiler-rt's __divdc3: logb/scalbn-scaled Smith division (cdiv below,
 *    restated from the published C99 Annex G algorithm);
 *  - cdexp is libm cexp, EXP is libm exp, x**2 is x*x.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/iris.h"

typedef struct { double re, im; } cx;

static inline cx cadd(cx a, cx b) { cx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cx csub(cx a, cx b) { cx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cx cmul_r(double s, cx a) { cx r = {s * a.re, s * a.im}; return r; }
static inline cx cmul(cx a, cx b) { cx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cx cdiv(cx a, cx b) {
  double c = b.re, d = b.im;
  int k = 0;
  double lw = logb(fmax(fabs(c), fabs(d)));
  if (isfinite(lw)) { k = (int)lw; c = scalbn(c, -k); d = scalbn(d, -k); }
  double den = c * c + d * d;
  cx r = {scalbn((a.re * c + a.im * d) / den, -k), scalbn((a.im * c - a.re * d) / den, -k)};
  return r;
}
static inline cx cexp_ref(cx z) {
  double complex w = cexp(CMPLX(z.re, z.im));
  cx r = {creal(w), cimag(w)};
  return r;
}

typedef struct oracle {
  iris_params p;
  int nx, ny, nz, sx, sy;      /* sx = nx+2, sy = ny+2 */
  int64_t ncell;               /* (nx+2)(ny+2)(nz+2) */
  cx *e_cur, *e_prev, *h_cur, *h_prev;  /* (3, cell) */
  cx *d_cur[2], *d_prev[2], *b_cur[2], *b_prev[2]; /* (3,nx,ny,zlen) */
  int zlen[2], z0[2];
  double *fpzmn, *fmzmn, *fpzmx, *fmzmx;
  cx *eps_inv;                 /* (0:nx+1,0:ny+1,hi-1:hf+1) */
  double *fw;
  /* tfsf_data for the two planes used: comps 1..6 at z = iz_min-1 and iz_min,
     x in [-1, nx+2] (y-independent: u%y = 0) */
  cx *tfsf;                    /* [plane 0/1][comp 0..5][ix+1] */
  int tfsf_nxs;
  /* a tabulated incident field in place of tfsf_gauss (oracle_set_incident):
     (nt+1) x 2 planes x 3 comps at ix = 0, times exp(i phix ix) at ix */
  cx *inc_tab;
  double phix;
  /* Fourier boxes */
  int npts[2][3];
  int *pto[2][3];              /* (npts,4): comp, x, y, z */
  cx *emg[2][3];               /* (npts, nof) */
  int it;
} oracle;

#define F(o, arr, c, ix, iy, iz) (arr)[(c) + 3 * ((int64_t)(ix) + (int64_t)(o)->sx * ((int64_t)(iy) + (int64_t)(o)->sy * (int64_t)(iz)))]
#define PML(o, arr, k, c, ix, iy, iz) (arr)[(c) + 3 * ((int64_t)((ix)-1) + (int64_t)(o)->nx * ((int64_t)((iy)-1) + (int64_t)(o)->ny * (int64_t)((iz) - (o)->z0[k])))]
#define EPSI(o, ix, iy, iz) (o)->eps_inv[(int64_t)(ix) + (int64_t)(o)->sx * ((int64_t)(iy) + (int64_t)(o)->sy * (int64_t)((iz) - ((o)->p.hi - 1)))]

static void bounds_of(const oracle* o, int nobox, int* b) { memcpy(b, o->p.box[nobox - 1], 6 * sizeof(int)); }

/* code/dataprocess.f90:793-903 */
static void init_store_pts(oracle* o, int nobox) {
  int b[6];
  bounds_of(o, nobox, b);
  int bxi = b[0], bxf = b[1], byi = b[2], byf = b[3], bzi = b[4], bzf = b[5];
  int nxb = bxf + 1 - bxi, nyb = byf + 1 - byi, nzb = bzf + 1 - bzi;
  int k = nobox - 1;
  o->npts[k][0] = (o->nx != 1) ? 2 * 6 * nyb * nzb : 1;
  o->npts[k][1] = (o->ny != 1) ? 2 * 6 * nxb * nzb : 1;
  o->npts[k][2] = 2 * 6 * nxb * nyb;
  for (int j = 0; j < 3; j++) {
    o->pto[k][j] = (int*)malloc(sizeof(int) * 4 * (size_t)o->npts[k][j]);
    for (int i = 0; i < 4 * o->npts[k][j]; i++) o->pto[k][j][i] = 1;
    o->emg[k][j] = (cx*)calloc((size_t)o->npts[k][j] * (size_t)o->p.nof, sizeof(cx));
  }
#define SET(j, i, c, x, y, z) do { int* q_ = o->pto[k][j] + 4 * (size_t)((i)-1); q_[0] = (c); q_[1] = (x); q_[2] = (y); q_[3] = (z); } while (0)
  if (o->nx != 1) {
    for (int iy = 1; iy <= nyb; iy++)
      for (int iz = 1; iz <= nzb; iz++)
        for (int ip = 1; ip <= 6; ip++) {
          int i = ip + 6 * (iy - 1 + nyb * (iz - 1));
          SET(0, i, ip, bxi, iy + byi - 1, iz + bzi - 1);
          SET(0, 6 * nyb * nzb + i, ip, bxf, iy + byi - 1, iz + bzi - 1);
        }
  }
  if (o->ny != 1) {
    for (int ix = 1; ix <= nxb; ix++)
      for (int iz = 1; iz <= nzb; iz++)
        for (int ip = 1; ip <= 6; ip++) {
          int i = ip + 6 * (ix - 1 + nxb * (iz - 1));
          SET(1, i, ip, ix + bxi - 1, byi, iz + bzi - 1);
          SET(1, 6 * nxb * nzb + i, ip, ix + bxi - 1, byf, iz + bzi - 1);
        }
  }
  for (int ix = 1; ix <= nxb; ix++)
    for (int iy = 1; iy <= nyb; iy++)
      for (int ip = 1; ip <= 6; ip++) {
        int i = ip + 6 * (ix - 1 + nxb * (iy - 1));
        SET(2, i, ip, ix + bxi - 1, iy + byi - 1, bzi);
        SET(2, 6 * nxb * nyb + i, ip, ix + bxi - 1, iy + byi - 1, bzf);
      }
#undef SET
}

oracle* oracle_create(const iris_params* p, const double* eps_inv, const double* fpzmn,
                      const double* fmzmn, const double* fpzmx, const double* fmzmx,
                      const double* fw) {
  oracle* o = (oracle*)calloc(1, sizeof(oracle));
  o->p = *p;
  o->nx = p->nx; o->ny = p->ny; o->nz = p->nz;
  o->sx = p->nx + 2; o->sy = p->ny + 2;
  o->ncell = (int64_t)o->sx * o->sy * (p->nz + 2);
  o->e_cur = (cx*)calloc(3 * (size_t)o->ncell, sizeof(cx));
  o->e_prev = (cx*)calloc(3 * (size_t)o->ncell, sizeof(cx));
  o->h_cur = (cx*)calloc(3 * (size_t)o->ncell, sizeof(cx));
  o->h_prev = (cx*)calloc(3 * (size_t)o->ncell, sizeof(cx));
  o->z0[0] = 1; o->zlen[0] = p->pmlzi;
  o->z0[1] = p->pmlzf; o->zlen[1] = p->nz - p->pmlzf + 1;
  for (int k = 0; k < 2; k++) {
    size_t n = 3 * (size_t)p->nx * p->ny * (size_t)(o->zlen[k] > 0 ? o->zlen[k] : 0);
    o->d_cur[k] = (cx*)calloc(n + 1, sizeof(cx)); o->d_prev[k] = (cx*)calloc(n + 1, sizeof(cx));
    o->b_cur[k] = (cx*)calloc(n + 1, sizeof(cx)); o->b_prev[k] = (cx*)calloc(n + 1, sizeof(cx));
  }
  size_t nmn = 3 * (size_t)p->pmlzi, nmx = 3 * (size_t)o->zlen[1];
  o->fpzmn = (double*)malloc(sizeof(double) * (nmn + 1)); memcpy(o->fpzmn, fpzmn, sizeof(double) * nmn);
  o->fmzmn = (double*)malloc(sizeof(double) * (nmn + 1)); memcpy(o->fmzmn, fmzmn, sizeof(double) * nmn);
  o->fpzmx = (double*)malloc(sizeof(double) * (nmx + 1)); memcpy(o->fpzmx, fpzmx, sizeof(double) * nmx);
  o->fmzmx = (double*)malloc(sizeof(double) * (nmx + 1)); memcpy(o->fmzmx, fmzmx, sizeof(double) * nmx);
  size_t ne = (size_t)o->sx * o->sy * (size_t)(p->hf - p->hi + 3);
  o->eps_inv = (cx*)malloc(sizeof(cx) * ne);
  memcpy(o->eps_inv, eps_inv, sizeof(cx) * ne);
  o->fw = (double*)malloc(sizeof(double) * (size_t)p->nof);
  memcpy(o->fw, fw, sizeof(double) * (size_t)p->nof);
  o->tfsf_nxs = p->nx + 4;
  o->tfsf = (cx*)calloc(2 * 6 * (size_t)o->tfsf_nxs, sizeof(cx));
  init_store_pts(o, 1);
  if (p->nbox == 2) init_store_pts(o, 2);
  o->it = 0;
  return o;
}

/* Take the incident field from a table ((nt+1) x 2 x 3 complex, the layout
 * of iris_incident_table) instead of the analytic pulse; the corrections are
 * those of 3D grids (code/tfsf.f90:76-119) on every grid. */
void oracle_set_incident(oracle* o, const double* table) {
  const size_t n = ((size_t)o->p.nt + 1) * 6;
  free(o->inc_tab);
  o->inc_tab = (cx*)malloc(sizeof(cx) * n);
  memcpy(o->inc_tab, table, sizeof(cx) * n);
  o->phix = o->p.ak[0][0] / o->p.nx;
}

void oracle_destroy(oracle* o) {
  if (!o) return;
  free(o->e_cur); free(o->e_prev); free(o->h_cur); free(o->h_prev);
  for (int k = 0; k < 2; k++) { free(o->d_cur[k]); free(o->d_prev[k]); free(o->b_cur[k]); free(o->b_prev[k]); }
  free(o->fpzmn); free(o->fmzmn); free(o->fpzmx); free(o->fmzmx);
  free(o->eps_inv); free(o->fw); free(o->tfsf); free(o->inc_tab);
  for (int k = 0; k < 2; k++)
    for (int j = 0; j < 3; j++) { free(o->pto[k][j]); free(o->emg[k][j]); }
  free(o);
}

/* code/fdtd.f90:78-83 -- curl of H with backward differences */
static inline void curl_h(const oracle* o, const cx* h, int ix, int iy, int iz, cx* c) {
  c[0] = cadd(csub(csub(F(o, h, 2, ix, iy, iz), F(o, h, 2, ix, iy - 1, iz)), F(o, h, 1, ix, iy, iz)), F(o, h, 1, ix, iy, iz - 1));
  c[1] = cadd(csub(csub(F(o, h, 0, ix, iy, iz), F(o, h, 0, ix, iy, iz - 1)), F(o, h, 2, ix, iy, iz)), F(o, h, 2, ix - 1, iy, iz));
  c[2] = cadd(csub(csub(F(o, h, 1, ix, iy, iz), F(o, h, 1, ix - 1, iy, iz)), F(o, h, 0, ix, iy, iz)), F(o, h, 0, ix, iy - 1, iz));
}

/* code/fdtd.f90:182-187 -- curl of E with forward differences */
static inline void curl_e(const oracle* o, const cx* e, int ix, int iy, int iz, cx* c) {
  c[0] = cadd(csub(csub(F(o, e, 2, ix, iy + 1, iz), F(o, e, 2, ix, iy, iz)), F(o, e, 1, ix, iy, iz + 1)), F(o, e, 1, ix, iy, iz));
  c[1] = cadd(csub(csub(F(o, e, 0, ix, iy, iz + 1), F(o, e, 0, ix, iy, iz)), F(o, e, 2, ix + 1, iy, iz)), F(o, e, 2, ix, iy, iz));
  c[2] = cadd(csub(csub(F(o, e, 1, ix + 1, iy, iz), F(o, e, 1, ix, iy, iz)), F(o, e, 0, ix, iy + 1, iz)), F(o, e, 0, ix, iy, iz));
}

/* code/fdtd.f90:57-153 */
static void e_fdtd(oracle* o) {
  const iris_params* p = &o->p;
  for (int iz = p->pmlzi + 1; iz <= p->pmlzf - 1; iz++) {
    int region = (iz < p->hi) ? 0 : (iz <= p->hf ? -1 : 1);
    for (int iy = 1; iy <= o->ny; iy++)
      for (int ix = 1; ix <= o->nx; ix++) {
        cx c[3];
        curl_h(o, o->h_prev, ix, iy, iz, c);
        if (region < 0) {
          cx ei = EPSI(o, ix, iy, iz);
          for (int i = 0; i < 3; i++) F(o, o->e_cur, i, ix, iy, iz) = cadd(F(o, o->e_prev, i, ix, iy, iz), cmul(ei, c[i]));
        } else {
          for (int i = 0; i < 3; i++) {
            const double* a = p->epsa_inv[region][i];
            cx v = F(o, o->e_prev, i, ix, iy, iz);
            v = cadd(v, cmul_r(a[0], c[0]));
            v = cadd(v, cmul_r(a[1], c[1]));
            v = cadd(v, cmul_r(a[2], c[2]));
            F(o, o->e_cur, i, ix, iy, iz) = v;
          }
        }
      }
  }
}

/* code/fdtd.f90:157-248 (mu_inv_bbox = 1 in the structure: dtcq2*1.0 = dtcq2) */
static void h_fdtd(oracle* o, double dtcq2) {
  const iris_params* p = &o->p;
  int zmx = (p->pmlzf < p->hf) ? p->hf : p->pmlzf - 1;
  for (int iz = p->pmlzi; iz <= zmx; iz++)
    for (int iy = 1; iy <= o->ny; iy++)
      for (int ix = 1; ix <= o->nx; ix++) {
        cx c[3];
        curl_e(o, o->e_cur, ix, iy, iz, c);
        for (int i = 0; i < 3; i++) F(o, o->h_cur, i, ix, iy, iz) = csub(F(o, o->h_prev, i, ix, iy, iz), cmul_r(dtcq2, c[i]));
      }
}

/* code/boundary_II.f90:60-138 */
static void upml_e(oracle* o) {
  const iris_params* p = &o->p;
  const double e02 = p->eps0 * p->eps0;
  for (int k = 0; k < 2; k++) {
    int zlo = k ? p->pmlzf : 1, zhi = k ? p->nz : p->pmlzi;
    double ep = k ? p->eps[2] : p->eps[0];
    const double* fp = k ? o->fpzmx : o->fpzmn;
    const double* fm = k ? o->fmzmx : o->fmzmn;
    for (int iz = zlo; iz <= zhi; iz++) {
      int t = 3 * (iz - o->z0[k]);
      double fp1 = fp[t], fp2 = fp[t + 1], fp3 = fp[t + 2];
      double fm1 = fm[t], fm2 = fm[t + 1], fm3 = fm[t + 2];
      for (int iy = 1; iy <= o->ny; iy++)
        for (int ix = 1; ix <= o->nx; ix++) {
          cx c[3];
          curl_h(o, o->h_prev, ix, iy, iz, c);
          for (int i = 0; i < 3; i++) c[i] = cmul_r(e02, c[i]);
          cx dp1 = PML(o, o->d_prev[k], k, 0, ix, iy, iz), dp2 = PML(o, o->d_prev[k], k, 1, ix, iy, iz), dp3 = PML(o, o->d_prev[k], k, 2, ix, iy, iz);
          cx d1 = cadd(cmul_r(fm2 / fp2, dp1), cmul_r((2.0 * ep) / fp2, c[0]));
          cx d2 = cadd(cmul_r(fm3 / fp3, dp2), cmul_r((2.0 * ep) / fp3, c[1]));
          cx d3 = cadd(cmul_r(fm1 / fp1, dp3), cmul_r((2.0 * ep) / fp1, c[2]));
          PML(o, o->d_cur[k], k, 0, ix, iy, iz) = d1;
          PML(o, o->d_cur[k], k, 1, ix, iy, iz) = d2;
          PML(o, o->d_cur[k], k, 2, ix, iy, iz) = d3;
          F(o, o->e_cur, 0, ix, iy, iz) = cadd(cmul_r(fm3 / fp3, F(o, o->e_prev, 0, ix, iy, iz)),
                                               cmul_r(1.0 / (p->eps0 * ep * fp3), csub(cmul_r(fp1, d1), cmul_r(fm1, dp1))));
          F(o, o->e_cur, 1, ix, iy, iz) = cadd(cmul_r(fm1 / fp1, F(o, o->e_prev, 1, ix, iy, iz)),
                                               cmul_r(1.0 / (p->eps0 * ep * fp1), csub(cmul_r(fp2, d2), cmul_r(fm2, dp2))));
          F(o, o->e_cur, 2, ix, iy, iz) = cadd(cmul_r(fm2 / fp2, F(o, o->e_prev, 2, ix, iy, iz)),
                                               cmul_r(1.0 / (p->eps0 * ep * fp2), csub(cmul_r(fp3, d3), cmul_r(fm3, dp3))));
        }
    }
  }
}

/* code/boundary_II.f90:142-220 */
static void upml_h(oracle* o) {
  const iris_params* p = &o->p;
  const double dtq2 = (p->dt / p->q) * (p->dt / p->q);
  for (int k = 0; k < 2; k++) {
    int zlo = k ? p->pmlzf : 1, zhi = k ? p->nz : p->pmlzi;
    double ep = k ? p->eps[2] : p->eps[0];
    const double* fp = k ? o->fpzmx : o->fpzmn;
    const double* fm = k ? o->fmzmx : o->fmzmn;
    for (int iz = zlo; iz <= zhi; iz++) {
      int t = 3 * (iz - o->z0[k]);
      double fp1 = fp[t], fp2 = fp[t + 1], fp3 = fp[t + 2];
      double fm1 = fm[t], fm2 = fm[t + 1], fm3 = fm[t + 2];
      for (int iy = 1; iy <= o->ny; iy++)
        for (int ix = 1; ix <= o->nx; ix++) {
          cx c[3];
          curl_e(o, o->e_cur, ix, iy, iz, c);
          cx bp1 = PML(o, o->b_prev[k], k, 0, ix, iy, iz), bp2 = PML(o, o->b_prev[k], k, 1, ix, iy, iz), bp3 = PML(o, o->b_prev[k], k, 2, ix, iy, iz);
          cx b1 = csub(cmul_r(fm2 / fp2, bp1), cmul_r(((2.0 * ep) / fp2) * dtq2, c[0]));
          cx b2 = csub(cmul_r(fm3 / fp3, bp2), cmul_r(((2.0 * ep) / fp3) * dtq2, c[1]));
          cx b3 = csub(cmul_r(fm1 / fp1, bp3), cmul_r(((2.0 * ep) / fp1) * dtq2, c[2]));
          PML(o, o->b_cur[k], k, 0, ix, iy, iz) = b1;
          PML(o, o->b_cur[k], k, 1, ix, iy, iz) = b2;
          PML(o, o->b_cur[k], k, 2, ix, iy, iz) = b3;
          F(o, o->h_cur, 0, ix, iy, iz) = cadd(cmul_r(fm3 / fp3, F(o, o->h_prev, 0, ix, iy, iz)),
                                               cmul_r(1.0 / (p->mu0 * fp3), csub(cmul_r(fp1, b1), cmul_r(fm1, bp1))));
          F(o, o->h_cur, 1, ix, iy, iz) = cadd(cmul_r(fm1 / fp1, F(o, o->h_prev, 1, ix, iy, iz)),
                                               cmul_r(1.0 / (p->mu0 * fp1), csub(cmul_r(fp2, b2), cmul_r(fm2, bp2))));
          F(o, o->h_cur, 2, ix, iy, iz) = cadd(cmul_r(fm2 / fp2, F(o, o->h_prev, 2, ix, iy, iz)),
                                               cmul_r(1.0 / (p->mu0 * fp2), csub(cmul_r(fp3, b3), cmul_r(fm3, bp3))));
        }
    }
  }
}

static cx bloch_factor(const oracle* o, int axis, int sign) {
  /* cdexp(+-ci*ak): ci*ak = (0,1)*(a,b) = (-b, a) */
  cx ak = {o->p.ak[axis][0], o->p.ak[axis][1]};
  cx ci = {0.0, 1.0};
  cx arg = cmul(ci, ak);
  if (sign < 0) { arg.re = -arg.re; arg.im = -arg.im; } /* -(ci*ak) */
  return cexp_ref(arg);
}

/* code/boundary.f90:193-216, 257-284, 321-348 */
static void bc_max(oracle* o) {
  cx fx = bloch_factor(o, 0, +1), fy = bloch_factor(o, 1, +1), fz = bloch_factor(o, 2, +1);
  int nx = o->nx, ny = o->ny, nz = o->nz;
  for (int iy = 1; iy <= ny; iy++)
    for (int iz = 1; iz <= nz; iz++)
      for (int i = 0; i < 3; i++) {
        F(o, o->e_cur, i, nx + 1, iy, iz) = cmul(fx, F(o, o->e_cur, i, 1, iy, iz));
        F(o, o->h_cur, i, nx + 1, iy, iz) = cmul(fx, F(o, o->h_cur, i, 1, iy, iz));
      }
  for (int ix = 1; ix <= nx; ix++)
    for (int iz = 1; iz <= nz; iz++)
      for (int i = 0; i < 3; i++) {
        F(o, o->e_cur, i, ix, ny + 1, iz) = cmul(fy, F(o, o->e_cur, i, ix, 1, iz));
        F(o, o->h_cur, i, ix, ny + 1, iz) = cmul(fy, F(o, o->h_cur, i, ix, 1, iz));
      }
  for (int ix = 1; ix <= nx; ix++)
    for (int iy = 1; iy <= ny; iy++)
      for (int i = 0; i < 3; i++) {
        F(o, o->e_cur, i, ix, iy, nz + 1) = cmul(fz, F(o, o->e_cur, i, ix, iy, 1));
        F(o, o->h_cur, i, ix, iy, nz + 1) = cmul(fz, F(o, o->h_cur, i, ix, iy, 1));
      }
}

/* code/boundary.f90:163-187, 224-251, 288-315 */
static void bc_min(oracle* o) {
  cx fx = bloch_factor(o, 0, -1), fy = bloch_factor(o, 1, -1), fz = bloch_factor(o, 2, -1);
  int nx = o->nx, ny = o->ny, nz = o->nz;
  for (int iy = 1; iy <= ny; iy++)
    for (int iz = 1; iz <= nz; iz++)
      for (int i = 0; i < 3; i++) {
        F(o, o->e_cur, i, 0, iy, iz) = cmul(fx, F(o, o->e_cur, i, nx, iy, iz));
        F(o, o->h_cur, i, 0, iy, iz) = cmul(fx, F(o, o->h_cur, i, nx, iy, iz));
      }
  for (int ix = 1; ix <= nx; ix++)
    for (int iz = 1; iz <= nz; iz++)
      for (int i = 0; i < 3; i++) {
        F(o, o->e_cur, i, ix, 0, iz) = cmul(fy, F(o, o->e_cur, i, ix, ny, iz));
        F(o, o->h_cur, i, ix, 0, iz) = cmul(fy, F(o, o->h_cur, i, ix, ny, iz));
      }
  for (int ix = 1; ix <= nx; ix++)
    for (int iy = 1; iy <= ny; iy++)
      for (int i = 0; i < 3; i++) {
        F(o, o->e_cur, i, ix, iy, 0) = cmul(fz, F(o, o->e_cur, i, ix, iy, nz));
        F(o, o->h_cur, i, ix, iy, 0) = cmul(fz, F(o, o->h_cur, i, ix, iy, nz));
      }
}

#define TF(o, plane, comp, ix) (o)->tfsf[((plane) * 6 + (comp)) * (o)->tfsf_nxs + ((ix) + 1)]

/* code/tfsf.f90:231-355 (plane_wave=1, FTwindow=0), evaluated on the two
 * z-planes the corrections read: plane 0 = iz_min-1, plane 1 = iz_min.  The
 * incident field does not depend on iy (u%y = 0). */
static void tfsf_gauss(oracle* o, int it) {
  const iris_params* p = &o->p;
  const int ito = 0;
  double Dk = p->gpump * p->q;
  double ko = (2.0 * p->pi) / (p->wlpump * p->f_abohr);
  double eps_ms, sign_kz;
  if (p->iz_min > p->hi) { eps_ms = p->eps[2]; sign_kz = -1.0; } else { eps_ms = p->eps[0]; sign_kz = 1.0; }
  double ux = p->ux, uz = p->uz, Q = p->q;
  double c0dt_it = (p->c0 * p->dt) * (double)(it - ito);
  double sq = sqrt(eps_ms);
  memset(o->tfsf, 0, sizeof(cx) * 2 * 6 * (size_t)o->tfsf_nxs);
  for (int plane = 0; plane < 2; plane++) {
    int iz = p->iz_min - 1 + plane;
    for (int ix = -1; ix <= o->nx + 2; ix++) {
      double A = (Q * (double)(ix - p->ixo)) * ux + (Q * ((double)iz - p->zopump)) * uz;
      double kzz = ko * ((sq * sign_kz) * A - c0dt_it);
      double tt = (A - c0dt_it / sq) / Dk;
      double env = exp(-(tt * tt));
      cx ci = {0.0, 1.0};
      cx kz = {kzz, 0.0};
      cx ph = cexp_ref(cmul(ci, kz));
      cx one = {1.0, 0.0};
      double amp = p->eamp * Q;
      double hamp = amp * ((p->dt * p->c0) / Q);
      if (p->i_pol == 2) {
        cx ex = cmul_r(uz, cmul_r(env, cmul(cmul_r(sign_kz, one), ph)));
        cx ez = cmul_r(ux, cmul_r(env, cmul(one, ph)));
        ez.re = -ez.re; ez.im = -ez.im;
        cx hy = cmul(one, csub(cmul_r(uz * sign_kz, ex), cmul_r(ux, ez)));
        if (p->cosine == 1) { ex.im = 0.0; ez.im = 0.0; hy.im = 0.0; }
        TF(o, plane, 0, ix) = cadd(TF(o, plane, 0, ix), cmul_r(amp, ex));
        TF(o, plane, 2, ix) = cadd(TF(o, plane, 2, ix), cmul_r(amp, ez));
        TF(o, plane, 4, ix) = cadd(TF(o, plane, 4, ix), cmul_r(hamp, hy));
      } else if (p->i_pol == 1) {
        cx hx = cmul_r(uz, cmul_r(env, cmul(cmul_r(sign_kz, one), ph)));
        hx.re = -hx.re; hx.im = -hx.im;
        cx hz = cmul_r(ux, cmul_r(env, cmul(one, ph)));
        cx ey = cmul(one, csub(cmul_r(ux, hz), cmul_r(uz * sign_kz, hx)));
        if (p->cosine == 1) { ey.im = 0.0; hz.im = 0.0; hx.im = 0.0; }
        TF(o, plane, 3, ix) = cadd(TF(o, plane, 3, ix), cmul_r(hamp, hx));
        TF(o, plane, 5, ix) = cadd(TF(o, plane, 5, ix), cmul_r(hamp, hz));
        TF(o, plane, 1, ix) = cadd(TF(o, plane, 1, ix), cmul_r(amp, ey));
      }
    }
  }
  if (p->win != 1.0)
    for (size_t i = 0; i < 2 * 6 * (size_t)o->tfsf_nxs; i++) o->tfsf[i] = cmul_r(p->win, o->tfsf[i]);
}

/* Tabulated incident field (oracle_set_incident): the table's components
 * 1, 3, 5 (i_pol = 2) or 2, 4, 6 (i_pol = 1) at ix = 0 of step it on both
 * planes, times exp(i phix ix) at ix, with no other change. */
static void tfsf_table(oracle* o, int it) {
  const iris_params* p = &o->p;
  const int c0 = p->i_pol == 2 ? 0 : 1;
  memset(o->tfsf, 0, sizeof(cx) * 2 * 6 * (size_t)o->tfsf_nxs);
  for (int plane = 0; plane < 2; plane++) {
    const cx* v = o->inc_tab + ((size_t)it * 2 + plane) * 3;
    for (int ix = -1; ix <= o->nx + 2; ix++) {
      cx arg = {0.0, o->phix * (double)ix};
      cx ph = cexp_ref(arg);
      for (int k = 0; k < 3; k++) TF(o, plane, c0 + 2 * k, ix) = cmul(v[k], ph);
    }
  }
}

/* code/tfsf.f90:76-119 (the far-side block is disabled, :122) */
static void tfsf_update_3d(oracle* o) {
  const iris_params* p = &o->p;
  double dtcq2 = ((p->dt * p->c0) / p->q) * ((p->dt * p->c0) / p->q);
  double ein = p->eps_in[0];
  int ixmin = 0, ixmax = o->nx + 1, iymin = 0, iymax = o->ny + 1, z = p->iz_min;
  for (int iy = iymin; iy <= iymax; iy++)
    for (int ix = ixmin; ix <= ixmax - 1; ix++)
      F(o, o->e_cur, 0, ix, iy, z) = cadd(F(o, o->e_cur, 0, ix, iy, z), cmul_r(ein, TF(o, 0, 4, ix)));
  for (int iy = iymin; iy <= iymax - 1; iy++)
    for (int ix = ixmin; ix <= ixmax; ix++)
      F(o, o->e_cur, 1, ix, iy, z) = csub(F(o, o->e_cur, 1, ix, iy, z), cmul_r(ein, TF(o, 0, 3, ix)));
  for (int iy = iymin; iy <= iymax - 1; iy++)
    for (int ix = ixmin; ix <= ixmax; ix++)
      F(o, o->h_cur, 0, ix, iy, z - 1) = csub(F(o, o->h_cur, 0, ix, iy, z - 1), cmul_r(dtcq2 * 1.0, TF(o, 1, 1, ix)));
  for (int iy = iymin; iy <= iymax; iy++)
    for (int ix = ixmin; ix <= ixmax - 1; ix++)
      F(o, o->h_cur, 1, ix, iy, z - 1) = cadd(F(o, o->h_cur, 1, ix, iy, z - 1), cmul_r(dtcq2 * 1.0, TF(o, 1, 0, ix)));
}

/* code/tfsf.f90:163-227 (used when nx==1 or ny==1 with i_pol=2) */
static void tfsf_update_2d(oracle* o) {
  const iris_params* p = &o->p;
  double dtcq2 = ((p->dt * p->c0) / p->q) * ((p->dt * p->c0) / p->q);
  double eps_ms = (p->iz_min > p->hi) ? p->eps[2] : p->eps[0];
  cx em = {eps_ms, 0.0};
  int z = p->iz_min;
  for (int iy = 0; iy <= o->ny + 1; iy++)
    for (int ix = 0; ix <= o->nx + 1; ix++) {
      F(o, o->e_cur, 0, ix, iy, z) = cadd(F(o, o->e_cur, 0, ix, iy, z), cdiv(TF(o, 1, 4, ix), em));
      F(o, o->e_cur, 1, ix, iy, z) = csub(F(o, o->e_cur, 1, ix, iy, z), cdiv(TF(o, 0, 3, ix), em));
      F(o, o->h_cur, 0, ix, iy, z - 1) = csub(F(o, o->h_cur, 0, ix, iy, z - 1), cmul_r(dtcq2 * 1.0, TF(o, 1, 1, ix)));
      F(o, o->h_cur, 1, ix, iy, z) = cadd(F(o, o->h_cur, 1, ix, iy, z), cmul_r(dtcq2 * 1.0, TF(o, 1, 0, ix)));
    }
}

/* code/dataprocess.f90:109-209 (system_model = 0) */
static void scattering_fourier(oracle* o, int nobox, int n) {
  int k = nobox - 1;
  int use[3];
  if (o->nx != 1 && o->ny != 1) { use[0] = use[1] = use[2] = 1; }
  else if (o->ny == 1 && o->nx != 1) { use[0] = 1; use[1] = 0; use[2] = 1; }
  else if (o->nx == 1 && o->ny != 1) { use[0] = 0; use[1] = 1; use[2] = 1; }
  else { use[0] = 0; use[1] = 0; use[2] = 1; }
  double dt = o->p.dt;
  for (int iw = 0; iw < o->p.nof; iw++) {
    double w = o->fw[iw];
    cx arg = {0.0, w * (double)n};
    cx ph = cexp_ref(arg);
    for (int j = 0; j < 3; j++) {
      if (!use[j]) continue;
      int np = o->npts[k][j];
      cx* acc = o->emg[k][j] + (size_t)iw * np;
      const int* pt = o->pto[k][j];
      for (int i = 0; i < np; i++) {
        int c = pt[4 * i], x = pt[4 * i + 1], y = pt[4 * i + 2], z = pt[4 * i + 3];
        cx f = (c < 4) ? F(o, o->e_cur, c - 1, x, y, z) : F(o, o->h_cur, c - 4, x, y, z);
        acc[i] = cadd(acc[i], cmul(cmul_r(dt, f), ph));
      }
    }
  }
}

void oracle_step(oracle* o, int it_first, int it_last) {
  const iris_params* p = &o->p;
  double dtcq2 = ((p->dt * p->c0) / p->q) * ((p->dt * p->c0) / p->q);
  for (int it = it_first; it <= it_last; it++) {
    cx* t;
    t = o->e_cur; o->e_cur = o->e_prev; o->e_prev = t;
    t = o->h_cur; o->h_cur = o->h_prev; o->h_prev = t;
    for (int k = 0; k < 2; k++) {
      t = o->d_cur[k]; o->d_cur[k] = o->d_prev[k]; o->d_prev[k] = t;
      t = o->b_cur[k]; o->b_cur[k] = o->b_prev[k]; o->b_prev[k] = t;
    }
    e_fdtd(o);
    upml_e(o);
    bc_max(o);
    h_fdtd(o, dtcq2);
    upml_h(o);
    bc_min(o);
    if (p->em_source && o->inc_tab) {
      tfsf_table(o, it);
      tfsf_update_3d(o);
    } else if (p->em_source) {
      tfsf_gauss(o, it);
      if (o->ny != 1 && o->nx != 1) tfsf_update_3d(o);
      else if (p->i_pol == 2) tfsf_update_2d(o);
      else if (p->i_pol == 1) tfsf_update_3d(o);
    }
    if (it >= p->to_fourier && (it % p->sample_fourier) == 0) {
      scattering_fourier(o, 1, it + 1 - p->to_fourier);
      if (p->nbox == 2) scattering_fourier(o, 2, it + 1 - p->to_fourier);
    }
    o->it = it;
  }
}

/* Weight eps_i of component i at cell (x,y,z): the inverse of the factor the
 * E update applies to the curl there (code/fdtd.f90:86-147), real part for a
 * complex eps_inv as the reference's dble(1/eps_inv) (code/energy.f90:184). */
static double eps_weight(const oracle* o, int i, int ix, int iy, int iz) {
  const iris_params* p = &o->p;
  if (iz < p->hi) return 1.0 / p->epsa_inv[0][i][i];
  if (iz > p->hf) return 1.0 / p->epsa_inv[1][i][i];
  cx e = EPSI(o, ix, iy, iz);
  return e.re / (e.re * e.re + e.im * e.im);
}

/* Field monitors over x in 1..nx, y in 1..ny, z in zi..zf, restating the
 * reference's uncalled diagnostics (code/energy.f90):
 *   out[0] = sum Re(eps_i conj(E_i^n) E_i^(n-1))          calc_energy_density :166-201
 *   out[1] = sum |H_i^(n-1)|^2 / dtcq2                      (mu_inv = 1)
 *   out[2] = max |sum_i eps_i E_i^n(c) - eps_i E_i^n(c - e_i)|   calc_div_D :110-162
 *   out[3] = max |sum_i H_i^n(c + e_i) - H_i^n(c)|               calc_div_B :57-106
 * The reference's calc_energy_density indexes eps_inv(i,i,...) on an array
 * whose first two extents are 1 and outside its z range, and calc_div_D drops
 * eps and the imaginary parts ("valid only in vacuum", :139); here eps is the
 * weight of the E update and the fields are complex.  With these weights
 * out[0]+out[1] and the two divergences are invariants of the update in a
 * source-free, lossless region.  x/y neighbours across the period take the
 * Bloch factor exp(-+i ak) (code/boundary.f90:163-348). */
void oracle_diagnostics(const oracle* o, int zi, int zf, double* out) {
  const iris_params* p = &o->p;
  const double dtcq2 = ((p->dt * p->c0) / p->q) * ((p->dt * p->c0) / p->q);
  const cx fxm = bloch_factor(o, 0, -1), fym = bloch_factor(o, 1, -1);
  const cx fxp = bloch_factor(o, 0, +1), fyp = bloch_factor(o, 1, +1);
  double we = 0.0, wh = 0.0, dd = 0.0, db = 0.0;
  for (int iz = zi; iz <= zf; iz++)
    for (int iy = 1; iy <= o->ny; iy++)
      for (int ix = 1; ix <= o->nx; ix++) {
        cx divd = {0.0, 0.0}, divb = {0.0, 0.0};
        for (int i = 0; i < 3; i++) {
          const double w = eps_weight(o, i, ix, iy, iz);
          const cx ec = F(o, o->e_cur, i, ix, iy, iz), ep = F(o, o->e_prev, i, ix, iy, iz);
          const cx hp = F(o, o->h_prev, i, ix, iy, iz);
          we += w * (ec.re * ep.re + ec.im * ep.im);
          wh += (hp.re * hp.re + hp.im * hp.im) / dtcq2;
          /* backward neighbour of E_i and forward neighbour of H_i along axis i */
          int bx = ix, by = iy, bz = iz, fx = ix, fy = iy, fz = iz;
          cx fb = {1.0, 0.0}, ff = {1.0, 0.0};
          if (i == 0) {
            bx = ix - 1; fx = ix + 1;
            if (bx == 0) { bx = o->nx; fb = fxm; }
            if (fx == o->nx + 1) { fx = 1; ff = fxp; }
          } else if (i == 1) {
            by = iy - 1; fy = iy + 1;
            if (by == 0) { by = o->ny; fb = fym; }
            if (fy == o->ny + 1) { fy = 1; ff = fyp; }
          } else {
            bz = iz - 1; fz = iz + 1;
          }
          const cx eb = cmul(fb, F(o, o->e_cur, i, bx, by, bz));
          divd = cadd(divd, csub(cmul_r(w, ec), cmul_r(eps_weight(o, i, bx, by, bz), eb)));
          const cx hf = cmul(ff, F(o, o->h_cur, i, fx, fy, fz));
          divb = cadd(divb, csub(hf, F(o, o->h_cur, i, ix, iy, iz)));
        }
        const double ad = hypot(divd.re, divd.im), ab = hypot(divb.re, divb.im);
        if (ad > dd) dd = ad;
        if (ab > db) db = ab;
      }
  out[0] = we;
  out[1] = wh;
  out[2] = dd;
  out[3] = db;
}

/* code/dataprocess.f90:241-277 */
void oracle_flux(const oracle* o, int nobox, double* jabs) {
  int k = nobox - 1, nof = o->p.nof;
  for (int i = 0; i < 6 * nof; i++) jabs[i] = 0.0;
  for (int iw = 0; iw < nof; iw++) {
    for (int j = 0; j < 3; j++) {
      int np = o->npts[k][j];
      const cx* a = o->emg[k][j] + (size_t)iw * np;
      for (int i = 0; i < np; i += 6) {
        int first_half = (i + 1) <= np / 2;
        double v;
        if (j == 0) {
          if (o->nx == 1) continue;
          cx h5 = {a[i + 5].re, -a[i + 5].im}, h4 = {a[i + 4].re, -a[i + 4].im};
          v = cmul(a[i + 1], h5).re - cmul(a[i + 2], h4).re;
          if (first_half) jabs[iw * 6 + 0] -= v; else jabs[iw * 6 + 1] += v;
        } else if (j == 1) {
          if (o->ny == 1) continue;
          cx h3 = {a[i + 3].re, -a[i + 3].im}, h5 = {a[i + 5].re, -a[i + 5].im};
          v = cmul(a[i + 2], h3).re - cmul(a[i + 0], h5).re;
          if (first_half) jabs[iw * 6 + 2] -= v; else jabs[iw * 6 + 3] += v;
        } else {
          cx h4 = {a[i + 4].re, -a[i + 4].im}, h3 = {a[i + 3].re, -a[i + 3].im};
          v = cmul(a[i + 0], h4).re - cmul(a[i + 1], h3).re;
          if (first_half) jabs[iw * 6 + 4] -= v; else jabs[iw * 6 + 5] += v;
        }
      }
    }
  }
  double s = (o->p.eps0 * o->p.q) / (2.0 * o->p.dt);
  for (int i = 0; i < 6 * nof; i++) jabs[i] = s * jabs[i];
}

int64_t oracle_dft_points(const oracle* o, int nobox, int pair) { return o->npts[nobox - 1][pair - 1]; }

void oracle_get_dft(const oracle* o, int nobox, int pair, double* out) {
  memcpy(out, o->emg[nobox - 1][pair - 1], sizeof(cx) * (size_t)o->npts[nobox - 1][pair - 1] * (size_t)o->p.nof);
}

void oracle_get_field(const oracle* o, int which, double* out) {
  memcpy(out, which ? o->h_cur : o->e_cur, sizeof(cx) * 3 * (size_t)o->ncell);
}

void oracle_set_field(oracle* o, int which, const double* in) {
  memcpy(which ? o->h_cur : o->e_cur, in, sizeof(cx) * 3 * (size_t)o->ncell);
}

int oracle_current_step(const oracle* o) { return o->it; }
