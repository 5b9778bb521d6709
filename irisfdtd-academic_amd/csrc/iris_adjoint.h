// iris_adjoint.h -- discrete adjoint of the time step, for gradients of a
// face-flux objective with respect to the structure's eps_inv.
//
// The reference has no adjoint.  This is the exact transpose of the engine's
// own step (the operations of reference code/driver.f90:355-450): with the
// state u = (E, H, D, B) one step is u^n = A u^(n-1) + s^n, where A is
// linear (Yee curls, z-UPML, Bloch wraps) and s^n collects the TF/SF
// corrections, which do not depend on the state or on eps_inv.  For an
// objective J(u^1 .. u^N) the adjoint runs backwards,
//   u^(n-1)^ = A^T u^n^ + dJ/du^(n-1),
// and dJ/dW(c) = sum_n Re(conj(E^n^(c)) . curl_b H^(n-1)(c)) for the
// per-cell factor W = eps_inv of the structure's E update (and with i*curl
// for its imaginary part).  Complex values are pairs of reals: the adjoint of
// multiplying by a complex factor f is multiplying by conj(f).
//
// Transpose of one step, in reverse order of its parts:
//   DFT samples (after the step)  add dt * sum_w G(w) exp(-i w n') at the
//                                 face points (G = dJ/d(DFT field));
//   H min-side wraps              H^(src) += conj(f) H^(ghost), ghost := 0;
//   H update (code/fdtd.f90:157-248, UPML code/boundary_II.f90:142-220)
//                                 bulk: E^ += -dtcq2 curl_b(H^);
//                                 UPML: Bt = B^ + gh fp H^, c = -gb Bt,
//                                 B^ = cd Bt - gh fm H^, H^ = ce H^,
//                                 E^ += curl_b(c)  (curl_f^T = curl_b);
//   E max-side wraps              E^(src) += conj(f) E^(ghost), ghost := 0;
//   gradient                      on the structure cells;
//   E update (code/fdtd.f90:57-153, UPML code/boundary_II.f90:60-138)
//                                 region I/III: H^ += curl_f(a E^);
//                                 structure: H^ += curl_f(conj(W) E^);
//                                 UPML: Dt = D^ + ge fp E^, c = gd e02 Dt,
//                                 D^ = cd Dt - ge fm E^, E^ = ce E^,
//                                 H^ += curl_f(c)  (curl_b^T = curl_f).
#pragma once
#include "iris_kernels.h"

namespace iris {

// coordinates of thread t over an extended box x in [x0, x0+nxr), y in
// [y0, y0+nyr), z in [z0, z0+nzr)
struct Range3 {
  int x0, y0, z0, nxr, nyr, nzr;
  __device__ int64_t count() const { return (int64_t)nxr * nyr * nzr; }
  __device__ void at(int64_t t, int* x, int* y, int* z) const {
    *x = x0 + (int)(t % nxr);
    *y = y0 + (int)((t / nxr) % nyr);
    *z = z0 + (int)(t / ((int64_t)nxr * nyr));
  }
};

template <typename C> __host__ __device__ inline C cconj(C a) { a.y = -a.y; return a; }

// G = dJ/d(DFT field) for the points of pair j of a box, in the emg layout
// (npts, nof): J = sum_{iw, face} w[iw*6+face] * s * sign * sum_p
// Re(A conj B - C conj D) (the face flux of reference
// code/dataprocess.f90:241-277; sign -1 on the first face of the pair).
__global__ void k_adj_g(BoxGeom g, int j, const double2* __restrict__ emg, int nof, const double* __restrict__ w,
                        double s, double2* __restrict__ G) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t np = g.npts[j];
  if (r >= np) return;
  const int64_t half = np / 2;
  const bool second = r >= half;
  const int64_t base = r - (r % 6);  // first component of this point
  const int comp = (int)(r % 6);
  int i1, i2, i3, i4;  // A, B, C, D components (Ex..Hz = 0..5)
  if (j == 0) { i1 = 1; i2 = 5; i3 = 2; i4 = 4; }
  else if (j == 1) { i1 = 2; i2 = 3; i3 = 0; i4 = 5; }
  else { i1 = 0; i2 = 4; i3 = 1; i4 = 3; }
  const int face = 2 * j + (second ? 1 : 0);
  const double sg = second ? 1.0 : -1.0;
  const double2* e = emg + g.start[j] * (int64_t)nof;  // pair block, (np, nof)
  double2* out = G + g.start[j] * (int64_t)nof;
  for (int iw = 0; iw < nof; ++iw) {
    const double k = w[iw * 6 + face] * s * sg;
    double2 v = make_double2(0.0, 0.0);
    const double2* pt = e + (int64_t)iw * np + base;
    if (comp == i1) v = make_double2(k * pt[i2].x, k * pt[i2].y);
    else if (comp == i2) v = make_double2(k * pt[i1].x, k * pt[i1].y);
    else if (comp == i3) v = make_double2(-k * pt[i4].x, -k * pt[i4].y);
    else if (comp == i4) v = make_double2(-k * pt[i3].x, -k * pt[i3].y);
    out[(int64_t)iw * np + r] = v;
  }
}

// Adjoint source of one DFT sample: for the face half (pair j, second) add
// dt * sum_w G(w) conj(ph[w]) (ph = exp(i w n')) to the adjoint field of the
// point's component.  One launch per face half, so no two threads of a
// launch touch the same cell.
template <typename T>
__global__ void k_adj_source(BoxGeom g, int j, int second, const double2* __restrict__ G, int nof,
                             const double2* __restrict__ ph, double dt, cplx<T>* ae, cplx<T>* ah, int64_t cs,
                             int64_t sx, int64_t sxy, int xoff) {
  const int64_t np = g.npts[j];
  const int64_t half = np / 2;
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r0 >= half) return;
  const int64_t r = r0 + (second ? half : 0);
  int c, x, y, z;
  box_point(g, g.start[j] + r, &c, &x, &y, &z);
  const double2* Gp = G + g.start[j] * (int64_t)nof + r;
  double sr = 0.0, si = 0.0;
  for (int iw = 0; iw < nof; ++iw) {
    const double2 v = Gp[(int64_t)iw * np];
    const double2 p = ph[iw];
    // v * conj(p)
    sr += v.x * p.x + v.y * p.y;
    si += v.y * p.x - v.x * p.y;
  }
  if (sr == 0.0 && si == 0.0) return;
  const int64_t idx = (int64_t)(x + xoff) + sx * y + sxy * z + (int64_t)(c % 3) * cs;
  cplx<T>* f = (c < 3) ? ae : ah;
  f[idx].x += (T)(dt * sr);
  f[idx].y += (T)(dt * si);
}

// Transposed wrap along one axis: f(src) += conj(ph) * f(ghost) over the
// face's interior range, then every cell of the ghost plane is cleared.
// axis 0/1/2 = x/y/z; ghost = 0 (min side, src = n) or n+1 (max side, src = 1).
template <typename T>
__global__ void k_adj_wrap(cplx<T>* f, int64_t cs, int axis, int ghost, int src, cplx<T> ph, int nx, int ny, int nz,
                           int64_t sx, int64_t sxy, int xoff) {
  // the ghost plane is indexed by the other two coordinates, 0..n+1 each
  const int na = (axis == 0) ? ny + 2 : nx + 2;
  const int nb = (axis == 2) ? ny + 2 : nz + 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)na * nb * 3) return;
  const int comp = (int)(t / ((int64_t)na * nb));
  const int64_t q = t % ((int64_t)na * nb);
  const int ia = (int)(q % na), ib = (int)(q / na);
  int x, y, z, xs, ys, zs;
  bool inner;
  if (axis == 0) { x = ghost; y = ia; z = ib; xs = src; ys = y; zs = z; inner = y >= 1 && y <= ny && z >= 1 && z <= nz; }
  else if (axis == 1) { x = ia; y = ghost; z = ib; xs = x; ys = src; zs = z; inner = x >= 1 && x <= nx && z >= 1 && z <= nz; }
  else { x = ia; y = ib; z = ghost; xs = x; ys = y; zs = src; inner = x >= 1 && x <= nx && y >= 1 && y <= ny; }
  const int64_t gi = (int64_t)(x + xoff) + sx * y + sxy * z + comp * cs;
  const cplx<T> v = f[gi];
  if (inner) {
    const int64_t si = (int64_t)(xs + xoff) + sx * ys + sxy * zs + comp * cs;
    f[si] = cadd(f[si], cmul(cconj(ph), v));
  }
  f[gi] = mk<T>(0, 0);
}

// Local part of the transposed H update on the interior cells: writes the
// curl argument c into tmp and updates H^ and B^ in place.
template <typename T>
__global__ void k_adj_h_local(StepArgs<T> a, cplx<T>* ah, cplx<T>* ab, cplx<T>* tmp) {
  using C = cplx<T>;
  const Range3 R{1, 1, 1, a.nx, a.ny, a.nz};
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.count()) return;
  int x, y, z;
  R.at(t, &x, &y, &z);
  const int64_t col = (int64_t)(x + a.xoff) + a.sx * y;
  const int64_t idx = col + a.sxy * z;
  if (z <= a.pmlzi || z >= a.pmlzf) {
    const int k = z >= a.pmlzf;
    const int64_t pidx = col + a.sxy * (k ? (z - a.pmlzf) : (z - 1));
    const PmlPlane<T>& P = a.pml[z];
    for (int c = 0; c < 3; ++c) {
      const int64_t ai = a.aux(k, c, pidx);
      const C hh = ah[idx + c * a.cs];
      const C bt = cadd(ab[ai], cscale(P.gh[c] * P.fp[c], hh));
      tmp[idx + c * a.cs] = cscale(-P.gb[c], bt);
      ab[ai] = csub(cscale(P.cd[c], bt), cscale(P.gh[c] * P.fm[c], hh));
      ah[idx + c * a.cs] = cscale(P.ce[c], hh);
    }
  } else {
    for (int c = 0; c < 3; ++c) tmp[idx + c * a.cs] = cscale(-a.dtcq2, ah[idx + c * a.cs]);
  }
}

// Local part of the transposed E update on the interior cells: the curl
// argument c into tmp, E^ and D^ updated in place (UPML).
template <typename T>
__global__ void k_adj_e_local(StepArgs<T> a, cplx<T>* ae, cplx<T>* ad, cplx<T>* tmp) {
  using C = cplx<T>;
  const Range3 R{1, 1, 1, a.nx, a.ny, a.nz};
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.count()) return;
  int x, y, z;
  R.at(t, &x, &y, &z);
  const int64_t col = (int64_t)(x + a.xoff) + a.sx * y;
  const int64_t idx = col + a.sxy * z;
  if (z <= a.pmlzi || z >= a.pmlzf) {
    const int k = z >= a.pmlzf;
    const int64_t pidx = col + a.sxy * (k ? (z - a.pmlzf) : (z - 1));
    const PmlPlane<T>& P = a.pml[z];
    for (int c = 0; c < 3; ++c) {
      const int64_t ai = a.aux(k, c, pidx);
      const C ee = ae[idx + c * a.cs];
      const C dt_ = cadd(ad[ai], cscale(P.ge[c] * P.fp[c], ee));
      tmp[idx + c * a.cs] = cscale(P.gd[c] * a.e02, dt_);
      ad[ai] = csub(cscale(P.cd[c], dt_), cscale(P.ge[c] * P.fm[c], ee));
      ae[idx + c * a.cs] = cscale(P.ce[c], ee);
    }
  } else if (z < a.hi || z > a.hf) {
    const T* ac = (z < a.hi) ? a.a1 : a.a3;
    for (int c = 0; c < 3; ++c) tmp[idx + c * a.cs] = cscale(ac[c], ae[idx + c * a.cs]);
  } else {
    const int64_t eidx = col + a.sxy * (z - a.hi);
    const C w = mk<T>(a.eps_re[eidx], a.eps_im ? a.eps_im[eidx] : (T)0);
    for (int c = 0; c < 3; ++c) tmp[idx + c * a.cs] = cmul(cconj(w), ae[idx + c * a.cs]);
  }
}

// E^ += curl_b(tmp) on E cells x 1..nx+1, y 1..ny+1, z 1..nz+1 (tmp is
// zero outside the interior): the transpose of the H update's curl_f.
template <typename T>
__global__ void k_adj_curl_b(StepArgs<T> a, const cplx<T>* __restrict__ v, cplx<T>* ae) {
  using C = cplx<T>;
  const Range3 R{1, 1, 1, a.nx + 1, a.ny + 1, a.nz + 1};
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.count()) return;
  int x, y, z;
  R.at(t, &x, &y, &z);
  const int64_t i = (int64_t)(x + a.xoff) + a.sx * y + a.sxy * z;
  const int64_t cs = a.cs;
  const C v0 = v[i], v1 = v[i + cs], v2 = v[i + 2 * cs];
  const C c0 = csub(csub(v2, v[i - a.sx + 2 * cs]), csub(v1, v[i - a.sxy + cs]));
  const C c1 = csub(csub(v0, v[i - a.sxy]), csub(v2, v[i - 1 + 2 * cs]));
  const C c2 = csub(csub(v1, v[i - 1 + cs]), csub(v0, v[i - a.sx]));
  ae[i] = cadd(ae[i], c0);
  ae[i + cs] = cadd(ae[i + cs], c1);
  ae[i + 2 * cs] = cadd(ae[i + 2 * cs], c2);
}

// H^ += curl_f(tmp) on H cells x 0..nx, y 0..ny, z 0..nz: the transpose of
// the E update's curl_b.
template <typename T>
__global__ void k_adj_curl_f(StepArgs<T> a, const cplx<T>* __restrict__ v, cplx<T>* ah) {
  using C = cplx<T>;
  const Range3 R{0, 0, 0, a.nx + 1, a.ny + 1, a.nz + 1};
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.count()) return;
  int x, y, z;
  R.at(t, &x, &y, &z);
  const int64_t i = (int64_t)(x + a.xoff) + a.sx * y + a.sxy * z;
  const int64_t cs = a.cs;
  const C v0 = v[i], v1 = v[i + cs], v2 = v[i + 2 * cs];
  const C c0 = csub(csub(v[i + a.sx + 2 * cs], v2), csub(v[i + a.sxy + cs], v1));
  const C c1 = csub(csub(v[i + a.sxy], v0), csub(v[i + 1 + 2 * cs], v2));
  const C c2 = csub(csub(v[i + 1 + cs], v1), csub(v[i + a.sx], v0));
  ah[i] = cadd(ah[i], c0);
  ah[i + cs] = cadd(ah[i + cs], c1);
  ah[i + 2 * cs] = cadd(ah[i + 2 * cs], c2);
}

// grad(c) += (Re(conj(E^).curl_b H), Re(conj(E^).i curl_b H)) summed over the
// components, on the structure cells; hrec holds H^(n-1) on planes hi-1..hf
// (the field's xy layout, three components nplane elements apart).
template <typename T>
__global__ void k_adj_grad(StepArgs<T> a, const cplx<T>* __restrict__ ae, const cplx<T>* __restrict__ hrec,
                           int64_t nplane, double2* __restrict__ grad) {
  using C = cplx<T>;
  const Range3 R{1, 1, a.hi, a.nx, a.ny, a.hf - a.hi + 1};
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= R.count()) return;
  int x, y, z;
  R.at(t, &x, &y, &z);
  const int64_t col = (int64_t)(x + a.xoff) + a.sx * y;
  const int64_t r = col + a.sxy * (z - (a.hi - 1));  // index in hrec
  const C h1 = hrec[r], h2 = hrec[r + nplane], h3 = hrec[r + 2 * nplane];
  const C c0 = cadd(csub(csub(h3, hrec[r - a.sx + 2 * nplane]), h2), hrec[r - a.sxy + nplane]);
  const C c1 = cadd(csub(csub(h1, hrec[r - a.sxy]), h3), hrec[r - 1 + 2 * nplane]);
  const C c2 = cadd(csub(csub(h2, hrec[r - 1 + nplane]), h1), hrec[r - a.sx]);
  const int64_t i = col + a.sxy * z;
  const C e0 = ae[i], e1 = ae[i + a.cs], e2 = ae[i + 2 * a.cs];
  double gr = 0.0, gi = 0.0;
  gr += (double)e0.x * c0.x + (double)e0.y * c0.y;
  gr += (double)e1.x * c1.x + (double)e1.y * c1.y;
  gr += (double)e2.x * c2.x + (double)e2.y * c2.y;
  gi += (double)e0.y * c0.x - (double)e0.x * c0.y;
  gi += (double)e1.y * c1.x - (double)e1.x * c1.y;
  gi += (double)e2.y * c2.x - (double)e2.x * c2.y;
  grad[t].x += gr;
  grad[t].y += gi;
}

// ---------------------------------------------------------------------------
// Fused adjoint step over double-buffered adjoint fields, with the tile shape
// and z-march of the forward fused pass.
//
// With Y = the curl argument of the transposed H update and X = that of the
// transposed E update, one adjoint step is
//   E^mid = E^ + curl_b(Y),   X = a E^mid (a = epsa_inv, or conj(eps_inv) on
//                             the structure, or gd e02 (D^ + ge fp E^mid) in
//                             the UPML),
//   H^new = H^loc + curl_f(X),
// where Y = -dtcq2 H^ (bulk) or -gb (B^ + gh fp H^) (UPML), H^loc = H^ or
// ce H^, and the transposed wraps make Y(0) = conj(exp(+i ak)) Y(n) and
// X(n+1) = conj(exp(-i ak)) X(1).  The state is kept per plane kind:
//   UPML and structure planes  E^ and H^ (and D^, B^ in the UPML);
//   the other planes           X = a E^ and Y = -dtcq2 H^.
// On the latter the step reads  X' = X + a curl_b(Y),  Y' = Y - dtcq2
// curl_f(X'): the forward Yee update with the same coefficients.  So the
// forward's lean pass and its two-step pass advance the adjoint on chunks of
// such planes unchanged, given the Bloch factors conj(exp(-+i ak)) in place
// of exp(+-i ak); this kernel takes the other chunks, with the operations of
// the forward kernels on those planes so that both routes agree bitwise.
// ---------------------------------------------------------------------------

enum { kAdjUpml = 0, kAdjStruct = 1, kAdjPlain = 2 };

template <typename T>
__device__ inline int adj_kind(const StepArgs<T>& a, int z) {
  if (z <= a.pmlzi || z >= a.pmlzf) return kAdjUpml;
  if (z >= a.hi && z <= a.hf) return kAdjStruct;
  return kAdjPlain;
}

// Y component c at field index i (aux index pidx in the UPML) of a plane of
// kind `kind`
template <typename T>
__device__ inline cplx<T> adj_y(const StepArgs<T>& a, int c, int64_t i, int kind, int k, int64_t pidx,
                                const PmlPlane<T>* P) {
  const cplx<T> hh = a.h[i + c * a.cs];
  if (kind == kAdjUpml) {
    const cplx<T> bt = cadd(a.b[a.aux(k, c, pidx)], cscale(P->gh[c] * P->fp[c], hh));
    return cscale(-P->gb[c], bt);
  }
  if (kind == kAdjStruct) return cscale(-a.dtcq2, hh);
  return hh;
}

struct AdjPlane {
  int kind;
  int k;
  int64_t pz;  // aux plane offset
};
template <typename T>
__device__ inline AdjPlane adj_plane(const StepArgs<T>& a, int z) {
  AdjPlane q{adj_kind(a, z), 0, 0};
  if (q.kind == kAdjUpml) {
    q.k = z >= a.pmlzf;
    q.pz = a.sxy * (q.k ? (z - a.pmlzf) : (z - 1));
  }
  return q;
}

// Y1, Y2 at source cell (xs, ys) of plane z; z = 0 is the image of plane nz
template <typename T>
__device__ inline void adj_y12(const StepArgs<T>& a, int xs, int ys, int z, cplx<T>* y1, cplx<T>* y2) {
  const bool img = z == 0;
  const int zz = img ? a.nz : z;
  const AdjPlane q = adj_plane<T>(a, zz);
  const int64_t col = (int64_t)(xs + a.xoff) + a.sx * ys;
  const int64_t i = col + a.sxy * zz;
  const PmlPlane<T>* P = &a.pml[zz];
  *y1 = adj_y<T>(a, 0, i, q.kind, q.k, col + q.pz, P);
  *y2 = adj_y<T>(a, 1, i, q.kind, q.k, col + q.pz, P);
  if (img) {
    *y1 = cmul(a.fzm, *y1);
    *y2 = cmul(a.fzm, *y2);
  }
}

// The E side of the adjoint step at source cell (xs, ys, zs) in 1..n: X, the
// own Y and H^loc (Y itself on a plain plane); stores E^new (X on a plain
// plane), D^new and B^new and adds the gradient when own.  y1m/y2m are Y1/Y2
// of the cell below.  a.fxm/a.fym are the factors of the Y ghosts.
template <typename T>
__device__ inline void adj_e_item(const StepArgs<T>& a, int xs, int ys, int zs, cplx<T> y1m, cplx<T> y2m,
                                  bool own, const cplx<T>* __restrict__ hrec, int64_t nplane,
                                  double2* __restrict__ grad, cplx<T> x[3], cplx<T> yown[3], cplx<T> hloc[3]) {
  using C = cplx<T>;
  const int64_t cs = a.cs;
  const int64_t col = (int64_t)(xs + a.xoff) + a.sx * ys;
  const int64_t idx = col + a.sxy * zs;
  const AdjPlane q = adj_plane<T>(a, zs);
  const PmlPlane<T>* P = &a.pml[zs];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const C hh = a.h[idx + c * cs];
    if (q.kind == kAdjUpml) {
      const int64_t ai = a.aux(q.k, c, col + q.pz);
      const C bt = cadd(a.b[ai], cscale(P->gh[c] * P->fp[c], hh));
      yown[c] = cscale(-P->gb[c], bt);
      hloc[c] = cscale(P->ce[c], hh);
      if (own) a.bn[ai] = csub(cscale(P->cd[c], bt), cscale(P->gh[c] * P->fm[c], hh));
    } else {
      yown[c] = (q.kind == kAdjStruct) ? cscale(-a.dtcq2, hh) : hh;
      hloc[c] = hh;
    }
  }
  // x-1 and y-1 neighbours.  On a plain plane index 0 is the stored ghost
  // (as the forward passes read it); elsewhere it is formed from index n.
  const bool plain = q.kind == kAdjPlain;
  const int xm = (xs == 1 && !plain) ? a.nx : xs - 1;
  const int ym = (ys == 1 && !plain) ? a.ny : ys - 1;
  const int64_t cxm = (int64_t)(xm + a.xoff) + a.sx * ys, cym = (int64_t)(xs + a.xoff) + a.sx * ym;
  C y3xm = adj_y<T>(a, 2, cxm + a.sxy * zs, q.kind, q.k, cxm + q.pz, P);
  C y2xm = adj_y<T>(a, 1, cxm + a.sxy * zs, q.kind, q.k, cxm + q.pz, P);
  C y3ym = adj_y<T>(a, 2, cym + a.sxy * zs, q.kind, q.k, cym + q.pz, P);
  C y1ym = adj_y<T>(a, 0, cym + a.sxy * zs, q.kind, q.k, cym + q.pz, P);
  if (xs == 1 && !plain) { y3xm = cmul(a.fxm, y3xm); y2xm = cmul(a.fxm, y2xm); }
  if (ys == 1 && !plain) { y3ym = cmul(a.fym, y3ym); y1ym = cmul(a.fym, y1ym); }
  // curl_b(Y) in the forward's order (code/fdtd.f90:78-83)
  const C c0 = cadd(csub(csub(yown[2], y3ym), yown[1]), y2m);
  const C c1 = cadd(csub(csub(yown[0], y1m), yown[2]), y3xm);
  const C c2 = cadd(csub(csub(yown[1], y2xm), yown[0]), y1ym);
  const C e0 = a.e[idx], e1 = a.e[idx + cs], e2 = a.e[idx + 2 * cs];
  C en[3];
  if (q.kind == kAdjPlain) {
    // X' = X + a curl_b(Y), as the forward E update of region I / III
    const T* ac = (zs < a.hi) ? a.a1 : a.a3;
    x[0] = cmadd<T>(e0, ac[0], c0);
    x[1] = cmadd<T>(e1, ac[1], c1);
    x[2] = cmadd<T>(e2, ac[2], c2);
    en[0] = x[0]; en[1] = x[1]; en[2] = x[2];
  } else {
    const C em[3] = {cadd(e0, c0), cadd(e1, c1), cadd(e2, c2)};
    if (q.kind == kAdjUpml) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int64_t ai = a.aux(q.k, c, col + q.pz);
        const C dt_ = cadd(a.d[ai], cscale(P->ge[c] * P->fp[c], em[c]));
        x[c] = cscale(P->gd[c] * a.e02, dt_);
        if (own) a.dn[ai] = csub(cscale(P->cd[c], dt_), cscale(P->ge[c] * P->fm[c], em[c]));
        en[c] = cscale(P->ce[c], em[c]);
      }
    } else {
      const int64_t eidx = col + a.sxy * (zs - a.hi);
      const C w = mk<T>(a.eps_re[eidx], a.eps_im ? a.eps_im[eidx] : (T)0);
#pragma unroll
      for (int c = 0; c < 3; ++c) { x[c] = cmul(cconj(w), em[c]); en[c] = em[c]; }
      if (own) {
        // gradient with H of the step before, recorded on planes hi-1..hf
        const int64_t r = col + a.sxy * (zs - (a.hi - 1));
        const C h1 = hrec[r], h2 = hrec[r + nplane], h3 = hrec[r + 2 * nplane];
        const C d0 = cadd(csub(csub(h3, hrec[r - a.sx + 2 * nplane]), h2), hrec[r - a.sxy + nplane]);
        const C d1 = cadd(csub(csub(h1, hrec[r - a.sxy]), h3), hrec[r - 1 + 2 * nplane]);
        const C d2 = cadd(csub(csub(h2, hrec[r - 1 + nplane]), h1), hrec[r - a.sx]);
        double gr = 0.0, gi = 0.0;
        gr += (double)em[0].x * d0.x + (double)em[0].y * d0.y;
        gr += (double)em[1].x * d1.x + (double)em[1].y * d1.y;
        gr += (double)em[2].x * d2.x + (double)em[2].y * d2.y;
        gi += (double)em[0].y * d0.x - (double)em[0].x * d0.y;
        gi += (double)em[1].y * d1.x - (double)em[1].x * d1.y;
        gi += (double)em[2].y * d2.x - (double)em[2].x * d2.y;
        const int64_t t = (int64_t)(xs - 1) + (int64_t)a.nx * ((ys - 1) + (int64_t)a.ny * (zs - a.hi));
        double2 g = grad[t];
        g.x += gr;
        g.y += gi;
        grad[t] = g;
      }
    }
  }
  if (own) {
    a.en[idx] = en[0];
    a.en[idx + cs] = en[1];
    a.en[idx + 2 * cs] = en[2];
  }
}

// The adjoint step over the chunks listed (the general chunks of the forward
// fused pass, or the chunks a two-step pass leaves to single steps); chunks,
// tm and the launch shape are the forward fused pass's.  a holds the adjoint
// state of step n in e/h/d/b and receives that of step n-1 in en/hn/dn/bn;
// a's Bloch factors are the adjoint ones (fxm: Y ghosts, fxp: X images).
template <typename T, int TY>
__global__ __launch_bounds__(kFTX * (TY + 1), 1) void k_adj_gen(
    StepArgs<T> a, const int* chunks, TileMap tm, const cplx<T>* __restrict__ hrec, int64_t nplane,
    double2* __restrict__ grad) {
  using C = cplx<T>;
  __shared__ C sxl[3][TY + 1][kFTX + 1];
  const int lane = threadIdx.x, wy = threadIdx.y;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t cs = a.cs;
  int tbx, tby, tbz;
  tile_of(tm, blockIdx.x, &tbx, &tby, &tbz);
  const int x0 = tbx * kFTX + 1, y0 = tby * TY + 1;
  const int cw = chunks[tbz];
  const int z0 = cw & kChunkZMask, zl = (cw >> kChunkLenShift) & kChunkLenMask;
  const int z1 = min(z0 + (zl ? zl : a.zchunk) - 1, nz);
  const bool halo = wy == TY;
  const int xa = x0 + lane, ya = y0 + wy;
  bool va = xa <= nx + 1 && ya <= ny + 1 && !(xa == nx + 1 && ya == ny + 1);
  if (halo) va = va && xa <= nx;
  const bool owna = !halo && xa <= nx && ya <= ny;
  const int xsa = (xa == nx + 1) ? 1 : xa, ysa = (ya == ny + 1) ? 1 : ya;
  const int xb = x0 + kFTX, yb = y0 + lane;
  const bool vb = halo && lane < TY && xb <= nx + 1 && yb <= ny;
  const int xsb = (xb == nx + 1) ? 1 : xb;
  const int64_t cola = (int64_t)(xa + a.xoff) + a.sx * ya;

  C a1m{}, a2m{}, b1m{}, b2m{}, xp[3] = {}, hl[3] = {};
  if (va) adj_y12<T>(a, xsa, ysa, z0 - 1, &a1m, &a2m);
  if (vb) adj_y12<T>(a, xsb, yb, z0 - 1, &b1m, &b2m);

  for (int z = z0; z <= z1 + 1; ++z) {
    const bool real_plane = z <= z1;
    const bool image_z = z == nz + 1;
    const int zs = image_z ? 1 : z;
    C na[3], nb[3], ya_[3], yb_[3], ha[3], hb[3];
    const bool doa = va && (real_plane || owna);
    if (doa) {
      C z1m = a1m, z2m = a2m;
      if (image_z) {
        // Y of plane 0 is the image of plane nz, carried in a1m/a2m
        z1m = cmul(a.fzm, a1m);
        z2m = cmul(a.fzm, a2m);
      }
      adj_e_item<T>(a, xsa, ysa, zs, z1m, z2m, owna && real_plane, hrec, nplane, grad, na, ya_, ha);
      if (xa == nx + 1) phase3(a.fxp, na);
      if (ya == ny + 1) phase3(a.fyp, na);
      if (image_z) phase3(a.fzp, na);
    }
    const bool dob = vb && real_plane;
    if (dob) {
      adj_e_item<T>(a, xsb, yb, zs, b1m, b2m, false, hrec, nplane, grad, nb, yb_, hb);
      if (xb == nx + 1) phase3(a.fxp, nb);
    }
    // H side of plane z-1 from X of plane z-1 (own in xp, +x/+y in LDS) and
    // X of plane z (own in na)
    if (owna && z > z0) {
      const int zh = z - 1;
      const C x3yp = sxl[2][wy + 1][lane], x1yp = sxl[0][wy + 1][lane];
      const C x3xp = sxl[2][wy][lane + 1], x2xp = sxl[1][wy][lane + 1];
      const C c0 = cadd(csub(csub(x3yp, xp[2]), na[1]), xp[1]);
      const C c1 = cadd(csub(csub(na[0], xp[0]), x3xp), xp[2]);
      const C c2 = cadd(csub(csub(x2xp, xp[1]), x1yp), xp[0]);
      const int64_t ih = cola + a.sxy * zh;
      C hn[3];
      if (adj_kind(a, zh) == kAdjPlain) {
        // Y' = Y - dtcq2 curl_f(X), as the forward H update, with the
        // min-side ghosts the forward passes read
        hn[0] = cmsub<T>(hl[0], a.dtcq2, c0);
        hn[1] = cmsub<T>(hl[1], a.dtcq2, c1);
        hn[2] = cmsub<T>(hl[2], a.dtcq2, c2);
        if (xa == nx) store3_phase(a.hn, ih - nx, cs, a.fxm, hn);
        if (ya == ny) store3_phase(a.hn, ih - a.sx * ny, cs, a.fym, hn);
      } else {
        hn[0] = cadd(hl[0], c0);
        hn[1] = cadd(hl[1], c1);
        hn[2] = cadd(hl[2], c2);
      }
      store3(a.hn, ih, cs, hn);
    }
    if (doa) { a1m = ya_[0]; a2m = ya_[1]; hl[0] = ha[0]; hl[1] = ha[1]; hl[2] = ha[2]; }
    if (dob) { b1m = yb_[0]; b2m = yb_[1]; }
    if (!real_plane) break;
    __syncthreads();
    if (doa) { sxl[0][wy][lane] = na[0]; sxl[1][wy][lane] = na[1]; sxl[2][wy][lane] = na[2]; }
    if (dob) { sxl[0][lane][kFTX] = nb[0]; sxl[1][lane][kFTX] = nb[1]; sxl[2][lane][kFTX] = nb[2]; }
    __syncthreads();
    if (owna) { xp[0] = na[0]; xp[1] = na[1]; xp[2] = na[2]; }
  }
}

// Objective source at the face points in the mixed representation: on a
// plain plane the E part goes in as a E^ (X) and the H part as -dtcq2 H^ (Y),
// and the stored min-side ghosts of Y follow their cells.
// which = 0 / 1: the first / second face of pair j; 2: both (faces that
// share no cell)
template <typename T>
__global__ void k_adj_source_mixed(StepArgs<T> a, BoxGeom g, int j, int which, const double2* __restrict__ G,
                                   int nof, const double2* __restrict__ ph, double dt) {
  const int64_t np = g.npts[j];
  const int64_t half = np / 2;
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r0 >= (which == 2 ? np : half)) return;
  const int64_t r = r0 + (which == 1 ? half : 0);
  int c, x, y, z;
  box_point(g, g.start[j] + r, &c, &x, &y, &z);
  const double2* Gp = G + g.start[j] * (int64_t)nof + r;
  double sr = 0.0, si = 0.0;
  for (int iw = 0; iw < nof; ++iw) {
    const double2 v = Gp[(int64_t)iw * np];
    const double2 p = ph[iw];
    sr += v.x * p.x + v.y * p.y;
    si += v.y * p.x - v.x * p.y;
  }
  if (sr == 0.0 && si == 0.0) return;
  double f = dt;
  if (adj_kind(a, z) == kAdjPlain)
    f = (c < 3) ? dt * (double)((z < a.hi) ? a.a1[c] : a.a3[c % 3]) : -dt * (double)a.dtcq2;
  const int64_t idx = (int64_t)(x + a.xoff) + a.sx * y + a.sxy * z + (int64_t)(c % 3) * a.cs;
  cplx<T>* fld = (c < 3) ? a.e : a.h;
  const cplx<T> d = mk<T>((T)(f * sr), (T)(f * si));
  fld[idx] = cadd(fld[idx], d);
  if (c >= 3 && adj_kind(a, z) == kAdjPlain) {
    if (x == a.nx) fld[idx - a.nx] = cadd(fld[idx - a.nx], cmul(a.fxm, d));
    if (y == a.ny) fld[idx - a.sx * a.ny] = cadd(fld[idx - a.sx * a.ny], cmul(a.fym, d));
  }
}

// H on the record planes of one step: three components, n16 16-byte words
// each, component c at src + c*cs16 words; the record holds them back to back
__global__ __launch_bounds__(256) void k_record(const int4* __restrict__ src, int64_t cs16, int64_t n16,
                                                int4* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * n16) return;
  const int64_t c = t / n16;
  dst[t] = src[c * cs16 + (t - c * n16)];
}

}  // namespace iris
