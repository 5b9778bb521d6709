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

template <typename C> __device__ inline C cconj(C a) { a.y = -a.y; return a; }

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

}  // namespace iris
