// iris_device.h -- device-side helpers and the kernel argument block of the
// MI355X FDTD engine.  Complex values are T2 (float2/double2) with the
// reference's evaluation order for every operation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace iris {

template <typename T> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename T> using cplx = typename vec2<T>::type;

template <typename T> __host__ __device__ inline cplx<T> mk(T re, T im) { cplx<T> r; r.x = re; r.y = im; return r; }
template <typename C> __host__ __device__ inline C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C> __host__ __device__ inline C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// real * complex: the reference promotes the real to (r,0); for finite
// operands that equals (r*re, r*im)
template <typename C, typename T> __host__ __device__ inline C cscale(T s, C a) { C r; r.x = s * a.x; r.y = s * a.y; return r; }
// The Yee updates e + s*c and h - s*c (real s): with FP32 fields one fused
// multiply-add per component (one rounding; every FP32 pass forms them this
// way, so the passes stay bitwise equal to each other), with FP64 fields the
// reference's separate multiply and add.
template <typename T> __device__ inline cplx<T> cmadd(cplx<T> e, T s, cplx<T> c) {
  if constexpr (sizeof(T) == 4) return mk<T>(__builtin_fmaf(s, c.x, e.x), __builtin_fmaf(s, c.y, e.y));
  else return cadd(e, cscale(s, c));
}
template <typename T> __device__ inline cplx<T> cmsub(cplx<T> h, T s, cplx<T> c) {
  if constexpr (sizeof(T) == 4) return mk<T>(__builtin_fmaf(-s, c.x, h.x), __builtin_fmaf(-s, c.y, h.y));
  else return csub(h, cscale(s, c));
}
// complex * complex as (ac-bd, ad+bc)
template <typename C> __host__ __device__ inline C cmul(C a, C b) {
  C r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
// complex / complex with logb/scalbn scaling (C99 Annex G / compiler-rt
// __divdc3 algorithm), which is what the reference build uses for
// complex/real division
__device__ inline double2 cdiv_scaled(double2 a, double2 b) {
  double c = b.x, d = b.y;
  int k = 0;
  double lw = logb(fmax(fabs(c), fabs(d)));
  if (isfinite(lw)) { k = (int)lw; c = scalbn(c, -k); d = scalbn(d, -k); }
  double den = c * c + d * d;
  double2 r;
  r.x = scalbn((a.x * c + a.y * d) / den, -k);
  r.y = scalbn((a.y * c - a.x * d) / den, -k);
  return r;
}

// Per-plane UPML coefficients, each one the FP64 value of a subexpression
// of reference code/boundary_II.f90:86-99 and :170-183 evaluated on the host.
// Index 0..2 follow the field component (1..3 in the reference).
template <typename T> struct PmlPlane {
  T fp[3], fm[3];   // fp_i, fm_i at this plane
  T cd[3];          // D/B decay: fm2/fp2, fm3/fp3, fm1/fp1
  T gd[3];          // D gain:    (2eps)/fp2, (2eps)/fp3, (2eps)/fp1
  T gb[3];          // B gain:    ((2eps)/fp2)*dtq2, ...
  T ce[3];          // E/H decay: fm3/fp3, fm1/fp1, fm2/fp2
  T ge[3];          // E gain:    1/(eps0*eps*fp3), 1/(eps0*eps*fp1), 1/(eps0*eps*fp2)
  T gh[3];          // H gain:    1/(mu0*fp3), 1/(mu0*fp1), 1/(mu0*fp2)
};

// Everything the Yee passes need, passed by value.  Each field is one
// allocation holding its three components cs elements apart; the UPML aux
// fields D (E side) and B (H side) hold the bottom slab [1,pmlzi] and then
// the top slab [pmlzf,nz], three components each (component stride ps0/ps1,
// top slab at offset pb1).  e/h/d/b are the previous step's values, en/hn/
// dn/bn receive the new ones (the same arrays for the in-place passes, the
// other buffer of each pair for the fused pass).
template <typename T> struct StepArgs {
  using C = cplx<T>;
  C* e;
  C* h;
  C* en;
  C* hn;
  C* d;
  C* b;
  C* dn;
  C* bn;
  int64_t cs;       // component stride of E and H
  int64_t ps0, ps1; // component stride of the bottom / top UPML slab
  int64_t pb1;      // offset of the top slab in D and B
  const T* eps_re;  // structure planes hi..hf (xy plane layout)
  const T* eps_im;  // nullptr when eps_inv is real
  const PmlPlane<T>* pml;  // indexed by iz (0..nz+1)
  int nx, ny, nz;
  int64_t sx, sxy;  // x pitch and plane pitch of the field arrays
  int xoff;         // storage offset of ix = 0 within a row
  int pmlzi, pmlzf, hi, hf;
  T a1[3], a3[3];   // diagonal of epsa_inv for regions I and III
  T e02;            // eps0**2
  T dtcq2;
  C fxp, fyp, fzp;  // exp(+i ak) Bloch factors
  C fxm, fym, fzm;  // exp(-i ak)
  int zchunk;       // planes per thread in the z-marching passes
  int nt_out = 0;   // two-step pass: stream its E and H of step n+2 past the L2 (1) or not (0)
  // two cells a lane: planes rec_z0..rec_z1 of each chunk's own planes also
  // get H of step n+1 into the second buffers (a gradient run's records)
  int rec_z0 = 1, rec_z1 = 0;
  // offset of component c of UPML slab k at in-slab index pidx
  __device__ int64_t aux(int k, int c, int64_t pidx) const {
    return (k ? pb1 + c * ps1 : c * ps0) + pidx;
  }
};

// TF/SF source constants (reference code/tfsf.f90:231-290), in device
// memory; the step index enters as c0dt*it.
struct TfsfConst {
  double q, ux, uz, zopump, ko, dk, sq, sign_kz, c0dt, amp, hamp, win;
  double ein, dtcq2, eps_ms;
  int ixo, iz_min, i_pol, cosine, mode2d, on;
  // steps added to every step index the kernels are given: 0, except while a
  // captured graph of steps replays at a later step than it was captured at
  int shift;
  // oblique incidence: the incident field matched to the Bloch condition
  // (iris_incident.h), (nt + 1) x 2 planes (iz_min - 1, iz_min) x 3
  // components at ix = 0, times e^{i phix ix} at ix; null for the
  // reference's analytic pulse
  const double2* tab;
  double phix;
};

}  // namespace iris
