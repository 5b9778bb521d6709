// iris_kernels.h -- gfx950 kernels of the MI355X FDTD engine: the Yee step
// (iris_yee.h) and the on-the-fly DFT and face fluxes
// (reference code/dataprocess.f90:109-304).  Samples of the box faces are
// buffered and folded into the FP64 accumulators every K samples, in the
// reference's accumulation order.
#pragma once
#include "iris_yee.h"

namespace iris {

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

// whether point index i of box g lies on an x face (the first pair block)
__device__ inline bool on_xface(const BoxGeom& g, int64_t i) {
  return g.npts[0] && i < g.start[0] + g.npts[0];
}

// copy the box-face samples of this step into slot `slot` of the buffer
// (xmask set: not the x-face points on the planes z with xmask[z] != 0,
// which the two-step kernels have stored)
template <typename T>
__global__ __launch_bounds__(256) void k_dft_sample(StepArgs<T> a, BoxGeom g, cplx<T>* samp, int slot,
                                                   const unsigned char* __restrict__ xmask) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.ntot) return;
  int c, x, y, z;
  box_point(g, i, &c, &x, &y, &z);
  if (xmask && xmask[z] && on_xface(g, i)) return;
  const int64_t idx = (int64_t)(x + a.xoff) + a.sx * y + a.sxy * z;
  samp[(int64_t)slot * g.ntot + i] = (c < 3) ? a.e[idx + c * a.cs] : a.h[idx + (c - 3) * a.cs];
}

// Batched sampling: entry e (one box of one instance of a batch) owns the
// concatenated point indices start[e] .. start[e] + g.ntot - 1 and reads the
// fields of its instance's state a[par].  Every start[e] is a multiple of the
// 256-thread block, so a block works for one entry: the lookup and the
// entry's arguments are uniform (scalar loads).
template <typename T> struct BatchSamp {
  StepArgs<T> a[2];
  BoxGeom g;
  cplx<T>* samp;
  const unsigned char* zmask;  // when set: skip the planes z with zmask[z] != 0
  const unsigned char* xmask;  // with xm: skip the x-face points of the planes z with xmask[z] != 0
};
template <typename T>
__global__ __launch_bounds__(256) void k_dft_sample_batch(const BatchSamp<T>* __restrict__ ent,
                                                          const int64_t* __restrict__ start, int nent, int par,
                                                          int slot, int xm) {
  const int64_t b0 = (int64_t)blockIdx.x * blockDim.x;
  if (b0 >= start[nent]) return;
  int lo = 0, hi = nent;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= b0) lo = mid; else hi = mid;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  const BatchSamp<T>& E = ent[lo];
  const StepArgs<T>& a = E.a[par];
  const int64_t r = b0 + threadIdx.x - start[lo];
  if (r >= E.g.ntot) return;
  int c, x, y, z;
  box_point(E.g, r, &c, &x, &y, &z);
  if (E.zmask && E.zmask[z]) return;
  if (xm && E.xmask && E.xmask[z] && on_xface(E.g, r)) return;
  const int64_t idx = (int64_t)(x + a.xoff) + a.sx * y + a.sxy * z;
  E.samp[(int64_t)slot * E.g.ntot + r] = (c < 3) ? a.e[idx + c * a.cs] : a.h[idx + (c - 3) * a.cs];
}

// phase table ph[k][iw] = scale*exp(i*fw(iw)*n_k), n_k = n0 + k*stride
// (reference code/dataprocess.f90:145: cdexp(ci*w*it) = (cos, sin)(w*it));
// scale is 1 for the exact fold and dt for the fused-multiply-add fold.
// Rows nk..nrows-1 are zero.
__global__ void k_dft_phase(const double* fw, int nof, int nk, int nrows, int n0, int stride, double scale,
                            double2* ph, const TfsfConst* __restrict__ tp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nof * nrows) return;
  if (tp) n0 += tp->shift;  // sample indices count steps
  const int k = t / nof, iw = t % nof;
  if (k >= nk) {
    ph[t] = make_double2(0.0, 0.0);
    return;
  }
  const double arg = fw[iw] * (double)(n0 + k * stride);
  double s, c;
  sincos(arg, &s, &c);
  ph[t] = (scale == 1.0) ? make_double2(c, s) : make_double2(scale * c, scale * s);
}

// Sets the step shift of a handle's TF/SF constants (one lane).
__global__ void k_set_shift(TfsfConst* tp, int shift) {
  if (blockIdx.x == 0 && threadIdx.x == 0) tp->shift = shift;
}

// emg(i,iw) += dt*F_k(i)*exp(i w n_k) for the buffered samples k < nk, with
// FP64 fields.  Thread = one point-component; its samples stay in registers
// for the sweep over all wavelengths, so each sample is read once per fold
// and each accumulator is read and written once per fold.  The phase table
// is staged through LDS in chunks of kDftChunk wavelengths: every lane reads
// the same phase (an LDS broadcast), and the reads do not wait on the L2
// latency of scalar loads.  Rows k >= nk of the phase table are zero, so a
// partial fold runs the same branch-free sweep and the extra terms add exact
// zeros.  Per sample the reference's update acc = acc + f*exp(i w n) with
// f = dt*F and separate multiplies and adds, the running sum in a register --
// the reference's rounding sequence.  (FP32 fields: k_dft_fold_mfma below.)
constexpr int kDftChunk = 16;

template <typename S, int KMAX>
__global__ __launch_bounds__(256) void k_dft_fold(BoxGeom g, const S* __restrict__ samp, int nk,
                                                  const double2* __restrict__ ph, int nof, double dt,
                                                  double2* __restrict__ emg) {
  static_assert(sizeof(S) == 16, "FP64 samples; FP32 fields fold with k_dft_fold_mfma");
  __shared__ double2 sph[KMAX * kDftChunk];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < g.ntot;
  const int64_t ic = valid ? i : 0;
  int j = 0;
  if (g.npts[1] && ic >= g.start[1]) j = 1;
  if (g.npts[2] && ic >= g.start[2]) j = 2;
  const int64_t np = g.npts[j];
  double2* acc_base = emg + g.start[j] * (int64_t)nof + (ic - g.start[j]);  // pair block: (np, nof)
  // every load is issued without a branch (slots past nk read slot 0 and are
  // zeroed), so the KMAX loads are in flight together
  S f[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) f[k] = samp[(int64_t)(k < nk ? k : 0) * g.ntot + ic];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) f[k] = (k < nk) ? make_double2(dt * f[k].x, dt * f[k].y) : make_double2(0.0, 0.0);
  auto mac = [&](double2 acc, int k, const double2 p) -> double2 { return cadd(acc, cmul(f[k], p)); };
  for (int c0 = 0; c0 < nof; c0 += kDftChunk) {
    const int cn = min(kDftChunk, nof - c0);
    __syncthreads();  // the previous chunk's phases are no longer read
    for (int t = threadIdx.x; t < KMAX * kDftChunk; t += blockDim.x) {
      const int k = t / kDftChunk, u = t % kDftChunk;
      sph[t] = (u < cn) ? ph[k * nof + c0 + u] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (!valid) continue;
    int u0 = 0;
    for (; u0 + 4 <= cn; u0 += 4) {
      double2 acc[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] = acc_base[(int64_t)(c0 + u0 + u) * np];
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u] = mac(acc[u], k, sph[k * kDftChunk + u0 + u]);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) acc_base[(int64_t)(c0 + u0 + u) * np] = acc[u];
    }
    for (; u0 < cn; ++u0) {
      double2 acc = acc_base[(int64_t)(c0 + u0) * np];
#pragma unroll
      for (int k = 0; k < KMAX; ++k) acc = mac(acc, k, sph[k * kDftChunk + u0]);
      acc_base[(int64_t)(c0 + u0) * np] = acc;
    }
  }
}

// The fold with FP32 samples as an FP64 matrix product on the matrix
// cores: per pair block, C[w][p] += sum_k P[k][w] * S[p][k] over the
// buffered samples, complex, written as two real products over the inner
// index k' = 2k + c (c = 0 real part of S_k, 1 imaginary part):
//   Re C += B_r A,  Im C += B_i A,  A[k'][p] = S_c[p][k],
//   B_r[w][k'] = (Re P, -Im P)[c],  B_i[w][k'] = (Im P, Re P)[c].
// A wave takes 16 point-components p (one pair block; columns past its end
// are zero) and all wavelengths in groups of 16.  v_mfma_f64_16x16x4f64
// lanes hold its first operand at [l % 16][l / 16], its second at
// [l / 16][l % 16] and rows 4r + l / 16 of column l % 16 of the result in
// register r: the phases are the first operand (staged in LDS per group),
// the samples the second (in registers for the whole sweep), and a result
// register of the 16 lanes of one row group is 16 consecutive points of one
// wavelength, so the accumulators move in contiguous 256-byte pieces.  P
// carries dt (k_dft_phase with scale dt); slots k >= nk add zero.
//
// The phases of up to GMAX wavelength groups are staged at once, so the
// workgroup passes a barrier once per GMAX groups, not twice per group, and
// its waves drift apart: each wave sweeps the staged groups starting at its
// own (wave mod groups), and loads the next group's accumulators before the
// matrix products of the current one.  While some waves wait on accumulator
// loads, others keep the matrix cores busy.
typedef double iris_d4 __attribute__((ext_vector_type(4)));

template <int KMAX, int WAVES, int GMAX>
__global__ __launch_bounds__(64 * WAVES) void k_dft_fold_mfma(BoxGeom g, const float2* __restrict__ samp, int nk,
                                                       const double2* __restrict__ ph, int nof,
                                                       double2* __restrict__ emg) {
  static_assert((KMAX & (KMAX - 1)) == 0, "KMAX is a power of two");
  __shared__ double2 sph[GMAX][KMAX][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;  // point (second operand column), inner index within a step
  // pair block j and the tile of 16 point-components in it
  int64_t tl = (int64_t)blockIdx.x * WAVES + wave;
  int j = -1;
  for (int q = 0; q < 3; ++q) {
    const int64_t tq = (g.npts[q] + 15) / 16;
    if (tl < tq) {
      j = q;
      break;
    }
    tl -= tq;
  }
  const bool active = j >= 0;
  const int64_t np = active ? g.npts[j] : 0;
  const int64_t p0 = tl * 16;
  const bool colok = active && p0 + col < np;
  const int64_t ic = colok ? g.start[j] + p0 + col : 0;
  double sv[KMAX / 2];
#pragma unroll
  for (int q = 0; q < KMAX / 2; ++q) {
    const int k = 2 * q + (grp >> 1);
    float v = 0.f;
    if (colok && k < nk) {
      const float2 f = samp[(int64_t)k * g.ntot + ic];
      v = (grp & 1) ? f.y : f.x;
    }
    sv[q] = (double)v;
  }
  double2* const cbase = active ? emg + g.start[j] * (int64_t)nof + p0 + col : emg;
  // accumulators of wavelength group c0 (register r: wavelength c0 + 4r + grp)
  auto load = [&](int c0, double2 v[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int w = c0 + 4 * r + grp;
      v[r] = (colok && w < nof) ? cbase[(int64_t)w * np] : make_double2(0.0, 0.0);
    }
  };
  const int ng = (nof + 15) / 16;
  for (int g0 = 0; g0 < ng; g0 += GMAX) {
    const int gn = min(GMAX, ng - g0);
    if (g0) __syncthreads();  // the previous block's phases are no longer read
    for (int t = threadIdx.x; t < gn * KMAX * 16; t += blockDim.x) {
      const int gi = t / (KMAX * 16), k = (t >> 4) & (KMAX - 1), u = t & 15;
      const int c = (g0 + gi) * 16 + u;
      sph[gi][k][u] = (c < nof) ? ph[k * nof + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (!active) continue;
    int gi = wave % gn;
    double2 nx[4];
    load((g0 + gi) * 16, nx);
    for (int s = 0; s < gn; ++s) {
      const int c0 = (g0 + gi) * 16;
      iris_d4 cr, ci;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cr[r] = nx[r].x;
        ci[r] = nx[r].y;
      }
      const int gnext = gi + 1 == gn ? 0 : gi + 1;
      if (s + 1 < gn) load((g0 + gnext) * 16, nx);
#pragma unroll
      for (int q = 0; q < KMAX / 2; ++q) {
        const double2 P = sph[gi][2 * q + (grp >> 1)][col];  // first operand row = wavelength c0 + col
        const double br = (grp & 1) ? -P.y : P.x;
        const double bi = (grp & 1) ? P.x : P.y;
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(br, sv[q], cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(bi, sv[q], ci, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int w = c0 + 4 * r + grp;
        if (colok && w < nof) cbase[(int64_t)w * np] = make_double2(cr[r], ci[r]);
      }
      gi = gnext;
    }
  }
}

// The same fold with every wavelength group staged at once (nof <= 16 GMAX)
// and persistent workgroups: a workgroup stages the phases once, then each
// of its waves takes point tiles wave + WAVES * (block + grid * i) on its
// own, with no barrier after the staging, so one wave's sample and
// accumulator loads overlap the other waves' matrix products across tiles as
// well as across groups.  Every tile's operations, and so its results, are
// those of k_dft_fold_mfma.
template <int KMAX, int WAVES, int GMAX>
__global__ __launch_bounds__(64 * WAVES) void k_dft_fold_mfma_p(BoxGeom g, const float2* __restrict__ samp, int nk,
                                                         const double2* __restrict__ ph, int nof,
                                                         double2* __restrict__ emg, int64_t ntiles) {
  static_assert((KMAX & (KMAX - 1)) == 0, "KMAX is a power of two");
  __shared__ double2 sph[GMAX][KMAX][16];
  // the wave index in a scalar register: the tile walk stays scalar
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & 15, grp = lane >> 4;
  const int ng = (nof + 15) / 16;  // <= GMAX (host check)
  for (int t = threadIdx.x; t < ng * KMAX * 16; t += blockDim.x) {
    const int gi = t / (KMAX * 16), k = (t >> 4) & (KMAX - 1), u = t & 15;
    const int c = gi * 16 + u;
    sph[gi][k][u] = (c < nof) ? ph[k * nof + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int64_t tile = (int64_t)blockIdx.x * WAVES + wave; tile < ntiles; tile += (int64_t)gridDim.x * WAVES) {
    // pair block j and the tile of 16 point-components in it
    int64_t tl = tile;
    int j = 0;
    for (; j < 2; ++j) {
      const int64_t tq = (g.npts[j] + 15) / 16;
      if (tl < tq) break;
      tl -= tq;
    }
    const int64_t np = g.npts[j];
    const int64_t p0 = tl * 16;
    const bool colok = p0 + col < np;
    const int64_t ic = colok ? g.start[j] + p0 + col : 0;
    float sv[KMAX / 2];
    // slots k = 2q + (grp >> 1), walked by a pointer (no per-slot offsets held)
    const float2* sp = samp + ic + (int64_t)(grp >> 1) * g.ntot;
    const int64_t sstep = 2 * g.ntot;
#pragma unroll
    for (int q = 0; q < KMAX / 2; ++q) {
      const int k = 2 * q + (grp >> 1);
      float v = 0.f;
      if (colok && k < nk) {
        const float2 f = *sp;
        v = (grp & 1) ? f.y : f.x;
      }
      sv[q] = v;
      sp += sstep;
    }
    double2* const cbase = emg + g.start[j] * (int64_t)nof + p0 + col;
    auto load = [&](int c0, double2 v[4]) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int w = c0 + 4 * r + grp;
        v[r] = (colok && w < nof) ? cbase[(int64_t)w * np] : make_double2(0.0, 0.0);
      }
    };
    int gi = wave % ng;
    double2 nx[4];
    load(gi * 16, nx);
    for (int s = 0; s < ng; ++s) {
      const int c0 = gi * 16;
      iris_d4 cr, ci;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cr[r] = nx[r].x;
        ci[r] = nx[r].y;
      }
      const int gnext = gi + 1 == ng ? 0 : gi + 1;
      if (s + 1 < ng) load(gnext * 16, nx);
#pragma unroll
      for (int q = 0; q < KMAX / 2; ++q) {
        const double2 P = sph[gi][2 * q + (grp >> 1)][col];
        const double br = (grp & 1) ? -P.y : P.x;
        const double bi = (grp & 1) ? P.x : P.y;
        float sf = sv[q];
        asm volatile("" : "+v"(sf));  // keeps the sample in FP32 between groups (converted here, per use)
        const double sq = (double)sf;
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(br, sq, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(bi, sq, ci, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int w = c0 + 4 * r + grp;
        if (colok && w < nof) cbase[(int64_t)w * np] = make_double2(cr[r], ci[r]);
      }
      gi = gnext;
    }
  }
}

// One field component on one plane: out[a + na*b] = f(x, y, z) with the
// fixed coordinate `index` on `axis` (1: x, 2: y, 3: z) and the two free
// coordinates a0 + a, b0 + b in (x, y, z) order.
template <typename T>
__global__ void k_plane(const cplx<T>* __restrict__ f, int axis, int index, int a0, int na, int b0, int nb,
                        int64_t sx, int64_t sxy, int xoff, cplx<T>* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)na * nb) return;
  const int ia = a0 + (int)(t % na), ib = b0 + (int)(t / na);
  int x, y, z;
  if (axis == 1) { x = index; y = ia; z = ib; }
  else if (axis == 2) { x = ia; y = index; z = ib; }
  else { x = ia; y = ib; z = index; }
  out[t] = f[(int64_t)(x + xoff) + sx * y + sxy * z];
}

// ---------------------------------------------------------------------------
// Field monitors (restating the reference's uncalled diagnostics,
// code/energy.f90:57-201): over x 1..nx, y 1..ny, z zi..zf
//   W_E = sum Re(eps_i conj(E_i^n) E_i^(n-1)),  W_H = sum |H_i^(n-1)|^2/dtcq2,
//   max |div D| with div D = sum_i eps_i E_i^n(c) - eps_i E_i^n(c - e_i),
//   max |div B| with div B = sum_i H_i^n(c + e_i) - H_i^n(c),
// eps_i being the inverse of the factor the E update applies (real part for a
// complex eps_inv).  x/y neighbours across the period take the Bloch factor.
// Grid-stride over the cells with a fixed grid; each block writes its
// partial (sum, sum, max, max) and k_diag_final reduces them in a fixed
// order, so the result does not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int kDiagBlocks = 1024;

template <typename T>
__device__ inline double eps_weight(const StepArgs<T>& a, int i, int x, int y, int z) {
  if (z < a.hi) return 1.0 / (double)a.a1[i];
  if (z > a.hf) return 1.0 / (double)a.a3[i];
  const int64_t idx = (int64_t)(x + a.xoff) + a.sx * y + a.sxy * (z - a.hi);
  const double er = (double)a.eps_re[idx];
  const double ei = a.eps_im ? (double)a.eps_im[idx] : 0.0;
  return er / (er * er + ei * ei);
}

template <typename T>
__global__ __launch_bounds__(256) void k_diag(StepArgs<T> a, const cplx<T>* __restrict__ ep,
                                              const cplx<T>* __restrict__ hp, int zi, int zf, double dtcq2,
                                              double* __restrict__ partial) {
  using C = cplx<T>;
  const int64_t n = (int64_t)a.nx * a.ny * (zf - zi + 1);
  double we = 0.0, wh = 0.0, dd = 0.0, db = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(t % a.nx) + 1;
    const int y = (int)((t / a.nx) % a.ny) + 1;
    const int z = (int)(t / ((int64_t)a.nx * a.ny)) + zi;
    const int64_t idx = (int64_t)(x + a.xoff) + a.sx * y + a.sxy * z;
    double dre = 0.0, dim = 0.0, bre = 0.0, bim = 0.0;
    for (int i = 0; i < 3; ++i) {
      const int64_t o = idx + i * a.cs;
      const double w = eps_weight(a, i, x, y, z);
      const C ec = a.e[o], epv = ep[o], hpv = hp[o], hc = a.h[o];
      we += w * ((double)ec.x * (double)epv.x + (double)ec.y * (double)epv.y);
      wh += ((double)hpv.x * (double)hpv.x + (double)hpv.y * (double)hpv.y) / dtcq2;
      int bx = x, by = y, bz = z, fx = x, fy = y, fz = z;
      C fb = mk<T>(1, 0), ff = mk<T>(1, 0);
      if (i == 0) {
        bx = x - 1; fx = x + 1;
        if (bx == 0) { bx = a.nx; fb = a.fxm; }
        if (fx == a.nx + 1) { fx = 1; ff = a.fxp; }
      } else if (i == 1) {
        by = y - 1; fy = y + 1;
        if (by == 0) { by = a.ny; fb = a.fym; }
        if (fy == a.ny + 1) { fy = 1; ff = a.fyp; }
      } else {
        bz = z - 1; fz = z + 1;
      }
      const C eb = cmul(fb, a.e[(int64_t)(bx + a.xoff) + a.sx * by + a.sxy * bz + i * a.cs]);
      const double wb = eps_weight(a, i, bx, by, bz);
      dre += w * (double)ec.x - wb * (double)eb.x;
      dim += w * (double)ec.y - wb * (double)eb.y;
      const C hf = cmul(ff, a.h[(int64_t)(fx + a.xoff) + a.sx * fy + a.sxy * fz + i * a.cs]);
      bre += (double)hf.x - (double)hc.x;
      bim += (double)hf.y - (double)hc.y;
    }
    dd = fmax(dd, hypot(dre, dim));
    db = fmax(db, hypot(bre, bim));
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = we;
  red[1][threadIdx.x] = wh;
  red[2][threadIdx.x] = dd;
  red[3][threadIdx.x] = db;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + w]);
      red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_diag_final(const double* __restrict__ partial, int nblk,
                                                    double* __restrict__ out) {
  __shared__ double red[4][256];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    v[0] += partial[b * 4 + 0];
    v[1] += partial[b * 4 + 1];
    v[2] = fmax(v[2], partial[b * 4 + 2]);
    v[3] = fmax(v[3], partial[b * 4 + 3]);
  }
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + w]);
      red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = red[threadIdx.x][0];
}

// Face fluxes, reference code/dataprocess.f90:241-271: block = (pair, iw,
// half); writes the unscaled sum with the reference's sign into
// out[iw*6 + face].  Each lane sums a strided share of the face's points;
// the 64 lanes of a wave reduce by shuffles (fixed butterfly order), the
// four wave sums in order through LDS, so the result does not depend on
// scheduling.
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
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = (red[0] + red[1]) + (red[2] + red[3]);
    out[iw * 6 + 2 * j + half] = half ? t : -t;
  }
}

}  // namespace iris

#include "iris_yee2.h"
