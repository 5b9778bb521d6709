// iris_incident.h -- incident field of an oblique plane wave, matched to the
// grid's Bloch condition (host code).
//
// The reference's TF/SF source (code/tfsf.f90:231-355) evaluates a Gaussian
// pulse travelling along u = (cos D, 0, sin D), D = RATIO_ANGLE*pi/2, at
// every (ix, iz, it).  The reference itself only runs RATIO_ANGLE = 1
// (code/files.f90:267).  Off normal incidence, every frequency of that
// pulse has its own transverse wavenumber k(w)*n*sin(theta), while a grid
// with a Bloch phase ak across its x period holds only kx = ak/(nx*Q): away
// from the carrier the injected field does not satisfy the boundary, and
// the TF/SF plane leaks.
//
// Here the incident field is instead the solution of the grid's own update
// for a mode exp(i*phi*ix), phi = ak/nx per cell, in the medium of the
// source plane: with x-neighbours E(x+1) = e^{i phi} E(x), H(x-1) =
// e^{-i phi} H(x) and no y-dependence, the Yee step (code/fdtd.f90:57-248)
// becomes a 1D update in z over three components (i_pol = 2: Ex, Ez, Hy;
// i_pol = 1: Ey, Hx, Hz).  A soft source below the TF/SF plane drives it
// with the reference pulse's time signal (carrier, width); split-field PML
// layers (Berenger: only the z-derivative part of each component is damped)
// absorb both ends.  Each frequency then has kx fixed by the Bloch phase, so
// it travels at its own angle, sin(theta(w)) = kx / (k(w) n), as a periodic
// structure's incidence at fixed kx does.
//
// The table holds, for every step it = 0..nt, the three source components
// at the two TF/SF planes (iz_min - 1, iz_min) at ix = 0 after step it's E
// and H updates -- the time level at which the reference evaluates its
// pulse for the corrections that follow those updates
// (code/driver.f90:413-442, code/tfsf.f90:76-153).  With that order a
// discrete wave of the grid passes the corrections exactly: the scattered-
// field side stays empty.  At ix the value is the table's times
// e^{i phi ix}.  The peak of the tangential E component at iz_min is scaled
// to the reference pulse's (Epump*Q*uz for i_pol = 2, Epump*Q for i_pol =
// 1, times its window factor).
//
// Cut-off.  At fixed kx a frequency below w_c = c kx / n has no propagating
// wave (it is evanescent), and frequencies just above it travel nearly
// along x.  The reference pulse is about one cycle long, so much of its
// energy lies there; injected, it excites the bound (guided) modes of a
// film, which ring without decay, and the finite DFT window leaks their
// ringing into the longest wavelengths.  The drive is therefore the pulse
// with its spectrum raised-cosine filtered: zero below 1.1 w_c, whole above
// 1.4 w_c (by FFT over the run zero-padded to four times its length, so
// that the filter's tails do not wrap around into it).  T and R are ratios per wavelength, so the
// shaping leaves them unchanged where the drive has power; wavelengths above
// lambda_c / 1.1 (lambda_c = 2 pi / kx) get none.
#pragma once
#include <cmath>
#include <complex>
#include <utility>
#include <vector>

#include "iris.h"

namespace iris {

using zcplx = std::complex<double>;

// Whether the engine takes the matched incident field for these parameters:
// a TF/SF pulse off normal incidence.
inline bool incident_matched(const iris_params& p) { return p.em_source && p.ratio_angle != 1.0; }

// Reasons the matched field does not apply (nullptr when it does).
inline const char* incident_unsupported(const iris_params& p) {
  if (p.plane_wave != 1) return "the matched oblique source needs the Gaussian pulse (plane_wave = 1)";
  if (p.iz_min > p.hi) return "the matched oblique source needs the TF/SF plane below the structure";
  if (p.ak[0][1] != 0.0 || p.ak[1][0] != 0.0 || p.ak[1][1] != 0.0)
    return "the matched oblique source takes a real Bloch phase in x only";
  return nullptr;
}

// In-place radix-2 FFT (a.size() a power of two); inverse without the 1/n.
inline void fft_pow2(std::vector<zcplx>& a, bool inverse) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = 2.0 * M_PI / (double)len * (inverse ? 1.0 : -1.0);
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const zcplx w = std::polar(1.0, ang * (double)k);
        const zcplx u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

// The drive g(it), it = 0..nt, with its spectrum raised-cosine filtered to
// zero below wlo and whole above whi (radians per step, by |w|).
inline std::vector<zcplx> filtered_drive(int nt, const std::vector<zcplx>& g, double wlo, double whi) {
  size_t n = 1;
  while (n < 4 * (size_t)(nt + 1)) n <<= 1;
  std::vector<zcplx> a(n, zcplx(0.0, 0.0));
  for (int it = 0; it <= nt; ++it) a[it] = g[it];
  fft_pow2(a, false);
  for (size_t k = 0; k < n; ++k) {
    const double w = std::abs(2.0 * M_PI * (double)(k <= n / 2 ? (double)k : (double)k - (double)n) / (double)n);
    double f = (w - wlo) / (whi - wlo);
    f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
    a[k] *= (0.5 - 0.5 * std::cos(M_PI * f)) / (double)n;
  }
  fft_pow2(a, true);
  return std::vector<zcplx>(a.begin(), a.begin() + nt + 1);
}

// Table of (nt + 1) x 2 planes x 3 components (see above).
inline std::vector<zcplx> incident_table(const iris_params& p) {
  const int nt = p.nt;
  const double eps_s = p.eps[0];
  const double a = p.eps_in[0];                                   // region I E-update factor
  const double d = ((p.dt * p.c0) / p.q) * ((p.dt * p.c0) / p.q);  // dtcq2
  const double phi = p.ak[0][0] / p.nx;
  const zcplx ip = std::exp(zcplx(0.0, phi)), im = std::exp(zcplx(0.0, -phi));
  const zcplx fx_b = 1.0 - im;  // H(x) - H(x-1) = (1 - e^{-i phi}) H(x)
  const zcplx fx_f = ip - 1.0;  // E(x+1) - E(x) = (e^{i phi} - 1) E(x)
  // grid: PML | buffer | source | S cells | iz_min-1, iz_min | buffer | PML
  const int L = 80, B = 8, S = 12;
  const int ks = L + B, k0 = ks + S;  // k0: plane iz_min - 1
  const int K = k0 + 2 + B + L;
  // damping per step at position zz (cells; E at k, H at k + 1/2)
  const double smax = 0.25;
  auto damp = [&](double zz) {
    double r = 0.0;
    if (zz < L) r = (L - zz) / L;
    else if (zz > K - 1 - L) r = (zz - (K - 1 - L)) / L;
    return smax * r * r * r;
  };
  std::vector<double> caE(K), cbE(K), caH(K), cbH(K);
  for (int k = 0; k < K; ++k) {
    const double se = damp(k), sh = damp(k + 0.5);
    caE[k] = (1.0 - se / 2) / (1.0 + se / 2);
    cbE[k] = 1.0 / (1.0 + se / 2);
    caH[k] = (1.0 - sh / 2) / (1.0 + sh / 2);
    cbH[k] = 1.0 / (1.0 + sh / 2);
  }
  // the reference pulse's time signal at (ix = 0, iz_min) (code/tfsf.f90:264-350)
  const double sq = std::sqrt(eps_s);
  const double ko = (2.0 * p.pi) / (p.wlpump * p.f_abohr);
  const double dk = p.gpump * p.q;
  const double A0 = (p.q * (0.0 - p.ixo)) * p.ux + (p.q * (p.iz_min - p.zopump)) * p.uz;
  const double tau = dk * sq / (p.c0 * p.dt);  // envelope width in steps
  const double cos_t = std::sqrt(std::max(0.0, 1.0 - std::pow(phi / (p.q * ko * sq), 2)));
  const double vz = (p.c0 * p.dt / p.q) / sq * std::max(cos_t, 0.2);  // cells per step along z
  // the drive peaks so that the pulse passes iz_min when the reference
  // pulse's envelope peaks there
  const double it_src = std::max(A0 * sq / (p.c0 * p.dt) - (S + 1) / vz, 0.0);
  const double w0 = ko * p.c0 * p.dt;  // carrier phase per step
  std::vector<zcplx> g(nt + 1, zcplx(0.0, 0.0));
  for (int it = 1; it <= nt; ++it) {
    const double t = (it - it_src) / tau;
    const double env = std::exp(-t * t);
    g[it] = p.cosine == 1 ? zcplx(env * std::cos(w0 * it), 0.0) : env * std::exp(zcplx(0.0, -w0 * it));
  }
  const double wc = (std::abs(phi) / p.q) * (p.c0 * p.dt) / sq;  // cut-off, radians per step
  if (wc > 0.0) {
    g = filtered_drive(nt, g, 1.1 * wc, 1.4 * wc);
    if (p.cosine == 1)
      for (zcplx& v : g) v = zcplx(v.real(), 0.0);
  }
  auto drive = [&](int it) { return g[it]; };
  std::vector<zcplx> tab((size_t)(nt + 1) * 6, zcplx(0.0, 0.0));
  // the three components per polarisation (below); a component with both a
  // z-derivative and an x-term is split, the x-term part in the *x array
  std::vector<zcplx> e1(K), e1x(K), f2(K), f3(K), f3x(K);
  const bool tm = p.i_pol == 2;
  for (int it = 1; it <= nt; ++it) {
    if (tm) {
      // i_pol = 2: e1 = Ex, f2 = Ez, f3 + f3x = Hy (z part, x part)
      for (int k = K - 1; k >= 1; --k) {
        const zcplx hy = f3[k] + f3x[k], hym = f3[k - 1] + f3x[k - 1];
        e1[k] = caE[k] * e1[k] - cbE[k] * (a * (hy - hym));
      }
      for (int k = 0; k < K; ++k) f2[k] += a * (fx_b * (f3[k] + f3x[k]));
      e1[ks] += drive(it);
      for (int k = 0; k < K - 1; ++k) {
        f3[k] = caH[k] * f3[k] - cbH[k] * (d * (e1[k + 1] - e1[k]));
        f3x[k] += d * (fx_f * f2[k]);
      }
      for (int pl = 0; pl < 2; ++pl) {
        const int k = k0 + pl;
        zcplx* o = &tab[((size_t)it * 2 + pl) * 3];
        o[0] = e1[k];
        o[1] = f2[k];
        o[2] = f3[k] + f3x[k];
      }
    } else {
      // i_pol = 1: e1 + e1x = Ey (z part, x part), f2 = Hx, f3 = Hz
      for (int k = K - 1; k >= 1; --k) {
        e1[k] = caE[k] * e1[k] + cbE[k] * (a * (f2[k] - f2[k - 1]));
        e1x[k] -= a * (fx_b * f3[k]);
      }
      e1x[ks] += drive(it);
      for (int k = 0; k < K - 1; ++k) {
        const zcplx ey = e1[k] + e1x[k], eyp = e1[k + 1] + e1x[k + 1];
        f2[k] = caH[k] * f2[k] + cbH[k] * (d * (eyp - ey));
      }
      for (int k = 0; k < K; ++k) f3[k] -= d * (fx_f * (e1[k] + e1x[k]));
      for (int pl = 0; pl < 2; ++pl) {
        const int k = k0 + pl;
        zcplx* o = &tab[((size_t)it * 2 + pl) * 3];
        o[0] = e1[k] + e1x[k];
        o[1] = f2[k];
        o[2] = f3[k];
      }
    }
  }
  // scale the peak tangential E at iz_min to the reference pulse's
  double peak = 0.0;
  for (int it = 1; it <= nt; ++it) peak = std::max(peak, std::abs(tab[((size_t)it * 2 + 1) * 3]));
  const double target = p.eamp * p.q * (tm ? p.uz : 1.0) * p.win;
  const double scale = peak > 0.0 ? target / peak : 0.0;
  for (zcplx& v : tab) v *= scale;
  return tab;
}

}  // namespace iris
