// Contact-mode reservoirs and the full-2D inflow right-hand side on the
// device, from host-solved transverse contact modes.
//
// For contact s with modes phi_m(y) (energy E_m, masses m_x,m, m_z,m) and
// Fermi level E_F (get_ContactModeReservoirRho_full2D.m:35-137):
//   f_m(kx)   = 2 deg/pi * trapz_kz 1/(1 + exp(q (E_m + hb^2 kx^2/2m_x + hb^2 kz^2/2m_z - E_F)/kT))
//               (or 2 deg/(1 + exp(...)) without the k_z continuum; the
//               exponent clipped to +-80; longitudinalOccupation :406-424)
//   F_m(rx)   = sum_kx cos(rx kx) f_m(kx) dkx / (2 pi)
//   rho(ry,rx; Y) = sum_m phi_m(Y + ry/2) conj(phi_m(Y - ry/2)) F_m(rx)      (contact rows only)
// and the inflow rhs of the X faces (get_Diff_full2D.m:393-452, the rhs
// parts of addSourceBoundary / addDrainBoundary):
//   b[(Y, a_face, ry, rx')] = sign * s_x * lift_a * sum_rx A(rx', rx) rho(ry, rx; Y)
// with A = A+ (source, sign +1, first X element) or A- (drain, sign -1,
// last X element).  Vector order (rho_x, rho_y, X dof, Y dof), rho_x fastest.
#pragma once

#include "kernels.hpp"

namespace wdg {

// physicalConstants of get_ContactModeReservoirRho_full2D.m:808-813
constexpr double kResHbar = 1.054571817e-34;
constexpr double kResQe = 1.602176634e-19;
constexpr double kResKb = 1.38064852e-23;
constexpr int kResMaxModes = 128;

enum { RES_MODES_BOX = 0, RES_MODES_FOURIER = 1 };

// One contact's modes on the device.
struct ResContact {
  int n_modes, kind, n_grid;
  const double* grid;     // BOX: ascending Y grid incl. the two zero ghost nodes; FOURIER: wave numbers
  const double* vals;     // BOX: mode m at grid point i in vals[m * n_grid + i]
  const double2* coef;    // FOURIER: coefficient i of mode m in coef[m * n_grid + i]
  double y0, inv_sqrt_w;  // FOURIER: phase origin and 1/sqrt(cell width)
  const double* energy;   // [eV]
  const double* mass_x;   // [kg]
  const double* mass_z;   // [kg]
  double ef;              // [eV]
  const int* mask;        // contact Y dofs
  const double2* achar;   // n_rho_x^2, row-major
  int face;               // 0: first X element (source), 1: last (drain)
  double sign;
};

struct ResGrid {
  int nkx, nkz, nrx, nry, nyd, nxd, nl;
  const double* kx;       // nkx
  const double* kz;       // nkz (nkz == 0: no k_z integration)
  const double* rx;       // rho_x cell centres
  const double* ry;       // rho_y cell centres
  const double* ynodes;   // Y dof coordinates
  const double* lift;     // nl values of the face lift (low face: lift_low, high: lift_high) per contact
  double dkx, temp, deg, sx;
};

// F_m(rho_x) of every (contact, mode): one workgroup per (contact, mode);
// f_m(kx) goes through LDS.  Dynamic LDS: nkx doubles.
__global__ __launch_bounds__(256) void res_profile_kernel(ResContact c0, ResContact c1, ResGrid g,
                                                          double* __restrict__ prof) {
  extern __shared__ double res_f[];
  const int s = blockIdx.y;
  const ResContact& c = s ? c1 : c0;
  const int m = blockIdx.x;
  if (m >= c.n_modes) return;
  const double kt = kResKb * g.temp;
  const double em = c.energy[m], mx = c.mass_x[m], mz = c.mass_z[m];
  for (int i = threadIdx.x; i < g.nkx; i += blockDim.x) {
    const double k = g.kx[i];
    const double ekx = kResHbar * kResHbar * k * k / (2.0 * mx * kResQe);
    double f;
    if (g.nkz > 0) {
      double acc = 0.0, prev = 0.0;
      for (int q = 0; q < g.nkz; ++q) {
        const double kz = g.kz[q];
        const double ekz = kResHbar * kResHbar * kz * kz / (2.0 * mz * kResQe);
        const double arg = fmin(80.0, fmax(-80.0, kResQe * (em + ekx + ekz - c.ef) / kt));
        const double v = 1.0 / (1.0 + exp(arg));
        if (q > 0) acc += 0.5 * (prev + v) * (kz - g.kz[q - 1]);
        prev = v;
      }
      f = 2.0 * g.deg * acc / M_PI;
    } else {
      const double arg = fmin(80.0, fmax(-80.0, kResQe * (em + ekx - c.ef) / kt));
      f = 2.0 * g.deg / (1.0 + exp(arg));
    }
    res_f[i] = f;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < g.nrx; r += blockDim.x) {
    const double x = g.rx[r];
    double acc = 0.0;
    for (int i = 0; i < g.nkx; ++i) acc += cos(x * g.kx[i]) * res_f[i];
    prof[((long)s * kResMaxModes + m) * g.nrx + r] = acc * g.dkx / (2.0 * M_PI);
  }
}

// phi_m(y) of a contact: BOX by linear interpolation with zero outside the
// grid (interp1(..., 'linear', 0)), FOURIER as sum_i c_i e^{i k_i (y - y0)} / sqrt(W).
__device__ __forceinline__ double2 res_mode_value(const ResContact& c, int m, double y) {
  if (c.kind == RES_MODES_FOURIER) {
    double re = 0.0, im = 0.0;
    for (int i = 0; i < c.n_grid; ++i) {
      double sn, cs;
      sincos(c.grid[i] * (y - c.y0), &sn, &cs);
      const double2 a = c.coef[(long)m * c.n_grid + i];
      re += a.x * cs - a.y * sn;
      im += a.x * sn + a.y * cs;
    }
    return mk(re * c.inv_sqrt_w, im * c.inv_sqrt_w);
  }
  const double* gy = c.grid;
  const int n = c.n_grid;
  if (!(y >= gy[0] && y <= gy[n - 1])) return mk(0.0, 0.0);
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (gy[mid] <= y) lo = mid;
    else hi = mid;
  }
  const double* v = c.vals + (long)m * n;
  const double t = (y - gy[lo]) / (gy[hi] - gy[lo]);
  return mk(v[lo] + t * (v[hi] - v[lo]), 0.0);
}

// One workgroup per (Y dof, rho_y cell, contact), one thread per rho_x cell:
// the reservoir row rho(ry, :; Y) (also written to res_out when given), then
// its characteristic projection lifted into the face element of b.
// Dynamic LDS: n_modes complex mode products + n_rho_x complex.
__global__ __launch_bounds__(256) void res_rhs_kernel(ResContact c0, ResContact c1, ResGrid g,
                                                      const double* __restrict__ prof, const double* __restrict__ lift0,
                                                      const double* __restrict__ lift1, double2* __restrict__ b,
                                                      double2* __restrict__ res0, double2* __restrict__ res1) {
  extern __shared__ double2 res_lds[];
  const int iy = blockIdx.x, ry = blockIdx.y, s = blockIdx.z;
  const ResContact& c = s ? c1 : c0;
  double2* res_out = s ? res1 : res0;
  const int nrx = g.nrx, nrel = g.nrx * g.nry;
  const bool on = c.mask[iy] != 0 && c.n_modes > 0;
  double2* pm = res_lds;                 // n_modes mode products
  double2* row = res_lds + kResMaxModes; // the reservoir row
  if (on) {
    const double y = g.ynodes[iy], half = 0.5 * g.ry[ry];
    for (int m = threadIdx.x; m < c.n_modes; m += blockDim.x) {
      const double2 p = res_mode_value(c, m, y + half), q = res_mode_value(c, m, y - half);
      pm[m] = mk(p.x * q.x + p.y * q.y, p.y * q.x - p.x * q.y);   // p conj(q)
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nrx; r += blockDim.x) {
    double re = 0.0, im = 0.0;
    if (on)
      for (int m = 0; m < c.n_modes; ++m) {
        const double f = prof[((long)s * kResMaxModes + m) * nrx + r];
        re += pm[m].x * f;
        im += pm[m].y * f;
      }
    row[r] = mk(re, im);
    if (res_out) res_out[(long)iy * nrel + (long)ry * nrx + r] = mk(re, im);
  }
  __syncthreads();
  if (!on) return;   // the rows of b were cleared before the launch
  const double* lift = s ? lift1 : lift0;
  for (int r = threadIdx.x; r < nrx; r += blockDim.x) {
    const double2* arow = c.achar + (long)r * nrx;
    double re = 0.0, im = 0.0;
    for (int q = 0; q < nrx; ++q) {
      const double2 a = arow[q], v = row[q];
      re += a.x * v.x - a.y * v.y;
      im += a.x * v.y + a.y * v.x;
    }
    for (int a = 0; a < g.nl; ++a) {
      const int col = c.face ? g.nxd - g.nl + a : a;
      const double w = c.sign * g.sx * lift[a];
      b[((long)iy * g.nxd + col) * nrel + (long)ry * nrx + r] = mk(w * re, w * im);
    }
  }
}

// Block partial sums of |b|^2 (one double per block).
__global__ __launch_bounds__(256) void norm2_partials_kernel(const double2* __restrict__ b, long N,
                                                             double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    s += b[i].x * b[i].x + b[i].y * b[i].y;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void fill_zero_kernel(double2* __restrict__ x, long N) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    x[i] = mk(0.0, 0.0);
}

}  // namespace wdg
