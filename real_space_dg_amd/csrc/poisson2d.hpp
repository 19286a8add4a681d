// Full-2D electrostatics on the device: the density of the transport
// contexts' solutions, its conservative projection onto the Poisson dual
// volumes (RTD/project_DG_density_RTD.m:1-69) and one FP64 Newton (or
// direct) update of the potential (RTD/solve_Poisson2D_rtd.m:85-130) by a
// banded LU of the 5-point operator.
//
// Orders: the host hands unique material nodes u = ix + nx*iy (x fastest,
// the reference's column-major [Nx, Ny]); the banded system runs in the
// y-fastest order k = iy + nu*ix, where every coupling of the 5-point
// stencil (x neighbours at +-nu, y neighbours at +-1, the periodic wrap at
// +-(nu-1)) lies within the bandwidth b = nu.
#pragma once

#include "kernels.hpp"

namespace wdg {

constexpr int kP2MaxWeights = 8;

struct P2W {
  int cnt;
  int idx[kP2MaxWeights];
  double w[kP2MaxWeights];
};

// out[i] (+)= scale * Re sum_j w_j rho[i*nrel + idx_j] over the centre DOFs
// (the zero-separation density of get_Observables_full2D.m:24-60).
__global__ void p2_density_kernel(const double2* __restrict__ rho, int ncen, int nrel, P2W w, double scale,
                                  int accumulate, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncen) return;
  double s = 0;
  for (int j = 0; j < w.cnt; ++j) s += w.w[j] * rho[(long)i * nrel + w.idx[j]].x;
  out[i] = (accumulate ? out[i] : 0.0) + scale * s;
}

// nmat[ix + nx*iu] = sum_{a,b} px[ix][a] ndg[a + xdof*b] py[iu][b] / vol,
// over the nonzero column range [pxa, pxb) of each projection row.
__global__ void p2_project_kernel(const double* __restrict__ ndg, int nx, int nu, int xdof, int ydof,
                                  const double* __restrict__ px, const int* __restrict__ pxr,
                                  const double* __restrict__ py, const int* __restrict__ pyr,
                                  const double* __restrict__ vol, double* __restrict__ nmat) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * nu) return;
  const int ix = t % nx, iu = t / nx;
  double s = 0;
  for (int b = pyr[2 * iu]; b < pyr[2 * iu + 1]; ++b) {
    const double wy = py[(long)iu * ydof + b];
    double r = 0;
    for (int a = pxr[2 * ix]; a < pxr[2 * ix + 1]; ++a) r += px[(long)ix * xdof + a] * ndg[a + (long)xdof * b];
    s += wy * r;
  }
  nmat[t] = s / vol[t];
}

struct P2Args {
  int nx, nu, n, b;
  const double* A;        // n x (2b+1) band of A, row k holds columns k-b..k+b
  const int* type;        // n (k order): 1 free, 2 Dirichlet
  const double* net;      // n (k order)
  const double* nmat;     // projected density (u order)
  const double* phi_in;   // u order
  const double* bc;       // u order
  double* J;              // n x (2b+1) work band
  double* v;              // n
  double* f;              // n: F, then the right-hand side, then delta
  double* chg;            // n
  double* phi_out;        // u order
  double* cand_out;       // u order
  double* info;           // 4
  double shift_coef;      // q^2 / kT (newton) or 0
  double alpha;
};

__device__ __forceinline__ int p2_u(int k, int nx, int nu) { return (k / nu) + nx * (k % nu); }

template <int NT>
__device__ __forceinline__ double p2_block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double m = red[0];
  for (int i = 1; i < NT / 64; ++i) m = fmax(m, red[i]);
  return m;
}

// One Poisson update in one workgroup (FP64 throughout).
template <int NT>
__global__ __launch_bounds__(NT) void p2_newton_kernel(P2Args a) {
  __shared__ double red[NT / 64];
  const int n = a.n, b = a.b, W = 2 * b + 1;
  const double q = 1.602176634e-19;
  // v = phi_in with the Dirichlet values; charge q (n - net)
  for (int k = threadIdx.x; k < n; k += NT) {
    const int u = p2_u(k, a.nx, a.nu);
    a.v[k] = a.type[k] == 2 ? a.bc[u] : a.phi_in[u];
    a.chg[k] = q * (a.nmat[u] - a.net[k]);
  }
  __syncthreads();
  // F = A v - charge; scale = max(|q n|, |q net|, |A v|) over free nodes, >= q
  double mx = q, mf = 0;
  for (int k = threadIdx.x; k < n; k += NT) {
    double av = 0;
    for (int d = 0; d < W; ++d) {
      const int c = k + d - b;
      if (c >= 0 && c < n) av += a.A[(long)k * W + d] * a.v[c];
    }
    const double F = av - a.chg[k];
    a.f[k] = F;
    if (a.type[k] == 1) {
      const int u = p2_u(k, a.nx, a.nu);
      mx = fmax(mx, fmax(fabs(q * a.nmat[u]), fmax(fabs(q * a.net[k]), fabs(av))));
      mf = fmax(mf, fabs(F));
    }
  }
  const double scale = p2_block_max<NT>(mx, red);
  const double fmaxv = p2_block_max<NT>(mf, red);
  // J = A - diag(shift) on free rows, identity on fixed rows; rhs = -F (free), 0
  for (long e = threadIdx.x; e < (long)n * W; e += NT) {
    const int k = (int)(e / W), d = (int)(e % W);
    double v;
    if (a.type[k] == 1) {
      v = a.A[e];
      if (d == b) v -= a.shift_coef * a.nmat[p2_u(k, a.nx, a.nu)];
    } else {
      v = (d == b) ? 1.0 : 0.0;
    }
    a.J[e] = v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += NT) a.f[k] = a.type[k] == 1 ? -a.f[k] : 0.0;
  __syncthreads();
  // banded LU without pivoting with the forward substitution folded in:
  // step k eliminates column k from rows k+1..k+b (each work item (i, j)
  // recomputes its multiplier J(i,k)/J(k,k); J(i,k) itself is not written)
  for (int k = 0; k < n; ++k) {
    const int nr = min(b, n - 1 - k);
    const double piv = a.J[(long)k * W + b];
    for (int t = threadIdx.x; t < nr * (nr + 1); t += NT) {
      const int i = k + 1 + t / (nr + 1), jj = t % (nr + 1);   // jj = nr: the right-hand side
      const double l = a.J[(long)i * W + (k - i + b)] / piv;
      if (jj < nr) {
        const int j = k + 1 + jj;
        a.J[(long)i * W + (j - i + b)] -= l * a.J[(long)k * W + (j - k + b)];
      } else {
        a.f[i] -= l * a.f[k];
      }
    }
    __syncthreads();
  }
  // back substitution: delta_k = (rhs_k - sum_{j>k} U(k,j) delta_j) / U(k,k),
  // then rows k-b..k-1 subtract their column-k term
  for (int k = n - 1; k >= 0; --k) {
    if (threadIdx.x == 0) a.f[k] /= a.J[(long)k * W + b];
    __syncthreads();
    const double dk = a.f[k];
    for (int i = k - 1 - (int)threadIdx.x; i >= max(0, k - b); i -= NT) a.f[i] -= a.J[(long)i * W + (k - i + b)] * dk;
    __syncthreads();
  }
  // linear residual J delta + F (free rows: A delta - shift delta), the
  // update and the physical residual after it
  double ml = 0, md = 0;
  for (int k = threadIdx.x; k < n; k += NT) {
    if (a.type[k] != 1) continue;
    double ad = 0;
    for (int d = 0; d < W; ++d) {
      const int c = k + d - b;
      if (c >= 0 && c < n && a.type[c] == 1) ad += a.A[(long)k * W + d] * a.f[c];
    }
    const int u = p2_u(k, a.nx, a.nu);
    ad -= a.shift_coef * a.nmat[u] * a.f[k];
    // F = -(rhs before the solve): recompute from v
    double av = 0;
    for (int d = 0; d < W; ++d) {
      const int c = k + d - b;
      if (c >= 0 && c < n) av += a.A[(long)k * W + d] * a.v[c];
    }
    ml = fmax(ml, fabs(ad + av - a.chg[k]));
    md = fmax(md, fabs(a.f[k]));
  }
  const double lin = p2_block_max<NT>(ml, red) / fmax(fmaxv, q);
  const double step = p2_block_max<NT>(md, red);
  for (int k = threadIdx.x; k < n; k += NT) {
    const int u = p2_u(k, a.nx, a.nu);
    const bool fr = a.type[k] == 1;
    const double dl = fr ? a.f[k] : 0.0;
    a.phi_out[u] = a.v[k] + a.alpha * dl;
    a.cand_out[u] = a.v[k] + dl;
  }
  __syncthreads();
  double ma = 0;
  for (int k = threadIdx.x; k < n; k += NT) {
    if (a.type[k] != 1) continue;
    double au = 0;
    for (int d = 0; d < W; ++d) {
      const int c = k + d - b;
      if (c >= 0 && c < n) au += a.A[(long)k * W + d] * a.phi_out[p2_u(c, a.nx, a.nu)];
    }
    ma = fmax(ma, fabs(au - a.chg[k]));
  }
  const double after = p2_block_max<NT>(ma, red);
  if (threadIdx.x == 0) {
    a.info[0] = fmaxv / scale;
    a.info[1] = after / scale;
    a.info[2] = lin;
    a.info[3] = step;
  }
}

}  // namespace wdg
