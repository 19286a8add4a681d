// Full-2D Wigner operator F(rho_x, rho_y, X, Y): device kernels.
//
// Vector order (DG_code/initParams_full2D.m:4-8,105-107): rho_x fastest, then
// rho_y, the X DG DOF and the Y DG DOF, i.e. DOF (i, r) at i*nrel + r with
// centre DOF i = ix + nxd*iy and relative cell r = rx + nrx*ry.  An operator
// part c*kron(C, R) (get_Diff_full2D.m:80-146) acts as Y = c R U C^T on
// U = reshape(u, nrel, ncen).
//
// Kernels:
//  f2_apply_kernel   one workgroup per (centre element, band of rho_y rows):
//                    central (-i D) and identity (Rusanov) transport parts of
//                    both axes, the FV-consistent potential
//                    (1/16) P^T diag(DeltaV) P projected over the element's
//                    tensor quadrature (get_Drift_full2D.m:331-364) or the
//                    collocated diagonal, and the CAP.  Writes y, or
//                    sa u + sb y for a shifted operator.
//                    With MODE_DIAG / MODE_ROWABS the same tiling produces the
//                    drift diagonal and row-absolute bound instead
//                    (get_Drift_full2D.m:466-523).
//  f2_dense_kernel   a part whose relative factor is a dense matrix along one
//                    axis (characteristic |A|, A+ R, A- R): banded centre
//                    factor (interior upwind) or a face lift.  Adds into y.
//  f2_kron_diag_kernel, f2_observables_kernel, f2_scale_kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"

namespace wdg {

constexpr int kF2MaxNl = 4;    // DG nodes per element axis (p <= 3)
constexpr int kF2MaxNq = 8;    // potential quadrature points per element axis
constexpr int kF2Block = 256;

enum F2Mode { F2_APPLY = 0, F2_DIAG = 1, F2_ROWABS = 2 };

struct F2Const {
  int nrx, nry, nlx, nly, nex, ney, nxd, nyd, nrel;
  int rows, tiles;              // rho_y rows per apply tile, tiles per element
  int pot_mode;                 // 0 fv-consistent, 1 collocated
  int nqx, nqy, nxg, nyg;       // quadrature points, material grid size
  int npx, npy;                 // relative sample points per axis (vertices or cells)
  int v_lds;                    // 1: the material grid is staged in LDS
  int periodic_y;               // 1: rho_y, Y elements and the rho_y vertex map wrap around
  // stored potential differences DeltaV(element, q = qx*nqy + qy, vertex), vertex
  // x fastest (storedPotentialDifference, get_Drift_full2D.m:609-618); null: on the fly
  const double* dv64;
  const float* dv32;
  double inv2dx, inv2dy;        // 1/(2 delta) of the central difference
  double pot_coef, cap_coef;    // drift coefficients / operator scale (pot multiplies i*DeltaV)
  // banded centre parts (coefficient and 1/s_A folded in); null when absent
  const double *bxc, *bxi, *byc, *byi;
  // potential samples: per (centre sample, relative sample, sign) the material
  // cell and its linear weight along one axis
  const int *tix, *tiy;
  const double *ttx, *tty;
  const double* v;              // material grid potential, x fastest
  const double *capx, *capy;
  double ipx[kF2MaxNq * kF2MaxNl], ipy[kF2MaxNq * kF2MaxNl];   // interpolation (q, b)
  double prx[kF2MaxNq * kF2MaxNl], pry[kF2MaxNq * kF2MaxNl];   // projection (q, a)
};

template <typename T>
struct F2Vec;
template <>
struct F2Vec<double> { using T2 = double2; };
template <>
struct F2Vec<float> { using T2 = float2; };

template <typename T2>
__device__ __forceinline__ T2 f2_zero() { T2 z; z.x = 0; z.y = 0; return z; }

template <typename T, typename T2>
__device__ __forceinline__ void f2_fma(T2& acc, T s, const T2& v) {
  acc.x += s * v.x;
  acc.y += s * v.y;
}

// Bilinear interpolation of the material-grid potential between cells
// (ix, iy) and (ix+1, iy+1) with weights tx, ty (griddedInterpolant 'linear').
__device__ __forceinline__ double f2_bilin(const double* v, int nxg, int ix, double tx, int iy, double ty) {
  const double* p0 = v + (size_t)iy * nxg + ix;
  const double* p1 = p0 + nxg;
  const double a = p0[0] + tx * (p0[1] - p0[0]);
  const double b = p1[0] + tx * (p1[1] - p1[0]);
  return a + ty * (b - a);
}

// DeltaV = V(x+, y+) - V(x-, y-) for table rows kx (x samples) and ky.
__device__ __forceinline__ double f2_delta_v(const F2Const& c, int kx, int ky) {
  const double vp = f2_bilin(c.v, c.nxg, c.tix[2 * kx], c.ttx[2 * kx], c.tiy[2 * ky], c.tty[2 * ky]);
  const double vm = f2_bilin(c.v, c.nxg, c.tix[2 * kx + 1], c.ttx[2 * kx + 1], c.tiy[2 * ky + 1], c.tty[2 * ky + 1]);
  return vp - vm;
}

// The same from LDS-staged sample tables (index pairs and weights for +/-)
// and the LDS or global material grid.
__device__ __forceinline__ double f2_bilin_s(const double* __restrict__ vl, const double* __restrict__ vg, bool lds,
                                             int nxg, int ix, double tx, int iy, double ty) {
  const int i0 = iy * nxg + ix, i1 = i0 + nxg;
  const double v00 = lds ? vl[i0] : vg[i0], v10 = lds ? vl[i0 + 1] : vg[i0 + 1];
  const double v01 = lds ? vl[i1] : vg[i1], v11 = lds ? vl[i1 + 1] : vg[i1 + 1];
  const double a = v00 + tx * (v10 - v00);
  const double b = v01 + tx * (v11 - v01);
  return a + ty * (b - a);
}

// Neighbour value with zero padding outside [0, n).
template <typename T2>
__device__ __forceinline__ T2 f2_ld(const T2* p, long off, bool ok) {
  return ok ? p[off] : f2_zero<T2>();
}

// LDS layout of the apply kernel: su[NL][rows+2][nrx] (own element, one halo
// row each side), then for the FV-consistent potential sz[NL][rows+1][nrx+1]
// (vertex values), the element's x sample table (nqx*npx*2 cells, weights),
// the tile's y sample table (nqy*(rows+1)*2) and, when small enough, the
// material grid.  Offsets are 16-byte aligned.
struct F2Lds {
  size_t sz, xi, xt, yi, yt, v, total;
};
__host__ __device__ inline size_t f2_al(size_t b) { return (b + 15) & ~(size_t)15; }
__host__ __device__ inline F2Lds f2_lds_layout(int nl, int rows, int nrx, size_t t2, bool fv, bool tables, int nqx,
                                                int nqy, int npx, int nv) {
  F2Lds L{};
  size_t o = f2_al((size_t)nl * (rows + 2) * nrx * t2);
  if (fv) {
    L.sz = o;
    o = f2_al(o + (size_t)nl * (rows + 1) * (nrx + 1) * t2);
  }
  if (tables) {
    L.xi = o;
    o = f2_al(o + (size_t)nqx * npx * 2 * sizeof(int));
    L.xt = o;
    o = f2_al(o + (size_t)nqx * npx * 2 * sizeof(double));
    L.yi = o;
    o = f2_al(o + (size_t)nqy * (rows + 1) * 2 * sizeof(int));
    L.yt = o;
    o = f2_al(o + (size_t)nqy * (rows + 1) * 2 * sizeof(double));
  }
  L.v = o;
  o = f2_al(o + (size_t)nv * sizeof(double));
  L.total = o;
  return L;
}


// Resident waves per SIMD asked of the compiler: the FP32 apply with at most
// 4 quadrature points per axis fits 6 waves (79 VGPRs, no spills); elsewhere
// the compiler's own choice.
template <typename T, int MODE, int NQ>
struct F2ApplyOcc {
  static constexpr int value = (std::is_same<T, float>::value && MODE == F2_APPLY && NQ <= 4) ? 6 : 1;
};

// NQ: compile-time bound on the potential quadrature points per axis (nqx,
// nqy <= NQ <= kF2MaxNq); it sizes the per-vertex DeltaV registers.
template <int NLX, int NLY, typename T, int MODE, int NQ = kF2MaxNq>
__global__ void __launch_bounds__(kF2Block) __attribute__((amdgpu_waves_per_eu(F2ApplyOcc<T, MODE, NQ>::value)))
f2_apply_kernel(const F2Const* __restrict__ cp,
                                                            const typename F2Vec<T>::T2* __restrict__ u,
                                                            typename F2Vec<T>::T2* __restrict__ y,
                                                            const GmresState* st, double sa, double sb) {
  using T2 = typename F2Vec<T>::T2;
  constexpr int NL = NLX * NLY;
  if (skip(st)) return;
  const F2Const& c = *cp;
  const int nrx = c.nrx, nry = c.nry, nrel = c.nrel;
  const int el = blockIdx.x / c.tiles, tile = blockIdx.x - el * c.tiles;
  const int ex = el % c.nex, ey = el / c.nex;
  const int ry0 = tile * c.rows;
  const int nrows = min(c.rows, nry - ry0);
  const int R2 = c.rows + 2;
  const bool fv = c.pot_mode == 0;
  extern __shared__ __align__(16) unsigned char f2_smem[];
  const bool tab = std::is_same<T, double>::value ? c.dv64 != nullptr : c.dv32 != nullptr;
  const bool vl = c.v_lds != 0 && !tab;
  const F2Lds L = f2_lds_layout(NL, c.rows, nrx, sizeof(T2), fv, fv && !tab, c.nqx, c.nqy, c.npx, vl ? c.nxg * c.nyg : 0);
  T2* su = reinterpret_cast<T2*>(f2_smem);
  T2* sz = reinterpret_cast<T2*>(f2_smem + L.sz);
  int* sxi = reinterpret_cast<int*>(f2_smem + L.xi);
  double* sxt = reinterpret_cast<double*>(f2_smem + L.xt);
  int* syi = reinterpret_cast<int*>(f2_smem + L.yi);
  double* syt = reinterpret_cast<double*>(f2_smem + L.yt);
  double* sv = reinterpret_cast<double*>(f2_smem + L.v);
  auto cdof = [&](int ax, int ay) { return (long)(ex * NLX + ax) + (long)c.nxd * (ey * NLY + ay); };
  if (vl)
    for (int i = threadIdx.x; i < c.nxg * c.nyg; i += blockDim.x) sv[i] = c.v[i];
  if (fv && !tab) {
    const int nxt = c.nqx * c.npx * 2;
    const size_t xo = (size_t)ex * nxt;
    for (int i = threadIdx.x; i < nxt; i += blockDim.x) {
      sxi[i] = c.tix[xo + i];
      sxt[i] = c.ttx[xo + i];
    }
    const int nyr = (nrows + 1) * 2;
    for (int i = threadIdx.x; i < c.nqy * nyr; i += blockDim.x) {
      const int qy = i / nyr, rem = i - qy * nyr;
      const size_t g = ((size_t)(ey * c.nqy + qy) * c.npy + ry0) * 2 + rem;
      syi[qy * (c.rows + 1) * 2 + rem] = c.tiy[g];
      syt[qy * (c.rows + 1) * 2 + rem] = c.tty[g];
    }
  }
  if (MODE != F2_APPLY) __syncthreads();

  // thread -> (relative x cell tx, tile row ty); rows*nrx <= blockDim
  const int tx = threadIdx.x % nrx, ty = threadIdx.x / nrx;
  const bool active = ty < c.rows;
  if (MODE == F2_APPLY) {
    if (active) {
#pragma unroll
      for (int a = 0; a < NL; ++a) {
        const T2* src = u + cdof(a % NLX, a / NLX) * nrel + tx;
        T2* dst = su + (size_t)a * R2 * nrx + tx;
        for (int rr = ty; rr < R2; rr += c.rows) {
          int ry = ry0 - 1 + rr;
          if (c.periodic_y) ry = ry < 0 ? ry + nry : (ry >= nry ? ry - nry : ry);
          const bool ok = ry >= 0 && ry < nry && rr <= nrows + 1;
          dst[(size_t)rr * nrx] = ok ? src[(long)ry * nrx] : f2_zero<T2>();
        }
      }
    }
    __syncthreads();
  }

  // ---- vertex phase of the FV-consistent potential ----
  if (fv) {
    const int nvx = nrx + 1;
    const int nvert = (nrows + 1) * nvx;
    for (int v = threadIdx.x; v < nvert; v += blockDim.x) {
      const int vyl = v / nvx, vx = v - vyl * nvx;
      const int vy = ry0 + vyl;
      T2 W[NL], Z[NL];
      T deg = 0;
      if (MODE == F2_APPLY) {
        // P: vertex (vx, vy) sums the cells (vx-1..vx) x (vy-1..vy)
        const bool xl = vx >= 1, xr = vx < nrx;
#pragma unroll
        for (int b = 0; b < NL; ++b) {
          const T2* s = su + (size_t)b * R2 * nrx;
          T2 w = f2_zero<T2>();
#pragma unroll
          for (int dr = 0; dr < 2; ++dr) {
            const T2* row = s + (size_t)(vyl + dr) * nrx;   // su rows hold zeros outside the domain
            if (xl) { w.x += row[vx - 1].x; w.y += row[vx - 1].y; }
            if (xr) { w.x += row[vx].x; w.y += row[vx].y; }
          }
          W[b] = w;
        }
      } else {
        const int cxn = (vx >= 1) + (vx < nrx), cyn = c.periodic_y ? 2 : (vy >= 1) + (vy < nry);
        deg = (T)(cxn * cyn);
      }
#pragma unroll
      for (int a = 0; a < NL; ++a) Z[a] = f2_zero<T2>();
      // stored DeltaV: one quadrature row (nqy values) in flight ahead of its use
      const size_t qs = (size_t)(nry + 1) * (nrx + 1);
      const size_t tb = (size_t)el * c.nqx * c.nqy * qs + (size_t)vy * (nrx + 1) + vx;
      T dcur[NQ], dnext[NQ];
      if (tab) {
#pragma unroll
        for (int qy = 0; qy < NQ; ++qy)
          if (qy < c.nqy) dcur[qy] = std::is_same<T, double>::value ? (T)c.dv64[tb + qy * qs] : (T)c.dv32[tb + qy * qs];
      }
      for (int qx = 0; qx < c.nqx; ++qx) {
        const int kx = 2 * (qx * c.npx + vx);
        const int ixp = tab ? 0 : sxi[kx], ixm = tab ? 0 : sxi[kx + 1];
        const double txp = tab ? 0.0 : sxt[kx], txm = tab ? 0.0 : sxt[kx + 1];
        if (tab && qx + 1 < c.nqx) {
          const size_t o = tb + (size_t)(qx + 1) * c.nqy * qs;
#pragma unroll
          for (int qy = 0; qy < NQ; ++qy)
            if (qy < c.nqy) dnext[qy] = std::is_same<T, double>::value ? (T)c.dv64[o + qy * qs] : (T)c.dv32[o + qy * qs];
        }
        T2 t1[NLY];
        if (MODE == F2_APPLY) {
#pragma unroll
          for (int by = 0; by < NLY; ++by) {
            T2 s = f2_zero<T2>();
#pragma unroll
            for (int bx = 0; bx < NLX; ++bx) f2_fma(s, (T)c.ipx[qx * kF2MaxNl + bx], W[bx + NLX * by]);
            t1[by] = s;
          }
        }
        T2 acc[NLY];
#pragma unroll
        for (int ay = 0; ay < NLY; ++ay) acc[ay] = f2_zero<T2>();
#pragma unroll
        for (int qy = 0; qy < NQ; ++qy) {
          if (qy >= c.nqy) break;
          T dv;
          if (tab) {
            dv = dcur[qy];
          } else {
            const int ky = 2 * (qy * (c.rows + 1) + vyl);
            const double vp = f2_bilin_s(sv, c.v, vl, c.nxg, ixp, txp, syi[ky], syt[ky]);
            const double vm = f2_bilin_s(sv, c.v, vl, c.nxg, ixm, txm, syi[ky + 1], syt[ky + 1]);
            dv = (T)(vp - vm);
          }
          if (MODE == F2_APPLY) {
            T2 wq = f2_zero<T2>();
#pragma unroll
            for (int by = 0; by < NLY; ++by) f2_fma(wq, (T)c.ipy[qy * kF2MaxNl + by], t1[by]);
            wq.x *= dv;
            wq.y *= dv;
#pragma unroll
            for (int ay = 0; ay < NLY; ++ay) f2_fma(acc[ay], (T)c.pry[qy * kF2MaxNl + ay], wq);
          } else if (MODE == F2_DIAG) {
#pragma unroll
            for (int ay = 0; ay < NLY; ++ay)
              acc[ay].x += dv * (T)(c.pry[qy * kF2MaxNl + ay] * c.ipy[qy * kF2MaxNl + ay]);
          } else {
            T sy = 0;
#pragma unroll
            for (int by = 0; by < NLY; ++by) sy += fabs((T)c.ipy[qy * kF2MaxNl + by]);
#pragma unroll
            for (int ay = 0; ay < NLY; ++ay) acc[ay].x += fabs(dv) * fabs((T)c.pry[qy * kF2MaxNl + ay]) * sy;
          }
        }
        T wx[NLX];
        if (MODE == F2_APPLY) {
#pragma unroll
          for (int ax = 0; ax < NLX; ++ax) wx[ax] = (T)c.prx[qx * kF2MaxNl + ax];
        } else if (MODE == F2_DIAG) {
#pragma unroll
          for (int ax = 0; ax < NLX; ++ax) wx[ax] = (T)(c.prx[qx * kF2MaxNl + ax] * c.ipx[qx * kF2MaxNl + ax]);
        } else {
          T sx = 0;
#pragma unroll
          for (int bx = 0; bx < NLX; ++bx) sx += fabs((T)c.ipx[qx * kF2MaxNl + bx]);
#pragma unroll
          for (int ax = 0; ax < NLX; ++ax) wx[ax] = fabs((T)c.prx[qx * kF2MaxNl + ax]) * sx;
        }
#pragma unroll
        for (int ay = 0; ay < NLY; ++ay)
#pragma unroll
          for (int ax = 0; ax < NLX; ++ax) f2_fma(Z[ax + NLX * ay], wx[ax], acc[ay]);
        if (tab) {
#pragma unroll
          for (int qy = 0; qy < NQ; ++qy) dcur[qy] = dnext[qy];
        }
      }
      if (MODE == F2_ROWABS) {
#pragma unroll
        for (int a = 0; a < NL; ++a) Z[a].x *= deg;
      }
#pragma unroll
      for (int a = 0; a < NL; ++a) sz[((size_t)a * (c.rows + 1) + vyl) * nvx + vx] = Z[a];
    }
    __syncthreads();
  }

  // ---- cell phase ----
  const T pc = (T)c.pot_coef, cc = (T)c.cap_coef;
  if (active && ty < nrows) {
    const int ryl = ty, rx = tx;
    const int ry = ry0 + ryl;
    const long r = (long)ry * nrx + rx;
    const T cap = (T)(c.capx[rx] + c.capy[ry]);
    T2 out[NL];
    // drift
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      const int ax = a % NLX, ay = a / NLX;
      T2 s = f2_zero<T2>();
      if (fv) {
        const int nvx = nrx + 1;
        const T2* zb = sz + (size_t)a * (c.rows + 1) * nvx;
        const T2 z00 = zb[(size_t)ryl * nvx + rx], z01 = zb[(size_t)ryl * nvx + rx + 1];
        const T2 z10 = zb[(size_t)(ryl + 1) * nvx + rx], z11 = zb[(size_t)(ryl + 1) * nvx + rx + 1];
        s.x = z00.x + z01.x + z10.x + z11.x;
        s.y = z00.y + z01.y + z10.y + z11.y;
        const T f = pc * (T)0.0625;
        if (MODE == F2_APPLY) {
          // i * pot * (1/16) P^T Z  +  cap * U
          const T2 uu = su[((size_t)a * R2 + ryl + 1) * nrx + rx];
          out[a].x = -f * s.y + cc * cap * uu.x;
          out[a].y = f * s.x + cc * cap * uu.y;
        } else if (MODE == F2_DIAG) {
          out[a].x = cc * cap;
          out[a].y = f * s.x;
        } else {
          out[a].x = fabs(f) * s.x + fabs(cc) * fabs(cap);
          out[a].y = 0;
        }
      } else {
        // collocated: i * pot * DeltaV(node, cell) + cap, diagonal
        const int kx = (ex * NLX + ax) * c.npx + rx;
        const int ky = (ey * NLY + ay) * c.npy + ry;
        const double vp = f2_bilin_s(sv, c.v, vl, c.nxg, c.tix[2 * kx], c.ttx[2 * kx], c.tiy[2 * ky], c.tty[2 * ky]);
        const double vm = f2_bilin_s(sv, c.v, vl, c.nxg, c.tix[2 * kx + 1], c.ttx[2 * kx + 1], c.tiy[2 * ky + 1],
                                     c.tty[2 * ky + 1]);
        const T dv = (T)(vp - vm);
        const T dre = cc * cap, dim = pc * dv;
        if (MODE == F2_APPLY) {
          const T2 uu = su[((size_t)a * R2 + ryl + 1) * nrx + rx];
          out[a].x = dre * uu.x - dim * uu.y;
          out[a].y = dre * uu.y + dim * uu.x;
        } else if (MODE == F2_DIAG) {
          out[a].x = dre;
          out[a].y = dim;
        } else {
          out[a].x = sqrt(dre * dre + dim * dim);
          out[a].y = 0;
        }
      }
    }
    if (MODE == F2_APPLY) {
      // ---- transport: X parts along rho_x, Y parts along rho_y ----
      const bool rxl = rx > 0, rxr = rx < nrx - 1;
      const bool pY = c.periodic_y != 0;
      const bool ryd = ry > 0 || pY, ryu = ry < nry - 1 || pY;
      const long offd = ry > 0 ? -(long)nrx : (long)(nry - 1) * nrx;    // rho_y - 1 (wrapped when periodic)
      const long offu = ry < nry - 1 ? (long)nrx : -(long)(nry - 1) * nrx;
      const int sX = NLX * NLX + 2 * NLX, sY = NLY * NLY + 2 * NLY;
#pragma unroll
      for (int ay = 0; ay < NLY; ++ay) {
        // neighbour X face nodes on this Y row: last node of ex-1, first of ex+1
        T2 lm = f2_zero<T2>(), l0 = lm, lp = lm, rm = lm, r0 = lm, rp = lm;
        const long row_off = r;
        if (ex > 0) {
          const T2* p = u + ((long)(ex * NLX - 1) + (long)c.nxd * (ey * NLY + ay)) * nrel + row_off;
          lm = f2_ld(p, -1, rxl); l0 = p[0]; lp = f2_ld(p, 1, rxr);
        }
        if (ex < c.nex - 1) {
          const T2* p = u + ((long)((ex + 1) * NLX) + (long)c.nxd * (ey * NLY + ay)) * nrel + row_off;
          rm = f2_ld(p, -1, rxl); r0 = p[0]; rp = f2_ld(p, 1, rxr);
        }
#pragma unroll
        for (int ax = 0; ax < NLX; ++ax) {
          T2 dif = f2_zero<T2>(), idn = f2_zero<T2>();
          if (c.bxc) {
            const double* B = c.bxc + (size_t)ex * sX;
#pragma unroll
            for (int bx = 0; bx < NLX; ++bx) {
              const T2* s = su + ((size_t)(bx + NLX * ay) * R2 + ryl + 1) * nrx;
              const T w = (T)B[ax * NLX + bx];
              const T2 up = rxr ? s[rx + 1] : f2_zero<T2>();
              const T2 dn = rxl ? s[rx - 1] : f2_zero<T2>();
              dif.x += w * (up.x - dn.x);
              dif.y += w * (up.y - dn.y);
            }
            const T wl = (T)B[NLX * NLX + ax], wr = (T)B[NLX * NLX + NLX + ax];
            dif.x += wl * (lp.x - lm.x) + wr * (rp.x - rm.x);
            dif.y += wl * (lp.y - lm.y) + wr * (rp.y - rm.y);
          }
          if (c.bxi) {
            const double* B = c.bxi + (size_t)ex * sX;
#pragma unroll
            for (int bx = 0; bx < NLX; ++bx)
              f2_fma(idn, (T)B[ax * NLX + bx], su[((size_t)(bx + NLX * ay) * R2 + ryl + 1) * nrx + rx]);
            f2_fma(idn, (T)B[NLX * NLX + ax], l0);
            f2_fma(idn, (T)B[NLX * NLX + NLX + ax], r0);
          }
          // -i/(2 delta) * dif
          const T h = (T)c.inv2dx;
          T2& o = out[ax + NLX * ay];
          o.x += h * dif.y + idn.x;
          o.y += -h * dif.x + idn.y;
        }
      }
#pragma unroll
      for (int ax = 0; ax < NLX; ++ax) {
        T2 dm = f2_zero<T2>(), d0 = dm, dp = dm, tm = dm, t0 = dm, tp = dm;
        if (ey > 0 || pY) {
          const int eb = ey > 0 ? ey - 1 : c.ney - 1;
          const T2* p = u + ((long)(ex * NLX + ax) + (long)c.nxd * (eb * NLY + NLY - 1)) * nrel + r;
          dm = f2_ld(p, offd, ryd); d0 = p[0]; dp = f2_ld(p, offu, ryu);
        }
        if (ey < c.ney - 1 || pY) {
          const int et = ey < c.ney - 1 ? ey + 1 : 0;
          const T2* p = u + ((long)(ex * NLX + ax) + (long)c.nxd * (et * NLY)) * nrel + r;
          tm = f2_ld(p, offd, ryd); t0 = p[0]; tp = f2_ld(p, offu, ryu);
        }
#pragma unroll
        for (int ay = 0; ay < NLY; ++ay) {
          T2 dif = f2_zero<T2>(), idn = f2_zero<T2>();
          if (c.byc) {
            const double* B = c.byc + (size_t)ey * sY;
#pragma unroll
            for (int by = 0; by < NLY; ++by) {
              const T2* s = su + (size_t)(ax + NLX * by) * R2 * nrx;
              const T w = (T)B[ay * NLY + by];
              const T2 up = s[(size_t)(ryl + 2) * nrx + rx];   // zero rows outside the domain
              const T2 dn = s[(size_t)ryl * nrx + rx];
              dif.x += w * (up.x - dn.x);
              dif.y += w * (up.y - dn.y);
            }
            const T wl = (T)B[NLY * NLY + ay], wr = (T)B[NLY * NLY + NLY + ay];
            dif.x += wl * (dp.x - dm.x) + wr * (tp.x - tm.x);
            dif.y += wl * (dp.y - dm.y) + wr * (tp.y - tm.y);
          }
          if (c.byi) {
            const double* B = c.byi + (size_t)ey * sY;
#pragma unroll
            for (int by = 0; by < NLY; ++by)
              f2_fma(idn, (T)B[ay * NLY + by], su[((size_t)(ax + NLX * by) * R2 + ryl + 1) * nrx + rx]);
            f2_fma(idn, (T)B[NLY * NLY + ay], d0);
            f2_fma(idn, (T)B[NLY * NLY + NLY + ay], t0);
          }
          const T h = (T)c.inv2dy;
          T2& o = out[ax + NLX * ay];
          o.x += h * dif.y + idn.x;
          o.y += -h * dif.x + idn.y;
        }
      }
    }
    // (sa I + sb A) u of a shifted operator (Crank-Nicolson): u from the tile
    const bool shifted = MODE == F2_APPLY && (sa != 0.0 || sb != 1.0);
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      const long i = cdof(a % NLX, a / NLX);
      if (MODE == F2_APPLY) {
        T2 o = out[a];
        if (shifted) {
          const T2 uu = su[((size_t)a * R2 + ryl + 1) * nrx + rx];
          o.x = (T)sa * uu.x + (T)sb * o.x;
          o.y = (T)sa * uu.y + (T)sb * o.y;
        }
        y[i * nrel + r] = o;
      } else {
        // diagonal / row sums accumulate onto the transport part
        T2 o = y[i * nrel + r];
        o.x += out[a].x;
        o.y += out[a].y;
        y[i * nrel + r] = o;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Dense relative factor along one axis.
// ---------------------------------------------------------------------------
struct F2Dense {
  int axis;          // 0: C along X, R along rho_x; 1: along Y / rho_y
  int face;          // -1: banded centre factor; 0 / 1: face lift at the low / high end;
                     // -2: plain y = R u along the axis (overwrites y)
  int nl, ne;        // nodes per element and elements of the centre axis
  int chunk;         // relative cells across the axis per workgroup (gridDim.y chunks)
  int periodic;      // banded: the end elements couple across the seam
  double coef;       // part coefficient / operator scale
  const double* blocks;   // banded: per element diag nl*nl, lower nl, upper nl
  const double* lift;     // face: nl
  const int* mask;        // face: per transverse centre DOF (null = all)
  const double2* R;       // n x n row-major, n = n_rho along the axis
  int rglobal;            // 1: R is read from global memory (L2-resident) instead of LDS
};

// LDS: R (n x (n+1), padded, unless it is read from global memory) and the
// input block (n along the axis x chunk across).
inline size_t f2_dense_lds(int n, int chunk, size_t t2, bool rglobal = false) {
  return ((rglobal ? (size_t)0 : (size_t)n * (n + 1)) + (size_t)n * chunk) * t2;
}

// Workgroup (line, chunk): line = a centre DOF (banded, plain) or a
// transverse DOF of the face (lift); chunk = a range of relative cells across
// the axis (rho_y for axis 0, rho_x for axis 1), which the dense factor does
// not couple.  W = C-combination of u in LDS, then out(m, c) = sum_k R(m, k)
// W(k, c), 4 outputs per thread sharing one R (or W) load.
template <typename T, bool RG>
__global__ void __launch_bounds__(kF2Block) f2_dense_kernel(const F2Const* __restrict__ cp, F2Dense d,
                                                            const typename F2Vec<T>::T2* __restrict__ u,
                                                            typename F2Vec<T>::T2* __restrict__ y,
                                                            const GmresState* st) {
  using T2 = typename F2Vec<T>::T2;
  if (skip(st)) return;
  const F2Const& c = *cp;
  const int nrx = c.nrx, nry = c.nry, nrel = c.nrel;
  const int n = d.axis == 0 ? nrx : nry;          // along the axis
  const int ncol = d.axis == 0 ? nry : nrx;       // across
  const int c0 = blockIdx.y * d.chunk;
  const int cn = min(d.chunk, ncol - c0);
  if (cn <= 0) return;
  const int line = blockIdx.x;
  extern __shared__ __align__(16) unsigned char f2_smem[];
  T2* sR = reinterpret_cast<T2*>(f2_smem);          // n x (n+1) (RG: not staged)
  T2* sW = RG ? sR : sR + (size_t)n * (n + 1);      // [k][c - c0], n x chunk
  // R(m, k) scaled by the part coefficient, from LDS or (RG) from global memory
  auto rval = [&](int m, int k) -> T2 {
    if constexpr (RG) {
      const double2 v = d.R[(size_t)m * n + k];
      T2 w;
      w.x = (T)(d.coef * v.x);
      w.y = (T)(d.coef * v.y);
      return w;
    } else {
      return sR[(size_t)m * (n + 1) + k];
    }
  };
  const int nxd = c.nxd;
  auto dof = [&](int k, int t) -> long { return d.axis == 0 ? (long)k + (long)nxd * t : (long)t + (long)nxd * k; };
  // relative index of (k along, cc across)
  auto rel = [&](int k, int cc) -> long { return d.axis == 0 ? (long)cc * nrx + k : (long)k * nrx + cc; };
  const int nalong = d.nl * d.ne;
  const bool plain = d.face == -2;
  int k_out = 0, t = 0;
  if (d.face < 0) {
    k_out = d.axis == 0 ? line % nxd : line / nxd;
    t = d.axis == 0 ? line / nxd : line % nxd;
  } else {
    t = line;
    if (d.mask && !d.mask[t]) return;
  }
  if constexpr (!RG) {
    // eight loads in flight per thread before their conversions and stores
    const int nn = n * n, bs = blockDim.x;
    for (int i0 = threadIdx.x; i0 < nn; i0 += 8 * bs) {
      double2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (i0 + q * bs < nn) ? d.R[i0 + q * bs] : make_double2(0.0, 0.0);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int idx = i0 + q * bs;
        if (idx < nn) {
          const int m = idx / n, k = idx - m * n;
          T2 w;
          w.x = (T)(d.coef * v[q].x);
          w.y = (T)(d.coef * v[q].y);
          sR[(size_t)m * (n + 1) + k] = w;
        }
      }
    }
  }
  const int nw = n * cn;
  // staging order: the index that is contiguous in memory runs fastest
  // (along the axis for axis 0, across it for axis 1)
  auto widx = [&](int idx, int& k, int& cc) {
    if (d.axis == 0) {
      k = idx % n;
      cc = idx / n;
    } else {
      cc = idx % cn;
      k = idx / cn;
    }
  };
  if (plain) {
    for (int idx = threadIdx.x; idx < nw; idx += blockDim.x) {
      int k, cc;
      widx(idx, k, cc);
      sW[(size_t)k * d.chunk + cc] = u[(long)line * nrel + rel(k, c0 + cc)];
    }
  } else if (d.face < 0) {
    const int e = k_out / d.nl, a = k_out - e * d.nl;
    const int sb = d.nl * d.nl + 2 * d.nl;
    const double* B = d.blocks + (size_t)e * sb;
    for (int idx = threadIdx.x; idx < nw; idx += blockDim.x) {
      int k, cc;
      widx(idx, k, cc);
      const long r = rel(k, c0 + cc);
      T2 w = f2_zero<T2>();
      for (int b = 0; b < d.nl; ++b) f2_fma(w, (T)B[a * d.nl + b], u[dof(e * d.nl + b, t) * nrel + r]);
      if (e > 0 || d.periodic)
        f2_fma(w, (T)B[d.nl * d.nl + a], u[dof((e > 0 ? e : d.ne) * d.nl - 1, t) * nrel + r]);
      if (e < d.ne - 1 || d.periodic)
        f2_fma(w, (T)B[d.nl * d.nl + d.nl + a], u[dof(e < d.ne - 1 ? (e + 1) * d.nl : 0, t) * nrel + r]);
      sW[(size_t)k * d.chunk + cc] = w;
    }
  } else {
    const int ktr = d.face == 0 ? 0 : nalong - 1;
    const long base = dof(ktr, t) * nrel;
    for (int idx = threadIdx.x; idx < nw; idx += blockDim.x) {
      int k, cc;
      widx(idx, k, cc);
      sW[(size_t)k * d.chunk + cc] = u[base + rel(k, c0 + cc)];
    }
  }
  __syncthreads();
  // items: m fastest (axis 0: m = rx is contiguous in memory) with 4 c per
  // thread, or c fastest (axis 1: c = rx) with 4 m per thread
  const int nb = d.axis == 0 ? (cn + 3) / 4 : (n + 3) / 4;
  const int items = d.axis == 0 ? n * nb : cn * nb;
  for (int it = threadIdx.x; it < items; it += blockDim.x) {
    T2 acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = f2_zero<T2>();
    int mm[4], cc[4];
    if (d.axis == 0) {
      const int m = it % n, cb = it / n;
#pragma unroll
      for (int j = 0; j < 4; ++j) { mm[j] = m; cc[j] = cb * 4 + j; }
      for (int k = 0; k < n; ++k) {
        const T2 rv = rval(m, k);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (cc[j] < cn) {
            const T2 w = sW[(size_t)k * d.chunk + cc[j]];
            acc[j].x += rv.x * w.x - rv.y * w.y;
            acc[j].y += rv.x * w.y + rv.y * w.x;
          }
      }
    } else {
      const int c1 = it % cn, mb = it / cn;
#pragma unroll
      for (int j = 0; j < 4; ++j) { mm[j] = mb * 4 + j; cc[j] = c1; }
      for (int k = 0; k < n; ++k) {
        const T2 w = sW[(size_t)k * d.chunk + c1];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (mm[j] < n) {
            const T2 rv = rval(mm[j], k);
            acc[j].x += rv.x * w.x - rv.y * w.y;
            acc[j].y += rv.x * w.y + rv.y * w.x;
          }
      }
    }
    // y (+)= the outputs.  Every (output, node) address belongs to this
    // thread alone, so all the loads of the read-modify-writes are issued
    // before the stores (the compiler cannot prove that they do not alias)
    bool ok[4];
    long r[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ok[j] = cc[j] < cn && mm[j] < n;
      r[j] = ok[j] ? rel(mm[j], c0 + cc[j]) : 0;
    }
    if (plain) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (ok[j]) y[(long)line * nrel + r[j]] = acc[j];
    } else if (d.face < 0) {
      const long ob = dof(k_out, t) * nrel;
      T2 v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = ok[j] ? y[ob + r[j]] : f2_zero<T2>();
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (ok[j]) {
          v[j].x += acc[j].x;
          v[j].y += acc[j].y;
          y[ob + r[j]] = v[j];
        }
    } else {
      T2 v[kF2MaxNl][4];
      long ob[kF2MaxNl];
      T l[kF2MaxNl];
#pragma unroll
      for (int a = 0; a < kF2MaxNl; ++a) {
        const int k = d.face == 0 ? a : nalong - d.nl + a;
        ob[a] = a < d.nl ? dof(k, t) * nrel : 0;
        l[a] = a < d.nl ? (T)d.lift[a] : (T)0;
#pragma unroll
        for (int j = 0; j < 4; ++j) v[a][j] = (a < d.nl && ok[j]) ? y[ob[a] + r[j]] : f2_zero<T2>();
      }
#pragma unroll
      for (int a = 0; a < kF2MaxNl; ++a)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (a < d.nl && ok[j]) {
            v[a][j].x += l[a] * acc[j].x;
            v[a][j].y += l[a] * acc[j].y;
            y[ob[a] + r[j]] = v[a][j];
          }
    }
  }
}

// Transport diagonal / row-abs bound of the Kronecker parts:
// sum_p coef_p * dC_p[i] * dR_p[r] (complex) into y (re, im), and the row
// bound sum_p |coef_p| aC_p[i] aR_p[r] into the real part when rowabs.
__global__ void f2_kron_diag_kernel(int nparts, int ncen, int nrel, const double* __restrict__ dc,
                                    const double2* __restrict__ dr, double2* __restrict__ y) {
  const long n = (long)ncen * nrel;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n; idx += (long)gridDim.x * blockDim.x) {
    const long i = idx / nrel, r = idx - i * nrel;
    double re = 0, im = 0;
    for (int p = 0; p < nparts; ++p) {
      const double cv = dc[(size_t)p * ncen + i];
      const double2 rv = dr[(size_t)p * nrel + r];
      re += cv * rv.x;
      im += cv * rv.y;
    }
    y[idx] = make_double2(re, im);
  }
}

// Observables at zero separation: per centre DOF, three sparse sums over the
// relative block (n: value at 0; d/drho_x; d/drho_y), get_Observables_full2D.m:24-60.
struct F2ObsW {
  int idx[24];
  double w[24];
  int cnt[3];
};

__global__ void f2_observables_kernel(const double2* __restrict__ rho, int ncen, int nrel, F2ObsW ow, double scale,
                                      double2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncen) return;
  const double2* p = rho + (size_t)i * nrel;
  for (int k = 0; k < 3; ++k) {
    double re = 0, im = 0;
    for (int j = 0; j < ow.cnt[k]; ++j) {
      const double2 v = p[ow.idx[8 * k + j]];
      re += ow.w[8 * k + j] * v.x;
      im += ow.w[8 * k + j] * v.y;
    }
    out[(size_t)k * ncen + i] = make_double2(scale * re, scale * im);
  }
}

// z = p o r (complex elementwise): Jacobi / row-scaling preconditioner.
template <typename T>
__global__ void f2_scale_kernel(const typename F2Vec<T>::T2* __restrict__ p, const typename F2Vec<T>::T2* __restrict__ r,
                                typename F2Vec<T>::T2* __restrict__ z, long n, const GmresState* st) {
  if (skip(st)) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const auto a = p[i], b = r[i];
    typename F2Vec<T>::T2 o;
    o.x = a.x * b.x - a.y * b.y;
    o.y = a.x * b.y + a.y * b.x;
    z[i] = o;
  }
}

// z = p r with a real scale p (row-sum scaling): 12 (FP32) or 24 (FP64)
// bytes per entry instead of 16 or 32.
template <typename T>
__global__ void f2_scale_real_kernel(const T* __restrict__ p, const typename F2Vec<T>::T2* __restrict__ r,
                                     typename F2Vec<T>::T2* __restrict__ z, long n, const GmresState* st) {
  if (skip(st)) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const T a = p[i];
    const auto b = r[i];
    typename F2Vec<T>::T2 o;
    o.x = a * b.x;
    o.y = a * b.y;
    z[i] = o;
  }
}

// The scaling preconditioners of the FP32 solver applied to a 16-bit Krylov
// basis vector read in place: z = p r with a real scale (REAL) or z = p o r
// with a complex one.  Two entries per thread and step: one 8-byte basis
// packet (FP16 pairs stored times 256, or BF16 pairs), the two scales, one
// 16-byte store.
template <typename VP>
__device__ __forceinline__ void f2_basis_pair(VP p, float2& a, float2& b);
template <>
__device__ __forceinline__ void f2_basis_pair<ushort4>(ushort4 p, float2& a, float2& b) {
  const float s = 1.0f / 256.0f;
  a = make_float2((float)__builtin_bit_cast(_Float16, p.x) * s, (float)__builtin_bit_cast(_Float16, p.y) * s);
  b = make_float2((float)__builtin_bit_cast(_Float16, p.z) * s, (float)__builtin_bit_cast(_Float16, p.w) * s);
}
template <>
__device__ __forceinline__ void f2_basis_pair<uint2>(uint2 p, float2& a, float2& b) {
  a = lv_unpack_bf16x2(p.x);
  b = lv_unpack_bf16x2(p.y);
}

template <typename VP, bool REAL>
__global__ void f2_scale_basis_kernel(const void* __restrict__ scale, const VP* __restrict__ v,
                                      float4* __restrict__ z, long npairs, const GmresState* st) {
  if (skip(st)) return;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < npairs; i += (long)gridDim.x * blockDim.x) {
    float2 a, b;
    f2_basis_pair<VP>(v[i], a, b);
    float4 o;
    if constexpr (REAL) {
      const float2 p = reinterpret_cast<const float2*>(scale)[i];
      o = make_float4(p.x * a.x, p.x * a.y, p.y * b.x, p.y * b.y);
    } else {
      const float4 p = reinterpret_cast<const float4*>(scale)[i];
      o = make_float4(p.x * a.x - p.y * a.y, p.x * a.y + p.y * a.x, p.z * b.x - p.w * b.y, p.z * b.y + p.w * b.x);
    }
    z[i] = o;
  }
}

// max |d| and the number of entries with |d| < floor (preconditioner floors).
__global__ void f2_absmax_kernel(const double2* __restrict__ d, long n, double* __restrict__ part) {
  __shared__ double sm[256];
  double m = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    m = fmax(m, sqrt(d[i].x * d[i].x + d[i].y * d[i].y));
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Builds the preconditioner vector from the (scaled) diagonal or row sums:
// jacobi (kind 1): p = 1/d, identity action (p = s_A) where |d| < floor;
// rowabs (kind 2): p = 1/s, identity action where s < floor, stored real
// (the first n doubles / floats of p64 / p32).  Counts the floored entries.
__global__ void f2_precond_build_kernel(const double2* __restrict__ d, long n, int kind, double floor_dev,
                                        double ident, double2* __restrict__ p64, float2* __restrict__ p32,
                                        unsigned long long* __restrict__ nsmall) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double2 v = d[i];
    double2 o;
    if (kind == 1) {
      const double m2 = v.x * v.x + v.y * v.y;
      if (sqrt(m2) < floor_dev) {
        o = make_double2(ident, 0.0);
        atomicAdd(nsmall, 1ull);
      } else {
        o = make_double2(v.x / m2, -v.y / m2);
      }
    } else {
      if (v.x < floor_dev) {
        o = make_double2(ident, 0.0);
        atomicAdd(nsmall, 1ull);
      } else {
        o = make_double2(1.0 / v.x, 0.0);
      }
    }
    if (kind == 1) {
      p64[i] = o;
      p32[i] = make_float2((float)o.x, (float)o.y);
    } else {
      reinterpret_cast<double*>(p64)[i] = o.x;
      reinterpret_cast<float*>(p32)[i] = (float)o.x;
    }
  }
}

// DeltaV(element, q, vertex) for every element of the FV-consistent potential,
// once per potential (FP64 and FP32 copies).
__global__ void f2_dv_table_kernel(const F2Const* __restrict__ cp, double* __restrict__ t64, float* __restrict__ t32) {
  const F2Const& c = *cp;
  const long nv = (long)(c.nrx + 1) * (c.nry + 1);
  const long nq = (long)c.nqx * c.nqy;
  const long total = (long)c.nex * c.ney * nq * nv;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
    const long v = i % nv, rest = i / nv;
    const int q = (int)(rest % nq);
    const long el = rest / nq;
    const int vx = (int)(v % (c.nrx + 1)), vy = (int)(v / (c.nrx + 1));
    const int qx = q / c.nqy, qy = q - qx * c.nqy;
    const int ex = (int)(el % c.nex), ey = (int)(el / c.nex);
    const int kx = (ex * c.nqx + qx) * c.npx + vx;
    const int ky = (ey * c.nqy + qy) * c.npy + vy;
    const double dv = f2_delta_v(c, kx, ky);
    t64[i] = dv;
    t32[i] = (float)dv;
  }
}

// ---------------------------------------------------------------------------
// Line-transverse preconditioner (get_LineTransversePreconditioner_full2D.m
// :199-239) in the eigenbasis of -i D_rhoX.  Vectors keep the global layout
// with the rho_x index replaced by the mode index k.
// ---------------------------------------------------------------------------

// Per (mode k, transverse t = rho_y + nry*Ydof) line: the upwind X sweep
// prev = Xinv_e (w_e - Xcoup_e prev) over the X elements in the direction of
// the mode's speed, in place.
template <int NL, typename T>
__global__ void __launch_bounds__(256) f2_line_sweep_kernel(const F2Const* __restrict__ cp,
                                                            typename F2Vec<T>::T2* __restrict__ w,
                                                            const T* __restrict__ xinv, const T* __restrict__ xcoup,
                                                            const int* __restrict__ fwd, const GmresState* st) {
  using T2 = typename F2Vec<T>::T2;
  if (skip(st)) return;
  const F2Const& c = *cp;
  const int nrx = c.nrx, nry = c.nry;
  const long lines = (long)nrx * nry * c.nyd;
  const long id = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (id >= lines) return;
  const int k = (int)(id % nrx);
  const long rest = id / nrx;
  const int ry = (int)(rest % nry), iyd = (int)(rest / nry);
  const long sx = (long)nrx * nry;                            // stride of one X DOF
  T2* base = w + k + (long)nrx * ry + sx * (long)c.nxd * iyd;
  const bool forward = fwd[k] != 0;
  T2 prev[NL];
#pragma unroll
  for (int a = 0; a < NL; ++a) prev[a] = f2_zero<T2>();
  for (int step = 0; step < c.nex; ++step) {
    const int e = forward ? step : c.nex - 1 - step;
    const T* Xi = xinv + ((size_t)k * c.nex + e) * NL * NL;
    const T* Xc = xcoup + ((size_t)k * c.nex + e) * NL * NL;
    T2 blk[NL];
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      T2 v = base[sx * (e * NL + a)];
#pragma unroll
      for (int b = 0; b < NL; ++b) {
        v.x -= Xc[a * NL + b] * prev[b].x;
        v.y -= Xc[a * NL + b] * prev[b].y;
      }
      blk[a] = v;
    }
#pragma unroll
    for (int a = 0; a < NL; ++a) {
      T2 v = f2_zero<T2>();
#pragma unroll
      for (int b = 0; b < NL; ++b) f2_fma(v, Xi[a * NL + b], blk[b]);
      prev[a] = v;
      base[sx * (e * NL + a)] = v;
    }
  }
}

// Per mode k: out(t, X) = sum_t' Tinv_k(t, t') v(t', X) (the dense transverse
// solve as its explicit inverse).  Tile 64 t x 16 X per workgroup, 4 t per
// thread, t' staged through LDS in chunks of 16.
template <typename T>
__global__ void __launch_bounds__(256) f2_transverse_kernel(const F2Const* __restrict__ cp,
                                                            const typename F2Vec<T>::T2* __restrict__ v,
                                                            typename F2Vec<T>::T2* __restrict__ out,
                                                            const typename F2Vec<T>::T2* __restrict__ tinv,
                                                            const GmresState* st) {
  using T2 = typename F2Vec<T>::T2;
  if (skip(st)) return;
  const F2Const& c = *cp;
  const int nrx = c.nrx, nry = c.nry, nxd = c.nxd;
  const int nT = nry * c.nyd;
  const int k = blockIdx.z;
  const int t0 = blockIdx.x * 64, x0 = blockIdx.y * 16;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;   // X column, group of 4 t rows
  __shared__ T2 sB[64][17];
  __shared__ T2 sV[16][17];
  const T2* Bk = tinv + (size_t)k * nT * nT;
  auto addr = [&](int t, int ixd) -> long {
    const int ry = t % nry, iyd = t / nry;
    return k + (long)nrx * (ry + (long)nry * (ixd + (long)nxd * iyd));
  };
  T2 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = f2_zero<T2>();
  for (int p0 = 0; p0 < nT; p0 += 16) {
    for (int idx = threadIdx.x; idx < 64 * 16; idx += 256) {
      const int rr = idx >> 4, cc = idx & 15;
      const int t = t0 + rr, tp = p0 + cc;
      sB[rr][cc] = (t < nT && tp < nT) ? Bk[(size_t)t * nT + tp] : f2_zero<T2>();
    }
    {
      const int rr = threadIdx.x >> 4, cc = threadIdx.x & 15;
      const int tp = p0 + rr, ixd = x0 + cc;
      sV[rr][cc] = (tp < nT && ixd < nxd) ? v[addr(tp, ixd)] : f2_zero<T2>();
    }
    __syncthreads();
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {
      const T2 vv = sV[q][tx];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const T2 b = sB[ty * 4 + j][q];
        acc[j].x += b.x * vv.x - b.y * vv.y;
        acc[j].y += b.x * vv.y + b.y * vv.x;
      }
    }
    __syncthreads();
  }
  const int ixd = x0 + tx;
  if (ixd >= nxd) return;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int t = t0 + ty * 4 + j;
    if (t < nT) out[addr(t, ixd)] = acc[j];
  }
}

}  // namespace wdg
