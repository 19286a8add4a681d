// Device kernels of the 1D-x x 1D-k Wigner operator for gfx950 (CDNA4).
//
// Index convention (reference permuted order, DG_code/core/get_Diff.m:242-259):
//   i = (e*n + a)*nk + j      e element, a node, j wave number (fastest)
// Vectors are interleaved complex: T2 = {re, im} with T = double or float.
//
// The device operator is A_hat = A / s_A (s_A = largest transport
// coefficient) so that FP32 never overflows; GMRES is invariant to this.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

namespace wdg {

constexpr int kMaxN = 8;       // max nodes per element
constexpr int kMaxPairs = kMaxN * (kMaxN + 1) / 2;
constexpr int kNqD = 10;
constexpr int kMaxRestart = 64;

template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };

__device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }

// Device-side state of one GMRES cycle (see the GMRES section below).
struct GmresState {
  double beta;       // residual norm at the start of the cycle (normalised units)
  double target;     // stop when |g_{j+1}| <= target
  double resid;      // current Givens estimate |g_{j+1}|
  double nrm0;       // ||w|| before orthogonalisation
  double resnorm;    // output of residual_kernel
  double hnext;      // ||w|| after orthogonalisation = H[j+1,j]
  int reorth_count;  // second Gram-Schmidt passes taken (diagnostic)
  double tol_abs;    // solve tolerance (normalised units)
  double inner_tol;  // per-cycle relative reduction target (0 = none)
  int single_pass;   // 1: classical Gram-Schmidt without the second pass
  int done;          // cycle converged
  int jlast;         // number of completed Arnoldi steps in this cycle
  int reorth;        // second CGS pass needed for the current step
  int breakdown;
  unsigned int ticket;   // unused (kept for layout stability)
  int its_total;     // Arnoldi steps of the whole solve (index into the residual history)
  int slow_norm;     // 1: this step's norm needs the explicit pass (arnoldi_update_pyth_kernel)
};

__device__ __forceinline__ bool skip(const GmresState* st) { return st && st->done; }

// ---------------------------------------------------------------------------
// Constant operator data shared by all kernels of one context.
// ---------------------------------------------------------------------------
struct OpConst {
  int n_chi, n, nk, n_xi, n_pairs;
  int chunk;                   // elements per line-solve chunk
  int n_chunks;
  double tq;                   // Q_diff / J / s_A
  double inv_len;              // 1 / L_xi
  double hprime;               // phi sample spacing
  double K1[kMaxN * kMaxN];    // element matrices (xi-scaled inverse mass)
  double K2[kMaxN * kMaxN];
  double invM0[kMaxN];         // invM(:,0)
  double invMn[kMaxN];         // invM(:,n-1)
  double cab[kMaxN * kMaxN];   // Q_drift*invM_ab*delta_xi/4/s_A
  double Mq[kMaxN * kMaxN];    // sum_q w_q W_qa W_qb
  double Sq[kNqD * kMaxPairs]; // w_q W_qa W_qb per (a<=b) pair
  int pa[kMaxPairs], pb[kMaxPairs];
  int pidx[kMaxN * kMaxN];     // (a,b) -> pair
};

// ---------------------------------------------------------------------------
// Potential tables
// ---------------------------------------------------------------------------

// V_transient (DG_code/core/V_transient.m:1-8): interp1 on x_mat for x < L,
// min(EM) for x >= L, EM(1) for x < 0.
__device__ __forceinline__ double v_transient(double x, const double* __restrict__ xg,
                                              const double* __restrict__ ug, int ng, double L,
                                              double umin) {
  if (x < 0.0) return ug[0];
  if (x >= L) return umin;
  if (x <= xg[0]) return ug[0];
  if (x >= xg[ng - 1]) return ug[ng - 1];
  int lo = 0, hi = ng - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (xg[mid] <= x) lo = mid; else hi = mid;
  }
  const double s = (ug[hi] - ug[lo]) / (xg[hi] - xg[lo]);
  return s * (x - xg[lo]) + ug[lo];
}

// Sbar[e][pair][l] = sum_q w_q W_qa W_qb (V(x_eq + xi_l/2) - V(x_eq - xi_l/2))
// for the n_xi+1 edges xi_l (get_Drift_accel_trial_permutedShell.m:64-66,
// with the DG projection weights of :200-215).  One thread per (e, l).
__global__ void drift_table_kernel(const OpConst* __restrict__ oc, const double* __restrict__ xquad,
                                   const double* __restrict__ edges, const double* __restrict__ xg,
                                   const double* __restrict__ ug, int ng, double L, double umin,
                                   double* __restrict__ sbar64, float* __restrict__ sbar32) {
  const int ne = oc->n_chi, nl = oc->n_xi + 1, np = oc->n_pairs;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ne * nl) return;
  const int e = (int)(t / nl), l = (int)(t % nl);
  const double half = 0.5 * edges[l];
  double acc[kMaxPairs];
  for (int p = 0; p < np; ++p) acc[p] = 0.0;
  for (int q = 0; q < kNqD; ++q) {
    const double x = xquad[e * kNqD + q];
    const double dv = v_transient(x + half, xg, ug, ng, L, umin) - v_transient(x - half, xg, ug, ng, L, umin);
    for (int p = 0; p < np; ++p) acc[p] += oc->Sq[q * kMaxPairs + p] * dv;
  }
  for (int p = 0; p < np; ++p) {
    const long o = ((long)e * np + p) * nl + l;
    sbar64[o] = acc[p];
    sbar32[o] = (float)acc[p];
  }
}

// Edge coefficient Bbar_ab(l) = c_ab (Sbar_ab(l) - i CAP(l) Mq_ab).
__device__ __forceinline__ double2 bbar(const OpConst* oc, int a, int b, double s, double cap) {
  const double c = oc->cab[a * kMaxN + b];
  return mk(c * s, -c * cap * oc->Mq[a * kMaxN + b]);
}

// Line-preconditioner sums per (e,a,b): S0 = sum_l (Bbar(l)+Bbar(l+1)),
// S1 = sum_l Bbar(l+1) over the off-diagonal.  diag_k(G) = (i/L)(S0 + 2cos(k h')S1).
__global__ void precond_sums_kernel(const OpConst* __restrict__ oc, const double* __restrict__ sbar64,
                                    const double* __restrict__ cap, double2* __restrict__ s0s1) {
  const int n = oc->n, nl = oc->n_xi + 1, np = oc->n_pairs;
  const int e = blockIdx.x / (n * n);
  const int ab = blockIdx.x % (n * n);
  const int a = ab / n, b = ab % n;
  const int p = oc->pidx[a * kMaxN + b];
  const double* s = sbar64 + ((long)e * np + p) * nl;
  double r0 = 0, i0 = 0, r1 = 0, i1 = 0;
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    const double2 v = bbar(oc, a, b, s[l], cap[l]);
    // Bbar(l) enters the diagonal sum twice except at the two end edges.
    const double wdiag = (l == 0 || l == nl - 1) ? 1.0 : 2.0;
    r0 += wdiag * v.x;
    i0 += wdiag * v.y;
    if (l >= 1 && l <= nl - 2) {
      r1 += v.x;
      i1 += v.y;
    }
  }
  __shared__ double red[4][256];
  red[0][threadIdx.x] = r0; red[1][threadIdx.x] = i0; red[2][threadIdx.x] = r1; red[3][threadIdx.x] = i1;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    s0s1[((long)e * n * n + ab) * 2 + 0] = mk(red[0][0], red[1][0]);
    s0s1[((long)e * n * n + ab) * 2 + 1] = mk(red[2][0], red[3][0]);
  }
}

// ---------------------------------------------------------------------------
// Transport: y = (Q/J/s_A)[k K1 u_e + |k| K2 u_e - (k+|k|)/2 invM(:,0) u_{e-1,n-1}
//                          + (k-|k|)/2 invM(:,n-1) u_{e+1,0}]
// (H1 (x) D + H2 (x) |D|, DG_code/core/get_Diff.m:243-249, full upwind).
// One thread per (e, j); the n nodes of (e, j) are handled together.
// ---------------------------------------------------------------------------
template <int NN, typename T>
__global__ __launch_bounds__(256) void transport_kernel(const OpConst* __restrict__ oc,
                                                        const double* __restrict__ kvec,
                                                        const typename cvec<T>::type* __restrict__ u,
                                                        typename cvec<T>::type* __restrict__ y,
                                                        const GmresState* st) {
  using T2 = typename cvec<T>::type;
  if (skip(st)) return;
  // One thread per (e, pair of wave numbers j, j+1): the n nodes of the
  // element and the upstream node are read once, the pair as one 16-byte
  // packet in FP32 (two in FP64); arithmetic in the working precision T.
  constexpr int n = NN;
  const int nk = oc->nk, ne = oc->n_chi;
  const int nk2 = nk >> 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ne * nk2) return;
  const int e = (int)(t / nk2), j = 2 * (int)(t % nk2);
  const T tq = (T)oc->tq;
  const T k[2] = {(T)kvec[j], (T)kvec[j + 1]};
  const T ak[2] = {k[0] < 0 ? -k[0] : k[0], k[1] < 0 ? -k[1] : k[1]};
  T2 uv[NN][2];
  const T2* ue = u + (long)e * n * nk + j;
#pragma unroll
  for (int b = 0; b < NN; ++b) {
    if constexpr (sizeof(T2) == 8) {   // FP32: the pair in one 16-byte load
      const float4 v = *reinterpret_cast<const float4*>(ue + (long)b * nk);
      uv[b][0].x = v.x; uv[b][0].y = v.y;
      uv[b][1].x = v.z; uv[b][1].y = v.w;
    } else {
      uv[b][0] = ue[(long)b * nk];
      uv[b][1] = ue[(long)b * nk + 1];
    }
  }
  // upstream inflow values (k > 0: last node of e-1; k < 0: first node of e+1)
  T2 cv[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    cv[q].x = 0;
    cv[q].y = 0;
    if (k[q] > 0 && e > 0) cv[q] = u[((long)(e - 1) * n + (n - 1)) * nk + j + q];
    else if (k[q] < 0 && e < ne - 1) cv[q] = u[((long)(e + 1) * n) * nk + j + q];
  }
  T2* ye = y + (long)e * n * nk + j;
#pragma unroll
  for (int a = 0; a < NN; ++a) {
    T2 o[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      T sr = 0, si = 0;
#pragma unroll
      for (int b = 0; b < NN; ++b) {
        const T c = k[q] * (T)oc->K1[a * kMaxN + b] + ak[q] * (T)oc->K2[a * kMaxN + b];
        sr += c * uv[b][q].x;
        si += c * uv[b][q].y;
      }
      // -(k+|k|)/2 invM(:,0) u_{e-1,n-1}  or  (k-|k|)/2 invM(:,n-1) u_{e+1,0}
      const T cf = (k[q] > 0) ? ((T)-0.5 * (k[q] + ak[q]) * (T)oc->invM0[a])
                              : ((T)0.5 * (k[q] - ak[q]) * (T)oc->invMn[a]);
      sr += cf * cv[q].x;
      si += cf * cv[q].y;
      o[q].x = tq * sr;
      o[q].y = tq * si;
    }
    if constexpr (sizeof(T2) == 8) {
      *reinterpret_cast<float4*>(ye + (long)a * nk) = make_float4(o[0].x, o[0].y, o[1].x, o[1].y);
    } else {
      ye[(long)a * nk] = o[0];
      ye[(long)a * nk + 1] = o[1];
    }
  }
}

// ---------------------------------------------------------------------------
// Theta combine: Z_{e,a}(l) = sum_b [dg_ab(l) T_b(l) + off_ab(l) T_b(l+1)
//                + off_ab(l-1) T_b(l-1)],  dg(l) = Bbar(l)+Bbar(l+1),
// off(l) = Bbar(l+1)  (tridiagonal C of get_Drift_accel_trial_permutedShell.m:102-135
// projected with invM o (W~'W), :200-219).  Tm is GEMM1's output T[c][l],
// c = e*n + b.  One thread per (e, l).
// ---------------------------------------------------------------------------
template <int NN, typename T, typename TS>
__global__ __launch_bounds__(256) void combine_kernel(const OpConst* __restrict__ oc,
                                                      const TS* __restrict__ sbar,
                                                      const double* __restrict__ cap,
                                                      const typename cvec<T>::type* __restrict__ Tm,
                                                      typename cvec<T>::type* __restrict__ Z,
                                                      const GmresState* st) {
  using T2 = typename cvec<T>::type;
  if (skip(st)) return;
  constexpr int n = NN, np = NN * (NN + 1) / 2;
  const int nx = oc->n_xi, ne = oc->n_chi, nl = nx + 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ne * nx) return;
  const int e = (int)(t / nx), l = (int)(t % nx);
  double tr[3][NN], ti[3][NN];
#pragma unroll
  for (int b = 0; b < NN; ++b) {
      const T2* row = Tm + ((long)e * n + b) * nx;
      T2 v0 = (l > 0) ? row[l - 1] : T2{0, 0};
      T2 v1 = row[l];
      T2 v2 = (l < nx - 1) ? row[l + 1] : T2{0, 0};
      tr[0][b] = v0.x; ti[0][b] = v0.y;
      tr[1][b] = v1.x; ti[1][b] = v1.y;
      tr[2][b] = v2.x; ti[2][b] = v2.y;
    }
  const double c0 = cap[l], c1 = cap[l + 1], c2 = (l + 2 < nl) ? cap[l + 2] : 0.0;
  const TS* sb = sbar + (long)e * np * nl;
#pragma unroll
  for (int a = 0; a < NN; ++a) {
      double zr = 0, zi = 0;
#pragma unroll
      for (int b = 0; b < NN; ++b) {
          // pair index of (min(a,b), max(a,b)) in row-major upper-triangle order
          const int lo = a < b ? a : b, hi = a < b ? b : a;
          const int p = lo * NN - lo * (lo - 1) / 2 + (hi - lo);
          const TS* s = sb + (long)p * nl;
          const double2 B0 = bbar(oc, a, b, (double)s[l], c0);      // edge l
          const double2 B1 = bbar(oc, a, b, (double)s[l + 1], c1);  // edge l+1
          // diagonal: B0 + B1
          const double dr = B0.x + B1.x, di = B0.y + B1.y;
          zr += dr * tr[1][b] - di * ti[1][b];
          zi += dr * ti[1][b] + di * tr[1][b];
          // super diagonal C(l,l+1) = Bbar(l+1)
          if (l < nx - 1) {
            zr += B1.x * tr[2][b] - B1.y * ti[2][b];
            zi += B1.x * ti[2][b] + B1.y * tr[2][b];
          }
          // sub diagonal C(l,l-1) = Bbar(l)
          if (l > 0) {
            zr += B0.x * tr[0][b] - B0.y * ti[0][b];
            zi += B0.x * ti[0][b] + B0.y * tr[0][b];
          }
        }
      (void)c2;
      T2 o;
      o.x = (T)zr;
      o.y = (T)zi;
      Z[((long)e * n + a) * nx + l] = o;
    }
}

// ---------------------------------------------------------------------------
// Realified complex GEMM on MFMA:
//   C[c][m'] (+)= sum_k' A[m'][k'] * B[c][k']
// A: M' x K' row-major (precomputed realified phi), B: Ncol x K' row-major
// (interleaved complex state, k' = (j, re/im)), C: Ncol x M' row-major.
// M' % BM == 0 and K' % BK == 0 are guaranteed by the host; Ncol is ragged.
// ---------------------------------------------------------------------------

// FP64: v_mfma_f64_16x16x4_f64.  Lane layout (checked on gfx950):
// src0[i=lane%16][k=lane/16], src1[k=lane/16][j=lane%16],
// D[i=lane/16+4r][j=lane%16].  src0 is the state tile (rows c) and src1 the
// basis tile (rows m'), so lanes 0..15 of a store write 16 consecutive m'.
typedef double dbl4 __attribute__((ext_vector_type(4)));

template <int BM, int BN, bool ACCUM>
__global__ __launch_bounds__(256) void gemm_f64_kernel(const double* __restrict__ A,
                                                       const double* __restrict__ B,
                                                       double* __restrict__ C, int M, int N,
                                                       int K, const GmresState* st) {
  if (skip(st)) return;
  constexpr int BK = 16;
  constexpr int LDA = BK + 1;
  constexpr int WM = BM / 2, WN = BN / 2;       // per-wave tile (2x2 waves)
  constexpr int TM = WM / 16, TN = WN / 16;     // MFMA tiles per wave
  __shared__ double As[BM * LDA];
  __shared__ double Bs[BN * LDA];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  dbl4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = dbl4{0, 0, 0, 0};
  // loader: each thread moves (BM*BK/256) doubles of A and (BN*BK/256) of B
  constexpr int APT = BM * BK / 256, BPT = BN * BK / 256;
  double ra[APT], rb[BPT];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int idx = tid + i * 256;
      const int r = idx / BK, c = idx % BK;
      ra[i] = A[(long)(m0 + r) * K + k0 + c];
    }
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      const int idx = tid + i * 256;
      const int r = idx / BK, c = idx % BK;
      rb[i] = (n0 + r < N) ? B[(long)(n0 + r) * K + k0 + c] : 0.0;
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int idx = tid + i * 256;
      As[(idx / BK) * LDA + idx % BK] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      const int idx = tid + i * 256;
      Bs[(idx / BK) * LDA + idx % BK] = rb[i];
    }
  };
  gload(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < K) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[(wm * WM + i * 16 + (lane & 15)) * LDA + ks + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[(wn * WN + j * 16 + (lane & 15)) * LDA + ks + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int row = m0 + wm * WM + i * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = n0 + wn * WN + j * 16 + (lane >> 4) + 4 * r;
        if (col >= N) continue;
        double* cp = C + (long)col * M + row;
        if (ACCUM) *cp += acc[i][j][r];
        else *cp = acc[i][j][r];
      }
    }
}

// Split-BF16: A in bf16 (precomputed), B in fp32 split on the fly into
// hi = bf16(b) and lo = bf16(b - hi); C += A*hi + A*lo accumulated in fp32 on
// v_mfma_f32_16x16x32_bf16.  Lane layout (checked on gfx950):
// src0[i=lane%16][k=8*(lane/16)+t], src1[k=8*(lane/16)+t][j=lane%16],
// D[i=4*(lane/16)+r][j=lane%16]; src0 = state tile, src1 = basis tile.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float flt4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bf16_bits(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

template <int BM, int BN, bool ACCUM, int WGM = 2, int WGN = 2, int BK = 64>
__global__ __launch_bounds__(64 * WGM * WGN) void gemm_bf16x2_kernel(const uint16_t* __restrict__ A,
                                                          const float* __restrict__ B,
                                                          float* __restrict__ C, int M, int N,
                                                          int K, const GmresState* st) {
  if (skip(st)) return;
  constexpr int NT = 64 * WGM * WGN;
  constexpr int LDK = BK + 8;                   // bf16 elements per LDS row (16 B pad)
  constexpr int WM = BM / WGM, WN = BN / WGN;   // per-wave tile
  constexpr int TM = WM / 16, TN = WN / 16;     // 16x16 MFMA tiles per wave
  constexpr int ACH = BM * BK / 8 / NT;         // 16-B chunks of A per thread
  constexpr int BCH = BN * BK / 4 / NT;         // 16-B chunks of B per thread
  __shared__ __attribute__((aligned(16))) uint16_t As[BM * LDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bh[BN * LDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bl[BN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WGM, wn = wave / WGM;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  flt4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = flt4{0, 0, 0, 0};
  uint4 ra[ACH];
  float4 rb[BCH];
  // prefetch of k-step 0
#pragma unroll
  for (int i = 0; i < ACH; ++i) {
    const int idx = tid + i * NT;
    const int r = idx / (BK / 8), c = (idx % (BK / 8)) * 8;
    ra[i] = *reinterpret_cast<const uint4*>(A + (long)(m0 + r) * K + c);
  }
#pragma unroll
  for (int i = 0; i < BCH; ++i) {
    const int idx = tid + i * NT;
    const int r = idx / (BK / 4), c = (idx % (BK / 4)) * 4;
    rb[i] = (n0 + r < N) ? *reinterpret_cast<const float4*>(B + (long)(n0 + r) * K + c)
                         : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  for (int k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ACH; ++i) {
      const int idx = tid + i * NT;
      const int r = idx / (BK / 8), c = (idx % (BK / 8)) * 8;
      *reinterpret_cast<uint4*>(&As[r * LDK + c]) = ra[i];
    }
#pragma unroll
    for (int i = 0; i < BCH; ++i) {
      const int idx = tid + i * NT;
      const int r = idx / (BK / 4), c = (idx % (BK / 4)) * 4;
      const uint16_t h0 = bf16_bits(rb[i].x), h1 = bf16_bits(rb[i].y);
      const uint16_t h2 = bf16_bits(rb[i].z), h3 = bf16_bits(rb[i].w);
      const uint16_t l0 = bf16_bits(rb[i].x - bf16_float(h0)), l1 = bf16_bits(rb[i].y - bf16_float(h1));
      const uint16_t l2 = bf16_bits(rb[i].z - bf16_float(h2)), l3 = bf16_bits(rb[i].w - bf16_float(h3));
      uint2 hv, lv;
      hv.x = (uint32_t)h0 | ((uint32_t)h1 << 16);
      hv.y = (uint32_t)h2 | ((uint32_t)h3 << 16);
      lv.x = (uint32_t)l0 | ((uint32_t)l1 << 16);
      lv.y = (uint32_t)l2 | ((uint32_t)l3 << 16);
      *reinterpret_cast<uint2*>(&Bh[r * LDK + c]) = hv;
      *reinterpret_cast<uint2*>(&Bl[r * LDK + c]) = lv;
    }
    __syncthreads();
    if (k0 + BK < K) {
#pragma unroll
      for (int i = 0; i < ACH; ++i) {
        const int idx = tid + i * NT;
        const int r = idx / (BK / 8), c = (idx % (BK / 8)) * 8;
        ra[i] = *reinterpret_cast<const uint4*>(A + (long)(m0 + r) * K + k0 + BK + c);
      }
#pragma unroll
      for (int i = 0; i < BCH; ++i) {
        const int idx = tid + i * NT;
        const int r = idx / (BK / 4), c = (idx % (BK / 4)) * 4;
        rb[i] = (n0 + r < N) ? *reinterpret_cast<const float4*>(B + (long)(n0 + r) * K + k0 + BK + c)
                             : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      const int kk = ks + 8 * (lane >> 4);
      bf16x8 a[TM], bh[TN], bl[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i)
        a[i] = *reinterpret_cast<const bf16x8*>(&As[(wm * WM + i * 16 + (lane & 15)) * LDK + kk]);
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        bh[j] = *reinterpret_cast<const bf16x8*>(&Bh[(wn * WN + j * 16 + (lane & 15)) * LDK + kk]);
        bl[j] = *reinterpret_cast<const bf16x8*>(&Bl[(wn * WN + j * 16 + (lane & 15)) * LDK + kk]);
      }
      // all hi products first, then all lo products: consecutive MFMAs never
      // share an accumulator
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], a[i], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], a[i], acc[i][j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int row = m0 + wm * WM + i * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = n0 + wn * WN + j * 16 + 4 * (lane >> 4) + r;
        if (col >= N) continue;
        float* cp = C + (long)col * M + row;
        if (ACCUM) *cp += acc[i][j][r];
        else *cp = acc[i][j][r];
      }
    }
}

// ---------------------------------------------------------------------------
// Line preconditioner: per wave number k an exact upwind block-bidiagonal
// solve along x of  D_e(k) z_e = r_e - c(k) * s_in  with
//   D_e(k) = (Q/J/s_A)(k K1 + |k| K2) + diag_k(G)_e,
// swept forward for k > 0 and backward for k < 0 (the x-line factor of
// DG_code/core/get_LineTransversePreconditioner_full2D.m:174-220, in the k
// basis where the rho_x eigen-transform is the identity).  s_in is the
// upstream neighbour's outflow node value; it obeys the affine recurrence
// s_e = alpha_e - beta_e s_{e-1} with alpha_e = (D^-1 r_e)_out and
// beta_e = (D^-1 c)_out.  Three phases:
//   line_aggregate: per (chunk, k) compose the chunk's map s_out = A + B s_in
//   line_carry:     per k, sweep the chunk maps to get each chunk's s_in
//   line_apply:     per (chunk, k) sweep the chunk again and write z.
// Block solves are Gaussian elimination with partial pivoting in FP64 on
// the augmented system, whatever the I/O precision; n is a template
// parameter so the block stays in registers.
// ---------------------------------------------------------------------------
template <int NN, int NR>
struct Aug {
  double r[NN][NN + NR], i[NN][NN + NR];
};

template <int NN, int NR>
__device__ __forceinline__ void build_aug(const OpConst* __restrict__ oc, const double2* __restrict__ s0s1, int e,
                                          double k, double cosk, Aug<NN, NR>& B) {
  const double ak = fabs(k);
  const double2* se = s0s1 + (long)e * NN * NN * 2;
#pragma unroll
  for (int a = 0; a < NN; ++a)
#pragma unroll
    for (int b = 0; b < NN; ++b) {
      const double2 s0 = se[(a * NN + b) * 2 + 0];
      const double2 s1 = se[(a * NN + b) * 2 + 1];
      // diag_k(G) = (i/L)(S0 + 2 cos(k h') S1)
      const double gr = s0.x + 2.0 * cosk * s1.x, gi = s0.y + 2.0 * cosk * s1.y;
      B.r[a][b] = oc->tq * (k * oc->K1[a * kMaxN + b] + ak * oc->K2[a * kMaxN + b]) - oc->inv_len * gi;
      B.i[a][b] = oc->inv_len * gr;
    }
}

// Gaussian elimination with partial pivoting; the last NR columns hold the
// right-hand sides and end up holding the solutions.
template <int NN, int NR>
__device__ __forceinline__ void gauss_solve(Aug<NN, NR>& B) {
  constexpr int NC = NN + NR;
#pragma unroll
  for (int c = 0; c < NN; ++c) {
    // pivot search
    double best = B.r[c][c] * B.r[c][c] + B.i[c][c] * B.i[c][c];
    int p = c;
#pragma unroll
    for (int r = c + 1; r < NN; ++r) {
      const double v = B.r[r][c] * B.r[r][c] + B.i[r][c] * B.i[r][c];
      if (v > best) { best = v; p = r; }
    }
#pragma unroll
    for (int r = c + 1; r < NN; ++r) {
      const bool sw = (r == p);
#pragma unroll
      for (int q = c; q < NC; ++q) {
        const double ar = B.r[c][q], ai = B.i[c][q], br = B.r[r][q], bi = B.i[r][q];
        B.r[c][q] = sw ? br : ar; B.i[c][q] = sw ? bi : ai;
        B.r[r][q] = sw ? ar : br; B.i[r][q] = sw ? ai : bi;
      }
    }
    const double dd = B.r[c][c] * B.r[c][c] + B.i[c][c] * B.i[c][c];
    const double ir = B.r[c][c] / dd, ii = -B.i[c][c] / dd;
#pragma unroll
    for (int q = c + 1; q < NC; ++q) {  // scale pivot row
      const double xr = B.r[c][q] * ir - B.i[c][q] * ii;
      const double xi = B.r[c][q] * ii + B.i[c][q] * ir;
      B.r[c][q] = xr;
      B.i[c][q] = xi;
    }
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      if (r == c) continue;
      const double fr = B.r[r][c], fi = B.i[r][c];
#pragma unroll
      for (int q = c + 1; q < NC; ++q) {
        B.r[r][q] -= fr * B.r[c][q] - fi * B.i[c][q];
        B.i[r][q] -= fr * B.i[c][q] + fi * B.r[c][q];
      }
    }
  }
}

// coupling coefficient tq*c_a of the upstream value for wave number k
__device__ __forceinline__ double line_coupling(const OpConst* __restrict__ oc, int a, double k) {
  const double ak = fabs(k);
  return oc->tq * ((k > 0) ? (-0.5 * (k + ak) * oc->invM0[a]) : (0.5 * (k - ak) * oc->invMn[a]));
}

template <int NN, typename T>
__global__ __launch_bounds__(256) void line_aggregate_kernel(const OpConst* __restrict__ oc,
                                                             const double* __restrict__ kvec,
                                                             const double2* __restrict__ s0s1,
                                                             const typename cvec<T>::type* __restrict__ r,
                                                             double2* __restrict__ agg, const GmresState* st) {
  using T2 = typename cvec<T>::type;
  if (skip(st)) return;
  const int nk = oc->nk, ne = oc->n_chi, S = oc->chunk, nc = oc->n_chunks;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nc * nk) return;
  const int c = (int)(t / nk), j = (int)(t % nk);
  const double k = kvec[j];
  const double cosk = cos(k * oc->hprime);
  const int e_lo = c * S, e_hi = min(ne, e_lo + S);
  const int out = (k > 0) ? NN - 1 : 0;
  double Ar = 0, Ai = 0, Br = 1, Bi = 0;
  for (int step = 0; step < e_hi - e_lo; ++step) {
    const int e = (k > 0) ? e_lo + step : e_hi - 1 - step;
    Aug<NN, 2> B;
    build_aug<NN, 2>(oc, s0s1, e, k, cosk, B);
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      const T2 v = r[((long)e * NN + a) * nk + j];
      B.r[a][NN] = v.x; B.i[a][NN] = v.y;
      B.r[a][NN + 1] = line_coupling(oc, a, k); B.i[a][NN + 1] = 0.0;
    }
    gauss_solve<NN, 2>(B);
    double al_r = 0, al_i = 0, be_r = 0, be_i = 0;
#pragma unroll
    for (int a = 0; a < NN; ++a)
      if (a == out) { al_r = B.r[a][NN]; al_i = B.i[a][NN]; be_r = B.r[a][NN + 1]; be_i = B.i[a][NN + 1]; }
    // s_out = alpha - beta * s_in, composed with the running map s_in = A + B s0
    const double nAr = al_r - (be_r * Ar - be_i * Ai), nAi = al_i - (be_r * Ai + be_i * Ar);
    const double nBr = -(be_r * Br - be_i * Bi), nBi = -(be_r * Bi + be_i * Br);
    Ar = nAr; Ai = nAi; Br = nBr; Bi = nBi;
  }
  agg[((long)c * nk + j) * 2 + 0] = mk(Ar, Ai);
  agg[((long)c * nk + j) * 2 + 1] = mk(Br, Bi);
}

// Per k: carry-in of every chunk (overwrites the chunk's A with its s_in).
// One wavefront per k: lane l composes the chunk maps at sweep positions
// [l*cpl, (l+1)*cpl), a Kogge-Stone scan over the lanes composes the
// prefixes (s_out = A + B s_in; later map after earlier), and each lane
// re-walks its chunks from the carry-in of its first one (s = 0 enters the
// first chunk of the sweep).
__device__ __forceinline__ void affine_after(double& Ar, double& Ai, double& Br, double& Bi, double pAr, double pAi,
                                             double pBr, double pBi) {
  // (A, B) after (pA, pB): s -> A + B (pA + pB s)
  const double nAr = Ar + (Br * pAr - Bi * pAi), nAi = Ai + (Br * pAi + Bi * pAr);
  const double nBr = Br * pBr - Bi * pBi, nBi = Br * pBi + Bi * pBr;
  Ar = nAr; Ai = nAi; Br = nBr; Bi = nBi;
}

__global__ __launch_bounds__(256) void line_carry_kernel(const OpConst* __restrict__ oc, const double* __restrict__ kvec,
                                                         double2* __restrict__ agg, const GmresState* st) {
  if (skip(st)) return;
  const int nk = oc->nk, nc = oc->n_chunks;
  const int j = (int)(((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= nk) return;   // uniform per wavefront
  const bool fwd = kvec[j] > 0;
  const int cpl = (nc + 63) >> 6;
  double Ar = 0, Ai = 0, Br = 1, Bi = 0;
  for (int q = 0; q < cpl; ++q) {
    const int pos = lane * cpl + q;
    if (pos >= nc) break;
    const int c = fwd ? pos : nc - 1 - pos;
    const double2 a = agg[((long)c * nk + j) * 2 + 0], b = agg[((long)c * nk + j) * 2 + 1];
    double cAr = a.x, cAi = a.y, cBr = b.x, cBi = b.y;
    affine_after(cAr, cAi, cBr, cBi, Ar, Ai, Br, Bi);
    Ar = cAr; Ai = cAi; Br = cBr; Bi = cBi;
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double pAr = __shfl_up(Ar, off, 64), pAi = __shfl_up(Ai, off, 64);
    const double pBr = __shfl_up(Br, off, 64), pBi = __shfl_up(Bi, off, 64);
    if (lane >= off) affine_after(Ar, Ai, Br, Bi, pAr, pAi, pBr, pBi);
  }
  // carry into this lane's first chunk: the inclusive prefix of lane-1 at s = 0
  double sr = __shfl_up(Ar, 1, 64), si = __shfl_up(Ai, 1, 64);
  if (lane == 0) {
    sr = 0.0;
    si = 0.0;
  }
  for (int q = 0; q < cpl; ++q) {
    const int pos = lane * cpl + q;
    if (pos >= nc) break;
    const int c = fwd ? pos : nc - 1 - pos;
    const long o = ((long)c * nk + j) * 2;
    const double2 a = agg[o], b = agg[o + 1];
    agg[o] = mk(sr, si);
    const double nr = a.x + b.x * sr - b.y * si, ni = a.y + b.x * si + b.y * sr;
    sr = nr;
    si = ni;
  }
}

template <int NN, typename T>
__global__ __launch_bounds__(256) void line_apply_kernel(const OpConst* __restrict__ oc,
                                                         const double* __restrict__ kvec,
                                                         const double2* __restrict__ s0s1,
                                                         const double2* __restrict__ agg,
                                                         const typename cvec<T>::type* __restrict__ r,
                                                         typename cvec<T>::type* __restrict__ z,
                                                         const GmresState* st) {
  using T2 = typename cvec<T>::type;
  if (skip(st)) return;
  const int nk = oc->nk, ne = oc->n_chi, S = oc->chunk, nc = oc->n_chunks;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nc * nk) return;
  const int c = (int)(t / nk), j = (int)(t % nk);
  const double k = kvec[j];
  const double cosk = cos(k * oc->hprime);
  const int e_lo = c * S, e_hi = min(ne, e_lo + S);
  const int out = (k > 0) ? NN - 1 : 0;
  const double2 s_in = agg[((long)c * nk + j) * 2 + 0];
  double sr = s_in.x, si = s_in.y;
  for (int step = 0; step < e_hi - e_lo; ++step) {
    const int e = (k > 0) ? e_lo + step : e_hi - 1 - step;
    Aug<NN, 1> B;
    build_aug<NN, 1>(oc, s0s1, e, k, cosk, B);
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      const T2 v = r[((long)e * NN + a) * nk + j];
      const double cf = line_coupling(oc, a, k);
      B.r[a][NN] = (double)v.x - cf * sr;
      B.i[a][NN] = (double)v.y - cf * si;
    }
    gauss_solve<NN, 1>(B);
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      T2 o;
      o.x = (T)B.r[a][NN];
      o.y = (T)B.i[a][NN];
      z[((long)e * NN + a) * nk + j] = o;
      if (a == out) { sr = B.r[a][NN]; si = B.i[a][NN]; }
    }
  }
}

// ---------------------------------------------------------------------------
// Line preconditioner of the FP32 solver from stored block inverses.  Each
// (element, k) block D and the coupling column c are fixed for a potential,
// so line_inverse_kernel factors every block once (FP64 Gaussian elimination
// with partial pivoting, as in the sweeps above) and stores, as rows of nk
// values, D^-1 as BF16 complex pairs, dinv[(e * n*n + a*n + b) * nk + j], and
// D^-1 c in FP32, dcpl[(e * n + a) * nk + j].  The sweeps then compose
// s_out = (D^-1 r)_out - (D^-1 c)_out s_in and z = D^-1 r - (D^-1 c) s_in: the
// coupling factors multiply along the whole line and stay FP32, while the
// rounding of D^-1 (2^-9 relative) only perturbs each element's own term, so
// the preconditioner stays a fixed linear operator close to the FP32 one.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lv_pack_bf16x2(float re, float im) {
  uint32_t ur = __float_as_uint(re), ui = __float_as_uint(im);
  ur += 0x7FFFu + ((ur >> 16) & 1u);   // round to nearest even
  ui += 0x7FFFu + ((ui >> 16) & 1u);
  return (ur >> 16) | (ui & 0xFFFF0000u);
}
__device__ __forceinline__ float2 lv_unpack_bf16x2(uint32_t p) {
  return make_float2(__uint_as_float(p << 16), __uint_as_float(p & 0xFFFF0000u));
}

// One complex value of the preconditioner input r: FP32, or read in place
// from a 16-bit Krylov basis vector (FP16 pairs stored times 256, or BF16
// pairs), whose entry i sits in 32-bit word i.
template <typename RIN>
__device__ __forceinline__ float2 line_in(const RIN* r, long i);
template <>
__device__ __forceinline__ float2 line_in<float2>(const float2* r, long i) { return r[i]; }
template <>
__device__ __forceinline__ float2 line_in<ushort2>(const ushort2* r, long i) {
  const ushort2 p = r[i];
  return make_float2((float)__builtin_bit_cast(_Float16, p.x) * (1.0f / 256.0f),
                     (float)__builtin_bit_cast(_Float16, p.y) * (1.0f / 256.0f));
}
template <>
__device__ __forceinline__ float2 line_in<uint32_t>(const uint32_t* r, long i) { return lv_unpack_bf16x2(r[i]); }

template <int NN>
__global__ __launch_bounds__(256) void line_inverse_kernel(const OpConst* __restrict__ oc,
                                                           const double* __restrict__ kvec,
                                                           const double2* __restrict__ s0s1,
                                                           uint32_t* __restrict__ dinv, float2* __restrict__ dcpl) {
  const int nk = oc->nk, ne = oc->n_chi;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ne * nk) return;
  const int e = (int)(t / nk), j = (int)(t % nk);
  const double k = kvec[j];
  Aug<NN, NN + 1> B;
  build_aug<NN, NN + 1>(oc, s0s1, e, k, cos(k * oc->hprime), B);
#pragma unroll
  for (int a = 0; a < NN; ++a) {
#pragma unroll
    for (int q = 0; q < NN; ++q) {
      B.r[a][NN + q] = (a == q) ? 1.0 : 0.0;
      B.i[a][NN + q] = 0.0;
    }
    B.r[a][2 * NN] = line_coupling(oc, a, k);
    B.i[a][2 * NN] = 0.0;
  }
  gauss_solve<NN, NN + 1>(B);
  uint32_t* di = dinv + (long)e * NN * NN * nk + j;
  float2* dc = dcpl + (long)e * NN * nk + j;
#pragma unroll
  for (int a = 0; a < NN; ++a) {
#pragma unroll
    for (int b = 0; b < NN; ++b)
      di[(long)(a * NN + b) * nk] = lv_pack_bf16x2((float)B.r[a][NN + b], (float)B.i[a][NN + b]);
    dc[(long)a * nk] = make_float2((float)B.r[a][2 * NN], (float)B.i[a][2 * NN]);
  }
}

// line_aggregate_kernel from the stored inverses (FP32 products, the chunk
// map composed in FP64).
template <int NN, typename RIN = float2>
__global__ __launch_bounds__(256) void line_aggregate_inv_kernel(const OpConst* __restrict__ oc,
                                                                 const double* __restrict__ kvec,
                                                                 const uint32_t* __restrict__ dinv,
                                                                 const float2* __restrict__ dcpl,
                                                                 const RIN* __restrict__ r,
                                                                 double2* __restrict__ agg, const GmresState* st) {
  if (skip(st)) return;
  const int nk = oc->nk, ne = oc->n_chi, S = oc->chunk, nc = oc->n_chunks;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nc * nk) return;
  const int c = (int)(t / nk), j = (int)(t % nk);
  const bool fwd = kvec[j] > 0;
  const int out = fwd ? NN - 1 : 0;
  const int e_lo = c * S, e_hi = min(ne, e_lo + S);
  double Ar = 0, Ai = 0, Br = 1, Bi = 0;
  // the next element's inputs are loaded while this one is composed (only
  // the running map crosses elements)
  const int ns = e_hi - e_lo;
  uint32_t mc[NN], mn[NN];
  float2 vc[NN], vn[NN], bc, bn;
  auto load = [&](int step, uint32_t (&m)[NN], float2 (&v)[NN], float2& be) {
    const int e = fwd ? e_lo + step : e_hi - 1 - step;
    const uint32_t* di = dinv + (long)e * NN * NN * nk + j;
    const long re = (long)e * NN * nk + j;
#pragma unroll
    for (int b = 0; b < NN; ++b) {
      m[b] = di[(long)(out * NN + b) * nk];
      v[b] = line_in(r, re + (long)b * nk);
    }
    be = dcpl[((long)e * NN + out) * nk + j];
  };
  if (ns > 0) load(0, mc, vc, bc);
  for (int step = 0; step < ns; ++step) {
    if (step + 1 < ns) load(step + 1, mn, vn, bn);
    float al_r = 0, al_i = 0;
#pragma unroll
    for (int b = 0; b < NN; ++b) {
      const float2 m = lv_unpack_bf16x2(mc[b]), v = vc[b];
      al_r += m.x * v.x - m.y * v.y;
      al_i += m.x * v.y + m.y * v.x;
    }
    const float2 be = bc;
    // s_out = alpha - beta * s_in, composed with the running map s_in = A + B s0
    const double nAr = al_r - (be.x * Ar - be.y * Ai), nAi = al_i - (be.x * Ai + be.y * Ar);
    const double nBr = -(be.x * Br - be.y * Bi), nBi = -(be.x * Bi + be.y * Br);
    Ar = nAr; Ai = nAi; Br = nBr; Bi = nBi;
    if (step + 1 < ns) {
#pragma unroll
      for (int b = 0; b < NN; ++b) {
        mc[b] = mn[b];
        vc[b] = vn[b];
      }
      bc = bn;
    }
  }
  agg[((long)c * nk + j) * 2 + 0] = mk(Ar, Ai);
  agg[((long)c * nk + j) * 2 + 1] = mk(Br, Bi);
}

// line_apply_kernel from the stored inverses: z_e = D^-1 r_e - (D^-1 c) s_in.
template <int NN, typename RIN = float2>
__global__ __launch_bounds__(256) void line_apply_inv_kernel(const OpConst* __restrict__ oc,
                                                             const double* __restrict__ kvec,
                                                             const uint32_t* __restrict__ dinv,
                                                             const float2* __restrict__ dcpl,
                                                             const double2* __restrict__ agg,
                                                             const RIN* __restrict__ r, float2* __restrict__ z,
                                                             const GmresState* st) {
  if (skip(st)) return;
  const int nk = oc->nk, ne = oc->n_chi, S = oc->chunk, nc = oc->n_chunks;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nc * nk) return;
  const int c = (int)(t / nk), j = (int)(t % nk);
  const bool fwd = kvec[j] > 0;
  const int out = fwd ? NN - 1 : 0;
  const int e_lo = c * S, e_hi = min(ne, e_lo + S);
  const double2 s_in = agg[((long)c * nk + j) * 2 + 0];
  float sr = (float)s_in.x, si = (float)s_in.y;
  // the next element's inputs are loaded while this one is swept (only the
  // outflow value crosses elements)
  const int ns = e_hi - e_lo;
  uint32_t mc[NN * NN], mn[NN * NN];
  float2 vc[NN], vn[NN], bcc[NN], bcn[NN];
  auto load = [&](int step, uint32_t (&m)[NN * NN], float2 (&v)[NN], float2 (&bcv)[NN]) {
    const int e = fwd ? e_lo + step : e_hi - 1 - step;
    const uint32_t* di = dinv + (long)e * NN * NN * nk + j;
    const float2* dc = dcpl + (long)e * NN * nk + j;
    const long re = (long)e * NN * nk + j;
#pragma unroll
    for (int b = 0; b < NN; ++b) {
      v[b] = line_in(r, re + (long)b * nk);
      bcv[b] = dc[(long)b * nk];
    }
#pragma unroll
    for (int q = 0; q < NN * NN; ++q) m[q] = di[(long)q * nk];
  };
  if (ns > 0) load(0, mc, vc, bcc);
  for (int step = 0; step < ns; ++step) {
    const int e = fwd ? e_lo + step : e_hi - 1 - step;
    if (step + 1 < ns) load(step + 1, mn, vn, bcn);
    float2 zo[NN];
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      const float2 bc = bcc[a];
      float zr = -(bc.x * sr - bc.y * si), zi = -(bc.x * si + bc.y * sr);
#pragma unroll
      for (int b = 0; b < NN; ++b) {
        const float2 m = lv_unpack_bf16x2(mc[a * NN + b]);
        zr += m.x * vc[b].x - m.y * vc[b].y;
        zi += m.x * vc[b].y + m.y * vc[b].x;
      }
      zo[a] = make_float2(zr, zi);
    }
#pragma unroll
    for (int a = 0; a < NN; ++a) {
      z[((long)e * NN + a) * nk + j] = zo[a];
      if (a == out) { sr = zo[a].x; si = zo[a].y; }
    }
    if (step + 1 < ns) {
#pragma unroll
      for (int q = 0; q < NN * NN; ++q) mc[q] = mn[q];
#pragma unroll
      for (int b = 0; b < NN; ++b) {
        vc[b] = vn[b];
        bcc[b] = bcn[b];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// GMRES device state and reductions.  Reductions write per-block partials
// that a single-block finalize kernel sums in block order, so results are
// deterministic and no device-scope fence (an L2 write-back per block on a
// multi-XCD part) is needed inside a kernel.  Every Arnoldi kernel returns at once when the
// cycle has converged (st->done), so a whole cycle can be enqueued (or
// captured in a graph) without host round trips.
// ---------------------------------------------------------------------------

constexpr int kDotGroup = 8;   // basis vectors per block row of the multi-dot

struct GmresBufs {
  double2* H;        // (m+1) x m, column-major
  double2* cs;       // m
  double2* sn;       // m
  double2* g;        // m+1
  double2* h1;       // m+1 (+1 slot for a norm)
  double2* h2;       // m+1 (+1)
  double2* y;        // m
  double2* part;     // partial sums: blocks x (kDotGroup+1)
  GmresState* st;
  int m;
  double* resv;      // residual history of the solve: [k] after Arnoldi step k (null: none;
                     // the host writes [0], the initial residual)
  int resv_cap;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// Block reduction of the kDotGroup+1 complex values of every thread into
// part_out[0..kDotGroup] (thread 0 of each wave, then thread < 9).
template <int NT>
__device__ __forceinline__ void block_reduce_store(const double2 (&vals)[kDotGroup + 1], double2* part_out) {
  __shared__ double2 red[kDotGroup + 1][NT / 64];
#pragma unroll
  for (int i = 0; i < kDotGroup + 1; ++i) {
    double2 v = vals[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v.x += __shfl_down(v.x, off, 64);
      v.y += __shfl_down(v.y, off, 64);
    }
    if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < kDotGroup + 1) {
    double2 s = red[threadIdx.x][0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) { s.x += red[threadIdx.x][w].x; s.y += red[threadIdx.x][w].y; }
    part_out[threadIdx.x] = s;
  }
}

// Second-stage reduction (single-block finalize kernels): res[i] = sum over blocks
// b < nblocks of part[(base + b)*stride + i] for all kDotGroup+1 slots at
// once (strided loads, wave shuffles, one LDS exchange).  res is valid in
// every thread on return.
template <int NT>
__device__ __forceinline__ void finalize_group(const double2* part, long base, int stride, int nblocks,
                                               double2 (&res)[kDotGroup + 1]) {
  __shared__ double2 red[kDotGroup + 1][NT / 64];
  double2 acc[kDotGroup + 1];
#pragma unroll
  for (int i = 0; i < kDotGroup + 1; ++i) acc[i] = mk(0, 0);
  for (int b = threadIdx.x; b < nblocks; b += NT) {
    const double2* pp = part + (base + b) * stride;
#pragma unroll
    for (int i = 0; i < kDotGroup + 1; ++i) {
      const double2 v = pp[i];
      acc[i].x += v.x;
      acc[i].y += v.y;
    }
  }
#pragma unroll
  for (int i = 0; i < kDotGroup + 1; ++i) {
    double2 v = acc[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v.x += __shfl_down(v.x, off, 64);
      v.y += __shfl_down(v.y, off, 64);
    }
    if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = v;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kDotGroup + 1; ++i) {
    double2 t = red[i][0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) { t.x += red[i][w].x; t.y += red[i][w].y; }
    res[i] = t;
  }
  __syncthreads();
}

// Finish Hessenberg column j (entries col[0..j+1]) and apply the Givens
// rotations; sets done when |g_{j+1}| <= target.  Called by all threads of
// one block: previous rotations are staged in LDS, thread 0 does the serial
// recurrence, the column is written back in parallel.
__device__ __forceinline__ void givens_column(GmresBufs gb, int j, const double2* col) {
  __shared__ double2 hc[kMaxRestart + 2], cs_s[kMaxRestart], sn_s[kMaxRestart];
  const int m = gb.m;
  GmresState* st = gb.st;
  for (int i = threadIdx.x; i <= j + 1; i += blockDim.x) hc[i] = col[i];
  for (int i = threadIdx.x; i < j; i += blockDim.x) { cs_s[i] = gb.cs[i]; sn_s[i] = gb.sn[i]; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < j; ++i) {
      const double2 c = cs_s[i], s = sn_s[i];
      const double2 a = hc[i], b = hc[i + 1];
      // G = [c s; -conj(s) conj(c)]
      const double2 ca = cmul(c, a), sb = cmul(s, b);
      const double2 sa = cmul(mk(s.x, -s.y), a), cb = cmul(mk(c.x, -c.y), b);
      hc[i] = mk(ca.x + sb.x, ca.y + sb.y);
      hc[i + 1] = mk(cb.x - sa.x, cb.y - sa.y);
    }
    const double2 a = hc[j], b = hc[j + 1];
    const double r = sqrt(a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y);
    double2 c, s;
    if (r == 0.0) { c = mk(1, 0); s = mk(0, 0); }
    else { c = mk(a.x / r, -a.y / r); s = mk(b.x / r, -b.y / r); }
    gb.cs[j] = c;
    gb.sn[j] = s;
    hc[j] = mk(r, 0);
    hc[j + 1] = mk(0, 0);
    const double2 gj = gb.g[j];
    const double2 t = cmul(mk(s.x, -s.y), gj);
    gb.g[j + 1] = mk(-t.x, -t.y);
    gb.g[j] = cmul(c, gj);
    const double gn = sqrt(t.x * t.x + t.y * t.y);
    st->resid = gn;
    st->jlast = j + 1;
    const int k = st->its_total + 1;
    st->its_total = k;
    if (gb.resv && k < gb.resv_cap) gb.resv[k] = gn;
    if (gn <= st->target || r == 0.0) st->done = 1;
    if (r == 0.0) st->breakdown = 1;
  }
  __syncthreads();
  double2* Hc = gb.H + (long)j * (m + 1);
  for (int i = threadIdx.x; i <= j + 1; i += blockDim.x) Hc[i] = hc[i];
}

// 16-byte packets of a complex vector: one double complex or two float
// complex values.
template <typename T> struct Pk;
template <> struct Pk<double> {
  using type = double2;
  static constexpr int per = 1;
  __device__ static __forceinline__ void dot_acc(double& ar, double& ai, type v, type w) { dot(ar, ai, v, w); }
  __device__ static __forceinline__ double nrm2_w(type w) { return w.x * w.x + w.y * w.y; }
  __device__ static __forceinline__ void dot(double& ar, double& ai, type v, type w) {
    ar += v.x * w.x + v.y * w.y;
    ai += v.x * w.y - v.y * w.x;
  }
  __device__ static __forceinline__ double nrm2(type w) { return w.x * w.x + w.y * w.y; }
  __device__ static __forceinline__ void axpy(type& w, double2 h, type v) {  // w -= h v
    w.x -= h.x * v.x - h.y * v.y;
    w.y -= h.x * v.y + h.y * v.x;
  }
  __device__ static __forceinline__ type scale(type v, double s) { return make_double2(v.x * s, v.y * s); }
};
template <> struct Pk<float> {
  using type = float4;
  static constexpr int per = 2;
  __device__ static __forceinline__ void dot_acc(float& ar, float& ai, type v, type w) {
    ar += v.x * w.x + v.y * w.y + v.z * w.z + v.w * w.w;
    ai += v.x * w.y - v.y * w.x + v.z * w.w - v.w * w.z;
  }
  __device__ static __forceinline__ float nrm2_w(type w) { return w.x * w.x + w.y * w.y + w.z * w.z + w.w * w.w; }
  __device__ static __forceinline__ void dot(double& ar, double& ai, type v, type w) {
    ar += (double)v.x * w.x + (double)v.y * w.y + (double)v.z * w.z + (double)v.w * w.w;
    ai += (double)v.x * w.y - (double)v.y * w.x + (double)v.z * w.w - (double)v.w * w.z;
  }
  __device__ static __forceinline__ double nrm2(type w) {
    return (double)w.x * w.x + (double)w.y * w.y + (double)w.z * w.z + (double)w.w * w.w;
  }
  __device__ static __forceinline__ void axpy(type& w, double2 h, type v) {
    const float hr = (float)h.x, hi = (float)h.y;
    w.x -= hr * v.x - hi * v.y;
    w.y -= hr * v.y + hi * v.x;
    w.z -= hr * v.z - hi * v.w;
    w.w -= hr * v.w + hi * v.z;
  }
  __device__ static __forceinline__ type scale(type v, double s) {
    const float f = (float)s;
    return make_float4(v.x * f, v.y * f, v.z * f, v.w * f);
  }
};

// Storage of the Krylov basis.  A basis packet holds the same complex values
// as one working packet Pk<T>::type: FP64/FP32 packets are stored as they are;
// the FP32 inner solver can instead keep its basis as BF16 packets (uint2 =
// 2 complex in 8 bytes, round to nearest even), halving the Gram-Schmidt
// traffic.  load() widens to the working packet, store() narrows.
__device__ __forceinline__ float bf16_bits_to_float(uint32_t h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ uint32_t float_to_bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}
template <typename VP>
struct BasisIO {
  using work = VP;
  static constexpr double raw_scale = 1.0;   // load() == raw_scale * load_raw()
  __device__ static __forceinline__ work load(VP v) { return v; }
  __device__ static __forceinline__ work load_raw(VP v) { return v; }
  __device__ static __forceinline__ VP store(work v) { return v; }
};
template <>
struct BasisIO<uint2> {
  using work = float4;
  static constexpr double raw_scale = 1.0;
  __device__ static __forceinline__ float4 load_raw(uint2 p) { return load(p); }
  __device__ static __forceinline__ float4 load(uint2 p) {
    return make_float4(bf16_bits_to_float(p.x & 0xFFFFu), bf16_bits_to_float(p.x >> 16),
                       bf16_bits_to_float(p.y & 0xFFFFu), bf16_bits_to_float(p.y >> 16));
  }
  __device__ static __forceinline__ uint2 store(float4 v) {
    uint2 p;
    p.x = float_to_bf16_bits(v.x) | (float_to_bf16_bits(v.y) << 16);
    p.y = float_to_bf16_bits(v.z) | (float_to_bf16_bits(v.w) << 16);
    return p;
  }
};

// FP16 basis packets: the unit-norm basis vectors are stored times 256
// (|entries| <= 256 < 65504, and entries down to 2.4e-7 stay normal).
template <>
struct BasisIO<ushort4> {
  using work = float4;
  static constexpr double raw_scale = 1.0 / 256.0;
  // the stored values times 256, unscaled: the multi-dot and the update fold
  // the 1/256 into their sums and coefficients
  __device__ static __forceinline__ float4 load_raw(ushort4 p) {
    return make_float4((float)__builtin_bit_cast(_Float16, p.x), (float)__builtin_bit_cast(_Float16, p.y),
                       (float)__builtin_bit_cast(_Float16, p.z), (float)__builtin_bit_cast(_Float16, p.w));
  }
  __device__ static __forceinline__ float4 load(ushort4 p) {
    const float s = 1.0f / 256.0f;
    return make_float4((float)__builtin_bit_cast(_Float16, p.x) * s, (float)__builtin_bit_cast(_Float16, p.y) * s,
                       (float)__builtin_bit_cast(_Float16, p.z) * s, (float)__builtin_bit_cast(_Float16, p.w) * s);
  }
  __device__ static __forceinline__ ushort4 store(float4 v) {
    ushort4 p;
    p.x = __builtin_bit_cast(unsigned short, (_Float16)(v.x * 256.0f));
    p.y = __builtin_bit_cast(unsigned short, (_Float16)(v.y * 256.0f));
    p.z = __builtin_bit_cast(unsigned short, (_Float16)(v.z * 256.0f));
    p.w = __builtin_bit_cast(unsigned short, (_Float16)(v.w * 256.0f));
    return p;
  }
};

// Two 8-byte basis packets (FP16 / BF16) read with one 16-byte load.
template <typename VP>
struct VPPair {
  VP a, b;
};

// Multi-dot: h[i] = <V_i, w> for i < nv (plus ||w||^2 when mode 0).
// Grid (blocks, ceil(nv/kDotGroup)); block row y handles basis vectors
// [G y, G y + G), G = kDotGroup.  mode 0 writes h1 and nrm0; mode 1 (only when st->reorth)
// writes h2.  NP = number of working packets per vector.  W = 2 (8-byte
// basis packets, NP even): every thread step reads two basis packets per
// vector with one 16-byte load and the two matching working packets of w.
template <typename T, typename VP, int NT, int W = 1>
__global__ __launch_bounds__(NT) void arnoldi_dot_kernel(GmresBufs gb, const VP* __restrict__ V,
                                                         const typename Pk<T>::type* __restrict__ w, long NP, int nv,
                                                         int mode) {
  using P = Pk<T>;
  using IO = BasisIO<VP>;
  if (gb.st->done || (mode == 1 && !gb.st->reorth)) return;
  const int i0 = blockIdx.y * kDotGroup;
  const int ng = min(kDotGroup, nv - i0);
  // per-thread sums in the working precision (a thread adds a few dozen
  // products), the block and grid sums in FP64
  T ar[kDotGroup + 1], ai[kDotGroup + 1];
#pragma unroll
  for (int i = 0; i < kDotGroup + 1; ++i) ar[i] = ai[i] = 0;
  const VP* V0 = V + (long)i0 * NP;
  if constexpr (W == 2) {
    static_assert(sizeof(VP) == 8 && sizeof(VPPair<VP>) == 16, "two 8-byte packets per 16-byte load");
    const long NQ = NP / 2;
    const uint4* V0q = reinterpret_cast<const uint4*>(V0);
    for (long x = (long)blockIdx.x * NT + threadIdx.x; x < NQ; x += (long)gridDim.x * NT) {
      const typename P::type w0 = w[2 * x], w1 = w[2 * x + 1];
      uint4 raw[kDotGroup];
#pragma unroll
      for (int i = 0; i < kDotGroup; ++i) raw[i] = (i < ng) ? V0q[(long)i * NQ + x] : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int i = 0; i < kDotGroup; ++i) {
        const VPPair<VP> pr = __builtin_bit_cast(VPPair<VP>, raw[i]);
        P::dot_acc(ar[i], ai[i], IO::load_raw(pr.a), w0);
        P::dot_acc(ar[i], ai[i], IO::load_raw(pr.b), w1);
      }
      ar[kDotGroup] += P::nrm2_w(w0) + P::nrm2_w(w1);
    }
  } else if (ng == kDotGroup) {
    // full group: the group's basis loads are issued together
    for (long x = (long)blockIdx.x * NT + threadIdx.x; x < NP; x += (long)gridDim.x * NT) {
      const typename P::type wv = w[x];
      VP raw[kDotGroup];
#pragma unroll
      for (int i = 0; i < kDotGroup; ++i) raw[i] = V0[(long)i * NP + x];
#pragma unroll
      for (int i = 0; i < kDotGroup; ++i) P::dot_acc(ar[i], ai[i], IO::load_raw(raw[i]), wv);
      ar[kDotGroup] += P::nrm2_w(wv);
    }
  } else {
    for (long x = (long)blockIdx.x * NT + threadIdx.x; x < NP; x += (long)gridDim.x * NT) {
      const typename P::type wv = w[x];
#pragma unroll
      for (int i = 0; i < kDotGroup; ++i)
        if (i < ng) P::dot_acc(ar[i], ai[i], IO::load_raw(V0[(long)i * NP + x]), wv);
      ar[kDotGroup] += P::nrm2_w(wv);
    }
  }
  double2 acc[kDotGroup + 1];
#pragma unroll
  for (int i = 0; i < kDotGroup; ++i) acc[i] = mk(IO::raw_scale * ar[i], IO::raw_scale * ai[i]);
  acc[kDotGroup] = mk(ar[kDotGroup], 0);
  const int stride = kDotGroup + 1;
  const long blin = (long)blockIdx.y * gridDim.x + blockIdx.x;
  block_reduce_store<NT>(acc, gb.part + blin * stride);
}

// Finalize of arnoldi_dot_kernel (one block per group of kDotGroup basis
// vectors): sums the group's per-block partials into h1 (+ nrm0) or h2.
template <int NT>
__global__ __launch_bounds__(NT) void arnoldi_dot_finalize_kernel(GmresBufs gb, int nblocks, int nv, int mode) {
  if (gb.st->done || (mode == 1 && !gb.st->reorth)) return;
  const int stride = kDotGroup + 1;
  double2* out = (mode == 0) ? gb.h1 : gb.h2;
  const int gy = blockIdx.x;
  const int gi0 = gy * kDotGroup;
  const int gn = min(kDotGroup, nv - gi0);
  double2 res[kDotGroup + 1];
  finalize_group<NT>(gb.part, (long)gy * nblocks, stride, nblocks, res);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < kDotGroup; ++i)
      if (i < gn) out[gi0 + i] = res[i];
    if (mode == 0 && gy == 0) gb.st->nrm0 = sqrt(res[kDotGroup].x);
  }
}

// w -= sum_i h_i V_i and ||w||^2.  mode 0 uses h1 and the last block either
// finishes the Hessenberg column (DGKS test ||w_new|| >= ||w||/sqrt(2)) or
// requests a second pass; mode 1 (only when st->reorth) uses h2 and finishes
// the column with h1 + h2.
template <typename T, typename VP, int NT, int W = 1>
__global__ __launch_bounds__(NT) void arnoldi_update_kernel(GmresBufs gb, const VP* __restrict__ V,
                                                            typename Pk<T>::type* __restrict__ w, long NP, int j,
                                                            int mode) {
  using P = Pk<T>;
  using IO = BasisIO<VP>;
  if (gb.st->done || (mode == 1 && !gb.st->reorth)) return;
  const int nv = j + 1;
  __shared__ double2 hs[kMaxRestart + 1];
  const double2* h = (mode == 0) ? gb.h1 : gb.h2;
  for (int i = threadIdx.x; i < nv; i += NT) hs[i] = mk(IO::raw_scale * h[i].x, IO::raw_scale * h[i].y);
  __syncthreads();
  double2 acc[kDotGroup + 1];
#pragma unroll
  for (int i = 0; i < kDotGroup + 1; ++i) acc[i] = mk(0, 0);
  if constexpr (W == 2) {
    // two 8-byte basis packets per 16-byte load (the dot kernel's W = 2)
    static_assert(sizeof(VP) == 8 && sizeof(VPPair<VP>) == 16, "two 8-byte packets per 16-byte load");
    const long NQ = NP / 2;
    const uint4* Vq = reinterpret_cast<const uint4*>(V);
    for (long x = (long)blockIdx.x * NT + threadIdx.x; x < NQ; x += (long)gridDim.x * NT) {
      typename P::type w0 = w[2 * x], w1 = w[2 * x + 1];
      int i = 0;
      for (; i + 8 <= nv; i += 8) {
        uint4 raw[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) raw[q] = Vq[(long)(i + q) * NQ + x];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const VPPair<VP> pr = __builtin_bit_cast(VPPair<VP>, raw[q]);
          P::axpy(w0, hs[i + q], IO::load_raw(pr.a));
          P::axpy(w1, hs[i + q], IO::load_raw(pr.b));
        }
      }
      for (; i < nv; ++i) {
        const VPPair<VP> pr = __builtin_bit_cast(VPPair<VP>, Vq[(long)i * NQ + x]);
        P::axpy(w0, hs[i], IO::load_raw(pr.a));
        P::axpy(w1, hs[i], IO::load_raw(pr.b));
      }
      w[2 * x] = w0;
      w[2 * x + 1] = w1;
      acc[0].x += P::nrm2(w0) + P::nrm2(w1);
    }
    block_reduce_store<NT>(acc, gb.part + (long)blockIdx.x * (kDotGroup + 1));
    return;
  }
  for (long x = (long)blockIdx.x * NT + threadIdx.x; x < NP; x += (long)gridDim.x * NT) {
    typename P::type wv = w[x];
    int i = 0;
    for (; i + 8 <= nv; i += 8) {   // eight basis loads in flight
      VP raw[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) raw[q] = V[(long)(i + q) * NP + x];
#pragma unroll
      for (int q = 0; q < 8; ++q) P::axpy(wv, hs[i + q], IO::load_raw(raw[q]));
    }
    for (; i + 4 <= nv; i += 4) {
      const typename P::type v0 = IO::load_raw(V[(long)(i + 0) * NP + x]);
      const typename P::type v1 = IO::load_raw(V[(long)(i + 1) * NP + x]);
      const typename P::type v2 = IO::load_raw(V[(long)(i + 2) * NP + x]);
      const typename P::type v3 = IO::load_raw(V[(long)(i + 3) * NP + x]);
      P::axpy(wv, hs[i + 0], v0);
      P::axpy(wv, hs[i + 1], v1);
      P::axpy(wv, hs[i + 2], v2);
      P::axpy(wv, hs[i + 3], v3);
    }
    for (; i < nv; ++i) P::axpy(wv, hs[i], IO::load_raw(V[(long)i * NP + x]));
    w[x] = wv;
    acc[0].x += P::nrm2(wv);
  }
  const int stride = kDotGroup + 1;
  block_reduce_store<NT>(acc, gb.part + (long)blockIdx.x * stride);
}

// Finalize of arnoldi_update_kernel (one block): ||w||, then either the
// DGKS request for a second pass (mode 0) or the Hessenberg column with its
// Givens rotation.
template <int NT>
__global__ __launch_bounds__(NT) void arnoldi_update_finalize_kernel(GmresBufs gb, int nblocks, int j, int mode) {
  if (gb.st->done || (mode == 1 && !gb.st->reorth)) return;
  const int nv = j + 1;
  const int stride = kDotGroup + 1;
  double2 res[kDotGroup + 1];
  finalize_group<NT>(gb.part, 0, stride, nblocks, res);
  const double nrm = sqrt(res[0].x);
  GmresState* st = gb.st;
  __shared__ double2 col[kMaxRestart + 2];
  __shared__ int finish;
  if (threadIdx.x == 0) {
    finish = 1;
    if (mode == 0 && !st->single_pass && nrm < 0.7071067811865476 * st->nrm0) {
      st->reorth = 1;
      st->reorth_count += 1;
      finish = 0;
    } else if (mode == 0) {
      st->reorth = 0;
    }
    st->hnext = nrm;
  }
  __syncthreads();
  if (finish) {
    for (int i = threadIdx.x; i < nv; i += NT)
      col[i] = (mode == 0) ? gb.h1[i] : mk(gb.h1[i].x + gb.h2[i].x, gb.h1[i].y + gb.h2[i].y);
    if (threadIdx.x == 0) col[nv] = mk(nrm, 0);
    __syncthreads();
    givens_column(gb, j, col);
  }
}

// Single-pass Gram-Schmidt: the finalize of arnoldi_update_kernel fused
// with the normalisation.  Every block sums the per-block norm partials in
// the same order, so all blocks hold the same ||w||; block 0 also finishes
// the Hessenberg column with its Givens rotation (st->done becomes visible to
// the kernels after this one), and every block scales its slice of w into
// V_{j+1} (and vcur for a narrowed basis).
template <typename T, typename VP, int NT>
__global__ __launch_bounds__(NT) void arnoldi_finish_kernel(GmresBufs gb, int nblocks, int j,
                                                            const typename Pk<T>::type* __restrict__ w,
                                                            VP* __restrict__ vnext,
                                                            typename Pk<T>::type* __restrict__ vcur, long NP,
                                                            int only_slow) {
  GmresState* st = gb.st;
  if (st->done || (only_slow && !st->slow_norm)) return;
  const int stride = kDotGroup + 1;
  __shared__ double wsum[NT / 64];
  __shared__ double2 col[kMaxRestart + 2];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += NT) acc += gb.part[(long)b * stride].x;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
  __syncthreads();
  double tot = wsum[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) tot += wsum[i];
  const double nrm = sqrt(tot);
  if (blockIdx.x == 0) {
    const int nv = j + 1;
    if (threadIdx.x == 0) {
      st->reorth = 0;
      st->hnext = nrm;
      col[nv] = mk(nrm, 0);
    }
    for (int i = threadIdx.x; i < nv; i += NT) col[i] = gb.h1[i];
    __syncthreads();
    givens_column(gb, j, col);
  }
  const double sc = (nrm > 0) ? 1.0 / nrm : 0.0;
  for (long x = (long)blockIdx.x * NT + threadIdx.x; x < NP; x += (long)gridDim.x * NT) {
    const typename Pk<T>::type v = Pk<T>::scale(w[x], sc);
    vnext[x] = BasisIO<VP>::store(v);
    if (vcur) vcur[x] = v;
  }
}

// Single-pass Gram-Schmidt with the norm by Pythagoras: for an orthonormal
// V, ||w - V h||^2 = ||w||^2 - ||h||^2, with ||w|| from the multi-dot.  When
// the difference keeps at least 1 % of ||w||^2 (H[j+1,j] >= ||w||/10, nearly
// every step: the ratio sits near 0.5 on these operators), every block writes
// V_{j+1} = (w - V h)/H[j+1,j] at once and block 0 finishes the Hessenberg
// column: w is neither written back nor read again by a finish pass.  V_{j+1}
// times H[j+1,j] equals w - V h to rounding whatever the error of the norm,
// so the Arnoldi relation holds exactly; the norm only scales the new
// column.  Below the threshold (cancellation) the kernel does the two-pass
// form instead: w -= V h with the norm partials, and arnoldi_finish_kernel
// (run with only_slow) normalises.  Every block takes the same branch (the
// same h and ||w||).
template <typename T, typename VP, int NT, int W = 1>
__global__ __launch_bounds__(NT) void arnoldi_update_pyth_kernel(GmresBufs gb, const VP* __restrict__ V,
                                                                 typename Pk<T>::type* __restrict__ w,
                                                                 VP* __restrict__ vnext,
                                                                 typename Pk<T>::type* __restrict__ vcur, long NP,
                                                                 int j) {
  using P = Pk<T>;
  using IO = BasisIO<VP>;
  GmresState* st = gb.st;
  if (st->done) return;
  const int nv = j + 1;
  __shared__ double2 hs[kMaxRestart + 1];
  __shared__ double2 col[kMaxRestart + 2];
  double hh = 0.0;
  for (int i = 0; i < nv; ++i) {
    const double2 hi = gb.h1[i];
    hh += hi.x * hi.x + hi.y * hi.y;
  }
  for (int i = threadIdx.x; i < nv; i += NT) hs[i] = mk(IO::raw_scale * gb.h1[i].x, IO::raw_scale * gb.h1[i].y);
  const double w2 = st->nrm0 * st->nrm0;
  const double b2 = w2 - hh;
  const bool fast = w2 > 0 && b2 >= 0.01 * w2;
  const double beta = fast ? sqrt(b2) : 0.0;
  __syncthreads();
  if (fast && blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      st->slow_norm = 0;
      st->reorth = 0;
      st->hnext = beta;
      col[nv] = mk(beta, 0);
    }
    for (int i = threadIdx.x; i < nv; i += NT) col[i] = gb.h1[i];
    __syncthreads();
    givens_column(gb, j, col);
  } else if (!fast && blockIdx.x == 0 && threadIdx.x == 0) {
    st->slow_norm = 1;
  }
  const double sc = fast ? 1.0 / beta : 0.0;
  double2 acc[kDotGroup + 1];
#pragma unroll
  for (int i = 0; i < kDotGroup + 1; ++i) acc[i] = mk(0, 0);
  if constexpr (W == 2) {
    static_assert(sizeof(VP) == 8 && sizeof(VPPair<VP>) == 16, "two 8-byte packets per 16-byte load");
    const long NQ = NP / 2;
    const uint4* Vq = reinterpret_cast<const uint4*>(V);
    for (long x = (long)blockIdx.x * NT + threadIdx.x; x < NQ; x += (long)gridDim.x * NT) {
      typename P::type w0 = w[2 * x], w1 = w[2 * x + 1];
      int i = 0;
      for (; i + 8 <= nv; i += 8) {
        uint4 raw[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) raw[q] = Vq[(long)(i + q) * NQ + x];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const VPPair<VP> pr = __builtin_bit_cast(VPPair<VP>, raw[q]);
          P::axpy(w0, hs[i + q], IO::load_raw(pr.a));
          P::axpy(w1, hs[i + q], IO::load_raw(pr.b));
        }
      }
      for (; i < nv; ++i) {
        const VPPair<VP> pr = __builtin_bit_cast(VPPair<VP>, Vq[(long)i * NQ + x]);
        P::axpy(w0, hs[i], IO::load_raw(pr.a));
        P::axpy(w1, hs[i], IO::load_raw(pr.b));
      }
      if (fast) {
        const typename P::type v0 = P::scale(w0, sc), v1 = P::scale(w1, sc);
        vnext[2 * x] = IO::store(v0);
        vnext[2 * x + 1] = IO::store(v1);
        if (vcur) {
          vcur[2 * x] = v0;
          vcur[2 * x + 1] = v1;
        }
      } else {
        w[2 * x] = w0;
        w[2 * x + 1] = w1;
        acc[0].x += P::nrm2(w0) + P::nrm2(w1);
      }
    }
  } else {
    for (long x = (long)blockIdx.x * NT + threadIdx.x; x < NP; x += (long)gridDim.x * NT) {
      typename P::type wv = w[x];
      int i = 0;
      for (; i + 8 <= nv; i += 8) {   // eight basis loads in flight
        VP raw[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) raw[q] = V[(long)(i + q) * NP + x];
#pragma unroll
        for (int q = 0; q < 8; ++q) P::axpy(wv, hs[i + q], IO::load_raw(raw[q]));
      }
      for (; i < nv; ++i) P::axpy(wv, hs[i], IO::load_raw(V[(long)i * NP + x]));
      if (fast) {
        const typename P::type v = P::scale(wv, sc);
        vnext[x] = IO::store(v);
        if (vcur) vcur[x] = v;
      } else {
        w[x] = wv;
        acc[0].x += P::nrm2(wv);
      }
    }
  }
  if (!fast) block_reduce_store<NT>(acc, gb.part + (long)blockIdx.x * (kDotGroup + 1));
}

// V_{j+1} = w / H[j+1,j] (and, for a narrowed basis, the working-precision
// copy vcur that the operator reads next); skipped once the cycle is done.
template <typename T, typename VP>
__global__ void arnoldi_normalize_kernel(GmresBufs gb, const typename Pk<T>::type* __restrict__ w,
                                         VP* __restrict__ vnext, typename Pk<T>::type* __restrict__ vcur, long NP) {
  if (gb.st->done) return;
  const double s = (gb.st->hnext > 0) ? 1.0 / gb.st->hnext : 0.0;
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < NP; x += (long)gridDim.x * blockDim.x) {
    const typename Pk<T>::type v = Pk<T>::scale(w[x], s);
    vnext[x] = BasisIO<VP>::store(v);
    if (vcur) vcur[x] = v;
  }
}

// v0 = r / ||r|| and Givens/rhs initialisation for a new cycle.
__global__ void solve_init_kernel(GmresState* st, double tol_abs, double inner_tol, int single_pass) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->tol_abs = tol_abs;
    st->inner_tol = inner_tol;
    st->single_pass = single_pass;
    st->its_total = 0;
  }
}

// Tolerances come from the state (set once per solve) so that a captured
// cycle graph is valid for any tolerance.
__global__ void cycle_init_kernel(GmresBufs gb) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  GmresState* st = gb.st;
  const double tol_abs = st->tol_abs;
  st->beta = st->resnorm;
  st->target = fmax(tol_abs, st->inner_tol * st->resnorm);
  st->resid = st->resnorm;
  st->done = (st->resnorm <= tol_abs) ? 1 : 0;
  st->jlast = 0;
  st->reorth = 0;
  st->breakdown = 0;
  st->ticket = 0;
  st->reorth_count = 0;
  st->nrm0 = st->resnorm;
  st->hnext = st->resnorm;
  for (int i = 0; i <= gb.m; ++i) gb.g[i] = mk(0, 0);
  gb.g[0] = mk(st->resnorm, 0);
}

// y = H(1:jl,1:jl) \ g(1:jl): H staged in LDS by the block, solved by thread 0.
__global__ void hess_solve_kernel(GmresBufs gb) {
  __shared__ double2 Hs[(kMaxRestart + 1) * kMaxRestart];
  __shared__ double2 ys[kMaxRestart];
  const int m = gb.m, jl = gb.st->jlast;
  for (int i = threadIdx.x; i < (m + 1) * jl; i += blockDim.x) Hs[i] = gb.H[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int r = jl - 1; r >= 0; --r) {
      double2 s = gb.g[r];
      for (int c = r + 1; c < jl; ++c) {
        const double2 p = cmul(Hs[c * (m + 1) + r], ys[c]);
        s.x -= p.x; s.y -= p.y;
      }
      const double2 d = Hs[r * (m + 1) + r];
      const double dd = d.x * d.x + d.y * d.y;
      ys[r] = (dd > 0) ? mk((s.x * d.x + s.y * d.y) / dd, (s.y * d.x - s.x * d.y) / dd) : mk(0, 0);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < jl; i += blockDim.x) gb.y[i] = ys[i];
}

// s = sum_{i < jlast} y_i V_i
// s = V y over the basis packets (FP64 accumulation).
__device__ __forceinline__ void mac_packet(double (&acc)[4], double2 y, float4 v) {
  acc[0] += y.x * (double)v.x - y.y * (double)v.y;
  acc[1] += y.x * (double)v.y + y.y * (double)v.x;
  acc[2] += y.x * (double)v.z - y.y * (double)v.w;
  acc[3] += y.x * (double)v.w + y.y * (double)v.z;
}
__device__ __forceinline__ void mac_packet(double (&acc)[4], double2 y, double2 v) {
  acc[0] += y.x * v.x - y.y * v.y;
  acc[1] += y.x * v.y + y.y * v.x;
}
__device__ __forceinline__ void put_packet(float4& o, const double (&acc)[4]) {
  o = make_float4((float)acc[0], (float)acc[1], (float)acc[2], (float)acc[3]);
}
__device__ __forceinline__ void put_packet(double2& o, const double (&acc)[4]) { o = make_double2(acc[0], acc[1]); }

template <typename T, typename VP>
__global__ void lincomb_kernel(GmresBufs gb, const VP* __restrict__ V, typename Pk<T>::type* __restrict__ s,
                               long NP) {
  const int jl = gb.st->jlast;
  __shared__ double2 ys[kMaxRestart];
  for (int i = threadIdx.x; i < jl; i += blockDim.x) ys[i] = gb.y[i];
  __syncthreads();
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < NP; x += (long)gridDim.x * blockDim.x) {
    double acc[4] = {0, 0, 0, 0};
    for (int i = 0; i < jl; ++i) mac_packet(acc, ys[i], BasisIO<VP>::load(V[(long)i * NP + x]));
    put_packet(s[x], acc);
  }
}

// x64 += alpha * d   (d in T)
template <typename T>
__global__ void axpy_to64_kernel(double2* __restrict__ x, const typename cvec<T>::type* __restrict__ d,
                                 double alpha, long N) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const auto v = d[i];
    x[i].x += alpha * (double)v.x;
    x[i].y += alpha * (double)v.y;
  }
}

// r = b - y (FP64) and ||r|| into st->resnorm; v0 = r/||r|| is written by
// residual_normalize_kernel afterwards.
template <int NT>
__global__ __launch_bounds__(NT) void residual_kernel(const double2* __restrict__ b, const double2* __restrict__ y,
                                                      double2* __restrict__ r, long N, double2* part,
                                                      GmresState* st) {
  double2 acc[kDotGroup + 1];
#pragma unroll
  for (int i = 0; i < kDotGroup + 1; ++i) acc[i] = mk(0, 0);
  for (long i = (long)blockIdx.x * NT + threadIdx.x; i < N; i += (long)gridDim.x * NT) {
    const double2 v = mk(b[i].x - y[i].x, b[i].y - y[i].y);
    r[i] = v;
    acc[0].x += v.x * v.x + v.y * v.y;
  }
  block_reduce_store<NT>(acc, part + (long)blockIdx.x * (kDotGroup + 1));
}

template <int NT>
__global__ __launch_bounds__(NT) void residual_finalize_kernel(const double2* part, int nblocks, GmresState* st) {
  double2 res[kDotGroup + 1];
  finalize_group<NT>(part, 0, kDotGroup + 1, nblocks, res);
  if (threadIdx.x == 0) st->resnorm = sqrt(res[0].x);
}

// v0 = r / ||r|| in T precision.
// v0 = r/||r|| into basis slot 0 (and its working-precision copy vcur).
__device__ __forceinline__ void scaled_packet(const double2* r, long x, double s, double2& o) {
  o = make_double2(r[x].x * s, r[x].y * s);
}
__device__ __forceinline__ void scaled_packet(const double2* r, long x, double s, float4& o) {
  const double2 a = r[2 * x], b = r[2 * x + 1];
  o = make_float4((float)(a.x * s), (float)(a.y * s), (float)(b.x * s), (float)(b.y * s));
}
template <typename T, typename VP>
__global__ void residual_normalize_kernel(const double2* __restrict__ r, VP* __restrict__ v0,
                                          typename Pk<T>::type* __restrict__ vcur, long NP, const GmresState* st) {
  const double s = (st->resnorm > 0) ? 1.0 / st->resnorm : 0.0;
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < NP; x += (long)gridDim.x * blockDim.x) {
    typename Pk<T>::type v;
    scaled_packet(r, x, s, v);
    v0[x] = BasisIO<VP>::store(v);
    if (vcur) vcur[x] = v;
  }
}

// ---------------------------------------------------------------------------
// Observables at xi = 0: the two phi rows around the centre (n_xi even)
// rho_u(l, node) = sum_j phi(l, j) rho(node, j).  One block per node.
// ---------------------------------------------------------------------------
__global__ void observables_kernel(const double2* __restrict__ rho, const double2* __restrict__ phi_rows,
                                   int nk, double2* __restrict__ out) {
  const int node = blockIdx.x;
  double2 a0 = mk(0, 0), a1 = mk(0, 0);
  for (int j = threadIdx.x; j < nk; j += blockDim.x) {
    const double2 v = rho[(long)node * nk + j];
    const double2 p0 = phi_rows[j], p1 = phi_rows[nk + j];
    const double2 t0 = cmul(p0, v), t1 = cmul(p1, v);
    a0.x += t0.x; a0.y += t0.y; a1.x += t1.x; a1.y += t1.y;
  }
  __shared__ double2 r0[256], r1[256];
  r0[threadIdx.x] = a0;
  r1[threadIdx.x] = a1;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      r0[threadIdx.x].x += r0[threadIdx.x + st].x; r0[threadIdx.x].y += r0[threadIdx.x + st].y;
      r1[threadIdx.x].x += r1[threadIdx.x + st].x; r1[threadIdx.x].y += r1[threadIdx.x + st].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * node] = r0[0];
    out[2 * node + 1] = r1[0];
  }
}

// ---------------------------------------------------------------------------
// Electrostatics (RTD Gummel loop).  One block: DG nodal density from the
// two centre phi rows, conservative projection onto the Poisson dual cells
// (CSR rows of the exact polynomial integrals), then one FP64 Poisson update
// M phi = q (n - Nd + Na) with Dirichlet contacts, 'direct' or Newton with
// the local Boltzmann Jacobian J = A_ff - diag(q^2 n/(kB T))
// (RTD/solve_Poisson2D_rtd.m:85-130, RTD/project_DG_density_RTD.m:1-37).
// The tridiagonal free-node system is solved by the Thomas algorithm.
// ---------------------------------------------------------------------------
struct PoissonDev {
  int nmat, nodes, method;
  double temp, q, kb, scale_n;   // scale_n: device solution -> density units
  const double2* obs;            // 2 complex per DG node (centre rho_u rows)
  const int* t_ptr;              // CSR of the projection (nmat + 1)
  const int* t_col;
  const double* t_val;
  const double* widths;
  const double* lo;              // M(k, k-1)
  const double* up;              // M(k, k+1)
  const double* dg;              // M(k, k)
  const double* doping;          // Nd - Na
  const double* phi_in;
  double bias;
  double* n_dg;                  // scratch, nodes
  double* n_cv;                  // projected density, nmat
  double* work;                  // 3 * nmat scratch (F, c', d')
  double* phi_out;
  double* info;                  // [relres_before, unmixed_step, charge_integral, dg_integral]
};

__global__ void poisson_newton_kernel(PoissonDev P) {
  const int t = threadIdx.x, nt = blockDim.x;
  for (int i = t; i < P.nodes; i += nt) {
    const double2 a = P.obs[2 * i], b = P.obs[2 * i + 1];
    P.n_dg[i] = P.scale_n * 0.5 * (a.x + b.x);
  }
  __syncthreads();
  for (int r = t; r < P.nmat; r += nt) {
    double s = 0.0;
    for (int p = P.t_ptr[r]; p < P.t_ptr[r + 1]; ++p) s += P.t_val[p] * P.n_dg[P.t_col[p]];
    P.n_cv[r] = s / P.widths[r];
  }
  __syncthreads();
  const int nm = P.nmat;
  double* F = P.work;
  double* cp = P.work + nm;
  double* dp = P.work + 2 * nm;
  __shared__ double red[4][256];
  double fmx = 0, qn = 0, qd = 0, av = 0, csum = 0;
  for (int k = t; k < nm; k += nt) {
    double v = P.phi_in[k];
    if (k == 0) v = 0.0;
    if (k == nm - 1) v = P.bias;
    csum += P.n_cv[k] * P.widths[k];
    if (k == 0 || k == nm - 1) { F[k] = 0.0; continue; }
    const double vl = (k - 1 == 0) ? 0.0 : P.phi_in[k - 1];
    const double vr = (k + 1 == nm - 1) ? P.bias : P.phi_in[k + 1];
    const double Av = P.lo[k] * vl + P.dg[k] * v + P.up[k] * vr;
    const double f = Av - P.q * (P.n_cv[k] - P.doping[k]);
    F[k] = f;
    fmx = fmax(fmx, fabs(f));
    qn = fmax(qn, fabs(P.q * P.n_cv[k]));
    qd = fmax(qd, fabs(P.q * P.doping[k]));
    av = fmax(av, fabs(Av));
  }
  red[0][t] = fmx; red[1][t] = qn; red[2][t] = qd; red[3][t] = av;
  __shared__ double red_sum[256];
  red_sum[t] = csum;
  __syncthreads();
  if (t == 0) {
    double m0 = 0, m1 = 0, m2 = 0, m3 = 0, cs = 0;
    for (int i = 0; i < nt; ++i) {
      m0 = fmax(m0, red[0][i]); m1 = fmax(m1, red[1][i]); m2 = fmax(m2, red[2][i]); m3 = fmax(m3, red[3][i]);
      cs += red_sum[i];
    }
    double dgs = 0.0;
    for (int r = 0; r < nm; ++r)
      for (int p = P.t_ptr[r]; p < P.t_ptr[r + 1]; ++p) dgs += P.t_val[p] * P.n_dg[P.t_col[p]];
    const double scale = fmax(fmax(m1, m2), fmax(m3, P.q));
    P.info[0] = m0 / scale;
    P.info[2] = cs;
    P.info[3] = dgs;
    // Thomas algorithm on the free nodes 1..nm-2: J delta = -F
    const double beta = (P.method == 1) ? P.q * P.q / (P.kb * P.temp) : 0.0;
    double dmax = 0.0;
    for (int k = 1; k <= nm - 2; ++k) {
      const double a = (k > 1) ? P.lo[k] : 0.0;
      const double b = P.dg[k] - beta * P.n_cv[k];
      const double c = (k < nm - 2) ? P.up[k] : 0.0;
      const double rhs = -F[k];
      if (k == 1) {
        cp[k] = c / b;
        dp[k] = rhs / b;
      } else {
        const double den = b - a * cp[k - 1];
        cp[k] = c / den;
        dp[k] = (rhs - a * dp[k - 1]) / den;
      }
    }
    double x = 0.0;
    P.phi_out[nm - 1] = P.bias;
    P.phi_out[0] = 0.0;
    for (int k = nm - 2; k >= 1; --k) {
      x = dp[k] - ((k < nm - 2) ? cp[k] * x : 0.0);
      P.phi_out[k] = P.phi_in[k] + x;
      dmax = fmax(dmax, fabs(x));
    }
    P.info[1] = dmax;
  }
}

__global__ void scale_complex_kernel(double2* __restrict__ x, double s, long N) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    x[i].x *= s;
    x[i].y *= s;
  }
}

// ---------------------------------------------------------------------------
// Shifted operator alpha I + beta A (Crank-Nicolson steps): y = alpha u + beta y.
template <typename T>
__global__ void shift_kernel(const typename cvec<T>::type* __restrict__ u, typename cvec<T>::type* __restrict__ y,
                             double alpha, double beta, long N, const GmresState* st) {
  if (skip(st)) return;
  const T a = (T)alpha, b = (T)beta;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const auto v = u[i];
    auto w = y[i];
    w.x = a * v.x + b * w.x;
    w.y = a * v.y + b * w.y;
    y[i] = w;
  }
}

// Crank-Nicolson right-hand side c = 2 x - (I + beta A) x + gamma b, with
// per-block partial sums of |c|^2.
template <int BS>
__global__ void cn_rhs_kernel(const double2* __restrict__ x, const double2* __restrict__ ax,
                              const double2* __restrict__ b, double gamma, double2* __restrict__ c, long N,
                              double* __restrict__ part) {
  __shared__ double sm[BS];
  double s = 0;
  for (long i = (long)blockIdx.x * BS + threadIdx.x; i < N; i += (long)gridDim.x * BS) {
    const double2 xv = x[i], av = ax[i], bv = b[i];
    const double2 cv = make_double2(2 * xv.x - av.x + gamma * bv.x, 2 * xv.y - av.y + gamma * bv.y);
    c[i] = cv;
    s += cv.x * cv.x + cv.y * cv.y;
  }
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int k = BS / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) sm[threadIdx.x] += sm[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Diagonal / row-abs bound of alpha I + beta A from those of A.
__global__ void shift_diag_kernel(double2* __restrict__ d, long N, double alpha, double beta, int rowabs) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    double2 v = d[i];
    if (rowabs) v = make_double2(fabs(alpha) + fabs(beta) * v.x, 0.0);
    else v = make_double2(alpha + beta * v.x, beta * v.y);
    d[i] = v;
  }
}

// ---------------------------------------------------------------------------
// Explicit RK4 of d rho/dt = c_b rhs - c_A A rho (DG_code/core/transient.m:12-16,
// DG_code/solve_transport_DG_2D_transient_SC_new.m:145-158) in device units:
// with rho = sigma x, sigma = ||b||/s_A, the ODE is dx/dt' = c_b b^ - c_A A^ x
// for t' = s_A t.  Stage s gets y = A^ tmp_s and forms k_s = c_b b^ - c_A y;
//   s = 1: acc = k1,    tmp = x + h/2 k1
//   s = 2: acc += 2k2,  tmp = x + h/2 k2
//   s = 3: acc += 2k3,  tmp = x + h k3
//   s = 4: x += h/6 (acc + k4)
// ---------------------------------------------------------------------------
__global__ void rk4_stage_kernel(int stage, const double2* __restrict__ b, const double2* __restrict__ y,
                                 double2* __restrict__ x, double2* __restrict__ tmp, double2* __restrict__ acc,
                                 double cb, double ca, double h, long N) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x) {
    const double2 bi = b[i], yi = y[i];
    const double kr = cb * bi.x - ca * yi.x, ki = cb * bi.y - ca * yi.y;
    if (stage == 4) {
      const double2 a = acc[i];
      double2 xi = x[i];
      xi.x += h / 6.0 * (a.x + kr);
      xi.y += h / 6.0 * (a.y + ki);
      x[i] = xi;
      continue;
    }
    double2 a;
    if (stage == 1) {
      a = mk(kr, ki);
    } else {
      a = acc[i];
      a.x += 2.0 * kr;
      a.y += 2.0 * ki;
    }
    acc[i] = a;
    const double c = (stage == 3) ? h : 0.5 * h;
    const double2 xi = x[i];
    tmp[i] = mk(xi.x + c * kr, xi.y + c * ki);
  }
}

// Precision conversion helpers.
__global__ void d2f_kernel(const double2* __restrict__ a, float2* __restrict__ b, double s, long N) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
    b[i] = make_float2((float)(a[i].x * s), (float)(a[i].y * s));
}

}  // namespace wdg
