// Full-2D operator: host side of the wdg2_* entry points (included by wdg.hip
// after the context definition).  Setup runs once per context / potential;
// every apply is a fixed sequence of kernel launches on the context stream,
// so it can be captured inside the GMRES cycle graphs.
#pragma once

#include "full2d.hpp"

struct F2Data {
  wdg2_desc d{};
  int nxd = 0, nyd = 0, nrel = 0, ncen = 0;
  double sA = 1;                        // device operator = A / sA
  // banded central / identity parts merged per axis (coefficients folded in, / sA)
  std::vector<double> hxc, hxi, hyc, hyi;
  DevBuf<double> d_bxc, d_bxi, d_byc, d_byi;
  // dense parts
  struct Dense {
    F2Dense k{};
    int n = 0;                          // relative size along the axis
    int lines = 0;                      // workgroups per launch
    DevBuf<double> blocks, lift;
    DevBuf<int> mask;
    DevBuf<double2> R;
  };
  std::vector<std::unique_ptr<Dense>> dense;
  // Kronecker diagonal / row-abs factors of every part (/ sA)
  int n_kron = 0;
  DevBuf<double> d_dc, d_ac;
  DevBuf<double2> d_dr, d_ar;
  // potential
  F2Const hc{};
  DevBuf<F2Const> d_c;
  DevBuf<int> d_tix, d_tiy;
  DevBuf<double> d_ttx, d_tty, d_v, d_capx, d_capy;
  DevBuf<double> d_dv64;                // stored potential differences (FV-consistent, WDG2_DV_TABLE)
  int dv_mode = WDG2_DV_TABLE;          // wdg2_set_dv_mode
  long nv = 0;                          // material grid points
  DevBuf<float> d_dv32;
  bool fv = true;
  size_t lds_apply64 = 0, lds_apply32 = 0, lds_dense64 = 0, lds_dense32 = 0;
  // preconditioner
  int precond = WDG2_PRECOND_NONE;
  DevBuf<double2> d_p64;
  DevBuf<float2> d_p32;
  DevBuf<unsigned long long> d_count;
  bool have_rhs = false;
  // line-transverse preconditioner (mode transforms, X sweeps, transverse inverses)
  bool have_line = false;
  int nT = 0;
  DevBuf<int> d_lfwd;
  DevBuf<double> d_lxi64, d_lxc64;
  DevBuf<float> d_lxi32, d_lxc32;
  DevBuf<double2> d_ltinv64, d_la64, d_lb64;
  DevBuf<float2> d_ltinv32, d_la32, d_lb32;
  Dense lt_fwd64, lt_bwd64, lt_fwd32, lt_bwd32;   // U^H and U sA/sigma along rho_x
};

namespace {

void clear_graphs(wdg_ctx* c);

constexpr size_t kF2MaxLds = 160 * 1024;

inline bool f2_ctx(const wdg_ctx* c) { return c && c->f2; }

void f2_require(const wdg_ctx* c) {
  if (!f2_ctx(c)) throw std::invalid_argument("not a full-2D context");
}

// ---- kernel dispatch over the element node counts ----
template <typename T, int MODE, int NLX, int NLY, int NQ>
void f2_launch_apply_q(wdg_ctx* c, const typename F2Vec<T>::T2* u, typename F2Vec<T>::T2* y,
                       const GmresState* st, size_t lds, double sa, double sb) {
  F2Data& f = *c->f2;
  auto k = f2_apply_kernel<NLX, NLY, T, MODE, NQ>;
  static bool attr = false;   // per instantiation: raise the dynamic LDS limit once
  if (!attr) {
    HIPCHK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kF2MaxLds));
    attr = true;
  }
  const unsigned grid = (unsigned)(f.d.ne_x * f.d.ne_y * f.hc.tiles);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kF2Block), lds, c->stream, f.d_c.p, u, y, st, sa, sb);
}

template <typename T, int MODE, int NLX, int NLY>
void f2_launch_apply_n(wdg_ctx* c, const typename F2Vec<T>::T2* u, typename F2Vec<T>::T2* y,
                       const GmresState* st, size_t lds, double sa, double sb) {
  // at most 4 quadrature points per axis (the default up to p = 2): the
  // register-lean instance; the setup-time diagonal and row-sum modes use the
  // general one
  if constexpr (MODE == F2_APPLY) {
    if (c->f2->hc.nqx <= 4 && c->f2->hc.nqy <= 4) {
      f2_launch_apply_q<T, MODE, NLX, NLY, 4>(c, u, y, st, lds, sa, sb);
      return;
    }
  }
  f2_launch_apply_q<T, MODE, NLX, NLY, kF2MaxNq>(c, u, y, st, lds, sa, sb);
}

template <typename T, int MODE, int NLX>
void f2_launch_apply_x(wdg_ctx* c, const typename F2Vec<T>::T2* u, typename F2Vec<T>::T2* y,
                       const GmresState* st, size_t lds, double sa, double sb) {
  switch (c->f2->d.nl_y) {
    case 2: f2_launch_apply_n<T, MODE, NLX, 2>(c, u, y, st, lds, sa, sb); break;
    case 3: f2_launch_apply_n<T, MODE, NLX, 3>(c, u, y, st, lds, sa, sb); break;
    case 4: f2_launch_apply_n<T, MODE, NLX, 4>(c, u, y, st, lds, sa, sb); break;
    default: throw std::invalid_argument("nl_y outside 2..4");
  }
}

// sa, sb: y = sa u + sb (apply part) for a shifted operator (F2_APPLY only)
template <typename T, int MODE>
void f2_launch_apply(wdg_ctx* c, const typename F2Vec<T>::T2* u, typename F2Vec<T>::T2* y, const GmresState* st,
                     double sa = 0.0, double sb = 1.0) {
  const size_t lds = std::is_same<T, double>::value ? c->f2->lds_apply64 : c->f2->lds_apply32;
  switch (c->f2->d.nl_x) {
    case 2: f2_launch_apply_x<T, MODE, 2>(c, u, y, st, lds, sa, sb); break;
    case 3: f2_launch_apply_x<T, MODE, 3>(c, u, y, st, lds, sa, sb); break;
    case 4: f2_launch_apply_x<T, MODE, 4>(c, u, y, st, lds, sa, sb); break;
    default: throw std::invalid_argument("nl_x outside 2..4");
  }
}

// Relative cells across the axis per workgroup of a dense part: enough
// workgroups to fill the chip, but at least enough columns that the product
// phase gives every thread of the workgroup an item (each workgroup stages the
// whole n x n factor, so thin chunks re-read it for little work), within the
// LDS budget.  *rglobal = 1 when the factor does not fit in LDS beside one
// column and is read from global memory.
int f2_dense_chunk(int n, int ncol, int lines, int axis, int* rglobal) {
  const int want = std::max(1, (2048 + lines - 1) / lines);
  int chunk = std::max(1, (ncol + want - 1) / want);
  // product items: axis 0 has n rows x (chunk/4) column quads, axis 1 chunk
  // columns x (n/4) row quads
  const int fill = axis == 0 ? 4 * ((kF2Block + n - 1) / n) : (kF2Block + (n + 3) / 4 - 1) / ((n + 3) / 4);
  chunk = std::max(chunk, std::min(ncol, std::max(4, fill)));
  *rglobal = f2_dense_lds(n, 1, sizeof(double2)) > kF2MaxLds ? 1 : 0;
  while (chunk > 1 && f2_dense_lds(n, chunk, sizeof(double2), *rglobal) > kF2MaxLds) chunk = (chunk + 1) / 2;
  if (f2_dense_lds(n, chunk, sizeof(double2), *rglobal) > kF2MaxLds)
    throw std::invalid_argument("n_rho too large for one relative line in one workgroup's LDS");
  return chunk;
}

template <typename T>
void f2_launch_dense(wdg_ctx* c, const F2Data::Dense& dp, const typename F2Vec<T>::T2* u, typename F2Vec<T>::T2* y,
                     const GmresState* st, double scale = 1.0) {
  F2Data& f = *c->f2;
  static bool attr = false;
  if (!attr) {
    HIPCHK(hipFuncSetAttribute((const void*)f2_dense_kernel<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kF2MaxLds));
    HIPCHK(hipFuncSetAttribute((const void*)f2_dense_kernel<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kF2MaxLds));
    attr = true;
  }
  const int ncol = dp.k.axis == 0 ? f.d.n_rho_y : f.d.n_rho_x;
  const size_t lds = f2_dense_lds(dp.n, dp.k.chunk, sizeof(typename F2Vec<T>::T2), dp.k.rglobal != 0);
  const dim3 grid(dp.lines, (ncol + dp.k.chunk - 1) / dp.k.chunk);
  F2Dense k = dp.k;
  k.coef *= scale;
  if (k.rglobal)
    hipLaunchKernelGGL((f2_dense_kernel<T, true>), grid, dim3(kF2Block), lds, c->stream, f.d_c.p, k, u, y, st);
  else
    hipLaunchKernelGGL((f2_dense_kernel<T, false>), grid, dim3(kF2Block), lds, c->stream, f.d_c.p, k, u, y, st);
}

// y = A_hat u, or (sa I + sb A_hat) u: the shift rides on the apply kernel
// (which holds u) and on the dense parts' coefficients
template <typename T>
void f2_apply(wdg_ctx* c, const typename F2Vec<T>::T2* u, typename F2Vec<T>::T2* y, const GmresState* st,
              double sa = 0.0, double sb = 1.0) {
  F2Data& f = *c->f2;
  const double vb = (double)c->N * sizeof(typename F2Vec<T>::T2);
  {
    ProfScope ps(c, P_F2_APPLY, 2.0 * vb);
    f2_launch_apply<T, F2_APPLY>(c, u, y, st, sa, sb);
  }
  for (auto& dp : f.dense) {
    // a face part reads one trace line and updates nl lines per transverse DOF
    const double lines = dp->k.face < 0 ? (double)f.ncen * 3 : (double)dp->lines * (1 + 2 * dp->k.nl);
    ProfScope ps(c, P_F2_DENSE, lines * f.nrel * sizeof(typename F2Vec<T>::T2));
    f2_launch_dense<T>(c, *dp, u, y, st, sb);
  }
}

template <int NL, typename T>
void f2_line_sweep(wdg_ctx* c, typename F2Vec<T>::T2* a, const GmresState* st) {
  F2Data& f = *c->f2;
  const long lines = (long)f.d.n_rho_x * f.d.n_rho_y * f.nyd;
  const T* xi = std::is_same<T, double>::value ? (const T*)f.d_lxi64.p : (const T*)f.d_lxi32.p;
  const T* xc = std::is_same<T, double>::value ? (const T*)f.d_lxc64.p : (const T*)f.d_lxc32.p;
  hipLaunchKernelGGL((f2_line_sweep_kernel<NL, T>), dim3((unsigned)((lines + 255) / 256)), dim3(256), 0, c->stream,
                     f.d_c.p, a, xi, xc, f.d_lfwd.p, st);
}

// z = s_A M^-1 r (device units) for the line-transverse M.
template <typename T>
void f2_line_apply(wdg_ctx* c, const typename F2Vec<T>::T2* r, typename F2Vec<T>::T2* z, const GmresState* st) {
  using T2 = typename F2Vec<T>::T2;
  F2Data& f = *c->f2;
  const bool dbl = std::is_same<T, double>::value;
  T2* a = dbl ? (T2*)f.d_la64.p : (T2*)f.d_la32.p;
  T2* b = dbl ? (T2*)f.d_lb64.p : (T2*)f.d_lb32.p;
  const T2* tinv = dbl ? (const T2*)f.d_ltinv64.p : (const T2*)f.d_ltinv32.p;
  const double vb = (double)c->N * sizeof(T2);
  {
    ProfScope ps(c, P_F2_LINE, 2.0 * vb);
    f2_launch_dense<T>(c, dbl ? f.lt_fwd64 : f.lt_fwd32, r, a, st);
  }
  {
    ProfScope ps(c, P_F2_LINE, 2.0 * vb);
    switch (f.d.nl_x) {
      case 2: f2_line_sweep<2, T>(c, a, st); break;
      case 3: f2_line_sweep<3, T>(c, a, st); break;
      default: f2_line_sweep<4, T>(c, a, st); break;
    }
  }
  {
    ProfScope ps(c, P_F2_LINE, 2.0 * vb + (double)f.d.n_rho_x * f.nT * f.nT * sizeof(T2));
    const dim3 grid((f.nT + 63) / 64, (f.nxd + 15) / 16, f.d.n_rho_x);
    hipLaunchKernelGGL(f2_transverse_kernel<T>, grid, dim3(256), 0, c->stream, f.d_c.p, a, b, tinv, st);
  }
  {
    ProfScope ps(c, P_F2_LINE, 2.0 * vb);
    f2_launch_dense<T>(c, dbl ? f.lt_bwd64 : f.lt_bwd32, b, z, st);
  }
}

template <typename T>
void f2_precond(wdg_ctx* c, const typename F2Vec<T>::T2* r, typename F2Vec<T>::T2* z, const GmresState* st) {
  F2Data& f = *c->f2;
  if (f.precond == WDG2_PRECOND_LINE) {
    f2_line_apply<T>(c, r, z, st);
    return;
  }
  const void* p;
  if (f.precond == WDG2_PRECOND_NONE) {
    // identity: the solve skips the preconditioner; the host entry point copies
    HIPCHK(hipMemcpyAsync(z, r, (size_t)c->N * sizeof(typename F2Vec<T>::T2), hipMemcpyDeviceToDevice, c->stream));
    return;
  }
  if (std::is_same<T, double>::value) p = f.d_p64.p;
  else p = f.d_p32.p;
  if (f.precond == WDG2_PRECOND_ROWABS) {   // real scale
    hipLaunchKernelGGL(f2_scale_real_kernel<T>, dim3(c->red_blocks), dim3(256), 0, c->stream, (const T*)p, r, z, c->N,
                       st);
    return;
  }
  hipLaunchKernelGGL(f2_scale_kernel<T>, dim3(c->red_blocks), dim3(256), 0, c->stream,
                     (const typename F2Vec<T>::T2*)p, r, z, c->N, st);
}

// Whether the full-2D preconditioner can read a 16-bit Krylov basis vector in
// place (the elementwise scalings), so that the FP32 solver keeps no FP32
// copy of the newest basis vector.
bool f2_precond_reads_basis(const wdg_ctx* c) {
  return c->f2 && c->N % 2 == 0 &&
         (c->f2->precond == WDG2_PRECOND_ROWABS || c->f2->precond == WDG2_PRECOND_JACOBI);
}

template <typename VP>
void f2_precond_basis(wdg_ctx* c, const VP* v, float2* z, const GmresState* st) {
  F2Data& f = *c->f2;
  // a basis packet holds two entries (f2_precond_reads_basis requires an even N)
  const long npairs = c->N / 2;
  if (f.precond == WDG2_PRECOND_ROWABS)
    hipLaunchKernelGGL((f2_scale_basis_kernel<VP, true>), dim3(c->red_blocks), dim3(256), 0, c->stream,
                       (const void*)f.d_p32.p, v, reinterpret_cast<float4*>(z), npairs, st);
  else
    hipLaunchKernelGGL((f2_scale_basis_kernel<VP, false>), dim3(c->red_blocks), dim3(256), 0, c->stream,
                       (const void*)f.d_p32.p, v, reinterpret_cast<float4*>(z), npairs, st);
}

// ---- setup ----
std::vector<double> f2_merge_banded(const std::vector<const wdg2_part*>& ps, int nl, int ne, double sA) {
  const int sb = nl * nl + 2 * nl;
  std::vector<double> out((size_t)ne * sb, 0.0);
  for (const wdg2_part* p : ps)
    for (int e = 0; e < ne; ++e) {
      for (int i = 0; i < nl * nl; ++i) out[(size_t)e * sb + i] += p->coef * p->c_diag[(size_t)e * nl * nl + i] / sA;
      for (int a = 0; a < nl; ++a) {
        out[(size_t)e * sb + nl * nl + a] += p->coef * p->c_lower[(size_t)e * nl + a] / sA;
        out[(size_t)e * sb + nl * nl + nl + a] += p->coef * p->c_upper[(size_t)e * nl + a] / sA;
      }
    }
  return out;
}

void f2_create_impl(wdg_ctx* c, const wdg2_desc& d, const wdg2_part* parts, int np) {
  if (d.n_rho_x < 2 || d.n_rho_y < 2) throw std::invalid_argument("n_rho_x and n_rho_y must be at least 2");
  if (d.nl_x < 2 || d.nl_x > kF2MaxNl || d.nl_y < 2 || d.nl_y > kF2MaxNl)
    throw std::invalid_argument("nl_x and nl_y must be in 2..4");
  if (d.ne_x < 1 || d.ne_y < 1) throw std::invalid_argument("ne_x and ne_y must be positive");
  if (!(d.delta_x > 0) || !(d.delta_y > 0)) throw std::invalid_argument("delta_x and delta_y must be positive");
  if (d.periodic_y && (d.ne_y < 2 || d.n_rho_y < 3))
    throw std::invalid_argument("periodic Y needs at least two Y elements and three rho_y cells");
  if (np < 0 || (np > 0 && !parts)) throw std::invalid_argument("parts missing");
  auto f = std::make_shared<F2Data>();
  f->d = d;
  f->nxd = d.nl_x * d.ne_x;
  f->nyd = d.nl_y * d.ne_y;
  f->nrel = d.n_rho_x * d.n_rho_y;
  f->ncen = f->nxd * f->nyd;
  const long N = (long)f->nrel * f->ncen;
  if (N > (1L << 40)) throw std::invalid_argument("problem too large");
  // the dense kernel keeps one relative line in LDS (and the n x n factor when it fits)
  const int nmax = std::max(d.n_rho_x, d.n_rho_y);
  if (f2_dense_lds(nmax, 1, sizeof(double2), true) > kF2MaxLds)
    throw std::invalid_argument("n_rho too large: one relative line exceeds one workgroup's LDS");
  // operator scale: the largest |coef| * max|C| * max|R| over the parts
  double sA = 0;
  std::vector<const wdg2_part*> xc, xi, yc, yi;
  for (int i = 0; i < np; ++i) {
    const wdg2_part& p = parts[i];
    if (p.axis != 0 && p.axis != 1) throw std::invalid_argument("part axis must be X or Y");
    const int nl = p.axis == 0 ? d.nl_x : d.nl_y, ne = p.axis == 0 ? d.ne_x : d.ne_y;
    const int nr = p.axis == 0 ? d.n_rho_x : d.n_rho_y;
    if (!std::isfinite(p.coef)) throw std::invalid_argument("part coefficient must be finite");
    double cmax = 0, rmax = 0;
    if (p.center == WDG2_CENTER_BANDED) {
      if (!p.c_diag || !p.c_lower || !p.c_upper) throw std::invalid_argument("banded part without blocks");
      for (long k = 0; k < (long)ne * nl * nl; ++k) cmax = std::max(cmax, std::abs(p.c_diag[k]));
      for (long k = 0; k < (long)ne * nl; ++k)
        cmax = std::max(cmax, std::max(std::abs(p.c_lower[k]), std::abs(p.c_upper[k])));
    } else if (p.center == WDG2_CENTER_FACE) {
      if (!p.c_diag) throw std::invalid_argument("face part without lift");
      if (p.face != 0 && p.face != 1) throw std::invalid_argument("face must be 0 or 1");
      if (p.rel != WDG2_REL_DENSE) throw std::invalid_argument("face parts need a dense relative factor");
      for (int a = 0; a < nl; ++a) cmax = std::max(cmax, std::abs(p.c_diag[a]));
    } else {
      throw std::invalid_argument("unknown centre factor");
    }
    if (p.rel == WDG2_REL_CENTRAL) rmax = 0.5 / (p.axis == 0 ? d.delta_x : d.delta_y);
    else if (p.rel == WDG2_REL_IDENTITY) rmax = 1;
    else if (p.rel == WDG2_REL_DENSE) {
      if (!p.r_dense) throw std::invalid_argument("dense part without matrix");
      for (long k = 0; k < 2L * nr * nr; ++k) {
        if (!std::isfinite(p.r_dense[k])) throw std::invalid_argument("dense factor must be finite");
        rmax = std::max(rmax, std::abs(p.r_dense[k]));
      }
    } else {
      throw std::invalid_argument("unknown relative factor");
    }
    sA = std::max(sA, std::abs(p.coef) * cmax * rmax);
    if (p.center == WDG2_CENTER_BANDED && p.rel == WDG2_REL_CENTRAL) (p.axis == 0 ? xc : yc).push_back(&p);
    if (p.center == WDG2_CENTER_BANDED && p.rel == WDG2_REL_IDENTITY) (p.axis == 0 ? xi : yi).push_back(&p);
  }
  if (!(sA > 0)) sA = 1;
  f->sA = sA;
  c->g.scale = sA;
  c->f2 = f;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  auto upd = [&](DevBuf<double>& b, const std::vector<double>& v) {
    b.alloc(v.size());
    HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  };
  if (!xc.empty()) { f->hxc = f2_merge_banded(xc, d.nl_x, d.ne_x, sA); upd(f->d_bxc, f->hxc); }
  if (!xi.empty()) { f->hxi = f2_merge_banded(xi, d.nl_x, d.ne_x, sA); upd(f->d_bxi, f->hxi); }
  if (!yc.empty()) { f->hyc = f2_merge_banded(yc, d.nl_y, d.ne_y, sA); upd(f->d_byc, f->hyc); }
  if (!yi.empty()) { f->hyi = f2_merge_banded(yi, d.nl_y, d.ne_y, sA); upd(f->d_byi, f->hyi); }
  // dense parts and the Kronecker diagonal / row-abs factors of every part
  const int nrx = d.n_rho_x, nry = d.n_rho_y, nrel = f->nrel, ncen = f->ncen, nxd = f->nxd;
  // dense face parts with the same axis, face, lift and mask run as one launch
  // with the sum of their (scaled) relative factors
  std::vector<int> owner(np, -1);
  std::vector<std::vector<double2>> mergedR(np);
  for (int i = 0; i < np; ++i) {
    const wdg2_part& p = parts[i];
    if (p.rel != WDG2_REL_DENSE) continue;
    const int nl = p.axis == 0 ? d.nl_x : d.nl_y, nr = p.axis == 0 ? nrx : nry;
    const int ntr = p.axis == 0 ? f->nyd : f->nxd;
    owner[i] = i;
    if (p.center == WDG2_CENTER_FACE)
      for (int j = 0; j < i; ++j) {
        const wdg2_part& q = parts[j];
        if (owner[j] != j || q.center != WDG2_CENTER_FACE || q.axis != p.axis || q.face != p.face) continue;
        if (!std::equal(p.c_diag, p.c_diag + nl, q.c_diag)) continue;
        if ((p.mask == nullptr) != (q.mask == nullptr)) continue;
        if (p.mask && !std::equal(p.mask, p.mask + ntr, q.mask)) continue;
        owner[i] = j;
        break;
      }
    auto& R = mergedR[owner[i]];
    R.resize((size_t)nr * nr, make_double2(0, 0));
    const double cs = p.coef / sA;
    for (size_t k = 0; k < R.size(); ++k) {
      R[k].x += cs * p.r_dense[2 * k];
      R[k].y += cs * p.r_dense[2 * k + 1];
    }
  }
  std::vector<double> dc, ac;
  std::vector<double2> dr, ar;
  for (int i = 0; i < np; ++i) {
    const wdg2_part& p = parts[i];
    const int ax = p.axis;
    const int nl = ax == 0 ? d.nl_x : d.nl_y, ne = ax == 0 ? d.ne_x : d.ne_y;
    const int nr = ax == 0 ? nrx : nry;
    const int ntr = ax == 0 ? f->nyd : f->nxd;     // transverse centre DOFs
    if (p.rel == WDG2_REL_DENSE && owner[i] == i) {
      auto dp = std::make_unique<F2Data::Dense>();
      dp->n = nr;
      dp->k.axis = ax;
      dp->k.face = p.center == WDG2_CENTER_FACE ? p.face : -1;
      dp->k.nl = nl;
      dp->k.ne = ne;
      dp->k.coef = 1.0;
      dp->k.periodic = (ax == 1 && d.periodic_y) ? 1 : 0;
      if (d.periodic_y && ax == 1 && p.center == WDG2_CENTER_FACE)
        throw std::invalid_argument("a periodic Y axis has no faces");
      const std::vector<double2>& R = mergedR[i];
      dp->R.alloc(R.size());
      HIPCHK(hipMemcpyAsync(dp->R.p, R.data(), R.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
      dp->k.R = dp->R.p;
      if (p.center == WDG2_CENTER_BANDED) {
        const int sb = nl * nl + 2 * nl;
        std::vector<double> b((size_t)ne * sb);
        for (int e = 0; e < ne; ++e) {
          for (int k = 0; k < nl * nl; ++k) b[(size_t)e * sb + k] = p.c_diag[(size_t)e * nl * nl + k];
          for (int a = 0; a < nl; ++a) {
            b[(size_t)e * sb + nl * nl + a] = p.c_lower[(size_t)e * nl + a];
            b[(size_t)e * sb + nl * nl + nl + a] = p.c_upper[(size_t)e * nl + a];
          }
        }
        upd(dp->blocks, b);
        dp->k.blocks = dp->blocks.p;
        dp->lines = ncen;
        dp->k.chunk = f2_dense_chunk(nr, ax == 0 ? nry : nrx, dp->lines, ax, &dp->k.rglobal);
      } else {
        upd(dp->lift, std::vector<double>(p.c_diag, p.c_diag + nl));
        dp->k.lift = dp->lift.p;
        if (p.mask) {
          std::vector<int> m(p.mask, p.mask + ntr);
          dp->mask.alloc(m.size());
          HIPCHK(hipMemcpyAsync(dp->mask.p, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
          dp->k.mask = dp->mask.p;
        }
        dp->lines = ntr;
        dp->k.chunk = f2_dense_chunk(nr, ax == 0 ? nry : nrx, dp->lines, ax, &dp->k.rglobal);
      }
      HIPCHK(hipStreamSynchronize(c->stream));   // host vectors above go out of scope
      f->dense.push_back(std::move(dp));
    }
    // centre diagonal / row sums
    std::vector<double> vdc(ncen, 0.0), vac(ncen, 0.0);
    for (int i_c = 0; i_c < ncen; ++i_c) {
      const int ixd = i_c % nxd, iyd = i_c / nxd;
      const int k = ax == 0 ? ixd : iyd, t = ax == 0 ? iyd : ixd;
      const int e = k / nl, a = k - e * nl;
      if (p.center == WDG2_CENTER_BANDED) {
        vdc[i_c] = p.c_diag[(size_t)e * nl * nl + a * nl + a];
        double s = 0;
        for (int b = 0; b < nl; ++b) s += std::abs(p.c_diag[(size_t)e * nl * nl + a * nl + b]);
        const bool wrap = ax == 1 && d.periodic_y;
        if (e > 0 || wrap) s += std::abs(p.c_lower[(size_t)e * nl + a]);
        if (e < ne - 1 || wrap) s += std::abs(p.c_upper[(size_t)e * nl + a]);
        vac[i_c] = s;
      } else {
        const bool in_face = p.face == 0 ? e == 0 : e == ne - 1;
        const bool on = in_face && (!p.mask || p.mask[t]);
        if (on) {
          const int trace = p.face == 0 ? 0 : nl - 1;
          if (a == trace) vdc[i_c] = p.c_diag[a];
          vac[i_c] = std::abs(p.c_diag[a]);
        }
      }
      vdc[i_c] *= p.coef / sA;
      vac[i_c] *= std::abs(p.coef) / sA;
    }
    std::vector<double2> vdr(nrel), var(nrel);
    for (int r = 0; r < nrel; ++r) {
      const int rx = r % nrx, ry = r / nrx;
      const int k = ax == 0 ? rx : ry;
      double2 dv = make_double2(0, 0);
      double av = 0;
      if (p.rel == WDG2_REL_CENTRAL) {
        av = 0.5 / (ax == 0 ? d.delta_x : d.delta_y) * ((ax == 1 && d.periodic_y) ? 2 : (k > 0) + (k < nr - 1));
      } else if (p.rel == WDG2_REL_IDENTITY) {
        dv = make_double2(1, 0);
        av = 1;
      } else {
        dv = make_double2(p.r_dense[2 * ((size_t)k * nr + k)], p.r_dense[2 * ((size_t)k * nr + k) + 1]);
        for (int j = 0; j < nr; ++j)
          av += std::hypot(p.r_dense[2 * ((size_t)k * nr + j)], p.r_dense[2 * ((size_t)k * nr + j) + 1]);
      }
      vdr[r] = dv;
      var[r] = make_double2(av, 0);
    }
    dc.insert(dc.end(), vdc.begin(), vdc.end());
    ac.insert(ac.end(), vac.begin(), vac.end());
    dr.insert(dr.end(), vdr.begin(), vdr.end());
    ar.insert(ar.end(), var.begin(), var.end());
  }
  f->n_kron = np;
  upd(f->d_dc, dc);
  upd(f->d_ac, ac);
  f->d_dr.alloc(std::max<size_t>(dr.size(), 1));
  f->d_ar.alloc(std::max<size_t>(ar.size(), 1));
  if (!dr.empty()) {
    HIPCHK(hipMemcpyAsync(f->d_dr.p, dr.data(), dr.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(f->d_ar.p, ar.data(), ar.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  }
  // solve vectors
  c->N = N;
  c->ncols = ncen;
  c->red_blocks = (int)std::min<long>(1024, std::max<long>(1, N / (256 * 8)));
  c->d_b.alloc(N);
  c->d_x.alloc(N);
  c->d_r.alloc(N);
  c->d_y.alloc(N);
  HIPCHK(hipMemsetAsync(c->d_b.p, 0, (size_t)N * sizeof(double2), c->stream));
  c->d_st.alloc(1);
  HIPCHK(hipMemsetAsync(c->d_st.p, 0, sizeof(GmresState), c->stream));
  c->d_part.alloc((size_t)c->red_blocks * 2 * (kDotGroup + 1));
  f->d_count.alloc(1);
  f->d_c.alloc(1);
  HIPCHK(hipStreamSynchronize(c->stream));
}

// Per (centre sample, relative sample, sign): the material cell and the
// linear weight of coordinate centre +- rel/2 after clamping to the grid.
void f2_sample_table(const double* centre, int nc, const double* rel, int nr, const std::vector<double>& g,
                     std::vector<int>& idx, std::vector<double>& t, bool periodic = false) {
  const int ng = (int)g.size();
  idx.resize((size_t)nc * nr * 2);
  t.resize(idx.size());
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nr; ++j)
      for (int s = 0; s < 2; ++s) {
        double q = centre[i] + (s == 0 ? 0.5 : -0.5) * rel[j];
        if (periodic) {   // g0 + mod(q - g0, W), mod as floor-based remainder in [0, W)
          const double w = g[ng - 1] - g[0];
          const double r = q - g[0];
          q = g[0] + (r - std::floor(r / w) * w);
        }
        q = std::min(std::max(q, g[0]), g[ng - 1]);
        int k = (int)(std::upper_bound(g.begin(), g.end(), q) - g.begin()) - 1;
        k = std::min(std::max(k, 0), ng - 2);
        const size_t o = ((size_t)i * nr + j) * 2 + s;
        idx[o] = k;
        t[o] = (q - g[k]) / (g[k + 1] - g[k]);
      }
}

// Where the FV-consistent apply gets Delta V from (c->f2->dv_mode): a table
// per (element, quadrature point, relative vertex) in HBM that the apply
// streams (TABLE, the default: one load per sample), or the material grid,
// interpolated twice per sample inside the apply from LDS (INTERP: nothing
// streamed beyond u and y, more work per sample).  The choice depends on the
// mode alone, never on free memory.
// Sets the apply tile (rows of rho_y per workgroup, LDS) for that choice and
// uploads the constants.
void f2_configure_dv(wdg_ctx* c) {
  F2Data& f = *c->f2;
  const wdg2_desc& d = f.d;
  F2Const& h = f.hc;
  const bool fv = f.fv;
  const int nrx = d.n_rho_x, nry = d.n_rho_y;
  const int npx = h.npx;
  const bool tab = fv && f.dv_mode == WDG2_DV_TABLE;
  if (tab) {
    const size_t cnt = (size_t)d.ne_x * d.ne_y * h.nqx * h.nqy * (nrx + 1) * (nry + 1);
    f.d_dv64.alloc(cnt);
    f.d_dv32.alloc(cnt);
  } else {   // the table's memory goes back
    f.d_dv64.release();
    f.d_dv32.release();
  }
  // rho_y rows per apply tile: as many as fit 256 threads and the LDS budget;
  // the material grid joins the tile in LDS when it is small
  const int nl = d.nl_x * d.nl_y;
  const int nvl = (size_t)f.nv * sizeof(double) <= 48 * 1024 ? (int)f.nv : 0;
  h.v_lds = (!tab && nvl > 0) ? 1 : 0;
  auto lds = [&](int rows, size_t t2) {
    return f2_lds_layout(nl, rows, nrx, t2, fv, fv && !tab, h.nqx, h.nqy, npx, tab ? 0 : nvl).total;
  };
  int rows = std::max(1, std::min(nry, kF2Block / nrx));
  while (rows > 1 && lds(rows, sizeof(double2)) > 96 * 1024) --rows;
  if (lds(rows, sizeof(double2)) > kF2MaxLds) throw std::invalid_argument("n_rho_x too large for one apply tile");
  h.rows = rows;
  h.tiles = (nry + rows - 1) / rows;
  f.lds_apply64 = lds(rows, sizeof(double2));
  f.lds_apply32 = lds(rows, sizeof(float2));
  h.dv64 = nullptr;
  h.dv32 = nullptr;
  HIPCHK(hipMemcpyAsync(f.d_c.p, &h, sizeof(F2Const), hipMemcpyHostToDevice, c->stream));
  if (tab) {
    const size_t cnt = f.d_dv64.n;
    const unsigned grid = (unsigned)std::min<size_t>(65536, (cnt + 255) / 256);
    hipLaunchKernelGGL(f2_dv_table_kernel, dim3(grid), dim3(256), 0, c->stream, f.d_c.p, f.d_dv64.p, f.d_dv32.p);
    HIPCHK(hipGetLastError());
    h.dv64 = f.d_dv64.p;
    h.dv32 = f.d_dv32.p;
    HIPCHK(hipMemcpyAsync(f.d_c.p, &h, sizeof(F2Const), hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  clear_graphs(c);
}

void f2_set_potential_impl(wdg_ctx* c, const wdg2_potential& p) {
  F2Data& f = *c->f2;
  const wdg2_desc& d = f.d;
  if (p.mode != WDG2_POT_FV_CONSISTENT && p.mode != WDG2_POT_COLLOCATED)
    throw std::invalid_argument("unknown potential mode");
  const bool fv = p.mode == WDG2_POT_FV_CONSISTENT;
  if (p.nx_grid < 2 || p.ny_grid < 2) throw std::invalid_argument("the material grid needs at least 2 points per axis");
  if (!p.x_grid || !p.y_grid || !p.v || !p.center_x || !p.center_y || !p.rel_x || !p.rel_y || !p.cap_x || !p.cap_y)
    throw std::invalid_argument("potential arrays missing");
  if (fv && (p.nq_x < 1 || p.nq_x > kF2MaxNq || p.nq_y < 1 || p.nq_y > kF2MaxNq))
    throw std::invalid_argument("nq_x and nq_y must be in 1..8");
  if (fv && (!p.interp_x || !p.interp_y || !p.proj_x || !p.proj_y))
    throw std::invalid_argument("quadrature factors missing");
  std::vector<double> gx(p.x_grid, p.x_grid + p.nx_grid), gy(p.y_grid, p.y_grid + p.ny_grid);
  for (int i = 1; i < p.nx_grid; ++i)
    if (!(gx[i] > gx[i - 1])) throw std::invalid_argument("x_grid must be strictly ascending");
  for (int i = 1; i < p.ny_grid; ++i)
    if (!(gy[i] > gy[i - 1])) throw std::invalid_argument("y_grid must be strictly ascending");
  const long nv = (long)p.nx_grid * p.ny_grid;
  for (long i = 0; i < nv; ++i)
    if (!std::isfinite(p.v[i])) throw std::invalid_argument("potential must be finite");
  if (!std::isfinite(p.pot_coef) || !std::isfinite(p.cap_coef)) throw std::invalid_argument("coefficients must be finite");
  const int nrx = d.n_rho_x, nry = d.n_rho_y;
  if (nrx > kF2Block) throw std::invalid_argument("n_rho_x above 256 is not supported by the apply tiles");
  const int ncx = fv ? d.ne_x * p.nq_x : f.nxd, ncy = fv ? d.ne_y * p.nq_y : f.nyd;
  const int npx = fv ? nrx + 1 : nrx, npy = fv ? nry + 1 : nry;
  std::vector<int> tix, tiy;
  std::vector<double> ttx, tty;
  f2_sample_table(p.center_x, ncx, p.rel_x, npx, gx, tix, ttx);
  f2_sample_table(p.center_y, ncy, p.rel_y, npy, gy, tiy, tty, d.periodic_y != 0);
  auto upi = [&](DevBuf<int>& b, const std::vector<int>& v) {
    b.alloc(v.size());
    HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  };
  auto upd = [&](DevBuf<double>& b, const double* v, size_t n) {
    b.alloc(n);
    HIPCHK(hipMemcpyAsync(b.p, v, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  };
  upi(f.d_tix, tix);
  upi(f.d_tiy, tiy);
  upd(f.d_ttx, ttx.data(), ttx.size());
  upd(f.d_tty, tty.data(), tty.size());
  upd(f.d_v, p.v, nv);
  upd(f.d_capx, p.cap_x, nrx);
  upd(f.d_capy, p.cap_y, nry);
  F2Const& h = f.hc;
  std::memset(&h, 0, sizeof(h));
  h.nrx = nrx; h.nry = nry; h.nlx = d.nl_x; h.nly = d.nl_y; h.nex = d.ne_x; h.ney = d.ne_y;
  h.nxd = f.nxd; h.nyd = f.nyd; h.nrel = f.nrel;
  h.pot_mode = fv ? 0 : 1;
  h.nqx = fv ? p.nq_x : 0;
  h.nqy = fv ? p.nq_y : 0;
  h.nxg = p.nx_grid; h.nyg = p.ny_grid;
  h.npx = npx; h.npy = npy;
  h.periodic_y = d.periodic_y ? 1 : 0;
  h.inv2dx = 0.5 / d.delta_x;
  h.inv2dy = 0.5 / d.delta_y;
  h.pot_coef = p.pot_coef / f.sA;
  h.cap_coef = p.cap_coef / f.sA;
  h.bxc = f.d_bxc.p; h.bxi = f.d_bxi.p; h.byc = f.d_byc.p; h.byi = f.d_byi.p;
  h.tix = f.d_tix.p; h.tiy = f.d_tiy.p; h.ttx = f.d_ttx.p; h.tty = f.d_tty.p;
  h.v = f.d_v.p; h.capx = f.d_capx.p; h.capy = f.d_capy.p;
  if (fv)
    for (int q = 0; q < kF2MaxNq; ++q)
      for (int a = 0; a < kF2MaxNl; ++a) {
        const int k = q * kF2MaxNl + a;
        h.ipx[k] = (q < p.nq_x && a < d.nl_x) ? p.interp_x[q * d.nl_x + a] : 0.0;
        h.prx[k] = (q < p.nq_x && a < d.nl_x) ? p.proj_x[q * d.nl_x + a] : 0.0;
        h.ipy[k] = (q < p.nq_y && a < d.nl_y) ? p.interp_y[q * d.nl_y + a] : 0.0;
        h.pry[k] = (q < p.nq_y && a < d.nl_y) ? p.proj_y[q * d.nl_y + a] : 0.0;
      }
  f.fv = fv;
  f.nv = nv;
  f2_configure_dv(c);
  HIPCHK(hipStreamSynchronize(c->stream));
  clear_graphs(c);
  // the preconditioner depends on the potential: rebuild on the next request
  if (f.precond == WDG2_PRECOND_JACOBI || f.precond == WDG2_PRECOND_ROWABS) f.precond = -f.precond;
  c->have_potential = true;
}

// Diagonal (mode F2_DIAG) or row-abs bound (F2_ROWABS) of A_hat into dst.
void f2_diag_into(wdg_ctx* c, int mode, double2* dst) {
  F2Data& f = *c->f2;
  const bool rowabs = mode == F2_ROWABS;
  hipLaunchKernelGGL(f2_kron_diag_kernel, dim3(std::min<long>(4096, (c->N + 255) / 256)), dim3(256), 0, c->stream,
                     f.n_kron, f.ncen, f.nrel, rowabs ? f.d_ac.p : f.d_dc.p, rowabs ? f.d_ar.p : f.d_dr.p, dst);
  if (rowabs) f2_launch_apply<double, F2_ROWABS>(c, nullptr, dst, nullptr);
  else f2_launch_apply<double, F2_DIAG>(c, nullptr, dst, nullptr);
  HIPCHK(hipGetLastError());
}

void f2_build_precond(wdg_ctx* c, int kind, double rel_floor, bool fallback, int* used) {
  F2Data& f = *c->f2;
  if (kind == WDG2_PRECOND_NONE || kind == WDG2_PRECOND_LINE) {
    if (kind == WDG2_PRECOND_LINE && !f.have_line)
      throw std::logic_error("line-transverse factors have not been set (wdg2_set_line_precond)");
    f.precond = kind;
    *used = kind;
    clear_graphs(c);
    return;
  }
  if (kind != WDG2_PRECOND_JACOBI && kind != WDG2_PRECOND_ROWABS) throw std::invalid_argument("unknown preconditioner");
  f.d_p64.alloc(c->N);
  f.d_p32.alloc(c->N);
  for (int attempt = 0; attempt < 2; ++attempt) {
    f2_diag_into(c, kind == WDG2_PRECOND_JACOBI ? F2_DIAG : F2_ROWABS, c->d_y.p);
    if (c->shift_on)   // of alpha I + beta A_hat
      hipLaunchKernelGGL(shift_diag_kernel, dim3(c->red_blocks), dim3(256), 0, c->stream, c->d_y.p, c->N, c->shift_a,
                         c->shift_b, kind == WDG2_PRECOND_ROWABS ? 1 : 0);
    const int nb = c->red_blocks;
    hipLaunchKernelGGL(f2_absmax_kernel, dim3(nb), dim3(256), 0, c->stream, c->d_y.p, c->N, (double*)c->d_part.p);
    std::vector<double> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), c->d_part.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double mx = 0;
    for (double v : part) mx = std::max(mx, v);
    // floors in physical units (regularizePreconditionerDiagonal / regularizeRowAbsScale)
    const double floor_phys = rel_floor * std::max(mx * f.sA, 1.0);
    HIPCHK(hipMemsetAsync(f.d_count.p, 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(f2_precond_build_kernel, dim3(std::min<long>(4096, (c->N + 255) / 256)), dim3(256), 0,
                       c->stream, c->d_y.p, c->N, kind, floor_phys / f.sA, f.sA, f.d_p64.p, f.d_p32.p, f.d_count.p);
    unsigned long long small = 0;
    HIPCHK(hipMemcpyAsync(&small, f.d_count.p, sizeof(small), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (kind == WDG2_PRECOND_JACOBI && small > 0 && fallback) {
      kind = WDG2_PRECOND_ROWABS;
      continue;
    }
    break;
  }
  f.precond = kind;
  *used = kind;
  clear_graphs(c);
}

// The solve uses the preconditioner selected last; one whose potential
// changed since is rebuilt first.
void f2_prepare_solve(wdg_ctx* c) {
  F2Data& f = *c->f2;
  if (f.precond < 0) {
    int used = 0;
    f2_build_precond(c, -f.precond, 1e-12, true, &used);
  }
}

void f2_set_line_impl(wdg_ctx* c, const wdg2_line_factors& lf) {
  F2Data& f = *c->f2;
  const int nrx = f.d.n_rho_x, nl = f.d.nl_x, ne = f.d.ne_x;
  const int nT = f.d.n_rho_y * f.nyd;
  // the reference's factor-memory guard (get_LineTransversePreconditioner_full2D.m:33-43)
  const double factor_bytes = 32.0 * nT * nT * nrx + 32.0 * nl * nl * nrx * ne + 16.0 * nrx * nrx + 8.0 * nT * nrx;
  if (factor_bytes > c->line_max_factor_bytes) {
    char msg[256];
    std::snprintf(msg, sizeof msg, "Estimated factors %.3g GiB exceed full2D_lineMaxFactorBytes %.3g GiB.",
                  factor_bytes / 1073741824.0, c->line_max_factor_bytes / 1073741824.0);
    throw StatusError(WDG_ERR_LINE_MEMORY, msg);
  }
  if (!(lf.sigma > 0) || !std::isfinite(lf.sigma)) throw std::invalid_argument("sigma must be positive and finite");
  if (!lf.modes || !lf.forward || !lf.x_inverse || !lf.x_coupling || !lf.t_inverse)
    throw std::invalid_argument("line factor arrays missing");
  const size_t nxf = (size_t)nrx * ne * nl * nl;
  const size_t ntf = (size_t)nrx * nT * nT;
  for (size_t i = 0; i < nxf; ++i)
    if (!std::isfinite(lf.x_inverse[i]) || !std::isfinite(lf.x_coupling[i]))
      throw std::invalid_argument("X factors must be finite");
  for (size_t i = 0; i < 2 * ntf; ++i)
    if (!std::isfinite(lf.t_inverse[i])) throw std::invalid_argument("transverse factors must be finite");
  f.nT = nT;
  auto upd = [&](DevBuf<double>& b, const double* v, size_t n) {
    b.alloc(n);
    HIPCHK(hipMemcpyAsync(b.p, v, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  };
  std::vector<float> x32(nxf), c32(nxf);
  for (size_t i = 0; i < nxf; ++i) {
    x32[i] = (float)lf.x_inverse[i];
    c32[i] = (float)lf.x_coupling[i];
  }
  upd(f.d_lxi64, lf.x_inverse, nxf);
  upd(f.d_lxc64, lf.x_coupling, nxf);
  f.d_lxi32.alloc(nxf);
  f.d_lxc32.alloc(nxf);
  HIPCHK(hipMemcpyAsync(f.d_lxi32.p, x32.data(), nxf * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(f.d_lxc32.p, c32.data(), nxf * 4, hipMemcpyHostToDevice, c->stream));
  std::vector<int> fw(lf.forward, lf.forward + nrx);
  f.d_lfwd.alloc(nrx);
  HIPCHK(hipMemcpyAsync(f.d_lfwd.p, fw.data(), nrx * sizeof(int), hipMemcpyHostToDevice, c->stream));
  std::vector<float2> t32(ntf);
  for (size_t i = 0; i < ntf; ++i) t32[i] = make_float2((float)lf.t_inverse[2 * i], (float)lf.t_inverse[2 * i + 1]);
  f.d_ltinv64.alloc(ntf);
  f.d_ltinv32.alloc(ntf);
  HIPCHK(hipMemcpyAsync(f.d_ltinv64.p, lf.t_inverse, ntf * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(f.d_ltinv32.p, t32.data(), ntf * sizeof(float2), hipMemcpyHostToDevice, c->stream));
  // mode transforms: forward U^H, backward U s_A / sigma (z in device units)
  std::vector<double2> uf((size_t)nrx * nrx), ub((size_t)nrx * nrx);
  const double sb = f.sA / lf.sigma;
  for (int m = 0; m < nrx; ++m)
    for (int k = 0; k < nrx; ++k) {
      const double re = lf.modes[2 * ((size_t)k * nrx + m)], im = lf.modes[2 * ((size_t)k * nrx + m) + 1];
      uf[(size_t)m * nrx + k] = make_double2(re, -im);      // (U^H)[m][k] = conj(U[k][m])
      const double re2 = lf.modes[2 * ((size_t)m * nrx + k)], im2 = lf.modes[2 * ((size_t)m * nrx + k) + 1];
      ub[(size_t)m * nrx + k] = make_double2(sb * re2, sb * im2);
    }
  auto setup_dense = [&](F2Data::Dense& dp, const std::vector<double2>& R) {
    dp.n = nrx;
    dp.lines = f.ncen;
    dp.k = F2Dense{};
    dp.k.chunk = f2_dense_chunk(nrx, f.d.n_rho_y, dp.lines, 0, &dp.k.rglobal);
    dp.k.axis = 0;
    dp.k.face = -2;
    dp.k.nl = nl;
    dp.k.ne = ne;
    dp.k.coef = 1.0;
    dp.R.alloc(R.size());
    HIPCHK(hipMemcpyAsync(dp.R.p, R.data(), R.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    dp.k.R = dp.R.p;
  };
  setup_dense(f.lt_fwd64, uf);
  setup_dense(f.lt_bwd64, ub);
  setup_dense(f.lt_fwd32, uf);
  setup_dense(f.lt_bwd32, ub);
  f.d_la64.alloc(c->N);
  f.d_lb64.alloc(c->N);
  f.d_la32.alloc(c->N);
  f.d_lb32.alloc(c->N);
  HIPCHK(hipStreamSynchronize(c->stream));
  f.have_line = true;
  f.precond = WDG2_PRECOND_LINE;
  clear_graphs(c);
}

}  // namespace
