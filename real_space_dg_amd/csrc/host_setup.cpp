// Host-side setup: DG element matrices on Gauss-Lobatto nodes, the
// plane-wave basis and the inflow data of the FV-in-xi Wigner operator.
//
// Method (restated from the reference, computed independently here):
//  * GLL nodes by Newton iteration on (1-x^2) P'_N, weights 2/(N(N+1)P_N^2)
//    (DG_code/core/JacobiGL.m:44-50).
//  * Nodal Lagrange basis l_a on the element nodes r.  The reference builds
//    Mref = invV'*invV and Dr = Mref*Dmatrix1D from orthonormal Vandermonde
//    matrices (DG_code/initParams.m:188-191, DG_code/core/get_Diff.m:198);
//    both are the exact integrals int l_a l_b and int l_a l_b', computed here
//    by the 10-point GLL rule (exact to degree 17, so for n <= 9).
//  * W = V_tilde/V is Lagrange interpolation W_qa = l_a(x_q)
//    (get_Drift_accel_trial_permutedShell.m:183).
#include "host_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

namespace wdg {

namespace {

double legendre(int N, double x, double* dP = nullptr) {
  double p0 = 1.0, p1 = x;
  if (N == 0) {
    if (dP) *dP = 0;
    return 1.0;
  }
  for (int k = 1; k < N; ++k) {
    double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
    p0 = p1;
    p1 = p2;
  }
  if (dP) {
    if (std::abs(x) < 1.0)
      *dP = N * (x * p1 - p0) / (x * x - 1.0);
    else
      *dP = (x > 0 ? 1.0 : (N % 2 ? 1.0 : -1.0)) * 0.5 * N * (N + 1.0);
  }
  return p1;
}

// n GLL points on [-1,1] (degree N = n-1), ascending.
void gll(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  if (n == 1) {
    w[0] = 2.0;
    return;
  }
  const int N = n - 1;
  for (int i = 0; i < n; ++i) {
    double xi = -std::cos(M_PI * i / N);
    if (i > 0 && i < N) {
      for (int it = 0; it < 100; ++it) {
        // Newton on (1-x^2) P'_N(x) via the Hesthaven-Warburton update
        // x <- x - (x P_N - P_{N-1}) / (n P_N).
        double pN = legendre(N, xi);
        double pNm1 = legendre(N - 1, xi);
        double dx = (xi * pN - pNm1) / (n * pN);
        xi -= dx;
        if (std::abs(dx) < 1e-16) break;
      }
    }
    x[i] = xi;
  }
  x[0] = -1.0;
  x[N] = 1.0;
  for (int i = 0; i < n; ++i) {
    double p = legendre(N, x[i]);
    w[i] = 2.0 / (N * (N + 1.0) * p * p);
  }
}

double lagrange(const std::vector<double>& r, int a, double x) {
  double v = 1.0;
  for (size_t b = 0; b < r.size(); ++b)
    if ((int)b != a) v *= (x - r[b]) / (r[a] - r[b]);
  return v;
}

double lagrange_d(const std::vector<double>& r, int a, double x) {
  double s = 0.0;
  for (size_t m = 0; m < r.size(); ++m) {
    if ((int)m == a) continue;
    double t = 1.0 / (r[a] - r[m]);
    for (size_t b = 0; b < r.size(); ++b)
      if ((int)b != a && b != m) t *= (x - r[b]) / (r[a] - r[b]);
    s += t;
  }
  return s;
}

// Gauss-Jordan inverse with partial pivoting (small real matrices).
std::vector<double> inverse(const std::vector<double>& A, int n) {
  std::vector<double> a(A), inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::abs(a[r * n + c]) > std::abs(a[p * n + c])) p = r;
    if (a[p * n + c] == 0.0) throw std::runtime_error("singular element mass matrix");
    if (p != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[p * n + k], a[c * n + k]);
        std::swap(inv[p * n + k], inv[c * n + k]);
      }
    double d = 1.0 / a[c * n + c];
    for (int k = 0; k < n; ++k) {
      a[c * n + k] *= d;
      inv[c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[r * n + k] -= f * a[c * n + k];
        inv[r * n + k] -= f * inv[c * n + k];
      }
    }
  }
  return inv;
}

std::vector<double> matmul(const std::vector<double>& A, const std::vector<double>& B, int n) {
  std::vector<double> C(n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) C[i * n + j] += A[i * n + k] * B[k * n + j];
  return C;
}

std::vector<double> linspace(double a, double b, int n) {
  std::vector<double> v(n);
  if (n == 1) {
    v[0] = b;
    return v;
  }
  double step = (b - a) / (n - 1);
  for (int i = 0; i < n; ++i) v[i] = a + i * step;
  v[n - 1] = b;
  return v;
}

// Complex LU with partial pivoting, row-major, in place.
void lu_factor(std::vector<cplx>& a, std::vector<int>& piv, int n) {
  piv.resize(n);
  for (int c = 0; c < n; ++c) {
    int p = c;
    double best = std::abs(a[(size_t)c * n + c]);
    for (int r = c + 1; r < n; ++r) {
      double v = std::abs(a[(size_t)r * n + c]);
      if (v > best) {
        best = v;
        p = r;
      }
    }
    if (best == 0.0) throw std::runtime_error("plane-wave basis phi is singular");
    piv[c] = p;
    if (p != c)
      for (int k = 0; k < n; ++k) std::swap(a[(size_t)p * n + k], a[(size_t)c * n + k]);
    const cplx d = 1.0 / a[(size_t)c * n + c];
    const cplx* rowc = &a[(size_t)c * n];
#pragma omp parallel for schedule(static)
    for (int r = c + 1; r < n; ++r) {
      cplx* row = &a[(size_t)r * n];
      cplx f = row[c] * d;
      row[c] = f;
      for (int k = c + 1; k < n; ++k) row[k] -= f * rowc[k];
    }
  }
}

// One right-hand side in place.
void lu_solve1(const std::vector<cplx>& lu, const std::vector<int>& piv, int n, cplx* b) {
  for (int c = 0; c < n; ++c)
    if (piv[c] != c) std::swap(b[c], b[piv[c]]);
  for (int r = 1; r < n; ++r) {
    cplx s = b[r];
    const cplx* row = lu.data() + (size_t)r * n;
    for (int k = 0; k < r; ++k) s -= row[k] * b[k];
    b[r] = s;
  }
  for (int r = n - 1; r >= 0; --r) {
    cplx s = b[r];
    const cplx* row = lu.data() + (size_t)r * n;
    for (int k = r + 1; k < n; ++k) s -= row[k] * b[k];
    b[r] = s / row[r];
  }
}

// Two right-hand sides at once (one pass over the factors).
void lu_solve2(const std::vector<cplx>& lu, const std::vector<int>& piv, int n, std::vector<cplx>& b1,
               std::vector<cplx>& b2) {
  for (int c = 0; c < n; ++c)
    if (piv[c] != c) {
      std::swap(b1[c], b1[piv[c]]);
      std::swap(b2[c], b2[piv[c]]);
    }
  for (int r = 1; r < n; ++r) {
    cplx s1 = b1[r], s2 = b2[r];
    const cplx* row = lu.data() + (size_t)r * n;
    for (int k = 0; k < r; ++k) {
      s1 -= row[k] * b1[k];
      s2 -= row[k] * b2[k];
    }
    b1[r] = s1;
    b2[r] = s2;
  }
  for (int r = n - 1; r >= 0; --r) {
    cplx s1 = b1[r], s2 = b2[r];
    const cplx* row = lu.data() + (size_t)r * n;
    for (int k = r + 1; k < n; ++k) {
      s1 -= row[k] * b1[k];
      s2 -= row[k] * b2[k];
    }
    b1[r] = s1 / row[r];
    b2[r] = s2 / row[r];
  }
}

double cap_profile(double q, double l_xi, double delta_xi) {
  // DG_code/core/get_CAP.m:11-18 (W0 = 1, n = 1, width 0.4 L_xi)
  const double W0 = 1.0, nn = 1.0;
  const double L0 = l_xi / 2 - delta_xi;
  const double weite = l_xi * 0.40;
  double v = 0.0;
  if (q >= L0 - weite) v += std::pow(W0 / weite * (q - L0 + weite), 2 * nn);
  if (q <= -L0 + weite) v += std::pow(W0 / weite * (q + L0 - weite), 2 * nn);
  return v;
}

}  // namespace

void build_geometry(const wdg_desc& d, Geometry& g) {
  if (d.n_chi < 2) throw std::runtime_error("n_chi must be >= 2");
  if (d.n_k_chi < 2 || d.n_k_chi > 8) throw std::runtime_error("n_k_chi must be in 2..8");
  if (d.nk != d.n_xi) throw std::runtime_error("nk must equal n_xi (FV plane-wave basis is square)");
  if (d.nk < 32 || d.nk % 32 != 0) throw std::runtime_error("nk must be a positive multiple of 32");
  if (!(d.l_chi > 0) || !(d.l_xi > 0)) throw std::runtime_error("lengths must be positive");
  g.n_chi = d.n_chi;
  g.n = d.n_k_chi;
  g.n_xi = d.n_xi;
  g.nk = d.nk;
  g.l_chi = d.l_chi;
  g.l_xi = d.l_xi;
  const int n = g.n;
  g.q_diff = kHbarIP / kMElektronIP;  // initParams.m:106
  g.q_drift = kQIP / kHbarIP;         // initParams.m:107
  g.delta_chi = g.l_chi / g.n_chi;    // initParams.m:155
  g.jac = g.delta_chi / 2;            // get_Diff.m:199
  {
    auto xr = linspace(-g.l_xi / 2, g.l_xi / 2, g.n_xi);  // initParams.m:179-181
    g.delta_xi = xr[1] - xr[0];
  }
  g.y = linspace(-(g.l_xi - g.delta_xi) / 2, (g.l_xi - g.delta_xi) / 2, g.n_xi);  // :182
  gll(kNq, g.xq, g.wq);                                                         // :186
  std::vector<double> wr;
  gll(n, g.r, wr);                                                              // :187
  // Exact mass / stiffness of the nodal basis (see file header).
  g.Mref.assign(n * n, 0.0);
  g.S.assign(n * n, 0.0);
  g.Wq.assign(kNq * n, 0.0);
  for (int q = 0; q < kNq; ++q)
    for (int a = 0; a < n; ++a) g.Wq[q * n + a] = lagrange(g.r, a, g.xq[q]);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      double m = 0, s = 0;
      for (int q = 0; q < kNq; ++q) {
        m += g.wq[q] * g.Wq[q * n + a] * g.Wq[q * n + b];
        s += g.wq[q] * g.Wq[q * n + a] * lagrange_d(g.r, b, g.xq[q]);
      }
      g.Mref[a * n + b] = m;
      g.S[a * n + b] = (std::abs(s) < 1e-15) ? 0.0 : s;  // get_Diff.m:203-204
    }
  std::vector<double> Mx(n * n);
  for (int i = 0; i < n * n; ++i) Mx[i] = g.delta_xi / 2 * g.Mref[i];  // get_Diff.m:200
  g.invM = inverse(Mx, n);
  std::vector<double> P1(n * n, 0.0), P3(n * n, 0.0);
  P1[0] = 1.0;
  P3[(n - 1) * n + (n - 1)] = 1.0;
  std::vector<double> t1(n * n), t2L(n * n), t2R(n * n);
  for (int i = 0; i < n * n; ++i) {
    t1[i] = g.S[i] + 0.5 * P1[i] - 0.5 * P3[i];
    t2L[i] = 0.5 * P1[i];
    t2R[i] = 0.5 * P3[i];
  }
  g.K1 = matmul(g.invM, t1, n);   // get_Diff.m:214
  g.K2L = matmul(g.invM, t2L, n); // :216
  g.K2R = matmul(g.invM, t2R, n); // :217
  g.K2.assign(n * n, 0.0);
  for (int i = 0; i < n * n; ++i) g.K2[i] = g.K2L[i] + g.K2R[i];
  g.K3.assign(n * n, 0.0);
  g.K5.assign(n * n, 0.0);
  for (int a = 0; a < n; ++a) {
    g.K3[a * n + (n - 1)] = -0.5 * g.invM[a * n + 0];  // -0.5 invM P2, P2 = e1 en' (:218)
    g.K5[a * n + 0] = 0.5 * g.invM[a * n + (n - 1)];   // 0.5 invM P4, P4 = en e1' (:220)
  }
  // Wave numbers and basis (Basisfunktionen_Q_m.m:22-32).
  g.k.resize(g.nk);
  for (int j = 0; j < g.nk; ++j) g.k[j] = -2.0 * M_PI / g.l_xi * ((j + 1) - 0.5 * (g.nk + 1));
  {
    static std::mutex mu;
    static std::map<std::tuple<int, int, uint64_t>, std::shared_ptr<const BasisData>> cache;
    uint64_t lbits;
    std::memcpy(&lbits, &g.l_xi, sizeof(lbits));
    const auto key = std::make_tuple(g.n_xi, g.nk, lbits);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it == cache.end()) {
      auto bd = std::make_shared<BasisData>();
      const double isl = 1.0 / std::sqrt(g.l_xi);
      bd->phi.resize((size_t)g.n_xi * g.nk);
      bd->cosm.resize((size_t)g.n_xi * g.nk);
      for (int l = 0; l < g.n_xi; ++l)
        for (int j = 0; j < g.nk; ++j) {
          const double ph = g.k[j] * g.y[l];
          const double c = std::cos(ph), s = std::sin(ph);
          bd->phi[(size_t)l * g.nk + j] = cplx(isl * c, isl * s);
          bd->cosm[(size_t)l * g.nk + j] = c;
        }
      bd->lu = bd->phi;
      lu_factor(bd->lu, bd->piv, g.nk);
      if (g.nk <= kInflowMapMax) {
        // the inflow map phi^-1 cosm, one LU solve per column (once per basis)
        const int nk = g.nk;
        bd->gm.assign((size_t)nk * nk, cplx(0.0));
#pragma omp parallel for schedule(dynamic, 4)
        for (int j = 0; j < nk; ++j) {
          std::vector<cplx> col(nk);
          for (int l = 0; l < nk; ++l) col[l] = bd->cosm[(size_t)l * nk + j];
          lu_solve1(bd->lu, bd->piv, nk, col.data());
          for (int r = 0; r < nk; ++r) bd->gm[(size_t)r * nk + j] = col[r];
        }
      }
      if (cache.size() > 64) cache.clear();
      it = cache.emplace(key, std::move(bd)).first;
    }
    g.basis = it->second;
  }
  // Theta sample points (get_Drift_accel_trial_permutedShell.m:15-55).
  g.edges = linspace(-g.l_xi / 2, g.l_xi / 2, g.n_xi + 1);
  g.cap.resize(g.n_xi + 1);
  for (int l = 0; l <= g.n_xi; ++l) g.cap[l] = cap_profile(g.edges[l], g.l_xi, g.delta_xi);
  std::vector<double> chi_gl(kNq, 0.0);
  for (int i = 0; i + 1 < kNq; ++i) chi_gl[i + 1] = chi_gl[i] + (g.xq[i + 1] - g.xq[i]) / 2 * g.delta_chi;
  auto ends = linspace(0.0, g.l_chi, g.n_chi + 1);
  g.xquad.resize((size_t)g.n_chi * kNq);
  for (int e = 0; e < g.n_chi; ++e)
    for (int q = 0; q < kNq; ++q) g.xquad[(size_t)e * kNq + q] = ends[e] + chi_gl[q];
  g.chi.resize((size_t)g.n_chi * n);
  for (int e = 0; e < g.n_chi; ++e) {
    auto c = linspace(ends[e], ends[e + 1], n);  // initParams.m:157-159 (equispaced)
    for (int a = 0; a < n; ++a) g.chi[(size_t)e * n + a] = c[a];
  }
  // Device scale: the largest transport coefficient.
  double kmax = 0.0, k1max = 0.0;
  for (double kv : g.k) kmax = std::max(kmax, std::abs(kv));
  for (int i = 0; i < n * n; ++i) k1max = std::max(k1max, std::abs(g.K1[i]) + std::abs(g.K2[i]));
  g.scale = g.q_diff / g.jac * kmax * k1max;
}

namespace {

bool smooth23(int p) {
  while (p % 2 == 0) p /= 2;
  while (p % 3 == 0) p /= 3;
  return p == 1;
}

std::vector<int> fft_radices(int P) {
  std::vector<int> r;
  if (P == 1) return r;
  while (P % 8 == 0) { r.push_back(8); P /= 8; }
  while (P % 4 == 0) { r.push_back(4); P /= 4; }
  while (P % 2 == 0) { r.push_back(2); P /= 2; }
  while (P % 3 == 0) { r.push_back(3); P /= 3; }
  if (P != 1) throw std::runtime_error("FFT length must be of the form 2^p 3^q");
  return r;
}

// exp(i pi s t^2), the phase reduced modulo 2 pi before the trigonometry.
cplx chirp(double s, double t) {
  const double ph = std::fmod(s * t * t, 2.0);
  return cplx(std::cos(M_PI * ph), std::sin(M_PI * ph));
}

// Radix sequence (8, the radices of P/64, 8) and the Stockham base twiddles
// of the device transform of length P = 64 m.
void fft_plan_tables(int P, std::vector<int>& rad, std::vector<int>& tw_off, std::vector<cplx>& tw) {
  rad.clear();
  tw_off.clear();
  tw.clear();
  rad.push_back(8);
  for (int r : fft_radices(P / 64)) rad.push_back(r);
  rad.push_back(8);
  if (P / 64 == 1) rad = {8, 8};
  int Ns = 1;
  for (int R : rad) {
    tw_off.push_back((int)tw.size());
    for (int jm = 0; jm < Ns; ++jm) {
      const double ang = -2.0 * M_PI * (double)jm / ((double)Ns * R);
      tw.emplace_back(std::cos(ang), std::sin(ang));
    }
    Ns *= R;
  }
}

}  // namespace

void fft_host(std::vector<cplx>& x, bool inverse) {
  const int P = (int)x.size();
  const std::vector<int> rad = fft_radices(P);
  std::vector<cplx> y(P);
  const double sgn = inverse ? 1.0 : -1.0;
  int Ns = 1;
  for (int R : rad) {
    const int nb = P / R;
    std::vector<cplx> v(R), o(R);
    for (int b = 0; b < nb; ++b) {
      const int jm = b % Ns;
      for (int r = 0; r < R; ++r) {
        const double ang = sgn * 2.0 * M_PI * (double)jm * r / ((double)Ns * R);
        v[r] = x[b + r * nb] * cplx(std::cos(ang), std::sin(ang));
      }
      for (int s = 0; s < R; ++s) {
        cplx acc = 0;
        for (int r = 0; r < R; ++r) {
          const double ang = sgn * 2.0 * M_PI * (double)((r * s) % R) / R;
          acc += v[r] * cplx(std::cos(ang), std::sin(ang));
        }
        o[s] = acc;
      }
      const int base = (b / Ns) * Ns * R + jm;
      for (int r = 0; r < R; ++r) y[base + r * Ns] = o[r];
    }
    x.swap(y);
    Ns *= R;
  }
}

void build_chirp(const Geometry& g, ChirpPlan& cp) {
  cp = ChirpPlan();
  const int nk = g.nk, nx = g.n_xi;
  const double L = g.l_xi;
  const double hp = (g.l_xi - g.delta_xi) / (nx - 1);
  const double beta = hp / L;
  const double ck = 0.5 * (nk - 1), cx = 0.5 * (nx - 1);
  for (int j = 0; j < nk; ++j)
    if (std::abs(g.k[j] + 2.0 * M_PI / L * (j - ck)) > 1e-12 * 2.0 * M_PI / L * nk) return;
  for (int l = 0; l < nx; ++l)
    if (std::abs(g.y[l] - hp * (l - cx)) > 1e-12 * L) return;
  // P = 64 m (m = 2^a 3^b): the device transform starts and ends with a
  // radix-8 pass around the radices of m.
  int P = nk + nx - 1;
  while (P % 64 != 0 || !smooth23(P / 64)) ++P;
  cp.P = P;
  fft_plan_tables(P, cp.rad, cp.tw_off, cp.tw);
  const double norm = 1.0 / (std::sqrt(L) * P);
  cp.a_pre.resize(nk);
  cp.a_post.resize(nk);
  for (int j = 0; j < nk; ++j) {
    const double u = j - ck;
    cp.a_pre[j] = chirp(-beta, u);
    cp.a_post[j] = cplx(0.0, 1.0) * chirp(beta, u) * norm;
  }
  cp.w_post.resize(nx);
  cp.b_pre.resize(nx);
  for (int l = 0; l < nx; ++l) {
    const double v = l - cx;
    cp.w_post[l] = chirp(-beta, v) * norm;
    cp.b_pre[l] = chirp(beta, v);
  }
  const double dl = 0.5 * (nk - nx);
  cp.g1.assign(P, cplx(0.0, 0.0));
  for (int m = -(nk - 1); m <= nx - 1; ++m) cp.g1[(m + P) % P] = chirp(beta, m + dl);
  cp.g2.assign(P, cplx(0.0, 0.0));
  for (int m = -(nx - 1); m <= nk - 1; ++m) cp.g2[(m + P) % P] = chirp(-beta, m - dl);
  fft_host(cp.g1, false);
  fft_host(cp.g2, false);
  cp.ok = true;
}

void build_ndft(const Geometry& g, double trunc_tol, NdftPlan& np) {
  np = NdftPlan();
  const int N = g.nk;
  if (g.n_xi != N || N < 64 || N % 64 != 0 || !smooth23(N / 64)) return;
  const double L = g.l_xi;
  const double hp = (g.l_xi - g.delta_xi) / (N - 1);
  const double c = 0.5 * (N - 1);
  for (int j = 0; j < N; ++j)
    if (std::abs(g.k[j] + 2.0 * M_PI / L * (j - c)) > 1e-12 * 2.0 * M_PI / L * N) return;
  for (int l = 0; l < N; ++l)
    if (std::abs(g.y[l] - hp * (l - c)) > 1e-12 * L) return;
  // k_j y_l = -2 pi beta u v, beta = 1/N + delta
  const double delta = hp / L - 1.0 / N;
  const double s = -2.0 * M_PI * delta * c * c;
  int terms = 1;
  double r = std::abs(s);   // |s|^terms / terms!
  while (r > trunc_tol && terms < 4) {
    ++terms;
    r *= std::abs(s) / terms;
  }
  if (r > trunc_tol) return;
  np.P = N;
  np.terms = terms;
  np.s = s;
  np.c = c;
  fft_plan_tables(N, np.rad, np.tw_off, np.tw);
  np.fpre.resize(N);
  np.fpost.resize(N);
  np.bpre.resize(N);
  np.bpost.resize(N);
  // phases 2 pi c j / N and 2 pi c^2 / N reduced modulo 2 pi through
  // c j = (N - 1) j / 2 and c^2 = (N - 1)^2 / 4 in integer arithmetic
  const long long Nl = N, two_c = N - 1;
  auto ph = [&](long long num, long long den) {   // exp(2 pi i num / den)
    const long long q = ((num % den) + den) % den;
    const double a = 2.0 * M_PI * (double)q / (double)den;
    return cplx(std::cos(a), std::sin(a));
  };
  const cplx cc = ph(-two_c * two_c, 4 * Nl);   // exp(-2 pi i c^2 / N)
  const double isl = 1.0 / std::sqrt(L);
  for (int j = 0; j < N; ++j) {
    const cplx pj = ph(two_c * j, 2 * Nl);        // exp(2 pi i c j / N)
    np.fpre[j] = pj;
    np.fpost[j] = cc * pj * isl;
    np.bpre[j] = std::conj(pj);
    np.bpost[j] = cplx(0.0, 1.0) * std::conj(cc) * std::conj(pj) * isl;
  }
  np.ok = true;
}

void inflow(const Geometry& g, const wdg_potential& pot, std::vector<cplx>& fnl,
            std::vector<cplx>& fnr) {
  const double hb = kHVT / 2.0 / M_PI;
  const double q = kQVT, kB = kKBVT, T = pot.temp;
  const int nk = g.nk;
  const double dk = std::abs(g.k[1] - g.k[0]);
  const double V1 = pot.em[0], VN = pot.em[pot.n_mat - 1];
  std::vector<double> fL(nk), fR(nk);
  for (int j = 0; j < nk; ++j) {
    const double kk = g.k[j];
    const double EL = V1 + hb * hb * kk * kk / 2.0 / pot.m_left / q;   // verteilung.m:92 / :135
    const double ER = VN + hb * hb * kk * kk / 2.0 / pot.m_right / q;  // :93 / :136
    if (pot.fermi_model == WDG_FERMI_MSA_3D) {                          // :98-99
      fL[j] = 2.0 * pot.deg_factor / (1.0 + std::exp(q * (EL - pot.ef_left) / (kB * T)));
      fR[j] = 2.0 * pot.deg_factor / (1.0 + std::exp(q * (ER - pot.ef_right) / (kB * T)));
    } else if (pot.fermi_model == WDG_FERMI_TSU_ESAKI) {
      // m kT/(pi hbar^2) ln(1 + exp(q(EF - E)/kT)): verteilung.m:64-65 form
      // with the device's Fermi level, band edge and mass.
      const double aL = pot.m_left * kB * T / (M_PI * hb * hb);
      const double aR = pot.m_right * kB * T / (M_PI * hb * hb);
      const double xl = q * (pot.ef_left - EL) / (kB * T), xr = q * (pot.ef_right - ER) / (kB * T);
      fL[j] = aL * ((xl > 30) ? xl + std::log1p(std::exp(-xl)) : std::log1p(std::exp(xl)));
      fR[j] = aR * ((xr > 30) ? xr + std::log1p(std::exp(-xr)) : std::log1p(std::exp(xr)));
    } else {                                                            // :139-152
      double vmax = pot.em[0];
      for (int i = 0; i < pot.n_mat; ++i) vmax = std::max(vmax, pot.em[i]);
      const double kzmax = std::sqrt(2.0 * 2.0 * kM0VT * pot.me_z * q * (vmax + 0.3)) / hb;
      const int nkz = 1001;
      const double dkz = kzmax / (nkz - 1);
      double sl = 0.0, sr = 0.0;
      for (int z = 0; z < nkz; ++z) {
        const double kz = (z == nkz - 1) ? kzmax : z * dkz;
        const double Ekz = hb * hb * kz * kz / 2.0 / kM0VT / pot.me_z / q;
        sl += 1.0 / (1.0 + std::exp(q * (EL + Ekz - pot.ef_left) / (kB * T)));
        sr += 1.0 / (1.0 + std::exp(q * (ER + Ekz - pot.ef_right) / (kB * T)));
      }
      fL[j] = 2.0 / M_PI * dkz * sl;
      fR[j] = 2.0 / M_PI * dkz * sr;
    }
  }
  // Cosine transform to xi, then phi \ f (verteilung.m:109-114): with the
  // precomputed map phi^-1 cosm one row-parallel product per contact
  if (!g.basis->gm.empty()) {
    const cplx* G = g.basis->gm.data();
    const double sc = dk / 2.0 / M_PI;
    std::vector<cplx> xl(nk), xr(nk);
#pragma omp parallel for schedule(static) if (nk >= 256)
    for (int r = 0; r < nk; ++r) {
      const cplx* row = G + (size_t)r * nk;
      double lr = 0.0, li = 0.0, rr = 0.0, ri = 0.0;
      for (int j = 0; j < nk; ++j) {
        lr += row[j].real() * fL[j];
        li += row[j].imag() * fL[j];
        rr += row[j].real() * fR[j];
        ri += row[j].imag() * fR[j];
      }
      xl[r] = cplx(lr * sc, li * sc);
      xr[r] = cplx(rr * sc, ri * sc);
    }
    fnl = std::move(xl);
    fnr = std::move(xr);
    return;
  }
  std::vector<cplx> bl(g.n_xi), br(g.n_xi);
  const double* cm = g.basis->cosm.data();
  for (int l = 0; l < g.n_xi; ++l) {
    double sl = 0.0, sr = 0.0;
    const double* row = cm + (size_t)l * nk;
    for (int j = 0; j < nk; ++j) {
      sl += row[j] * fL[j];
      sr += row[j] * fR[j];
    }
    bl[l] = sl * dk / 2.0 / M_PI;
    br[l] = sr * dk / 2.0 / M_PI;
  }
  lu_solve2(g.basis->lu, g.basis->piv, nk, bl, br);
  fnl = std::move(bl);
  fnr = std::move(br);
}

void build_rhs_blocks(const Geometry& g, const std::vector<cplx>& fnl, const std::vector<cplx>& fnr,
                      std::vector<double>& first, std::vector<double>& last) {
  const int n = g.n, nk = g.nk;
  first.assign((size_t)2 * n * nk, 0.0);
  last.assign((size_t)2 * n * nk, 0.0);
  for (int a = 0; a < n; ++a)
    for (int j = 0; j < nk; ++j) {
      const double kk = g.k[j];
      const size_t i = (size_t)a * nk + j;
      if (kk > 0) {  // Q1 = invM(:,1)/detJx on element 1 (get_Diff.m:253-254,257)
        const cplx v = g.q_diff * g.invM[a * n + 0] / g.jac * std::abs(kk) * fnl[j];
        first[2 * i] = v.real();
        first[2 * i + 1] = v.imag();
      } else if (kk < 0) {  // Q2 = invM(:,n)/detJx on the last element (:255-256,258)
        const cplx v = g.q_diff * g.invM[a * n + (n - 1)] / g.jac * std::abs(kk) * fnr[j];
        last[2 * i] = v.real();
        last[2 * i + 1] = v.imag();
      }
    }
}

void build_rhs(const Geometry& g, const std::vector<cplx>& fnl, const std::vector<cplx>& fnr,
               std::vector<double>& rhs) {
  const int n = g.n, nk = g.nk, ne = g.n_chi;
  const size_t N = (size_t)ne * n * nk;
  rhs.assign(2 * N, 0.0);
  for (int a = 0; a < n; ++a)
    for (int j = 0; j < nk; ++j) {
      const double kk = g.k[j];
      if (kk > 0) {  // Q1 = invM(:,1)/detJx on element 1 (get_Diff.m:253-254,257)
        cplx v = g.q_diff * g.invM[a * n + 0] / g.jac * std::abs(kk) * fnl[j];
        size_t i = ((size_t)0 * n + a) * nk + j;
        rhs[2 * i] = v.real();
        rhs[2 * i + 1] = v.imag();
      } else if (kk < 0) {  // Q2 = invM(:,n)/detJx on the last element (:255-256,258)
        cplx v = g.q_diff * g.invM[a * n + (n - 1)] / g.jac * std::abs(kk) * fnr[j];
        size_t i = ((size_t)(ne - 1) * n + a) * nk + j;
        rhs[2 * i] = v.real();
        rhs[2 * i + 1] = v.imag();
      }
    }
}

}  // namespace wdg
