// Host-side setup of the 1D-x x 1D-k Wigner operator: grids, DG element
// matrices, plane-wave basis, inflow data.  Pure C++ (no HIP), so that it is
// usable and testable without a GPU.
#pragma once

#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "../../include/wigner_dg.h"

namespace wdg {

using cplx = std::complex<double>;

// Reference constants, as each reference file spells them.
constexpr double kHbarIP = 1.054571817e-34;           // DG_code/initParams.m:101
constexpr double kMElektronIP = 0.041 * 9.1094e-31;   // initParams.m:102 (hard-coded 0.041 m0)
constexpr double kQIP = 1.602176634e-19;              // initParams.m:103
constexpr double kHVT = 6.626e-34;                    // verteilung.m:6, solve_transport_DG_2D.m:6
constexpr double kM0VT = 9.109e-31;                   // verteilung.m:7
constexpr double kKBVT = 1.38e-23;                    // verteilung.m:9
constexpr double kQVT = 1.602e-19;                    // verteilung.m:10
constexpr int kNq = 10;                               // initParams.m:61 (Nqx)

struct Geometry {
  int n_chi = 0, n = 0, n_xi = 0, nk = 0;
  double l_chi = 0, l_xi = 0;
  double delta_chi = 0, delta_xi = 0, jac = 0;  // jac = delta_chi/2 (detJx)
  double q_diff = 0, q_drift = 0;
  double scale = 1;                             // s_A: device operator is A/s_A
  std::vector<double> r;                        // element GLL nodes (n)
  std::vector<double> xq, wq;                   // 10 GLL quadrature nodes/weights
  std::vector<double> Mref;                     // n*n mass (reference element)
  std::vector<double> S;                        // n*n stiffness (M*Dmatrix)
  std::vector<double> invM;                     // n*n inverse of (delta_xi/2)*Mref
  std::vector<double> K1, K2;                   // n*n, K2 = K2L + K2R
  std::vector<double> K2L, K2R, K3, K5;         // n*n
  std::vector<double> Wq;                       // kNq*n Lagrange interpolation W_qa
  std::vector<double> k;                        // nk wave numbers
  std::vector<double> y;                        // n_xi phi sample points (xi_Rechengebiet_Elem)
  std::vector<double> edges;                    // n_xi+1 B sample edges
  std::vector<double> cap;                      // CAP at the edges
  std::vector<double> xquad;                    // n_chi*kNq chi quadrature points
  std::vector<double> chi;                      // n_chi*n node coordinates (equispaced, as reference)
  // phi and what derives from it depend only on (n_xi, nk, L_xi): instances
  // of a sweep that share a geometry share one copy.
  std::shared_ptr<const struct BasisData> basis;
  const cplx* phi() const;                      // n_xi x nk, row-major phi[l*nk + j]
};

struct BasisData {
  std::vector<cplx> phi;     // n_xi x nk, row-major
  std::vector<cplx> lu;      // LU factors of phi (row-major, partial pivoting) for phi\f
  std::vector<int> piv;
  std::vector<double> cosm;  // n_xi x nk: cos(y_l k_j), the cosine transform of verteilung.m:109-110
  // phi^-1 cosm (nk x nk, row-major) for nk <= kInflowMapMax: the whole
  // f(k) -> phi \ (cosm f) inflow map as one matrix, empty above that size
  std::vector<cplx> gm;
};
constexpr int kInflowMapMax = 1024;

inline const cplx* Geometry::phi() const { return basis->phi.data(); }

// Chirp-z factorisation of the plane-wave basis:
//   phi[l][j] = exp(-i pi beta u_j^2) exp(-i pi beta v_l^2) exp(i pi beta (v_l-u_j)^2) / sqrt(L)
// with u_j = j-(nk-1)/2, v_l = l-(n_xi-1)/2, beta = h'/L, so that phi U and
// i phi^H Z are circular convolutions of length P (= 2^p 3^q >= nk+n_xi-1)
// between chirp-multiplied vectors.  All tables are FP64.
struct ChirpPlan {
  bool ok = false;            // the grids have the structure above
  int P = 0;
  std::vector<int> rad;       // radix sequence of the length-P FFT (8, 4, 2, 3)
  // Stockham base twiddles per pass: pass p (stride Ns, radix R) holds
  // exp(-2 pi i jm/(Ns R)) for jm in [0, Ns) at tw[tw_off[p] + jm].
  std::vector<int> tw_off;
  std::vector<cplx> tw;
  std::vector<cplx> a_pre;    // nk: exp(-i pi beta u^2)
  std::vector<cplx> w_post;   // n_xi: exp(-i pi beta v^2)/(sqrt(L) P)
  std::vector<cplx> g1;       // P: FFT of exp(i pi beta m^2), m in [-(nk-1), n_xi-1]
  std::vector<cplx> b_pre;    // n_xi: exp(+i pi beta v^2)
  std::vector<cplx> a_post;   // nk: i exp(+i pi beta u^2)/(sqrt(L) P)
  std::vector<cplx> g2;       // P: FFT of exp(-i pi beta m^2), m in [-(n_xi-1), nk-1]
};

// Throws std::runtime_error on invalid input.
void build_geometry(const wdg_desc& d, Geometry& g);

// Fills cp (cp.ok = false when the grids are not of the chirp form).
void build_chirp(const Geometry& g, ChirpPlan& cp);

// Near-DFT form of the same two products for square grids (nk = n_xi = N,
// N = 64 m, m = 2^a 3^b).  With c = (N-1)/2, ubar = (j - c)/c, vbar = (l - c)/c
// and s = -2 pi (beta - 1/N) c^2:
//   phi U     = fpost o sum_{m<terms} (i s)^m/m!  vbar^m o FFT_N(fpre o ubar^m o U)
//   i phi^H Z = bpost o sum_{m<terms} (-i s)^m/m! ubar^m o IFFT_N(bpre o vbar^m o Z)
// (unnormalised FFTs).  terms is the smallest count with |s|^terms/terms! <=
// trunc_tol (at most 4; ok = false if none is).
struct NdftPlan {
  bool ok = false;
  int P = 0, terms = 0;
  double s = 0, c = 0;
  std::vector<int> rad, tw_off;
  std::vector<cplx> tw;
  std::vector<cplx> fpre, fpost, bpre, bpost;   // N each
};
void build_ndft(const Geometry& g, double trunc_tol, NdftPlan& np);

// In-place FFT of length x.size() (a 2^p 3^q length): forward
// X_k = sum_m x_m exp(-2 pi i k m / P), or the unnormalised inverse.
void fft_host(std::vector<cplx>& x, bool inverse);

// Inflow distributions fnl, fnr (k basis), DG_code/core/verteilung.m:77-170.
void inflow(const Geometry& g, const wdg_potential& pot, std::vector<cplx>& fnl,
            std::vector<cplx>& fnr);

// Full right-hand side (interleaved complex, length 2N) in UNSCALED units
// (get_Diff.m:253-259).
void build_rhs(const Geometry& g, const std::vector<cplx>& fnl, const std::vector<cplx>& fnr,
               std::vector<double>& rhs);

// The only nonzero blocks of that right-hand side: element 0 (inflow of the
// k > 0 half) and element n_chi-1 (k < 0), each n*nk interleaved complex.
void build_rhs_blocks(const Geometry& g, const std::vector<cplx>& fnl, const std::vector<cplx>& fnr,
                      std::vector<double>& first, std::vector<double>& last);

}  // namespace wdg
