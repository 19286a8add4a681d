// libwdg: C ABI of the MI355X Wigner transport solver (see include/wigner_dg.h).
//
// One context = one stationary instance (grid + potential) with its own HIP
// stream and device buffers.  The solve runs restarted right-preconditioned
// GMRES entirely on the device: each restart cycle (Arnoldi steps, Givens
// rotations, Hessenberg solve, update) is enqueued without host round trips;
// kernels return immediately once the device-side convergence flag is set.
// The host synchronises once per cycle to read the FP64 true residual.
//
// Precision modes share that loop:
//   FP64  - every cycle in FP64 (FP64 MFMA GEMMs for Theta).
//   mixed - the residual r = b - A x and the update x += M^-1 V y are FP64;
//           each cycle's Krylov basis (stored in 16 bits by default),
//           preconditioner I/O and operator are FP32 (Theta by FP32 chirp-z
//           transforms, or split-BF16 MFMA GEMMs in the GEMM formulation).
//           This is GMRES-based iterative refinement with one cycle per
//           refinement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_setup.hpp"
#include "kernels.hpp"
#include "theta_fft.hpp"
#include "full2d.hpp"
#include "reservoir.hpp"
#include "poisson2d.hpp"

using namespace wdg;

namespace {

thread_local std::string g_create_error;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// A failure with its own status code (device memory exhausted, the
// reference's InvalidInitialRho / GmresMemoryLimit / LineMemoryLimit).
struct StatusError : std::runtime_error {
  int code;
  StatusError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess)                                                                  \
      throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));                      \
  } while (0)

// Workspace of wdg_precond_device_noalloc: the chunk maps of the 1D line
// sweeps (n_chunks * nk * 2 complex doubles), in a file-scope device array so
// that the entry point allocates nothing.  8 MiB, held by every GPU the
// library runs on until the process ends.  Host code never names the array
// itself (its host-side symbol is not a device address): it uses the address
// hipGetSymbolAddress returns for the current device, kept per device.
constexpr size_t kNoAllocScratch = (size_t)8 << 20;   // bytes
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double2 scratch[kNoAllocScratch / sizeof(double2)];
std::mutex g_scratch_mu;
double2* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of scratch, resolved once per device.
double2* scratch_for_current_device() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw HipError("device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double2*>(p);
  }
  return g_scratch_ptr[dev];
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    const hipError_t e = hipMalloc((void**)&p, bytes);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      p = nullptr;
      throw StatusError(WDG_ERR_MEMORY, "device memory exhausted (allocating " + std::to_string(bytes >> 20) + " MiB)");
    }
    if (e != hipSuccess) throw HipError(std::string("hipMalloc: ") + hipGetErrorString(e));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

enum ProfId {
  P_TRANSPORT = 0, P_GEMM1, P_COMBINE, P_GEMM2, P_LINE_AGG, P_LINE_APPLY, P_DOT, P_UPDATE, P_THETA_FFT, P_LINE_CARRY,
  P_F2_APPLY, P_F2_DENSE, P_F2_LINE,
  P_COUNT
};
const char* kProfNames[P_COUNT] = {"transport",   "theta_gemm1",    "theta_combine", "theta_gemm2",
                                   "line_aggregate", "line_apply",  "arnoldi_dot",   "arnoldi_update",
                                   "theta_fft",   "line_carry",     "f2_apply",      "f2_dense",
                                   "f2_line"};

}  // namespace

struct F2Data;   // full-2D operator data (full2d_host.hpp)

// Device data of the full-2D Poisson update (wdg2_poisson_setup).
struct P2Data {
  int nx = 0, nu = 0, n = 0, b = 0, xdof = 0, ydof = 0, ncen = 0, nrel = 0;
  double temp = 300;
  P2W w{};
  DevBuf<double> A, J, net, px, py, vol, ndg, nmat, v, f, chg, phi_in, bc, phi_out, cand, info;
  DevBuf<int> type, pxr, pyr;
};

struct wdg_ctx {
  std::shared_ptr<F2Data> f2;         // set for full-2D contexts
  std::shared_ptr<P2Data> p2;         // full-2D Poisson data (wdg2_poisson_setup)
  // shifted operator alpha I + beta A_hat (Crank-Nicolson steps)
  bool shift_on = false;
  double shift_a = 1.0, shift_b = 0.0;
  DevBuf<double2> d_b0;               // the normalised inflow RHS while b holds a step RHS
  Geometry g;
  std::string err;
  hipStream_t stream = nullptr;
  OpConst oc{};
  DevBuf<OpConst> d_oc;
  DevBuf<double> d_k, d_xquad, d_edges, d_cap, d_xg, d_ug;
  DevBuf<double> d_sbar64;
  DevBuf<float> d_sbar32;
  DevBuf<double2> d_s0s1, d_agg;
  DevBuf<uint32_t> d_linv16;          // line-preconditioner block inverses D^-1, BF16 complex (line_inverse_kernel)
  DevBuf<float2> d_lcpl32;            // and the FP32 coupling columns D^-1 c
  DevBuf<double> d_A1, d_A2;          // realified phi / i*phi^H, FP64 (GEMM formulation)
  DevBuf<uint16_t> d_A1b, d_A2b;      // same in BF16
  DevBuf<double2> d_T64, d_Z64;       // Theta intermediates (GEMM formulation)
  DevBuf<float2> d_T32, d_Z32;
  // Theta by chirp-z FFTs
  ChirpPlan chirp;
  bool fft_ok = false;                // grids of chirp form and the LDS plan fits
  bool use_fft = false;
  bool gemm_ready = false;
  FftPlan fplan{};
  DevBuf<double2> d_tab64;            // tw, a_pre, w_post, g1, b_pre, a_post, g2
  DevBuf<float2> d_tab32;
  ThetaTabs<double2> tt64{};
  ThetaTabs<float2> tt32{};
  size_t lds64 = 0, lds32 = 0, lds_phi = 0;
  int fft_block = 0, fft_groups = 0;  // phi_fft_kernel: threads per workgroup, transforms per workgroup
  int theta_block64 = 0, theta_block32 = 0;   // theta_fft_kernel threads per workgroup (256 or 1024)
  int theta32_base = 0, theta32_wide = 0;     // FP32 candidates: 256-thread block, wide block (0: unusable)
  size_t lds32_base = 0, lds32_wide = 0;
  bool theta32_wide_auto = false;             // the occupancy calculator prefers the wide block
  bool theta32_auto = true;                   // FP32 chirp-z block left to the autotuner (mode AUTO)
  bool theta32_tuned = false;                 // the autotuner has timed both blocks on this context
  // FP32 Theta in the near-DFT form (square grids): length-N transforms
  NdftPlan ndft;
  bool ndft_ok = false;                       // plan exists, n <= 5 and its LDS fits
  bool use_ndft = false;                      // the FP32 apply runs theta_ndft_kernel
  FftPlan nplan{};
  DevBuf<float2> d_ndft32;                    // tw, fpre, fpost, bpre, bpost
  NdftTabs ntabs{};
  int ndft_block = 0;                         // n * N / 8 threads
  size_t lds_ndft = 0;
  int apply_parts = WDG_PART_ALL;             // operator parts op_apply computes (wdg_apply_parts)
  double2* scratch_dev = nullptr;             // this device's scratch array (wdg_precond_device_noalloc)
  // wdg_set_memory_limits: GMRES basis estimate (16 N (m+1) bytes, 0: no
  // limit) and line-transverse factor bytes (the reference's 4 GiB default)
  double gmres_max_basis_bytes = 0.0;
  double line_max_factor_bytes = 4.0 * 1024 * 1024 * 1024;
  DevBuf<double2> d_b, d_x, d_r, d_y; // FP64 solve vectors (normalised units)
  DevBuf<double2> d_u64, d_v64;       // scratch for host-buffer entry points
  DevBuf<float2> d_u32, d_v32;
  DevBuf<double2> d_phic;             // two phi rows around xi = 0
  DevBuf<double2> d_obs;
  // Krylov storage (allocated per precision on demand)
  DevBuf<double2> d_V64, d_w64, d_z64, d_s64;
  DevBuf<float2> d_V32, d_w32, d_z32, d_s32;
  DevBuf<uint2> d_Vb16;               // 16-bit basis of the FP32 inner solver (2 complex per packet)
  DevBuf<float2> d_vcur32;            // FP32 copy of the newest 16-bit basis vector (when needed)
  DevBuf<double2> d_H, d_cs, d_sn, d_g, d_h1, d_h2, d_yv, d_part;
  DevBuf<GmresState> d_st;
  DevBuf<double> d_resv;   // residual history of the last solve (relative norms)
  int last_its = -1;       // Arnoldi steps of the last solve (-1: none yet)
  std::vector<cplx> fnl, fnr;        // inflow data of the current potential (k basis)
  double rhs_norm = 0;
  bool have_potential = false;
  bool have_solution = false;
  double sol_scale = 1;              // physical x = sol_scale * x_device
  long N = 0;
  int ncols = 0;
  int red_blocks = 0;
  int m_alloc = -1;
  // profiling
  unsigned prof_mask = 0;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Pending { int id; size_t ev; double units; };
  std::vector<Pending> ev_pending;
  double prof_ms[P_COUNT] = {0};
  long prof_n[P_COUNT] = {0};
  double prof_units[P_COUNT] = {0};   // algorithmic bytes (HBM kernels) or flops (GEMMs)
  std::map<long, hipGraphExec_t> graphs;  // captured cycles, keyed by (restart, precision, precond)
  GmresState* h_st = nullptr;        // pinned host copy of the device state
  std::vector<double> stage;         // host staging for warm starts
  int gemm_cfg32 = -1, gemm_cfg64 = -1;  // autotuned GEMM tile configurations
  // electrostatics
  bool have_poisson = false;
  int p_nmat = 0, p_method = 1;
  double p_temp = 300;
  DevBuf<int> d_tptr, d_tcol;
  DevBuf<double> d_tval, d_pw, d_plo, d_pup, d_pdg, d_pdop, d_pin, d_pout, d_pndg, d_pncv, d_pwork, d_pinfo;
  GmresBufs gb() {
    GmresBufs b;
    b.H = d_H.p; b.cs = d_cs.p; b.sn = d_sn.p; b.g = d_g.p; b.h1 = d_h1.p; b.h2 = d_h2.p;
    b.y = d_yv.p; b.part = d_part.p; b.st = d_st.p; b.m = m_alloc;
    b.resv = d_resv.p; b.resv_cap = (int)d_resv.n;
    return b;
  }
};

namespace {

int fail(wdg_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  else g_create_error = msg;
  return code;
}

// Failure of instance i of a batched call: the instance's own message and,
// for wdg_last_error(NULL), the same message naming the instance.
int batch_fail(wdg_ctx* c, int i, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_create_error = "instance " + std::to_string(i) + ": " + msg;
  return code;
}

template <typename F>
int guarded(wdg_ctx* c, F&& f) {
  try {
    return f();
  } catch (const HipError& e) {
    return fail(c, WDG_ERR_HIP, e.what());
  } catch (const StatusError& e) {
    return fail(c, e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(c, WDG_ERR_MEMORY, "host allocation failed");
  } catch (const std::invalid_argument& e) {
    return fail(c, WDG_ERR_ARG, e.what());
  } catch (const std::logic_error& e) {   // a call out of order (no potential, no solution)
    return fail(c, WDG_ERR_STATE, e.what());
  } catch (const std::runtime_error& e) {
    return fail(c, WDG_ERR_ARG, e.what());
  }
}

inline unsigned grid1(long work, int nt) { return (unsigned)std::max<long>(1, (work + nt - 1) / nt); }

// -------- profiling helpers --------
struct ProfScope {
  wdg_ctx* c;
  int id;
  double units;
  size_t i0 = (size_t)-1;
  ProfScope(wdg_ctx* c_, int id_, double units_) : c(c_), id(id_), units(units_) {
    if (!(c->prof_mask & (1u << id))) return;
    if (c->ev_used + 2 > c->ev_pool.size()) {
      for (int i = 0; i < 64; ++i) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        c->ev_pool.push_back(e);
      }
    }
    i0 = c->ev_used;
    c->ev_used += 2;
    HIPCHK(hipEventRecord(c->ev_pool[i0], c->stream));
  }
  ~ProfScope() {
    if (i0 == (size_t)-1) return;
    (void)hipEventRecord(c->ev_pool[i0 + 1], c->stream);
    c->ev_pending.push_back({id, i0, units});
  }
};

void prof_collect(wdg_ctx* c) {
  if (c->ev_pending.empty()) {
    c->ev_used = 0;
    return;
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto& pe : c->ev_pending) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev_pool[pe.ev], c->ev_pool[pe.ev + 1]));
    c->prof_ms[pe.id] += ms;
    c->prof_n[pe.id] += 1;
    c->prof_units[pe.id] += pe.units;
  }
  c->ev_pending.clear();
  c->ev_used = 0;
}

}  // namespace

#include "full2d_host.hpp"

namespace {

template <typename T>
void op_transport(wdg_ctx* c, const typename cvec<T>::type* u, typename cvec<T>::type* y, const GmresState* st) {
  ProfScope ps(c, P_TRANSPORT, 2.0 * c->N * sizeof(typename cvec<T>::type));
  const long work = (long)c->g.n_chi * (c->g.nk / 2);
  const dim3 grid(grid1(work, 256));
#define WDG_TRANSPORT(NN)                                                                                   \
  hipLaunchKernelGGL((transport_kernel<NN, T>), grid, dim3(256), 0, c->stream, c->d_oc.p, c->d_k.p, u, y, st); \
  break;
  switch (c->g.n) {
    case 2: WDG_TRANSPORT(2)
    case 3: WDG_TRANSPORT(3)
    case 4: WDG_TRANSPORT(4)
    case 5: WDG_TRANSPORT(5)
    case 6: WDG_TRANSPORT(6)
    case 7: WDG_TRANSPORT(7)
    case 8: WDG_TRANSPORT(8)
    default: throw std::invalid_argument("n_k_chi outside 2..8");
  }
#undef WDG_TRANSPORT
}

template <typename T, typename TS>
void op_combine(wdg_ctx* c, const TS* sbar, const typename cvec<T>::type* Tm, typename cvec<T>::type* Z,
                const GmresState* st) {
  const long work = (long)c->g.n_chi * c->g.n_xi;
  const dim3 grid(grid1(work, 256));
#define WDG_COMBINE(NN)                                                                                  \
  hipLaunchKernelGGL((combine_kernel<NN, T, TS>), grid, dim3(256), 0, c->stream, c->d_oc.p, sbar, c->d_cap.p, \
                     Tm, Z, st);                                                                         \
  break;
  switch (c->g.n) {
    case 2: WDG_COMBINE(2)
    case 3: WDG_COMBINE(3)
    case 4: WDG_COMBINE(4)
    case 5: WDG_COMBINE(5)
    case 6: WDG_COMBINE(6)
    case 7: WDG_COMBINE(7)
    case 8: WDG_COMBINE(8)
    default: throw std::invalid_argument("n_k_chi outside 2..8");
  }
#undef WDG_COMBINE
}

// GEMM tile configurations (chosen per shape by autotune()).
constexpr int kGemm64Cfgs = 3;   // (BM, BN) = (64,64), (128,64), (64,128)
constexpr int kGemm32Cfgs = 4;   // (BM, BN, waves) = (64,128,2x4), (128,128,2x2), (128,128,2x4), (64,64,2x2)

void launch_gemm64(wdg_ctx* c, int cfg, const double* A, const double* B, double* C, int M, int K, bool accum,
                   const GmresState* st) {
  const int N = c->ncols;
#define WDG_G64(BM, BN)                                                                                     \
  {                                                                                                         \
    dim3 grid(M / BM, (N + BN - 1) / BN);                                                                   \
    if (accum)                                                                                              \
      hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, true>), grid, dim3(256), 0, c->stream, A, B, C, M, N, K, st);  \
    else                                                                                                    \
      hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, false>), grid, dim3(256), 0, c->stream, A, B, C, M, N, K, st); \
  }
  if (cfg == 1 && M % 128 == 0) WDG_G64(128, 64)
  else if (cfg == 2) WDG_G64(64, 128)
  else WDG_G64(64, 64)
#undef WDG_G64
}

void launch_gemm32(wdg_ctx* c, int cfg, const uint16_t* A, const float* B, float* C, int M, int K, bool accum,
                   const GmresState* st) {
  const int N = c->ncols;
#define WDG_G32(BM, BN, WGM, WGN)                                                                            \
  {                                                                                                          \
    dim3 grid(M / BM, (N + BN - 1) / BN);                                                                    \
    if (accum)                                                                                               \
      hipLaunchKernelGGL((gemm_bf16x2_kernel<BM, BN, true, WGM, WGN>), grid, dim3(64 * WGM * WGN), 0, c->stream, \
                         A, B, C, M, N, K, st);                                                              \
    else                                                                                                     \
      hipLaunchKernelGGL((gemm_bf16x2_kernel<BM, BN, false, WGM, WGN>), grid, dim3(64 * WGM * WGN), 0, c->stream, \
                         A, B, C, M, N, K, st);                                                              \
  }
  if (cfg == 1 && M % 128 == 0) WDG_G32(128, 128, 2, 2)
  else if (cfg == 2 && M % 128 == 0) WDG_G32(128, 128, 2, 4)
  else if (cfg == 3) WDG_G32(64, 64, 2, 2)
  else WDG_G32(64, 128, 2, 4)
#undef WDG_G32
}

// Algorithmic flops of one complex GEMM: 8 real flops per complex MAC.
double gemm_flops(wdg_ctx* c, int M, int K) { return 8.0 * (M / 2) * (double)(K / 2) * c->ncols; }

void op_gemm64(wdg_ctx* c, int which, const double* A, const double* B, double* C, int M, int K, bool accum,
               const GmresState* st) {
  ProfScope ps(c, which, gemm_flops(c, M, K));
  launch_gemm64(c, c->gemm_cfg64 < 0 ? 0 : c->gemm_cfg64, A, B, C, M, K, accum, st);
}

void op_gemm32(wdg_ctx* c, int which, const uint16_t* A, const float* B, float* C, int M, int K, bool accum,
               const GmresState* st) {
  ProfScope ps(c, which, gemm_flops(c, M, K));
  launch_gemm32(c, c->gemm_cfg32 < 0 ? 0 : c->gemm_cfg32, A, B, C, M, K, accum, st);
}

// Tile autotuner: times every configuration on this context's GEMM shape
// (M' = 2 n_xi, K' = 2 nk, N = n_chi*n_k_chi) and keeps the fastest.
// Results are cached per shape for the process.
void autotune(wdg_ctx* c, bool fp64) {
  int& cfg = fp64 ? c->gemm_cfg64 : c->gemm_cfg32;
  if (cfg >= 0) return;
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, bool>, int> cache;
  const int M = 2 * c->g.n_xi, K = 2 * c->g.nk, N = c->ncols;
  const auto key = std::make_tuple(M, N, K, fp64);
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      cfg = it->second;
      return;
    }
  }
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  const int ncfg = fp64 ? kGemm64Cfgs : kGemm32Cfgs;
  int best = 0;
  float best_ms = 1e30f;
  for (int k = 0; k < ncfg; ++k) {
    auto launch = [&]() {
      if (fp64)
        launch_gemm64(c, k, c->d_A1.p, (const double*)c->d_r.p, (double*)c->d_T64.p, M, K, false, nullptr);
      else
        launch_gemm32(c, k, c->d_A1b.p, (const float*)c->d_Z32.p, (float*)c->d_T32.p, M, K, false, nullptr);
    };
    launch();
    HIPCHK(hipEventRecord(t0, c->stream));
    for (int r = 0; r < 3; ++r) launch();
    HIPCHK(hipEventRecord(t1, c->stream));
    HIPCHK(hipEventSynchronize(t1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, t0, t1));
    if (ms < best_ms) {
      best_ms = ms;
      best = k;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  cfg = best;
  std::lock_guard<std::mutex> lock(mu);
  cache[key] = best;
}

// Theta combine: read T, write Z (complex of size s), read the edge table.
double combine_bytes(wdg_ctx* c, int s, int ts) {
  return 2.0 * c->ncols * c->g.n_xi * s + (double)c->g.n_chi * c->oc.n_pairs * (c->g.n_xi + 1) * ts;
}

// y = (transport + Theta) u in FP32 by the near-DFT kernel (one workgroup of
// n * N/8 threads per element).
void op_theta_ndft(wdg_ctx* c, const float2* u, float2* y, const GmresState* st, int with_transport) {
  const dim3 grid(c->g.n_chi), block(c->ndft_block);
  const size_t lds = c->lds_ndft;
#define WDG_NDFT_P(NN, PC)                                                                                   \
  hipLaunchKernelGGL((theta_ndft_kernel<NN, PC>), grid, block, lds, c->stream, c->d_oc.p, c->nplan, c->ntabs, \
                     c->d_sbar32.p, c->d_cap.p, c->d_k.p, u, y, st, with_transport);
#define WDG_NDFT(NN)                                                                                          \
  if (c->nplan.P == 1024) {                                                                                   \
    WDG_NDFT_P(NN, 1024)                                                                                      \
  } else if (c->nplan.P == 512) {                                                                             \
    WDG_NDFT_P(NN, 512)                                                                                       \
  } else if (c->nplan.P == 256) {                                                                             \
    WDG_NDFT_P(NN, 256)                                                                                       \
  } else {                                                                                                    \
    WDG_NDFT_P(NN, 0)                                                                                         \
  }                                                                                                           \
  break;
  switch (c->g.n) {
    case 2: WDG_NDFT(2)
    case 3: WDG_NDFT(3)
    case 4: WDG_NDFT(4)
    case 5: WDG_NDFT(5)
    default: throw std::invalid_argument("the near-DFT Theta kernel covers n_k_chi <= 5");
  }
#undef WDG_NDFT
#undef WDG_NDFT_P
}

template <int NN>
void theta_ndft_attrs() {
  HIPCHK(hipFuncSetAttribute((const void*)theta_ndft_kernel<NN, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             kMaxLdsBytes));
  HIPCHK(hipFuncSetAttribute((const void*)theta_ndft_kernel<NN, 256>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             kMaxLdsBytes));
  HIPCHK(hipFuncSetAttribute((const void*)theta_ndft_kernel<NN, 512>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             kMaxLdsBytes));
  HIPCHK(hipFuncSetAttribute((const void*)theta_ndft_kernel<NN, 1024>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             kMaxLdsBytes));
}

// Near-DFT tables of the FP32 Theta kernel (square grids, n <= 5).
void setup_ndft(wdg_ctx* c) {
  const Geometry& g = c->g;
  c->ndft_ok = false;
  c->use_ndft = false;
  // truncation of the Taylor factor: below the FP32 transforms' own rounding
  build_ndft(g, 5e-6, c->ndft);
  const NdftPlan& np = c->ndft;
  if (!np.ok || g.n > 5 || (int)np.rad.size() > kFftMaxPasses) return;
  const int N = np.P, tt = N / 8;
  const size_t lds = (2 * (size_t)g.n * padded_len(N) + np.tw.size()) * sizeof(float2);
  if (lds > (size_t)kMaxLdsBytes || g.n * tt > 1024) return;
  FftPlan& pl = c->nplan;
  std::memset(&pl, 0, sizeof(pl));
  pl.P = N;
  pl.npass = (int)np.rad.size();
  pl.ntw = (int)np.tw.size();
  for (int i = 0; i < pl.npass; ++i) {
    pl.rad[i] = np.rad[i];
    pl.toff[i] = np.tw_off[i];
  }
  const std::vector<cplx>* parts[5] = {&np.tw, &np.fpre, &np.fpost, &np.bpre, &np.bpost};
  size_t tot = 0, off[5];
  for (int i = 0; i < 5; ++i) {
    off[i] = tot;
    tot += parts[i]->size();
  }
  std::vector<float2> h(tot);
  for (int i = 0; i < 5; ++i)
    for (size_t k = 0; k < parts[i]->size(); ++k)
      h[off[i] + k] = make_float2((float)(*parts[i])[k].real(), (float)(*parts[i])[k].imag());
  c->d_ndft32.alloc(tot);
  HIPCHK(hipMemcpy(c->d_ndft32.p, h.data(), tot * sizeof(float2), hipMemcpyHostToDevice));
  const float2* b = c->d_ndft32.p;
  c->ntabs = NdftTabs{b + off[0], b + off[1], b + off[2], b + off[3], b + off[4],
                      (float)np.s, (float)np.c, np.terms};
  c->ndft_block = g.n * tt;
  c->lds_ndft = lds;
  switch (g.n) {
    case 2: theta_ndft_attrs<2>(); break;
    case 3: theta_ndft_attrs<3>(); break;
    case 4: theta_ndft_attrs<4>(); break;
    case 5: theta_ndft_attrs<5>(); break;
    default: return;
  }
  c->ndft_ok = true;
}

// y += Theta u by the fused chirp-z kernel.
template <typename T>
void op_theta_fft(wdg_ctx* c, const typename cvec<T>::type* u, typename cvec<T>::type* y, const GmresState* st,
                  int with_transport) {
  constexpr bool f64 = std::is_same<T, double>::value;
  const size_t s = sizeof(typename cvec<T>::type), ts = f64 ? 8 : 4;
  // algorithmic HBM bytes: read u, write y, read the edge table
  ProfScope ps(c, P_THETA_FFT, 2.0 * c->N * s + (double)c->g.n_chi * c->oc.n_pairs * (c->g.n_xi + 1) * ts);
  if constexpr (!f64) {
    if (c->use_ndft) {
      op_theta_ndft(c, u, y, st, with_transport);
      return;
    }
  }
  const dim3 grid(c->g.n_chi), block(f64 ? c->theta_block64 : c->theta_block32);
  const size_t lds = f64 ? c->lds64 : c->lds32;
#define WDG_THETA_P(NN, PC)                                                                                    \
  if constexpr (f64)                                                                                           \
    hipLaunchKernelGGL((theta_fft_kernel<NN, double, double, 256, PC>), grid, block, lds, c->stream, c->d_oc.p, \
                       c->fplan, c->tt64, c->d_sbar64.p, c->d_cap.p, c->d_k.p, u, y, st, with_transport);                      \
  else if (c->theta_block32 > kFftThreads)                                                                     \
    hipLaunchKernelGGL((theta_fft_kernel<NN, float, float, 1024, PC>), grid, block, lds, c->stream, c->d_oc.p,  \
                       c->fplan, c->tt32, c->d_sbar32.p, c->d_cap.p, c->d_k.p, u, y, st, with_transport);                      \
  else                                                                                                         \
    hipLaunchKernelGGL((theta_fft_kernel<NN, float, float, 256, PC>), grid, block, lds, c->stream, c->d_oc.p,   \
                       c->fplan, c->tt32, c->d_sbar32.p, c->d_cap.p, c->d_k.p, u, y, st, with_transport);
// transform lengths with a compile-time plan (n <= 5), else the run-time plan
#define WDG_THETA(NN)                              \
  if (NN <= 5 && c->fplan.P == 2048) {             \
    WDG_THETA_P(NN, (NN <= 5 ? 2048 : 0))          \
  } else if (NN <= 5 && c->fplan.P == 1024) {      \
    WDG_THETA_P(NN, (NN <= 5 ? 1024 : 0))          \
  } else if (NN <= 5 && c->fplan.P == 512) {       \
    WDG_THETA_P(NN, (NN <= 5 ? 512 : 0))           \
  } else if (NN <= 5 && c->fplan.P == 128) {       \
    WDG_THETA_P(NN, (NN <= 5 ? 128 : 0))           \
  } else {                                         \
    WDG_THETA_P(NN, 0)                             \
  }                                                \
  break;
  switch (c->g.n) {
    case 2: WDG_THETA(2)
    case 3: WDG_THETA(3)
    case 4: WDG_THETA(4)
    case 5: WDG_THETA(5)
    case 6: WDG_THETA(6)
    case 7: WDG_THETA(7)
    case 8: WDG_THETA(8)
    default: throw std::invalid_argument("n_k_chi outside 2..8");
  }
#undef WDG_THETA
#undef WDG_THETA_P
}

// Workgroup autotuner of the FP32 chirp-z Theta kernel: times the 256-thread
// and the wide block on this context's grid (one warm-up and five applies
// each, HIP events on the context stream) and keeps the faster.  Results are
// cached per grid shape for the process.  Runs once per context, before its
// first mixed solve, while the block choice is AUTO.
void clear_graphs(wdg_ctx* c);

void autotune_theta32(wdg_ctx* c) {
  if (!c->use_fft || c->use_ndft || !c->theta32_auto || c->theta32_tuned || c->theta32_wide == 0) return;
  c->theta32_tuned = true;
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int>, bool> cache;
  const auto key = std::make_tuple(c->g.n_chi, c->g.n, c->g.nk, c->g.n_xi);
  int known = -1;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) known = it->second ? 1 : 0;
  }
  if (known < 0) {
    hipEvent_t t0, t1;
    HIPCHK(hipEventCreate(&t0));
    HIPCHK(hipEventCreate(&t1));
    const uint32_t mask = c->prof_mask;   // the trial applies stay out of the kernel statistics
    c->prof_mask = 0;
    HIPCHK(hipMemsetAsync(c->d_z32.p, 0, (size_t)c->N * sizeof(float2), c->stream));
    float ms[2] = {0, 0};
    for (int k = 0; k < 2; ++k) {
      c->theta_block32 = k ? c->theta32_wide : c->theta32_base;
      c->lds32 = k ? c->lds32_wide : c->lds32_base;
      op_theta_fft<float>(c, c->d_z32.p, c->d_w32.p, nullptr, 1);
      HIPCHK(hipEventRecord(t0, c->stream));
      for (int r = 0; r < 5; ++r) op_theta_fft<float>(c, c->d_z32.p, c->d_w32.p, nullptr, 1);
      HIPCHK(hipEventRecord(t1, c->stream));
      HIPCHK(hipEventSynchronize(t1));
      HIPCHK(hipEventElapsedTime(&ms[k], t0, t1));
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    c->prof_mask = mask;
    known = ms[1] < ms[0] ? 1 : 0;
    std::lock_guard<std::mutex> lock(mu);
    cache[key] = known != 0;
  }
  const bool wide = known != 0;
  c->theta_block32 = wide ? c->theta32_wide : c->theta32_base;
  c->lds32 = wide ? c->lds32_wide : c->lds32_base;
  clear_graphs(c);   // captured cycles hold the launch shape
}

// The attribute is process-wide per kernel, so it is set to the hardware
// maximum rather than to this context's plan (contexts of other grids share
// the kernel).
template <int NN, int PC>
void theta_fft_attrs_p() {
  HIPCHK(hipFuncSetAttribute((const void*)theta_fft_kernel<NN, double, double, 256, PC>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLdsBytes));
  HIPCHK(hipFuncSetAttribute((const void*)theta_fft_kernel<NN, float, float, 256, PC>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLdsBytes));
  HIPCHK(hipFuncSetAttribute((const void*)theta_fft_kernel<NN, float, float, 1024, PC>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLdsBytes));
}
template <int NN>
void theta_fft_attrs() {
  theta_fft_attrs_p<NN, 0>();
  if constexpr (NN <= 5) {
    theta_fft_attrs_p<NN, 128>();
    theta_fft_attrs_p<NN, 512>();
    theta_fft_attrs_p<NN, 1024>();
    theta_fft_attrs_p<NN, 2048>();
  }
}

// Resident workgroups per CU of the FP32 Theta kernel at the 256-thread and
// the wide block (0 where the calculator fails).
template <int NN>
void theta32_occupancy(int block, size_t lds, int wide, size_t lds_wide, int& nb_base, int& nb_wide) {
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_base, (const void*)theta_fft_kernel<NN, float, float, 256, 0>,
                                                   block, lds) != hipSuccess)
    nb_base = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb_wide,
                                                   (const void*)theta_fft_kernel<NN, float, float, 1024, 0>, wide,
                                                   lds_wide) != hipSuccess)
    nb_wide = 0;
  (void)hipGetLastError();
}

template <typename T>
void op_apply_base(wdg_ctx* c, const typename cvec<T>::type* u, typename cvec<T>::type* y, const GmresState* st);

// y = A_hat u in precision T, or (alpha I + beta A_hat) u while a shift is set.
template <typename T>
void op_apply(wdg_ctx* c, const typename cvec<T>::type* u, typename cvec<T>::type* y, const GmresState* st) {
  if (c->f2 && c->shift_on) {
    f2_apply<T>(c, u, y, st, c->shift_a, c->shift_b);   // shift fused into the apply
    return;
  }
  op_apply_base<T>(c, u, y, st);
  if (c->shift_on)
    hipLaunchKernelGGL(shift_kernel<T>, dim3(c->red_blocks), dim3(256), 0, c->stream, u, y, c->shift_a, c->shift_b,
                       c->N, st);
}

template <typename T>
void op_apply_base(wdg_ctx* c, const typename cvec<T>::type* u, typename cvec<T>::type* y, const GmresState* st) {
  if (c->f2) {
    f2_apply<T>(c, u, y, st);
    return;
  }
  const int parts = c->apply_parts;
  if (parts == WDG_PART_TRANSPORT) {
    op_transport<T>(c, u, y, st);
    return;
  }
  if (c->use_fft) {   // the chirp-z kernel writes transport + Theta
    op_theta_fft<T>(c, u, y, st, (parts & WDG_PART_TRANSPORT) ? 1 : 0);
    return;
  }
  if (parts & WDG_PART_TRANSPORT)
    op_transport<T>(c, u, y, st);
  else   // the potential part alone (never inside a captured cycle)
    HIPCHK(hipMemsetAsync(y, 0, c->N * sizeof(typename cvec<T>::type), c->stream));
  const int M1 = 2 * c->g.n_xi, K1 = 2 * c->g.nk;
  const long work = (long)c->g.n_chi * c->g.n_xi;
  if constexpr (std::is_same<T, double>::value) {
    op_gemm64(c, P_GEMM1, c->d_A1.p, (const double*)u, (double*)c->d_T64.p, M1, K1, false, st);
    {
      ProfScope ps(c, P_COMBINE, combine_bytes(c, 16, 8));
      op_combine<double, double>(c, c->d_sbar64.p, c->d_T64.p, c->d_Z64.p, st);
    }
    op_gemm64(c, P_GEMM2, c->d_A2.p, (const double*)c->d_Z64.p, (double*)y, 2 * c->g.nk, 2 * c->g.n_xi, true, st);
  } else {
    op_gemm32(c, P_GEMM1, c->d_A1b.p, (const float*)u, (float*)c->d_T32.p, M1, K1, false, st);
    {
      ProfScope ps(c, P_COMBINE, combine_bytes(c, 8, 4));
      op_combine<float, float>(c, c->d_sbar32.p, c->d_T32.p, c->d_Z32.p, st);
    }
    op_gemm32(c, P_GEMM2, c->d_A2b.p, (const float*)c->d_Z32.p, (float*)y, 2 * c->g.nk, 2 * c->g.n_xi, true, st);
  }
}

template <int NN, typename T>
void op_precond_n(wdg_ctx* c, const typename cvec<T>::type* r, typename cvec<T>::type* z, const GmresState* st,
                  double2* agg) {
  const long work = (long)c->oc.n_chunks * c->g.nk;
  if constexpr (std::is_same<T, float>::value) {
    // FP32: stream the stored block inverses
    const double rows = (double)c->N / NN;   // (element, k) blocks
    {
      ProfScope ps(c, P_LINE_AGG, 8.0 * c->N + rows * (4.0 * NN + 8.0));
      hipLaunchKernelGGL((line_aggregate_inv_kernel<NN>), dim3(grid1(work, 256)), dim3(256), 0, c->stream,
                         c->d_oc.p, c->d_k.p, c->d_linv16.p, c->d_lcpl32.p, r, agg, st);
    }
    {
      ProfScope ps(c, P_LINE_CARRY, 3.0 * c->oc.n_chunks * c->g.nk * 2 * sizeof(double2));
      hipLaunchKernelGGL(line_carry_kernel, dim3(grid1(64L * c->g.nk, 256)), dim3(256), 0, c->stream, c->d_oc.p,
                         c->d_k.p, agg, st);
    }
    {
      ProfScope ps(c, P_LINE_APPLY, 16.0 * c->N + rows * (4.0 * NN * NN + 8.0 * NN));
      hipLaunchKernelGGL((line_apply_inv_kernel<NN>), dim3(grid1(work, 256)), dim3(256), 0, c->stream, c->d_oc.p,
                         c->d_k.p, c->d_linv16.p, c->d_lcpl32.p, agg, r, z, st);
    }
    return;
  }
  {
    ProfScope ps(c, P_LINE_AGG, 1.0 * c->N * sizeof(typename cvec<T>::type));
    hipLaunchKernelGGL((line_aggregate_kernel<NN, T>), dim3(grid1(work, 256)), dim3(256), 0, c->stream, c->d_oc.p,
                       c->d_k.p, c->d_s0s1.p, r, agg, st);
  }
  {
    // chunk maps read twice, carries written once
    ProfScope ps(c, P_LINE_CARRY, 3.0 * c->oc.n_chunks * c->g.nk * 2 * sizeof(double2));
    hipLaunchKernelGGL(line_carry_kernel, dim3(grid1(64L * c->g.nk, 256)), dim3(256), 0, c->stream, c->d_oc.p,
                       c->d_k.p, agg, st);
  }
  {
    ProfScope ps(c, P_LINE_APPLY, 2.0 * c->N * sizeof(typename cvec<T>::type));
    hipLaunchKernelGGL((line_apply_kernel<NN, T>), dim3(grid1(work, 256)), dim3(256), 0, c->stream, c->d_oc.p,
                       c->d_k.p, c->d_s0s1.p, agg, r, z, st);
  }
}

// z = M^-1 r; agg is the chunk-map workspace of the 1D line sweeps
// (n_chunks * nk * 2 complex; the context's own unless given).
template <typename T>
void op_precond(wdg_ctx* c, const typename cvec<T>::type* r, typename cvec<T>::type* z, const GmresState* st,
                double2* agg = nullptr) {
  if (c->f2) {
    f2_precond<T>(c, r, z, st);
    return;
  }
  if (!agg) agg = c->d_agg.p;
  switch (c->g.n) {
    case 2: op_precond_n<2, T>(c, r, z, st, agg); break;
    case 3: op_precond_n<3, T>(c, r, z, st, agg); break;
    case 4: op_precond_n<4, T>(c, r, z, st, agg); break;
    case 5: op_precond_n<5, T>(c, r, z, st, agg); break;
    case 6: op_precond_n<6, T>(c, r, z, st, agg); break;
    case 7: op_precond_n<7, T>(c, r, z, st, agg); break;
    case 8: op_precond_n<8, T>(c, r, z, st, agg); break;
    default: throw std::invalid_argument("n_k_chi outside 2..8");
  }
}

// The FP32 line preconditioner applied to a 16-bit Krylov basis vector read
// in place (RIN = ushort2 for the FP16 basis, uint32_t for BF16): the same
// stored-inverse sweeps as op_precond_n<NN, float>, without an FP32 copy.
template <int NN, typename RIN>
void op_precond_basis_n(wdg_ctx* c, const RIN* r, float2* z, const GmresState* st) {
  const long work = (long)c->oc.n_chunks * c->g.nk;
  const double rows = (double)c->N / NN;
  {
    ProfScope ps(c, P_LINE_AGG, 4.0 * c->N + rows * (4.0 * NN + 8.0));
    hipLaunchKernelGGL((line_aggregate_inv_kernel<NN, RIN>), dim3(grid1(work, 256)), dim3(256), 0, c->stream,
                       c->d_oc.p, c->d_k.p, c->d_linv16.p, c->d_lcpl32.p, r, c->d_agg.p, st);
  }
  {
    ProfScope ps(c, P_LINE_CARRY, 3.0 * c->oc.n_chunks * c->g.nk * 2 * sizeof(double2));
    hipLaunchKernelGGL(line_carry_kernel, dim3(grid1(64L * c->g.nk, 256)), dim3(256), 0, c->stream, c->d_oc.p,
                       c->d_k.p, c->d_agg.p, st);
  }
  {
    ProfScope ps(c, P_LINE_APPLY, 12.0 * c->N + rows * (4.0 * NN * NN + 8.0 * NN));
    hipLaunchKernelGGL((line_apply_inv_kernel<NN, RIN>), dim3(grid1(work, 256)), dim3(256), 0, c->stream,
                       c->d_oc.p, c->d_k.p, c->d_linv16.p, c->d_lcpl32.p, c->d_agg.p, r, z, st);
  }
}

template <typename VP>
void op_precond_basis(wdg_ctx* c, const VP* v, float2* z, const GmresState* st) {
  using RIN = std::conditional_t<std::is_same<VP, ushort4>::value, ushort2, uint32_t>;
  const RIN* r = reinterpret_cast<const RIN*>(v);
  switch (c->g.n) {
    case 2: op_precond_basis_n<2, RIN>(c, r, z, st); break;
    case 3: op_precond_basis_n<3, RIN>(c, r, z, st); break;
    case 4: op_precond_basis_n<4, RIN>(c, r, z, st); break;
    case 5: op_precond_basis_n<5, RIN>(c, r, z, st); break;
    case 6: op_precond_basis_n<6, RIN>(c, r, z, st); break;
    case 7: op_precond_basis_n<7, RIN>(c, r, z, st); break;
    case 8: op_precond_basis_n<8, RIN>(c, r, z, st); break;
    default: throw std::invalid_argument("n_k_chi outside 2..8");
  }
}

void clear_graphs(wdg_ctx* c);

void ensure_krylov(wdg_ctx* c, int m, bool fp64, bool b16) {
  if (m > kMaxRestart) throw std::invalid_argument("restart exceeds the supported maximum of 64");
  // Captured graphs hold raw pointers: drop them before any buffer moves.
  const void* before[8] = {c->d_V64.p, c->d_V32.p, c->d_part.p, c->d_H.p,
                           c->d_w64.p, c->d_w32.p, c->d_Vb16.p, c->d_vcur32.p};
  if (c->m_alloc != m) {
    clear_graphs(c);
    c->d_H.alloc((size_t)(m + 1) * m);
    c->d_cs.alloc(m);
    c->d_sn.alloc(m);
    c->d_g.alloc(m + 1);
    c->d_h1.alloc(m + 2);
    c->d_h2.alloc(m + 2);
    c->d_yv.alloc(m);
    c->m_alloc = m;
  }
  const int ngrp = (m + kDotGroup) / kDotGroup + 1;
  c->d_part.alloc((size_t)c->red_blocks * ngrp * (kDotGroup + 1));
  if (!c->d_st.p) {
    c->d_st.alloc(1);
    HIPCHK(hipMemset(c->d_st.p, 0, sizeof(GmresState)));
  }
  const size_t N = c->N;
  if (fp64) {
    c->d_V64.alloc((size_t)(m + 1) * N);
    c->d_w64.alloc(N);
    c->d_z64.alloc(N);
    c->d_s64.alloc(N);
  } else {
    if (b16) {
      c->d_Vb16.alloc((size_t)(m + 1) * (N / 2));   // 2 complex per 8-byte packet
      c->d_vcur32.alloc(N);
    } else {
      c->d_V32.alloc((size_t)(m + 1) * N);
    }
    c->d_w32.alloc(N);
    c->d_z32.alloc(N);
    c->d_s32.alloc(N);
  }
  const void* after[8] = {c->d_V64.p, c->d_V32.p, c->d_part.p, c->d_H.p,
                          c->d_w64.p, c->d_w32.p, c->d_Vb16.p, c->d_vcur32.p};
  for (int i = 0; i < 8; ++i)
    if (before[i] && before[i] != after[i]) clear_graphs(c);
}

// r = b - A x in FP64; ||r|| into st->resnorm.
void residual64(wdg_ctx* c) {
  op_apply<double>(c, c->d_x.p, c->d_y.p, nullptr);
  hipLaunchKernelGGL((residual_kernel<256>), dim3(c->red_blocks), dim3(256), 0, c->stream, c->d_b.p, c->d_y.p,
                     c->d_r.p, c->N, c->d_part.p, c->d_st.p);
  hipLaunchKernelGGL((residual_finalize_kernel<256>), dim3(1), dim3(256), 0, c->stream, c->d_part.p, c->red_blocks,
                     c->d_st.p);
}

// One restart cycle in precision T (no host synchronisation).  VP is the
// basis packet: the working packet of T, or uint2 / ushort4 for a BF16 /
// FP16 basis of the FP32 solver.  With a 16-bit basis the 1D line
// preconditioner reads the stored vector in place; otherwise the newest
// vector is also kept in FP32 in vcur for the preconditioner and operator.
template <typename T, typename VP>
void gmres_cycle(wdg_ctx* c, int m, bool precond, bool single_pass) {
  using T2 = typename cvec<T>::type;
  using PT = typename Pk<T>::type;
  constexpr bool narrow = std::is_same<VP, uint2>::value || std::is_same<VP, ushort4>::value;
  T2 *w, *z, *s;
  VP* V;
  T2* vcur = nullptr;
  if constexpr (std::is_same<T, double>::value) {
    V = reinterpret_cast<VP*>(c->d_V64.p); w = c->d_w64.p; z = c->d_z64.p; s = c->d_s64.p;
  } else {
    w = c->d_w32.p; z = c->d_z32.p; s = c->d_s32.p;
    if constexpr (narrow) {
      V = reinterpret_cast<VP*>(c->d_Vb16.p);
      vcur = c->d_vcur32.p;
    } else {
      V = reinterpret_cast<VP*>(c->d_V32.p);
    }
  }
  const long N = c->N;
  const long NP = N * (long)sizeof(T2) / 16;        // working packets per vector
  const double vbytes = (double)NP * sizeof(VP);    // bytes of one stored basis vector
  const bool wide = sizeof(VP) == 8 && NP % 2 == 0;
  // the 1D line preconditioner and the full-2D scalings read a 16-bit basis
  // vector in place, so the newest vector needs no FP32 copy
  const bool direct16 = narrow && precond && (!c->f2 || f2_precond_reads_basis(c));
  PT* wp = reinterpret_cast<PT*>(w);
  PT* vcp = direct16 ? nullptr : reinterpret_cast<PT*>(vcur);
  GmresBufs gb = c->gb();
  const GmresState* st = c->d_st.p;
  hipLaunchKernelGGL(cycle_init_kernel, dim3(1), dim3(1), 0, c->stream, gb);
  hipLaunchKernelGGL((residual_normalize_kernel<T, VP>), dim3(c->red_blocks), dim3(256), 0, c->stream, c->d_r.p, V,
                     vcp, NP, st);
  for (int j = 0; j < m; ++j) {
    const T2* vj = narrow ? vcur : reinterpret_cast<const T2*>(V + (size_t)j * NP);
    const T2* opin = vj;
    if (precond) {
      bool done = false;
      if constexpr (narrow && std::is_same<T, float>::value) {
        if (direct16) {
          if (c->f2) f2_precond_basis<VP>(c, V + (size_t)j * NP, z, st);
          else op_precond_basis<VP>(c, V + (size_t)j * NP, z, st);
          done = true;
        }
      }
      if (!done) op_precond<T>(c, vj, z, st);
      opin = z;
    }
    op_apply<T>(c, opin, w, st);
    const int nv = j + 1;
    // 8-byte basis packets with an even packet count: the multi-dot and the
    // update read two packets per 16-byte load, the multi-dot on half the
    // blocks (fewer, longer blocks measured faster for it)
    const int nbd = wide ? std::max(1, c->red_blocks / 2) : c->red_blocks;
    const dim3 dg(nbd, (nv + kDotGroup - 1) / kDotGroup);
    const int nb = c->red_blocks;
    auto dot = [&](int mode) {
      if constexpr (sizeof(VP) == 8) {
        if (wide) {
          hipLaunchKernelGGL((arnoldi_dot_kernel<T, VP, 256, 2>), dg, dim3(256), 0, c->stream, gb, V, wp, NP, nv,
                             mode);
          return;
        }
      }
      hipLaunchKernelGGL((arnoldi_dot_kernel<T, VP, 256, 1>), dg, dim3(256), 0, c->stream, gb, V, wp, NP, nv, mode);
    };
    auto update = [&](int mode) {
      if constexpr (sizeof(VP) == 8) {
        if (wide) {
          hipLaunchKernelGGL((arnoldi_update_kernel<T, VP, 256, 2>), dim3(nb), dim3(256), 0, c->stream, gb, V, wp,
                             NP, j, mode);
          return;
        }
      }
      hipLaunchKernelGGL((arnoldi_update_kernel<T, VP, 256, 1>), dim3(nb), dim3(256), 0, c->stream, gb, V, wp, NP,
                         j, mode);
    };
    {
      ProfScope ps(c, P_DOT, nv * vbytes + N * sizeof(T2));
      dot(0);
    }
    hipLaunchKernelGGL((arnoldi_dot_finalize_kernel<256>), dim3(dg.y), dim3(256), 0, c->stream, gb, nbd, nv, 0);
    if (single_pass) {
      // Pythagorean norm: the update writes V_{j+1} directly; the finish
      // pass runs only for a step whose norm cancels (st->slow_norm)
      VP* vnext = V + (size_t)(j + 1) * NP;
      {
        ProfScope ps(c, P_UPDATE, nv * vbytes + N * sizeof(T2) + vbytes);
        if constexpr (sizeof(VP) == 8) {
          if (wide) {
            hipLaunchKernelGGL((arnoldi_update_pyth_kernel<T, VP, 256, 2>), dim3(nb), dim3(256), 0, c->stream, gb,
                               V, wp, vnext, vcp, NP, j);
          } else {
            hipLaunchKernelGGL((arnoldi_update_pyth_kernel<T, VP, 256, 1>), dim3(nb), dim3(256), 0, c->stream, gb,
                               V, wp, vnext, vcp, NP, j);
          }
        } else {
          hipLaunchKernelGGL((arnoldi_update_pyth_kernel<T, VP, 256, 1>), dim3(nb), dim3(256), 0, c->stream, gb, V,
                             wp, vnext, vcp, NP, j);
        }
      }
      // the fallback pass (a step whose norm cancels) is rare: a small grid
      // keeps the common, empty launch short (its loops are grid-stride)
      hipLaunchKernelGGL((arnoldi_finish_kernel<T, VP, 256>), dim3(std::min(nb, 128)), dim3(256), 0, c->stream,
                         gb, nb, j, wp, vnext, vcp, NP, 1);
      continue;
    }
    {
      ProfScope ps(c, P_UPDATE, nv * vbytes + 2.0 * N * sizeof(T2));
      update(0);
    }
    hipLaunchKernelGGL((arnoldi_update_finalize_kernel<256>), dim3(1), dim3(256), 0, c->stream, gb, nb, j, 0);
    // DGKS second pass, taken on the device only when requested (st->reorth)
    dot(1);
    hipLaunchKernelGGL((arnoldi_dot_finalize_kernel<256>), dim3(dg.y), dim3(256), 0, c->stream, gb, nbd, nv, 1);
    update(1);
    hipLaunchKernelGGL((arnoldi_update_finalize_kernel<256>), dim3(1), dim3(256), 0, c->stream, gb, nb, j, 1);
    hipLaunchKernelGGL((arnoldi_normalize_kernel<T, VP>), dim3(c->red_blocks), dim3(256), 0, c->stream, gb, wp,
                       V + (size_t)(j + 1) * NP, vcp, NP);
  }
  hipLaunchKernelGGL(hess_solve_kernel, dim3(1), dim3(256), 0, c->stream, gb);
  hipLaunchKernelGGL((lincomb_kernel<T, VP>), dim3(c->red_blocks), dim3(256), 0, c->stream, gb, V,
                     reinterpret_cast<PT*>(s), NP);
  const T2* corr = s;
  if (precond) {
    op_precond<T>(c, s, z, nullptr);
    corr = z;
  }
  hipLaunchKernelGGL((axpy_to64_kernel<T>), dim3(c->red_blocks), dim3(256), 0, c->stream, c->d_x.p, corr, 1.0, N);
}

// One refinement cycle followed by the FP64 true residual.  With graphs
// enabled (and no per-kernel event profiling) the whole sequence is captured
// once per (precision, restart, preconditioner) and replayed.
void run_cycle(wdg_ctx* c, bool fp64, int basis, int m, bool precond, bool single_pass, bool use_graph) {
  auto enqueue = [&]() {
    if (fp64) gmres_cycle<double, double2>(c, m, precond, single_pass);
    else if (basis == WDG_BASIS_BF16) gmres_cycle<float, uint2>(c, m, precond, single_pass);
    else if (basis == WDG_BASIS_FP16) gmres_cycle<float, ushort4>(c, m, precond, single_pass);
    else gmres_cycle<float, float4>(c, m, precond, single_pass);
    residual64(c);
  };
  if (!use_graph || c->prof_mask) {
    enqueue();
    return;
  }
  const long key = ((long)m << 6) | (c->shift_on ? 32 : 0) | (single_pass ? 16 : 0) | ((long)(fp64 ? 0 : basis) << 2) |
                   (fp64 ? 2 : 0) | (precond ? 1 : 0);
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue();
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(e);
    it = c->graphs.emplace(key, ge).first;
  }
  HIPCHK(hipGraphLaunch(it->second, c->stream));
}

void clear_graphs(wdg_ctx* c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
}

void upload_const(wdg_ctx* c) {
  HIPCHK(hipMemcpyAsync(c->d_oc.p, &c->oc, sizeof(OpConst), hipMemcpyHostToDevice, c->stream));
}

uint16_t host_bf16(double v) {
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Realified phi for GEMM1 (T = phi U) and i*phi^H for GEMM2 (Y = i phi^H Z),
// and the T/Z intermediates: only needed by the GEMM formulation of Theta.
void ensure_gemm(wdg_ctx* c) {
  if (c->gemm_ready) return;
  const Geometry& g = c->g;
  auto up = [&](DevBuf<double>& b, const std::vector<double>& v) {
    b.alloc(v.size());
    HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  };
  const int nx = g.n_xi, nk = g.nk;
  std::vector<double> A1((size_t)2 * nx * 2 * nk), A2((size_t)2 * nk * 2 * nx);
  for (int l = 0; l < nx; ++l)
    for (int j = 0; j < nk; ++j) {
      const double pr = g.phi()[(size_t)l * nk + j].real(), pi = g.phi()[(size_t)l * nk + j].imag();
      const size_t K1 = 2 * nk, K2 = 2 * nx;
      A1[(size_t)(2 * l) * K1 + 2 * j] = pr;
      A1[(size_t)(2 * l) * K1 + 2 * j + 1] = -pi;
      A1[(size_t)(2 * l + 1) * K1 + 2 * j] = pi;
      A1[(size_t)(2 * l + 1) * K1 + 2 * j + 1] = pr;
      // psi(j,l) = i conj(phi(l,j)) = pi + i pr
      A2[(size_t)(2 * j) * K2 + 2 * l] = pi;
      A2[(size_t)(2 * j) * K2 + 2 * l + 1] = -pr;
      A2[(size_t)(2 * j + 1) * K2 + 2 * l] = pr;
      A2[(size_t)(2 * j + 1) * K2 + 2 * l + 1] = pi;
    }
  up(c->d_A1, A1);
  up(c->d_A2, A2);
  std::vector<uint16_t> A1b(A1.size()), A2b(A2.size());
  for (size_t i = 0; i < A1.size(); ++i) A1b[i] = host_bf16(A1[i]);
  for (size_t i = 0; i < A2.size(); ++i) A2b[i] = host_bf16(A2[i]);
  c->d_A1b.alloc(A1b.size());
  c->d_A2b.alloc(A2b.size());
  HIPCHK(hipMemcpyAsync(c->d_A1b.p, A1b.data(), A1b.size() * 2, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_A2b.p, A2b.data(), A2b.size() * 2, hipMemcpyHostToDevice, c->stream));
  const size_t tz = (size_t)c->ncols * nx;
  c->d_T64.alloc(tz);
  c->d_Z64.alloc(tz);
  c->d_T32.alloc(tz);
  c->d_Z32.alloc(tz);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->gemm_ready = true;
}

// Chirp-z tables and the LDS plan of the fused Theta kernel.  Theta runs on
// FFTs whenever the grids allow it and the plan fits in LDS, else on GEMMs.
void setup_theta(wdg_ctx* c) {
  const Geometry& g = c->g;
  build_chirp(g, c->chirp);
  const ChirpPlan& cp = c->chirp;
  c->fft_ok = false;
  if (cp.ok && (int)cp.rad.size() <= kFftMaxPasses) {
    const int tt = cp.P / 8;
    const int ntw = (int)cp.tw.size();
    c->fft_groups = std::max(1, kFftThreads / tt);
    c->fft_block = tt * c->fft_groups;
    c->lds_phi = (2 * (size_t)c->fft_groups * padded_len(cp.P) + cp.tw.size()) * sizeof(double2);
    c->theta_block64 = c->fft_block;
    c->lds64 = (size_t)theta_lds_words(g.n, g.n_xi, cp.P, c->fft_groups, ntw) * sizeof(double2);
    c->theta_block32 = c->fft_block;
    c->lds32 = (size_t)theta_lds_words(g.n, g.n_xi, cp.P, c->fft_groups, ntw) * sizeof(float2);
    c->fft_ok = tt <= kFftThreads && (cp.P % 64) == 0 && c->lds64 <= (size_t)kMaxLdsBytes;
  }
  if (c->fft_ok) {
    FftPlan& pl = c->fplan;
    std::memset(&pl, 0, sizeof(pl));
    pl.P = cp.P;
    pl.npass = (int)cp.rad.size();
    pl.ntw = (int)cp.tw.size();
    for (int i = 0; i < pl.npass; ++i) {
      pl.rad[i] = cp.rad[i];
      pl.toff[i] = cp.tw_off[i];
    }
    const std::vector<cplx>* parts[7] = {&cp.tw, &cp.a_pre, &cp.w_post, &cp.g1, &cp.b_pre, &cp.a_post, &cp.g2};
    size_t tot = 0, off[7];
    for (int i = 0; i < 7; ++i) {
      off[i] = tot;
      tot += parts[i]->size();
    }
    std::vector<double2> h64(tot);
    std::vector<float2> h32(tot);
    for (int i = 0; i < 7; ++i)
      for (size_t k = 0; k < parts[i]->size(); ++k) {
        const cplx v = (*parts[i])[k];
        h64[off[i] + k] = make_double2(v.real(), v.imag());
        h32[off[i] + k] = make_float2((float)v.real(), (float)v.imag());
      }
    c->d_tab64.alloc(tot);
    c->d_tab32.alloc(tot);
    HIPCHK(hipMemcpyAsync(c->d_tab64.p, h64.data(), tot * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_tab32.p, h32.data(), tot * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    const double2* b64 = c->d_tab64.p;
    const float2* b32 = c->d_tab32.p;
    c->tt64 = ThetaTabs<double2>{b64 + off[0], b64 + off[1], b64 + off[2], b64 + off[3],
                                 b64 + off[4], b64 + off[5], b64 + off[6]};
    c->tt32 = ThetaTabs<float2>{b32 + off[0], b32 + off[1], b32 + off[2], b32 + off[3],
                                b32 + off[4], b32 + off[5], b32 + off[6]};
    switch (g.n) {
      case 2: theta_fft_attrs<2>(); break;
      case 3: theta_fft_attrs<3>(); break;
      case 4: theta_fft_attrs<4>(); break;
      case 5: theta_fft_attrs<5>(); break;
      case 6: theta_fft_attrs<6>(); break;
      case 7: theta_fft_attrs<7>(); break;
      case 8: theta_fft_attrs<8>(); break;
      default: break;
    }
    // FP32 Theta: 256 threads, or a wide block of ~1024 threads (four or
    // more transforms in flight), whichever keeps more waves resident per CU
    // (LDS and VGPRs both counted by the occupancy calculator).  n <= 5: the
    // wide kernel's combine of more nodes would spill past 128 VGPRs.
    {
      const int tt = cp.P / 8, wide = tt * (1024 / tt);
      const size_t lw = (size_t)theta_lds_words(g.n, g.n_xi, cp.P, wide / tt, (int)cp.tw.size()) * sizeof(float2);
      int nb_base = 0, nb_wide = 0;
      switch (g.n) {
        case 2: theta32_occupancy<2>(c->fft_block, c->lds32, wide, lw, nb_base, nb_wide); break;
        case 3: theta32_occupancy<3>(c->fft_block, c->lds32, wide, lw, nb_base, nb_wide); break;
        case 4: theta32_occupancy<4>(c->fft_block, c->lds32, wide, lw, nb_base, nb_wide); break;
        case 5: theta32_occupancy<5>(c->fft_block, c->lds32, wide, lw, nb_base, nb_wide); break;
        default: break;
      }
      c->theta32_base = c->fft_block;
      c->lds32_base = c->lds32;
      if (lw <= (size_t)kMaxLdsBytes && g.n <= 5 && wide / tt >= 4) {
        c->theta32_wide = wide;
        c->lds32_wide = lw;
      }
      c->theta32_wide_auto =
          c->theta32_wide > 0 && nb_wide * ((wide + 63) / 64) > nb_base * ((c->fft_block + 63) / 64);
      if (c->theta32_wide_auto) {
        c->theta_block32 = wide;
        c->lds32 = lw;
      }
    }
    HIPCHK(hipFuncSetAttribute((const void*)phi_fft_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                               kMaxLdsBytes));
    c->use_fft = true;
    setup_ndft(c);
  } else {
    c->use_fft = false;
    ensure_gemm(c);
  }
}

void create_impl(wdg_ctx* c, const wdg_desc& d) {
  build_geometry(d, c->g);
  const Geometry& g = c->g;
  const int n = g.n;
  c->N = (long)g.n_chi * n * g.nk;
  c->ncols = g.n_chi * n;
  c->red_blocks = (int)std::min<long>(1024, std::max<long>(1, c->N / (256 * 8)));
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  OpConst& oc = c->oc;
  std::memset(&oc, 0, sizeof(oc));
  oc.n_chi = g.n_chi;
  oc.n = n;
  oc.nk = g.nk;
  oc.n_xi = g.n_xi;
  oc.n_pairs = n * (n + 1) / 2;
  // Line-solve chunks: enough (chunk, k) threads to fill the chip.
  {
    const long want = 131072;
    int S = (int)std::max<long>(2, ((long)g.n_chi * g.nk + want - 1) / want);
    S = std::min(S, 16);
    oc.chunk = std::min(S, g.n_chi);
    oc.n_chunks = (g.n_chi + oc.chunk - 1) / oc.chunk;
  }
  oc.tq = g.q_diff / g.jac / g.scale;
  oc.inv_len = 1.0 / g.l_xi;
  oc.hprime = g.y[1] - g.y[0];
  for (int a = 0; a < n; ++a) {
    for (int b = 0; b < n; ++b) {
      oc.K1[a * kMaxN + b] = g.K1[a * n + b];
      oc.K2[a * kMaxN + b] = g.K2[a * n + b];
      oc.cab[a * kMaxN + b] = g.q_drift * g.invM[a * n + b] * g.delta_xi / 4.0 / g.scale;
      double mq = 0;
      for (int q = 0; q < kNq; ++q) mq += g.wq[q] * g.Wq[q * n + a] * g.Wq[q * n + b];
      oc.Mq[a * kMaxN + b] = mq;
    }
    oc.invM0[a] = g.invM[a * n + 0];
    oc.invMn[a] = g.invM[a * n + (n - 1)];
  }
  int p = 0;
  for (int a = 0; a < n; ++a)
    for (int b = a; b < n; ++b) {
      oc.pa[p] = a;
      oc.pb[p] = b;
      oc.pidx[a * kMaxN + b] = p;
      oc.pidx[b * kMaxN + a] = p;
      for (int q = 0; q < kNq; ++q) oc.Sq[q * kMaxPairs + p] = g.wq[q] * g.Wq[q * n + a] * g.Wq[q * n + b];
      ++p;
    }
  c->d_oc.alloc(1);
  upload_const(c);
  auto up = [&](DevBuf<double>& b, const std::vector<double>& v) {
    b.alloc(v.size());
    HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  };
  up(c->d_k, g.k);
  up(c->d_xquad, g.xquad);
  up(c->d_edges, g.edges);
  up(c->d_cap, g.cap);
  setup_theta(c);
  const int nx = g.n_xi, nk = g.nk;
  // Centre phi rows for the observables.
  const int lc = nx / 2;
  std::vector<double2> ph(2 * nk);
  for (int j = 0; j < nk; ++j) {
    ph[j] = make_double2(g.phi()[(size_t)(lc - 1) * nk + j].real(), g.phi()[(size_t)(lc - 1) * nk + j].imag());
    ph[nk + j] = make_double2(g.phi()[(size_t)lc * nk + j].real(), g.phi()[(size_t)lc * nk + j].imag());
  }
  c->d_phic.alloc(2 * nk);
  HIPCHK(hipMemcpyAsync(c->d_phic.p, ph.data(), ph.size() * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  c->d_s0s1.alloc((size_t)g.n_chi * n * n * 2);
  c->d_agg.alloc((size_t)oc.n_chunks * nk * 2);
  c->d_linv16.alloc((size_t)g.n_chi * nk * n * n);
  c->d_lcpl32.alloc((size_t)g.n_chi * nk * n);
  c->d_sbar64.alloc((size_t)g.n_chi * oc.n_pairs * (nx + 1));
  c->d_sbar32.alloc((size_t)g.n_chi * oc.n_pairs * (nx + 1));
  c->d_b.alloc(c->N);
  c->d_x.alloc(c->N);
  c->d_r.alloc(c->N);
  c->d_y.alloc(c->N);
  c->d_obs.alloc((size_t)2 * c->ncols);
  c->d_st.alloc(1);
  HIPCHK(hipMemsetAsync(c->d_st.p, 0, sizeof(GmresState), c->stream));
  c->d_part.alloc((size_t)c->red_blocks * 2 * (kDotGroup + 1));
  HIPCHK(hipStreamSynchronize(c->stream));
}

void set_potential_impl(wdg_ctx* c, const wdg_potential& pot) {
  const Geometry& g = c->g;
  if (pot.n_mat != g.n_chi * (g.n - 1) + 1)
    throw std::invalid_argument("n_mat must equal n_chi*(n_k_chi-1)+1 (mat.Nx)");
  if (!pot.x_mat || !pot.em) throw std::invalid_argument("x_mat and em are required");
  for (int i = 1; i < pot.n_mat; ++i)
    if (!(pot.x_mat[i] > pot.x_mat[i - 1])) throw std::invalid_argument("x_mat must be strictly ascending");
  for (int i = 0; i < pot.n_mat; ++i)
    if (!std::isfinite(pot.em[i])) throw std::invalid_argument("em must be finite");
  if (!(pot.m_left > 0) || !(pot.m_right > 0) || !(pot.temp > 0))
    throw std::invalid_argument("masses and temperature must be positive");
  if (pot.fermi_model != WDG_FERMI_MSA_3D && pot.fermi_model != WDG_FERMI_MSA_2D &&
      pot.fermi_model != WDG_FERMI_TSU_ESAKI)
    throw std::invalid_argument("unknown fermi_model");
  if (pot.fermi_model == WDG_FERMI_MSA_2D && !(pot.me_z > 0)) throw std::invalid_argument("me_z must be positive");
  // inflow and rhs (host, FP64)
  std::vector<cplx> fnl, fnr;
  inflow(g, pot, fnl, fnr);
  const double old_norm = c->rhs_norm;
  // Only the boundary elements carry inflow: upload those two blocks of the
  // normalised RHS and clear the rest on the device.
  std::vector<double> first, last;
  build_rhs_blocks(g, fnl, fnr, first, last);
  double s = 0;
  for (double v : first) s += v * v;
  for (double v : last) s += v * v;
  c->rhs_norm = std::sqrt(s);
  const double inv = (c->rhs_norm > 0) ? 1.0 / c->rhs_norm : 0.0;
  for (double& v : first) v *= inv;
  for (double& v : last) v *= inv;
  c->stage.resize(2 * first.size());
  std::memcpy(c->stage.data(), first.data(), first.size() * sizeof(double));
  std::memcpy(c->stage.data() + first.size(), last.data(), last.size() * sizeof(double));
  const size_t blk = (size_t)g.n * g.nk;   // complex values per element
  HIPCHK(hipMemsetAsync(c->d_b.p, 0, (size_t)c->N * sizeof(double2), c->stream));
  HIPCHK(hipMemcpyAsync(c->d_b.p, c->stage.data(), blk * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_b.p + (size_t)(g.n_chi - 1) * blk, c->stage.data() + first.size(),
                        blk * sizeof(double2), hipMemcpyHostToDevice, c->stream));
  c->fnl = std::move(fnl);
  c->fnr = std::move(fnr);
  // potential tables on the device
  std::vector<double> xg(pot.x_mat, pot.x_mat + pot.n_mat), ug(pot.em, pot.em + pot.n_mat);
  c->d_xg.alloc(xg.size());
  c->d_ug.alloc(ug.size());
  HIPCHK(hipMemcpyAsync(c->d_xg.p, xg.data(), xg.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_ug.p, ug.data(), ug.size() * 8, hipMemcpyHostToDevice, c->stream));
  const double umin = *std::min_element(ug.begin(), ug.end());
  const long work = (long)g.n_chi * (g.n_xi + 1);
  hipLaunchKernelGGL(drift_table_kernel, dim3(grid1(work, 256)), dim3(256), 0, c->stream, c->d_oc.p, c->d_xquad.p,
                     c->d_edges.p, c->d_xg.p, c->d_ug.p, pot.n_mat, g.l_chi, umin, c->d_sbar64.p, c->d_sbar32.p);
  hipLaunchKernelGGL(precond_sums_kernel, dim3(g.n_chi * g.n * g.n), dim3(256), 0, c->stream, c->d_oc.p,
                     c->d_sbar64.p, c->d_cap.p, c->d_s0s1.p);
  // FP32 line-preconditioner block inverses (FP64 elimination, once per potential)
  {
    const dim3 gi(grid1((long)g.n_chi * g.nk, 256));
    switch (g.n) {
#define WDG_LINV(NN)                                                                                   \
  case NN:                                                                                             \
    hipLaunchKernelGGL((line_inverse_kernel<NN>), gi, dim3(256), 0, c->stream, c->d_oc.p, c->d_k.p,    \
                       c->d_s0s1.p, c->d_linv16.p, c->d_lcpl32.p);                                     \
    break;
      WDG_LINV(2) WDG_LINV(3) WDG_LINV(4) WDG_LINV(5) WDG_LINV(6) WDG_LINV(7) WDG_LINV(8)
#undef WDG_LINV
      default: throw std::invalid_argument("n_k_chi outside 2..8");
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  // A previous solution stays usable as a warm start: keep its physical value
  // (x_phys = ||b|| / s_A * x_device) under the new RHS normalisation.
  if (c->have_solution && old_norm > 0 && c->rhs_norm > 0 && old_norm != c->rhs_norm) {
    hipLaunchKernelGGL(scale_complex_kernel, dim3(c->red_blocks), dim3(256), 0, c->stream, c->d_x.p,
                       old_norm / c->rhs_norm, c->N);
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (c->rhs_norm > 0) c->sol_scale = c->rhs_norm / g.scale;
  c->have_potential = true;
}

void check_ready(wdg_ctx* c) {
  if (!c->have_potential) throw std::logic_error("wdg_set_potential has not been called");
}

// Drives one GMRES solve as a sequence of asynchronous restart cycles
// (solveStationarySystem gmres branch, DG_code/solve_transport_DG_full2D.m
// :276-334): normalised RHS, optional warm start (a correction solve with the
// absolute target tol*||b||), FP64 true residual after every cycle, flag
// overridden by the true residual.
struct SolveRun {
  wdg_ctx* c;
  wdg_solve_opts o;
  wdg_solve_info* info;
  bool fp64 = true, b16 = false, active = false, pending = false, precond = true;
  int basis = WDG_BASIS_WORKING;
  int m = 0, its = 0, cycles = 0, reorth = 0, max_outer = 0;
  double tol = 0, inner = 0, relres = 1, reported = 1, xs = 1, prev_relres = 1;
  int m_max = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  SolveRun(wdg_ctx* c_, const wdg_solve_opts& o_, wdg_solve_info* info_) : c(c_), o(o_), info(info_) {
    if (!(o.tol > 0) || o.restart < 1 || o.max_iter < 1)
      throw std::invalid_argument("tol must be positive, restart and max_iter at least 1");
    if (o.precision != WDG_PREC_FP64 && o.precision != WDG_PREC_MIXED)
      throw std::invalid_argument("unknown precision");
    std::memset(info, 0, sizeof(*info));
  }
  ~SolveRun() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  void begin(const double* x0) {
    const long N = c->N;
    fp64 = o.precision == WDG_PREC_FP64;
    m = (int)std::min<long>(o.restart, N);
    m_max = (int)std::min<long>(std::max(o.restart_max, o.restart), std::min<long>(N, kMaxRestart));
    if (o.krylov_basis != WDG_BASIS_WORKING && o.krylov_basis != WDG_BASIS_BF16 &&
        o.krylov_basis != WDG_BASIS_FP16)
      throw std::invalid_argument("unknown krylov_basis");
    basis = fp64 ? WDG_BASIS_WORKING : o.krylov_basis;
    b16 = basis != WDG_BASIS_WORKING;
    // the reference's estimate of the basis (complex double, restart + 1
    // vectors; solve_transport_DG_full2D.m:183-203)
    const double basis_bytes = 16.0 * (double)N * (double)(std::min<long>(m_max, N) + 1);
    if (c->gmres_max_basis_bytes > 0 && basis_bytes > c->gmres_max_basis_bytes) {
      char msg[256];
      std::snprintf(msg, sizeof msg,
                    "Estimated GMRES basis %.3g GiB exceeds full2D_gmresMaxBasisBytes. Reduce restart or explicitly "
                    "increase the memory budget.", basis_bytes / 1073741824.0);
      throw StatusError(WDG_ERR_GMRES_MEMORY, msg);
    }
    ensure_krylov(c, m, fp64, b16);
    // residual history, one value per Arnoldi step (captured cycles hold its
    // address: a larger buffer drops them)
    const size_t resv_cap = std::min<size_t>((size_t)o.max_iter + 1, (size_t)1 << 22);
    if (c->d_resv.n < resv_cap) {
      c->d_resv.alloc(resv_cap);
      clear_graphs(c);
    }
    c->last_its = -1;
    if (c->f2) f2_prepare_solve(c);
    else if (!c->use_fft) autotune(c, fp64);
    else if (!fp64) autotune_theta32(c);
    precond = o.use_precond != 0 && !(c->f2 && c->f2->precond == WDG2_PRECOND_NONE);
    if (!c->h_st) HIPCHK(hipHostMalloc((void**)&c->h_st, sizeof(GmresState), hipHostMallocDefault));
    xs = (c->rhs_norm > 0) ? c->rhs_norm / c->g.scale : 1.0;  // physical x = xs * x_device
    c->sol_scale = xs;
    if (x0) {
      c->stage.resize(2 * N);
      for (long i = 0; i < 2 * N; ++i) {
        if (!std::isfinite(x0[i])) throw StatusError(WDG_ERR_INITIAL_RHO, "rho0 must contain only finite values.");
        c->stage[i] = x0[i] / xs;
      }
      HIPCHK(hipMemcpy(c->d_x.p, c->stage.data(), 2 * N * 8, hipMemcpyHostToDevice));
    } else if (!(o.warm_device && c->have_solution)) {
      HIPCHK(hipMemsetAsync(c->d_x.p, 0, N * 16, c->stream));
    }
    tol = o.tol;
    inner = (o.inner_tol > 0) ? o.inner_tol : 1e-4;
    max_outer = (o.max_outer > 0) ? o.max_outer : 1000000;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, c->stream));
    hipLaunchKernelGGL(solve_init_kernel, dim3(1), dim3(1), 0, c->stream, c->d_st.p, tol, fp64 ? 0.0 : inner,
                       o.orthogonalization == 1 ? 1 : 0);
    residual64(c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_st, c->d_st.p, sizeof(GmresState), hipMemcpyDeviceToHost, c->stream));
    pending = true;
    active = true;
  }
  // Wait for the last enqueued cycle (or the initial residual) and decide.
  void sync_poll() {
    HIPCHK(hipStreamSynchronize(c->stream));
    pending = false;
    prof_collect(c);
    const GmresState& hs = *c->h_st;
    if (cycles == 0 && its == 0 && info->initial_true_relres == 0) {
      info->initial_true_relres = hs.resnorm;
      relres = reported = hs.resnorm;
    } else {
      its += hs.jlast;
      reorth += hs.reorth_count;
      reported = hs.resid;
      prev_relres = relres;
      relres = hs.resnorm;
      if (hs.jlast == 0) active = false;  // breakdown at the start of a cycle
      // Adaptive restart: a full cycle that cut the residual by less than 3x
      // gets a 1.5x longer Krylov space next time (up to restart_max).
      if (hs.jlast == m && m < m_max && relres > prev_relres / 3.0) {
        m = std::min(m_max, m + m / 2);
        ensure_krylov(c, m, fp64, b16);
      }
    }
    if (!(relres > tol) || its >= o.max_iter || cycles >= max_outer || !std::isfinite(relres)) active = false;
  }
  void step() {
    const int mc = std::min<int>(m, o.max_iter - its);
    run_cycle(c, fp64, basis, mc, precond, o.orthogonalization == 1, o.use_graph != 0 && mc == m);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c->h_st, c->d_st.p, sizeof(GmresState), hipMemcpyDeviceToHost, c->stream));
    cycles += 1;
    pending = true;
  }
  void end(double* x) {
    if (pending) sync_poll();
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    const long N = c->N;
    info->seconds = ms * 1e-3;
    info->iterations = its;
    info->restarts = cycles;
    info->reorth_steps = reorth;
    info->outer_iterations = fp64 ? 1 : cycles;
    info->true_relres = relres;
    info->reported_relres = reported;
    info->flag = (std::isfinite(relres) && relres <= tol) ? 0 : 1;
    c->have_solution = true;
    c->last_its = its;
    if (c->d_resv.n > 0)   // the history starts with the initial residual
      HIPCHK(hipMemcpy(c->d_resv.p, &info->initial_true_relres, sizeof(double), hipMemcpyHostToDevice));
    if (x) {
      HIPCHK(hipMemcpy(x, c->d_x.p, 2 * N * 8, hipMemcpyDeviceToHost));
      for (long i = 0; i < 2 * N; ++i) x[i] *= xs;
    }
  }
};

// Gauss-Legendre rule on [-1, 1] by Newton iteration on P_n.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 1; k < n; ++k) {
        const double p2 = ((2.0 * k + 1.0) * z * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = p2;
      }
      const double dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::abs(dz) < 1e-16) {
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
        break;
      }
      x[i] = z;
      w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
  }
}

void poisson_setup_impl(wdg_ctx* c, const wdg_poisson_desc& d) {
  const Geometry& g = c->g;
  const int nm = d.n_mat, n = g.n;
  if (nm != g.n_chi * (n - 1) + 1) throw std::invalid_argument("n_mat must equal n_chi*(n_k_chi-1)+1");
  if (!d.x_mat || !d.eps_r || !d.net_doping) throw std::invalid_argument("x_mat, eps_r and net_doping are required");
  if (d.method != 0 && d.method != 1) throw std::invalid_argument("method must be 0 (direct) or 1 (newton)");
  if (!(d.temp > 0)) throw std::invalid_argument("temperature must be positive");
  const double dx = (d.x_mat[nm - 1] - d.x_mat[0]) / (nm - 1);
  for (int i = 1; i < nm; ++i)
    if (std::abs(d.x_mat[i] - d.x_mat[i - 1] - dx) > 1e-8 * dx)
      throw std::invalid_argument("the Poisson grid must be uniform (RTD/eval_Poisson2D_rtd.m:202-209)");
  for (int i = 0; i < nm; ++i)
    if (!(d.eps_r[i] > 0)) throw std::invalid_argument("eps_r must be positive");
  // projection rows: exact integrals of each element's Lagrange basis over
  // the dual cells (8-point Gauss-Legendre, exact to degree 15)
  std::vector<double> edges(nm + 1), widths(nm);
  edges[0] = d.x_mat[0];
  edges[nm] = d.x_mat[nm - 1];
  for (int i = 1; i < nm; ++i) edges[i] = 0.5 * (d.x_mat[i - 1] + d.x_mat[i]);
  for (int i = 0; i < nm; ++i) widths[i] = edges[i + 1] - edges[i];
  std::vector<double> gx, gw;
  gauss_legendre(8, gx, gw);
  std::vector<int> ptr(nm + 1, 0), col;
  std::vector<double> val;
  const double h = g.l_chi / g.n_chi;
  for (int r = 0; r < nm; ++r) {
    for (int e = 0; e < g.n_chi; ++e) {
      const double left = e * h, right = (e + 1) * h;
      const double a = std::max(edges[r], left), b = std::min(edges[r + 1], right);
      if (!(b > a)) continue;
      for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int q = 0; q < 8; ++q) {
          const double xx = 0.5 * (a + b) + 0.5 * (b - a) * gx[q];
          const double rr = (xx - 0.5 * (left + right)) / (0.5 * h);
          double l = 1.0;
          for (int m = 0; m < n; ++m)
            if (m != j) l *= (rr - g.r[m]) / (g.r[j] - g.r[m]);
          s += gw[q] * l;
        }
        col.push_back(e * n + j);
        val.push_back(0.5 * (b - a) * s);
      }
    }
    ptr[r + 1] = (int)col.size();
  }
  // operator rows (harmonic-mean permittivity, identity contact rows)
  const double cx = 8.8541878128e-12 / (dx * dx);
  std::vector<double> lo(nm, 0.0), up(nm, 0.0), dg(nm, 1.0);
  for (int k = 1; k < nm - 1; ++k) {
    const double e0 = d.eps_r[k];
    lo[k] = cx * (2 * e0 * d.eps_r[k - 1] / (e0 + d.eps_r[k - 1]));
    up[k] = cx * (2 * e0 * d.eps_r[k + 1] / (e0 + d.eps_r[k + 1]));
    dg[k] = -(lo[k] + up[k]);
  }
  auto up_i = [&](DevBuf<int>& bb, const std::vector<int>& v) {
    bb.alloc(v.size());
    HIPCHK(hipMemcpy(bb.p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
  };
  auto up_d = [&](DevBuf<double>& bb, const double* v, size_t nn) {
    bb.alloc(nn);
    HIPCHK(hipMemcpy(bb.p, v, nn * sizeof(double), hipMemcpyHostToDevice));
  };
  up_i(c->d_tptr, ptr);
  up_i(c->d_tcol, col);
  up_d(c->d_tval, val.data(), val.size());
  up_d(c->d_pw, widths.data(), nm);
  up_d(c->d_plo, lo.data(), nm);
  up_d(c->d_pup, up.data(), nm);
  up_d(c->d_pdg, dg.data(), nm);
  up_d(c->d_pdop, d.net_doping, nm);
  c->d_pin.alloc(nm);
  c->d_pout.alloc(nm);
  c->d_pndg.alloc(c->ncols);
  c->d_pncv.alloc(nm);
  c->d_pwork.alloc(3 * (size_t)nm);
  c->d_pinfo.alloc(4);
  c->p_nmat = nm;
  c->p_method = d.method;
  c->p_temp = d.temp;
  c->have_poisson = true;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* wdg_last_error(const wdg_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return g_create_error.c_str();
}

int wdg_create(const wdg_desc* desc, wdg_ctx** out) {
  if (!desc || !out) return fail(nullptr, WDG_ERR_ARG, "null argument");
  *out = nullptr;
  auto c = std::make_unique<wdg_ctx>();
  int rc = guarded(nullptr, [&]() {
    create_impl(c.get(), *desc);
    // set-up of the allocation-free preconditioner entry point: this
    // device's address of the scratch array (resolved once per device).  A
    // failure here disables only that entry point (it then reports it).
    try {
      c->scratch_dev = scratch_for_current_device();
    } catch (const HipError& e) {
      c->scratch_dev = nullptr;
      c->err = e.what();
      (void)hipGetLastError();
    }
    return WDG_OK;
  });
  if (rc != WDG_OK) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
    return rc;
  }
  *out = c.release();
  return WDG_OK;
}

int wdg_destroy(wdg_ctx* ctx) {
  if (!ctx) return WDG_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
  clear_graphs(ctx);
  if (ctx->h_st) (void)hipHostFree(ctx->h_st);
  hipStream_t s = ctx->stream;
  delete ctx;
  if (s) (void)hipStreamDestroy(s);
  return WDG_OK;
}

int64_t wdg_size(const wdg_ctx* ctx) { return ctx ? ctx->N : 0; }

int wdg_set_potential(wdg_ctx* ctx, const wdg_potential* pot) {
  if (!ctx || !pot) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "full-2D context: use wdg2_set_potential");
  return guarded(ctx, [&]() {
    set_potential_impl(ctx, *pot);
    return WDG_OK;
  });
}

int wdg_rhs(wdg_ctx* ctx, double* rhs) {
  if (!ctx || !rhs) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  if (ctx->f2) {
    return guarded(ctx, [&]() {
      HIPCHK(hipMemcpy(rhs, ctx->d_b.p, (size_t)ctx->N * 16, hipMemcpyDeviceToHost));
      for (long i = 0; i < 2 * ctx->N; ++i) rhs[i] *= ctx->rhs_norm;
      return WDG_OK;
    });
  }
  std::vector<double> full;
  build_rhs(ctx->g, ctx->fnl, ctx->fnr, full);
  std::memcpy(rhs, full.data(), full.size() * sizeof(double));
  return WDG_OK;
}

int wdg_apply(wdg_ctx* ctx, const double* u, double* y, int32_t precision) {
  if (!ctx || !u || !y) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  if (precision != WDG_PREC_FP64 && precision != WDG_PREC_MIXED) return fail(ctx, WDG_ERR_ARG, "unknown precision");
  return guarded(ctx, [&]() {
    const long N = ctx->N;
    const double sA = ctx->g.scale;
    if (precision == WDG_PREC_FP64) {
      ctx->d_u64.alloc(N);
      ctx->d_v64.alloc(N);
      HIPCHK(hipMemcpyAsync(ctx->d_u64.p, u, N * 16, hipMemcpyHostToDevice, ctx->stream));
      op_apply<double>(ctx, ctx->d_u64.p, ctx->d_v64.p, nullptr);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(y, ctx->d_v64.p, N * 16, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      for (long i = 0; i < 2 * N; ++i) y[i] *= sA;
    } else {
      std::vector<float> u32(2 * N), y32(2 * N);
      for (long i = 0; i < 2 * N; ++i) u32[i] = (float)u[i];
      ctx->d_u32.alloc(N);
      ctx->d_v32.alloc(N);
      HIPCHK(hipMemcpyAsync(ctx->d_u32.p, u32.data(), N * 8, hipMemcpyHostToDevice, ctx->stream));
      op_apply<float>(ctx, ctx->d_u32.p, ctx->d_v32.p, nullptr);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(y32.data(), ctx->d_v32.p, N * 8, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      for (long i = 0; i < 2 * N; ++i) y[i] = sA * (double)y32[i];
    }
    prof_collect(ctx);
    return WDG_OK;
  });
}

int wdg_apply_parts(wdg_ctx* ctx, const double* u, double* y, int32_t precision, int32_t parts) {
  if (!ctx) return fail(nullptr, WDG_ERR_ARG, "null argument");
  if (parts != WDG_PART_TRANSPORT && parts != WDG_PART_POTENTIAL && parts != WDG_PART_ALL)
    return fail(ctx, WDG_ERR_ARG, "parts must be WDG_PART_TRANSPORT, WDG_PART_POTENTIAL or WDG_PART_ALL");
  if (ctx->f2 && parts != WDG_PART_ALL)
    return fail(ctx, WDG_ERR_UNSUPPORTED, "separate operator parts are not available for a full-2D context");
  if (ctx->shift_on) return fail(ctx, WDG_ERR_STATE, "a Crank-Nicolson shift is active");
  ctx->apply_parts = parts;
  const int rc = wdg_apply(ctx, u, y, precision);
  ctx->apply_parts = WDG_PART_ALL;
  return rc;
}

int wdg_apply_device(wdg_ctx* ctx, const void* d_u, void* d_y, int32_t precision) {
  if (!ctx || !d_u || !d_y) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  return guarded(ctx, [&]() {
    if (precision == WDG_PREC_FP64)
      op_apply<double>(ctx, (const double2*)d_u, (double2*)d_y, nullptr);
    else
      op_apply<float>(ctx, (const float2*)d_u, (float2*)d_y, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    return WDG_OK;
  });
}

int wdg_time_apply(wdg_ctx* ctx, int32_t precision, int32_t nrep, double* ms_per_apply) {
  if (!ctx || !ms_per_apply) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  if (precision != WDG_PREC_FP64 && precision != WDG_PREC_MIXED) return fail(ctx, WDG_ERR_ARG, "unknown precision");
  if (nrep < 1) return fail(ctx, WDG_ERR_ARG, "nrep must be at least 1");
  if (ctx->shift_on) return fail(ctx, WDG_ERR_STATE, "a Crank-Nicolson shift is active");
  return guarded(ctx, [&]() {
    const long N = ctx->N;
    const bool f64 = precision == WDG_PREC_FP64;
    if (f64) {
      ctx->d_u64.alloc(N);
      ctx->d_v64.alloc(N);
      HIPCHK(hipMemsetAsync(ctx->d_u64.p, 0, (size_t)N * 16, ctx->stream));
    } else {
      ctx->d_u32.alloc(N);
      ctx->d_v32.alloc(N);
      HIPCHK(hipMemsetAsync(ctx->d_u32.p, 0, (size_t)N * 8, ctx->stream));
    }
    const uint32_t mask = ctx->prof_mask;   // no per-kernel event pairs inside the timed group
    ctx->prof_mask = 0;
    auto apply = [&]() {
      if (f64) op_apply<double>(ctx, ctx->d_u64.p, ctx->d_v64.p, nullptr);
      else op_apply<float>(ctx, ctx->d_u32.p, ctx->d_v32.p, nullptr);
    };
    hipEvent_t t0 = nullptr, t1 = nullptr;
    float ms = 0;
    try {
      HIPCHK(hipEventCreate(&t0));
      HIPCHK(hipEventCreate(&t1));
      apply();   // warm-up
      HIPCHK(hipEventRecord(t0, ctx->stream));
      for (int r = 0; r < nrep; ++r) apply();
      HIPCHK(hipEventRecord(t1, ctx->stream));
      HIPCHK(hipEventSynchronize(t1));
      HIPCHK(hipEventElapsedTime(&ms, t0, t1));
    } catch (...) {
      ctx->prof_mask = mask;
      if (t0) (void)hipEventDestroy(t0);
      if (t1) (void)hipEventDestroy(t1);
      throw;
    }
    ctx->prof_mask = mask;
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    *ms_per_apply = (double)ms / nrep;
    return WDG_OK;
  });
}

int wdg_precond(wdg_ctx* ctx, const double* r, double* z) {
  if (!ctx || !r || !z) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  return guarded(ctx, [&]() {
    const long N = ctx->N;
    if (ctx->f2) f2_prepare_solve(ctx);
    if (ctx->f2 && ctx->f2->precond == WDG2_PRECOND_NONE) {   // identity in physical units
      std::memcpy(z, r, (size_t)N * 16);
      return WDG_OK;
    }
    ctx->d_u64.alloc(N);
    ctx->d_v64.alloc(N);
    HIPCHK(hipMemcpyAsync(ctx->d_u64.p, r, N * 16, hipMemcpyHostToDevice, ctx->stream));
    op_precond<double>(ctx, ctx->d_u64.p, ctx->d_v64.p, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(z, ctx->d_v64.p, N * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const double inv = 1.0 / ctx->g.scale;
    for (long i = 0; i < 2 * N; ++i) z[i] *= inv;
    prof_collect(ctx);
    return WDG_OK;
  });
}

int wdg_precond_device(wdg_ctx* ctx, const void* d_r, void* d_z, int32_t precision) {
  if (!ctx || !d_r || !d_z) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  return guarded(ctx, [&]() {
    if (ctx->f2) f2_prepare_solve(ctx);
    if (precision == WDG_PREC_FP64)
      op_precond<double>(ctx, (const double2*)d_r, (double2*)d_z, nullptr);
    else
      op_precond<float>(ctx, (const float2*)d_r, (float2*)d_z, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    return WDG_OK;
  });
}

int wdg_precond_device_noalloc(wdg_ctx* ctx, const void* d_r, void* d_z, int32_t precision) {
  if (!ctx || !d_r || !d_z) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not available for a full-2D context (1D line preconditioner only)");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  if (precision != WDG_PREC_FP64 && precision != WDG_PREC_MIXED) return fail(ctx, WDG_ERR_ARG, "unknown precision");
  const size_t need = (size_t)ctx->oc.n_chunks * ctx->g.nk * 2 * sizeof(double2);
  if (need > kNoAllocScratch) {
    char msg[200];
    std::snprintf(msg, sizeof msg, "the line sweeps of this grid need %zu bytes of workspace, above the %zu bytes of "
                  "the allocation-free entry point (use wdg_precond_device)", need, kNoAllocScratch);
    return fail(ctx, WDG_ERR_ARG, msg);
  }
  if (!ctx->scratch_dev)
    return fail(ctx, WDG_ERR_HIP, "hipGetSymbolAddress of the scratch array failed on this device: " + ctx->err);
  return guarded(ctx, [&]() {
    // no per-kernel event timing here (it would create events)
    const unsigned mask = ctx->prof_mask;
    ctx->prof_mask = 0;
    try {
      // every chunk map the sweeps read is written earlier in the same call,
      // so the array needs no clearing
      if (precision == WDG_PREC_FP64)
        op_precond<double>(ctx, (const double2*)d_r, (double2*)d_z, nullptr, ctx->scratch_dev);
      else
        op_precond<float>(ctx, (const float2*)d_r, (float2*)d_z, nullptr, ctx->scratch_dev);
    } catch (...) {
      ctx->prof_mask = mask;
      throw;
    }
    ctx->prof_mask = mask;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WDG_OK;
  });
}

int wdg_diagonal(wdg_ctx* ctx, double* d) {
  if (!ctx || !d) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  return guarded(ctx, [&]() {
    if (ctx->f2) {
      f2_diag_into(ctx, F2_DIAG, ctx->d_y.p);
      HIPCHK(hipMemcpyAsync(d, ctx->d_y.p, (size_t)ctx->N * 16, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      for (long i = 0; i < 2 * ctx->N; ++i) d[i] *= ctx->f2->sA;
      return WDG_OK;
    }
    const Geometry& g = ctx->g;
    const int n = g.n, nk = g.nk;
    std::vector<double2> s0s1((size_t)g.n_chi * n * n * 2);
    HIPCHK(hipMemcpyAsync(s0s1.data(), ctx->d_s0s1.p, s0s1.size() * sizeof(double2), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const double sA = g.scale, tq = ctx->oc.tq, il = ctx->oc.inv_len, hp = ctx->oc.hprime;
    for (int e = 0; e < g.n_chi; ++e)
      for (int a = 0; a < n; ++a)
        for (int j = 0; j < nk; ++j) {
          const double k = g.k[j];
          const double2 s0 = s0s1[((size_t)e * n * n + a * n + a) * 2];
          const double2 s1 = s0s1[((size_t)e * n * n + a * n + a) * 2 + 1];
          const double ck = std::cos(k * hp);
          const double gr = s0.x + 2 * ck * s1.x, gi = s0.y + 2 * ck * s1.y;
          const double re = tq * (k * g.K1[a * n + a] + std::abs(k) * g.K2[a * n + a]) - il * gi;
          const double im = il * gr;
          const size_t i = ((size_t)e * n + a) * nk + j;
          d[2 * i] = sA * re;
          d[2 * i + 1] = sA * im;
        }
    return WDG_OK;
  });
}

int wdg_solve(wdg_ctx* ctx, const wdg_solve_opts* opts, const double* x0, double* x, wdg_solve_info* info) {
  if (!ctx || !opts || !info) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  return guarded(ctx, [&]() {
    SolveRun r(ctx, *opts, info);
    r.begin(x0);
    r.sync_poll();
    while (r.active) {
      r.step();
      r.sync_poll();
    }
    r.end(x);
    return WDG_OK;
  });
}

int wdg_solution(wdg_ctx* ctx, double* x) {
  if (!ctx || !x) return fail(ctx, WDG_ERR_ARG, "null argument");
  return guarded(ctx, [&]() {
    if (!ctx->have_solution) throw std::logic_error("no solution available: call wdg_solve first");
    const long N = ctx->N;
    HIPCHK(hipMemcpy(x, ctx->d_x.p, (size_t)N * 16, hipMemcpyDeviceToHost));
    for (long i = 0; i < 2 * N; ++i) x[i] *= ctx->sol_scale;
    return WDG_OK;
  });
}

int wdg_solve_resvec(wdg_ctx* ctx, double* out, int32_t cap, int32_t* n_out) {
  if (!ctx || !n_out || cap < 0 || (cap > 0 && !out)) return fail(ctx, WDG_ERR_ARG, "null argument or negative cap");
  return guarded(ctx, [&]() {
    const long n = ctx->last_its < 0 ? 0 : std::min<long>((long)ctx->last_its + 1, (long)ctx->d_resv.n);
    *n_out = (int32_t)n;
    const long k = std::min<long>(n, cap);
    if (k > 0) {
      HIPCHK(hipMemcpyAsync(out, ctx->d_resv.p, (size_t)k * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    return WDG_OK;
  });
}

int wdg_solve_batch(wdg_ctx* const* ctxs, int32_t n, const wdg_solve_opts* opts, wdg_solve_info* infos) {
  if (!ctxs || n < 0 || !opts || !infos) return fail(nullptr, WDG_ERR_ARG, "null argument");
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return batch_fail(nullptr, i, WDG_ERR_ARG, "null context in batch");
    if (!ctxs[i]->have_potential) return batch_fail(ctxs[i], i, WDG_ERR_STATE, "wdg_set_potential has not been called");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return batch_fail(ctxs[i], i, WDG_ERR_ARG, "a context appears twice in the batch");
  }
  std::vector<std::unique_ptr<SolveRun>> runs;
  runs.reserve(n);
  int failed = -1;
  int rc = WDG_OK;
  // Every instance runs its own restart cycles on its own stream; the host
  // polls each instance once per cycle and relaunches it, so the GPU keeps
  // all unconverged instances in flight together.
  for (int i = 0; i < n && rc == WDG_OK; ++i) {
    rc = guarded(ctxs[i], [&]() {
      runs.emplace_back(new SolveRun(ctxs[i], *opts, &infos[i]));
      runs.back()->begin(nullptr);
      return WDG_OK;
    });
    if (rc != WDG_OK) failed = i;
  }
  bool any = true;
  while (rc == WDG_OK && any) {
    any = false;
    for (int i = 0; i < (int)runs.size() && rc == WDG_OK; ++i) {
      rc = guarded(ctxs[i], [&]() {
        SolveRun& r = *runs[i];
        if (r.pending) r.sync_poll();
        if (r.active) {
          r.step();
          any = true;
        }
        return WDG_OK;
      });
      if (rc != WDG_OK) failed = i;
    }
  }
  for (int i = 0; i < (int)runs.size() && rc == WDG_OK; ++i) {
    rc = guarded(ctxs[i], [&]() {
      runs[i]->end(nullptr);
      return WDG_OK;
    });
    if (rc != WDG_OK) failed = i;
  }
  if (rc != WDG_OK) {
    // other instances may still have restart cycles queued: let them drain
    // before their runs (events) are released
    for (size_t i = 0; i < runs.size(); ++i) (void)hipStreamSynchronize(ctxs[i]->stream);
    (void)hipGetLastError();
    if (failed >= 0) batch_fail(ctxs[failed], failed, rc, ctxs[failed]->err);
  }
  return rc;
}

int wdg_observables(wdg_ctx* ctx, const double* rho, const double* m_nodes, double* n_out, double* j_out,
                    double* rho_u) {
  if (!ctx || !m_nodes || !n_out || !j_out) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "full-2D context: use wdg2_observables");
  return guarded(ctx, [&]() {
    const Geometry& g = ctx->g;
    const long N = ctx->N;
    const double2* src;
    double scale;
    if (rho) {
      ctx->d_u64.alloc(N);
      HIPCHK(hipMemcpyAsync(ctx->d_u64.p, rho, N * 16, hipMemcpyHostToDevice, ctx->stream));
      src = ctx->d_u64.p;
      scale = 1.0;
    } else {
      if (!ctx->have_solution) throw std::logic_error("no solution available: call wdg_solve first");
      src = ctx->d_x.p;
      scale = ctx->sol_scale;
    }
    const int nodes = ctx->ncols;
    hipLaunchKernelGGL(observables_kernel, dim3(nodes), dim3(256), 0, ctx->stream, src, ctx->d_phic.p, g.nk,
                       ctx->d_obs.p);
    std::vector<double2> obs(2 * nodes);
    HIPCHK(hipMemcpyAsync(obs.data(), ctx->d_obs.p, obs.size() * sizeof(double2), hipMemcpyDeviceToHost,
                          ctx->stream));
    if (rho_u) {
      if (ctx->use_fft) {
        ctx->d_v64.alloc((size_t)nodes * g.n_xi);
        const int G = ctx->fft_groups;
        hipLaunchKernelGGL(phi_fft_kernel, dim3((nodes + G - 1) / G), dim3(ctx->fft_block), ctx->lds_phi,
                           ctx->stream, ctx->fplan, ctx->tt64, src, nodes, g.nk, g.n_xi, ctx->d_v64.p);
        HIPCHK(hipMemcpyAsync(rho_u, ctx->d_v64.p, (size_t)nodes * g.n_xi * 16, hipMemcpyDeviceToHost, ctx->stream));
      } else {
        op_gemm64(ctx, P_GEMM1, ctx->d_A1.p, (const double*)src, (double*)ctx->d_T64.p, 2 * g.n_xi, 2 * g.nk, false,
                  nullptr);
        HIPCHK(hipMemcpyAsync(rho_u, ctx->d_T64.p, (size_t)nodes * g.n_xi * 16, hipMemcpyDeviceToHost, ctx->stream));
      }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    // n = (Re rho_u(c-1) + Re rho_u(c))/2, j = q*hbar/m * Im(rho_u(c) - rho_u(c-1)) / delta_xi
    const double hb = kHVT / 2.0 / M_PI;
    for (int i = 0; i < nodes; ++i) {
      const double2 a = obs[2 * i], b = obs[2 * i + 1];
      n_out[i] = scale * 0.5 * (a.x + b.x);
      j_out[i] = scale * kQIP * hb / m_nodes[i] * (b.y - a.y) / g.delta_xi;
    }
    if (rho_u)
      for (long i = 0; i < (long)nodes * g.n_xi * 2; ++i) rho_u[i] *= scale;
    return WDG_OK;
  });
}

int wdg_poisson_setup(wdg_ctx* ctx, const wdg_poisson_desc* desc) {
  if (!ctx || !desc) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not available for a full-2D context");
  return guarded(ctx, [&]() {
    poisson_setup_impl(ctx, *desc);
    return WDG_OK;
  });
}

int wdg_poisson_step(wdg_ctx* ctx, const double* phi_in, double bias, double* phi_out, double* n_out,
                     wdg_poisson_info* info) {
  if (!ctx || !phi_in || !phi_out) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not available for a full-2D context");
  if (!ctx->have_poisson) return fail(ctx, WDG_ERR_STATE, "wdg_poisson_setup has not been called");
  if (!ctx->have_solution) return fail(ctx, WDG_ERR_STATE, "no solution available: call wdg_solve first");
  return guarded(ctx, [&]() {
    const int nm = ctx->p_nmat;
    for (int i = 0; i < nm; ++i)
      if (!std::isfinite(phi_in[i])) throw std::invalid_argument("phi_in must be finite");
    HIPCHK(hipMemcpyAsync(ctx->d_pin.p, phi_in, nm * 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(observables_kernel, dim3(ctx->ncols), dim3(256), 0, ctx->stream, ctx->d_x.p, ctx->d_phic.p,
                       ctx->g.nk, ctx->d_obs.p);
    PoissonDev P;
    P.nmat = nm;
    P.nodes = ctx->ncols;
    P.method = ctx->p_method;
    P.temp = ctx->p_temp;
    P.q = 1.602176634e-19;   // RTD/solve_Poisson2D_rtd.m:24
    P.kb = 1.380649e-23;     // RTD/solve_Poisson2D_rtd.m:25
    P.scale_n = ctx->sol_scale;
    P.obs = ctx->d_obs.p;
    P.t_ptr = ctx->d_tptr.p;
    P.t_col = ctx->d_tcol.p;
    P.t_val = ctx->d_tval.p;
    P.widths = ctx->d_pw.p;
    P.lo = ctx->d_plo.p;
    P.up = ctx->d_pup.p;
    P.dg = ctx->d_pdg.p;
    P.doping = ctx->d_pdop.p;
    P.phi_in = ctx->d_pin.p;
    P.bias = bias;
    P.n_dg = ctx->d_pndg.p;
    P.n_cv = ctx->d_pncv.p;
    P.work = ctx->d_pwork.p;
    P.phi_out = ctx->d_pout.p;
    P.info = ctx->d_pinfo.p;
    hipLaunchKernelGGL(poisson_newton_kernel, dim3(1), dim3(256), 0, ctx->stream, P);
    HIPCHK(hipGetLastError());
    double inf[4];
    HIPCHK(hipMemcpyAsync(phi_out, ctx->d_pout.p, nm * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (n_out) HIPCHK(hipMemcpyAsync(n_out, ctx->d_pncv.p, nm * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(inf, ctx->d_pinfo.p, 4 * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (info) {
      info->physical_relres_before = inf[0];
      info->unmixed_step = inf[1];
      info->charge_integral = inf[2];
      info->dg_integral = inf[3];
    }
    return WDG_OK;
  });
}

int wdg_host_setup(const wdg_desc* desc, const wdg_potential* pot, double* rhs, double* kmats) {
  if (!desc) return fail(nullptr, WDG_ERR_ARG, "null argument");
  return guarded(nullptr, [&]() {
    Geometry g;
    build_geometry(*desc, g);
    const int n = g.n;
    if (kmats) {
      const std::vector<double>* mats[5] = {&g.K1, &g.K2L, &g.K2R, &g.K3, &g.K5};
      for (int i = 0; i < 5; ++i) std::memcpy(kmats + (size_t)i * n * n, mats[i]->data(), n * n * sizeof(double));
    }
    if (rhs) {
      if (!pot) throw std::invalid_argument("rhs requires a potential");
      if (pot->n_mat != g.n_chi * (g.n - 1) + 1) throw std::invalid_argument("n_mat must equal mat.Nx");
      std::vector<cplx> fnl, fnr;
      inflow(g, *pot, fnl, fnr);
      std::vector<double> r;
      build_rhs(g, fnl, fnr, r);
      std::memcpy(rhs, r.data(), r.size() * sizeof(double));
    }
    return WDG_OK;
  });
}

int wdg_transient_rk4(wdg_ctx* ctx, const wdg_transient_opts* o, const double* rho0, double* rho_out,
                      const double* m_nodes, double* n_t, double* j_t, int32_t max_saves) {
  if (!ctx || !o || !m_nodes) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not available for a full-2D context");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg_set_potential has not been called");
  return guarded(ctx, [&]() {
    if (!(o->dt > 0) || o->nsteps < 0 || o->save_every < 1 || !std::isfinite(o->a_scale) ||
        !std::isfinite(o->b_scale))
      throw std::invalid_argument("dt must be positive, nsteps >= 0, save_every >= 1, finite scales");
    const int nsave = o->nsteps / o->save_every;
    if (nsave > 0 && (!n_t || !j_t)) throw std::invalid_argument("n_t and j_t are required when saving");
    if (nsave > max_saves) throw std::invalid_argument("nsteps/save_every exceeds max_saves");
    const Geometry& g = ctx->g;
    const long N = ctx->N;
    const int nodes = ctx->ncols;
    // device units: rho = sigma x with sigma = ||b|| / s_A
    const double sigma = (ctx->rhs_norm > 0) ? ctx->rhs_norm / g.scale : 1.0;
    if (rho0) {
      ctx->stage.resize(2 * N);
      for (long i = 0; i < 2 * N; ++i) {
        if (!std::isfinite(rho0[i])) throw StatusError(WDG_ERR_INITIAL_RHO, "rho0 must contain only finite values.");
        ctx->stage[i] = rho0[i] / sigma;
      }
      HIPCHK(hipMemcpy(ctx->d_x.p, ctx->stage.data(), 2 * N * 8, hipMemcpyHostToDevice));
    } else if (!ctx->have_solution) {
      throw std::logic_error("no initial state: pass rho0 or solve first");
    } else if (ctx->sol_scale != sigma) {
      hipLaunchKernelGGL(scale_complex_kernel, dim3(ctx->red_blocks), dim3(256), 0, ctx->stream, ctx->d_x.p,
                         ctx->sol_scale / sigma, N);
    }
    ctx->sol_scale = sigma;
    ctx->d_u64.alloc(N);   // tmp
    ctx->d_v64.alloc(N);   // acc
    DevBuf<double2> obs_t;
    if (nsave > 0) obs_t.alloc((size_t)nsave * 2 * nodes);
    const double h = o->dt * g.scale;
    const double cb = o->b_scale, ca = o->a_scale;
    const dim3 gr(ctx->red_blocks), bl(256);
    int saved = 0;
    for (int step = 1; step <= o->nsteps; ++step) {
      for (int stage = 1; stage <= 4; ++stage) {
        op_apply<double>(ctx, stage == 1 ? ctx->d_x.p : ctx->d_u64.p, ctx->d_y.p, nullptr);
        hipLaunchKernelGGL(rk4_stage_kernel, gr, bl, 0, ctx->stream, stage, ctx->d_b.p, ctx->d_y.p, ctx->d_x.p,
                           ctx->d_u64.p, ctx->d_v64.p, cb, ca, h, N);
      }
      if (step % o->save_every == 0 && saved < nsave) {
        hipLaunchKernelGGL(observables_kernel, dim3(nodes), dim3(256), 0, ctx->stream, ctx->d_x.p, ctx->d_phic.p,
                           g.nk, obs_t.p + (size_t)saved * 2 * nodes);
        ++saved;
      }
    }
    HIPCHK(hipGetLastError());
    std::vector<double2> obs((size_t)nsave * 2 * nodes);
    if (nsave > 0)
      HIPCHK(hipMemcpyAsync(obs.data(), obs_t.p, obs.size() * sizeof(double2), hipMemcpyDeviceToHost, ctx->stream));
    if (rho_out) HIPCHK(hipMemcpyAsync(rho_out, ctx->d_x.p, N * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_solution = true;
    if (rho_out)
      for (long i = 0; i < 2 * N; ++i) rho_out[i] *= sigma;
    const double hb = kHVT / 2.0 / M_PI;
    for (int sidx = 0; sidx < nsave; ++sidx)
      for (int i = 0; i < nodes; ++i) {
        const double2 a = obs[((size_t)sidx * 2 * nodes) + 2 * i], b = obs[((size_t)sidx * 2 * nodes) + 2 * i + 1];
        n_t[(size_t)sidx * nodes + i] = sigma * 0.5 * (a.x + b.x);
        j_t[(size_t)sidx * nodes + i] = sigma * kQIP * hb / m_nodes[i] * (b.y - a.y) / g.delta_xi;
      }
    return WDG_OK;
  });
}

namespace {

// One instance of a Crank-Nicolson run: the shifted operator, the per-step
// right-hand side and the restore of the unshifted state.  Steps of several
// instances interleave in wdg_transient_cn_batch.
struct CnRun {
  wdg_ctx* ctx;
  const wdg_cn_opts* o;
  wdg_cn_info* info;
  wdg_solve_opts so{};
  double sigma = 1, beta = 0, gamma = 0, cn = 1, max_rel = 0;
  int flag = 0;
  std::vector<double> part;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::unique_ptr<SolveRun> run;
  wdg_solve_info si{};

  CnRun(wdg_ctx* c, const wdg_cn_opts* opts, wdg_cn_info* inf) : ctx(c), o(opts), info(inf) {}
  ~CnRun() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  // Initial state, shift and preconditioner (device units: rho = sigma x,
  // sigma = ||b|| / s_A; b_hat = b / ||b||).
  void begin(const double* rho0) {
    std::memset(info, 0, sizeof(*info));
    if (!(o->dt > 0) || o->nsteps < 0 || !std::isfinite(o->a_scale) || !std::isfinite(o->b_scale) || !(o->tol > 0))
      throw std::invalid_argument("dt and tol must be positive, nsteps >= 0, finite scales");
    const long N = ctx->N;
    const double sA = ctx->g.scale;
    sigma = (ctx->rhs_norm > 0) ? ctx->rhs_norm / sA : 1.0 / sA;
    if (rho0) {
      ctx->stage.resize(2 * N);
      for (long i = 0; i < 2 * N; ++i) {
        if (!std::isfinite(rho0[i])) throw StatusError(WDG_ERR_INITIAL_RHO, "rho0 must contain only finite values.");
        ctx->stage[i] = rho0[i] / sigma;
      }
      HIPCHK(hipMemcpy(ctx->d_x.p, ctx->stage.data(), 2 * N * 8, hipMemcpyHostToDevice));
    } else if (!ctx->have_solution) {
      throw std::logic_error("no initial state: pass rho0 or solve first");
    } else if (ctx->sol_scale != sigma) {
      hipLaunchKernelGGL(scale_complex_kernel, dim3(ctx->red_blocks), dim3(256), 0, ctx->stream, ctx->d_x.p,
                         ctx->sol_scale / sigma, N);
    }
    ctx->sol_scale = sigma;
    ctx->have_solution = true;
    ctx->d_b0.alloc(N);
    HIPCHK(hipMemcpyAsync(ctx->d_b0.p, ctx->d_b.p, N * 16, hipMemcpyDeviceToDevice, ctx->stream));
    // (I + beta A_hat) x+ = (I - beta A_hat) x + gamma b_hat: the trapezoidal
    // rule for d rho/dt = b_scale rhs - a_scale A rho in device units
    beta = 0.5 * o->dt * o->a_scale * sA;
    gamma = (ctx->rhs_norm > 0) ? o->dt * o->b_scale * sA : 0.0;
    ctx->shift_on = true;
    ctx->shift_a = 1.0;
    ctx->shift_b = beta;
    clear_graphs(ctx);
    so.tol = o->tol;
    so.restart = o->restart > 0 ? o->restart : 30;
    so.max_iter = o->max_iter > 0 ? o->max_iter : 20000;
    so.precision = o->precision;
    so.inner_tol = 1e-4;
    // Graph replay of FP64 restart cycles on full-2D Crank-Nicolson contexts
    // stays off: at 33.5M unknowns such a run ended in a host-side fault
    // inside the runtime whose cause is not identified; the cycles are
    // launched kernel by kernel instead (same kernels, same results).
    so.use_graph = (o->use_graph && !(ctx->f2 && o->precision == WDG_PREC_FP64)) ? 1 : 0;
    so.orthogonalization = 1;
    so.warm_device = 1;
    so.krylov_basis = o->krylov_basis;
    // the 1D line preconditioner inverts A_hat's sweep, not the shifted operator's
    so.use_precond = (ctx->f2 && o->use_precond) ? 1 : 0;
    if (ctx->f2 && (ctx->f2->precond == WDG2_PRECOND_JACOBI || ctx->f2->precond == WDG2_PRECOND_ROWABS)) {
      int used = 0;
      f2_build_precond(ctx, ctx->f2->precond, 1e-12, true, &used);
    } else if (ctx->f2 && ctx->f2->precond == WDG2_PRECOND_LINE) {
      so.use_precond = 0;   // the line-transverse factors approximate A, not I + beta A
    }
    part.assign(ctx->red_blocks, 0.0);
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, ctx->stream));
  }
  // Step right-hand side c = 2x - (I + beta A_hat) x + gamma b_hat and its
  // norm partials (asynchronous on the instance's stream).
  void rhs_launch() {
    const long N = ctx->N;
    const int nb = ctx->red_blocks;
    op_apply<double>(ctx, ctx->d_x.p, ctx->d_y.p, nullptr);
    hipLaunchKernelGGL((cn_rhs_kernel<256>), dim3(nb), dim3(256), 0, ctx->stream, ctx->d_x.p, ctx->d_y.p,
                       ctx->d_b0.p, gamma, ctx->d_b.p, N, (double*)ctx->d_part.p);
    HIPCHK(hipMemcpyAsync(part.data(), ctx->d_part.p, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  // After rhs_launch: normalise the step system and start its solve.
  void rhs_finish() {
    const long N = ctx->N;
    const int nb = ctx->red_blocks;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double s2 = 0;
    for (double v : part) s2 += v;
    cn = std::sqrt(s2);
    if (!(cn > 0) || !std::isfinite(cn)) throw std::runtime_error("Crank-Nicolson right-hand side vanished or is not finite");
    hipLaunchKernelGGL(scale_complex_kernel, dim3(nb), dim3(256), 0, ctx->stream, ctx->d_b.p, 1.0 / cn, N);
    hipLaunchKernelGGL(scale_complex_kernel, dim3(nb), dim3(256), 0, ctx->stream, ctx->d_x.p, 1.0 / cn, N);
    run.reset(new SolveRun(ctx, so, &si));
    run->begin(nullptr);
  }
  void step_end(int step) {
    run->end(nullptr);
    run.reset();
    hipLaunchKernelGGL(scale_complex_kernel, dim3(ctx->red_blocks), dim3(256), 0, ctx->stream, ctx->d_x.p, cn,
                       ctx->N);
    info->iterations += si.iterations;
    info->max_step_iterations = std::max(info->max_step_iterations, si.iterations);
    max_rel = std::max(max_rel, si.true_relres);
    if (si.flag != 0) flag = 1;
    info->steps = step + 1;
  }
  void finish(double* rho_out) {
    const long N = ctx->N;
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    info->seconds = ms * 1e-3;
    info->max_relres = max_rel;
    info->flag = flag;
    // restore the inflow RHS and the solution scale
    HIPCHK(hipMemcpyAsync(ctx->d_b.p, ctx->d_b0.p, N * 16, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->sol_scale = sigma;
    if (rho_out) {
      HIPCHK(hipMemcpyAsync(rho_out, ctx->d_x.p, N * 16, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      for (long i = 0; i < 2 * N; ++i) rho_out[i] *= sigma;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  // The unshifted operator again (also after a failure).
  void restore() {
    if (ctx->shift_on) {
      ctx->shift_on = false;
      clear_graphs(ctx);
      if (ctx->f2 && (ctx->f2->precond == WDG2_PRECOND_JACOBI || ctx->f2->precond == WDG2_PRECOND_ROWABS))
        ctx->f2->precond = -ctx->f2->precond;   // rebuilt for the unshifted operator on the next solve
    }
  }
};

}  // namespace

int wdg_transient_cn_batch(wdg_ctx* const* ctxs, int32_t n, const wdg_cn_opts* o, const double* const* rho0,
                           double* const* rho_out, wdg_cn_info* infos) {
  if (!ctxs || n < 0 || !o || !infos) return fail(nullptr, WDG_ERR_ARG, "null argument");
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) return batch_fail(nullptr, i, WDG_ERR_ARG, "null context in batch");
    if (!ctxs[i]->have_potential) return batch_fail(ctxs[i], i, WDG_ERR_STATE, "the operator has no potential yet");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return batch_fail(ctxs[i], i, WDG_ERR_ARG, "a context appears twice in the batch");
  }
  std::vector<std::unique_ptr<CnRun>> runs;
  for (int i = 0; i < n; ++i) runs.emplace_back(new CnRun(ctxs[i], o, &infos[i]));
  int rc = WDG_OK, failed = -1;
  // run f on every instance in order; stop at the first failure
  auto each = [&](auto&& f) {
    for (int i = 0; i < n && rc == WDG_OK; ++i) {
      rc = guarded(ctxs[i], [&]() {
        f(*runs[i], i);
        return WDG_OK;
      });
      if (rc != WDG_OK) failed = i;
    }
  };
  each([&](CnRun& r, int i) { r.begin(rho0 ? rho0[i] : nullptr); });
  // per step: every instance's right-hand side, then all step solves with
  // their restart cycles interleaved on the instances' streams
  for (int step = 0; step < o->nsteps && rc == WDG_OK; ++step) {
    each([&](CnRun& r, int) { r.rhs_launch(); });
    each([&](CnRun& r, int) { r.rhs_finish(); });
    bool any = true;
    while (rc == WDG_OK && any) {
      any = false;
      each([&](CnRun& cr, int) {
        SolveRun& r = *cr.run;
        if (r.pending) r.sync_poll();
        if (r.active) {
          r.step();
          any = true;
        }
      });
    }
    each([&](CnRun& r, int) { r.step_end(step); });
  }
  each([&](CnRun& r, int i) { r.finish(rho_out ? rho_out[i] : nullptr); });
  if (rc != WDG_OK) {
    // instances after a failure may still have work queued: drain every
    // stream before the captured graphs and events are released
    for (int i = 0; i < n; ++i) (void)hipStreamSynchronize(ctxs[i]->stream);
    (void)hipGetLastError();
  }
  for (auto& r : runs) {
    try {
      r->restore();
    } catch (...) {
    }
  }
  if (rc != WDG_OK && failed >= 0) batch_fail(ctxs[failed], failed, rc, ctxs[failed]->err);
  return rc;
}

int wdg_transient_cn(wdg_ctx* ctx, const wdg_cn_opts* o, const double* rho0, double* rho_out, wdg_cn_info* info) {
  if (!ctx || !o || !info) return fail(ctx, WDG_ERR_ARG, "null argument");
  return wdg_transient_cn_batch(&ctx, 1, o, &rho0, &rho_out, info);
}

int wdg_set_theta_method(wdg_ctx* ctx, int32_t method) {
  if (!ctx) return fail(nullptr, WDG_ERR_ARG, "null argument");
  if (ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not available for a full-2D context");
  return guarded(ctx, [&]() {
    if (method != WDG_THETA_AUTO && method != WDG_THETA_FFT && method != WDG_THETA_GEMM)
      throw std::invalid_argument("unknown Theta method");
    if (method == WDG_THETA_FFT && !ctx->fft_ok)
      throw std::invalid_argument("the chirp-z FFT formulation does not apply to this grid (or exceeds LDS)");
    const bool fft = (method == WDG_THETA_FFT) || (method == WDG_THETA_AUTO && ctx->fft_ok);
    if (!fft) ensure_gemm(ctx);
    if (fft != ctx->use_fft) clear_graphs(ctx);   // captured cycles hold the other kernels
    ctx->use_fft = fft;
    return WDG_OK;
  });
}

int wdg_set_memory_limits(wdg_ctx* ctx, double gmres_max_basis_bytes, double line_max_factor_bytes) {
  if (!ctx) return fail(nullptr, WDG_ERR_ARG, "null argument");
  if (std::isnan(gmres_max_basis_bytes) || gmres_max_basis_bytes < 0)
    return fail(ctx, WDG_ERR_ARG, "gmres_max_basis_bytes must be >= 0 (0: no limit)");
  if (!(line_max_factor_bytes > 0) || !std::isfinite(line_max_factor_bytes))
    return fail(ctx, WDG_ERR_LINE_MEMORY, "lineMaxFactorBytes must be positive and finite.");
  ctx->gmres_max_basis_bytes = std::isinf(gmres_max_basis_bytes) ? 0.0 : gmres_max_basis_bytes;
  ctx->line_max_factor_bytes = line_max_factor_bytes;
  return WDG_OK;
}

int wdg_set_fft_block(wdg_ctx* ctx, int32_t mode) {
  if (!ctx) return fail(nullptr, WDG_ERR_ARG, "null argument");
  if (ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not available for a full-2D context");
  return guarded(ctx, [&]() {
    if (mode != WDG_FFT_BLOCK_AUTO && mode != WDG_FFT_BLOCK_256 && mode != WDG_FFT_BLOCK_WIDE &&
        mode != WDG_FFT_BLOCK_NDFT)
      throw std::invalid_argument("unknown FFT block mode");
    if (!ctx->fft_ok) throw std::invalid_argument("the chirp-z FFT formulation does not apply to this grid");
    if (mode == WDG_FFT_BLOCK_WIDE && ctx->theta32_wide == 0)
      throw std::invalid_argument("the wide FP32 Theta block does not fit this grid (n_k_chi > 5, fewer than "
                                  "four transforms per round, or LDS)");
    if (mode == WDG_FFT_BLOCK_NDFT && !ctx->ndft_ok)
      throw std::invalid_argument("the near-DFT FP32 Theta kernel does not apply to this grid (n_xi != nk, nk not "
                                  "64 * 2^a 3^b, n_k_chi > 5, or LDS)");
    const bool ndft = mode == WDG_FFT_BLOCK_NDFT;
    const bool wide = mode == WDG_FFT_BLOCK_WIDE || (mode == WDG_FFT_BLOCK_AUTO && ctx->theta32_wide_auto);
    const int blk = wide ? ctx->theta32_wide : ctx->theta32_base;
    // captured cycles hold the kernel and its launch shape
    if (blk != ctx->theta_block32 || ndft != ctx->use_ndft) clear_graphs(ctx);
    ctx->use_ndft = ndft;
    ctx->theta32_auto = mode == WDG_FFT_BLOCK_AUTO;
    ctx->theta32_tuned = false;
    ctx->theta_block32 = blk;
    ctx->lds32 = wide ? ctx->lds32_wide : ctx->lds32_base;
    return WDG_OK;
  });
}

int wdg_fft_block(const wdg_ctx* ctx) {
  if (!ctx || !ctx->fft_ok) return -1;
  return ctx->use_ndft ? ctx->ndft_block : ctx->theta_block32;
}

int wdg_theta_method(const wdg_ctx* ctx) {
  if (!ctx) return -1;
  return ctx->use_fft ? WDG_THETA_FFT : WDG_THETA_GEMM;
}

int wdg_set_profiling(wdg_ctx* ctx, uint32_t mask) {
  if (!ctx) return fail(nullptr, WDG_ERR_ARG, "null argument");
  ctx->prof_mask = mask;
  for (int i = 0; i < P_COUNT; ++i) {
    ctx->prof_ms[i] = 0;
    ctx->prof_n[i] = 0;
    ctx->prof_units[i] = 0;
  }
  return WDG_OK;
}

int wdg_profile(wdg_ctx* ctx, char* names, int32_t names_len, double* avg_ms, int64_t* launches, double* units,
                int32_t max_entries, int32_t* n_entries) {
  if (!ctx) return fail(nullptr, WDG_ERR_ARG, "null argument");
  std::string s;
  int k = 0;
  for (int i = 0; i < P_COUNT && k < max_entries; ++i, ++k) {
    if (i) s += ",";
    s += kProfNames[i];
    if (avg_ms) avg_ms[k] = ctx->prof_n[i] ? ctx->prof_ms[i] / ctx->prof_n[i] : 0.0;
    if (launches) launches[k] = ctx->prof_n[i];
    if (units) units[k] = ctx->prof_units[i];
  }
  if (n_entries) *n_entries = k;
  if (names && names_len > 0) {
    std::strncpy(names, s.c_str(), names_len - 1);
    names[names_len - 1] = 0;
  }
  return WDG_OK;
}

// ---- full-2D operator (include/wigner_dg.h) ----

int wdg2_create(const wdg2_desc* desc, const wdg2_part* parts, int32_t n_parts, wdg_ctx** out) {
  if (!desc || !out) return fail(nullptr, WDG_ERR_ARG, "null argument");
  *out = nullptr;
  auto c = std::make_unique<wdg_ctx>();
  int rc = guarded(nullptr, [&]() {
    f2_create_impl(c.get(), *desc, parts, n_parts);
    return WDG_OK;
  });
  if (rc != WDG_OK) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->f2.reset();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
    return rc;
  }
  *out = c.release();
  return WDG_OK;
}

int wdg2_set_potential(wdg_ctx* ctx, const wdg2_potential* pot) {
  if (!ctx || !pot) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not a full-2D context");
  return guarded(ctx, [&]() {
    f2_set_potential_impl(ctx, *pot);
    return WDG_OK;
  });
}

int wdg2_set_rhs(wdg_ctx* ctx, const double* rhs) {
  if (!ctx || !rhs) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not a full-2D context");
  return guarded(ctx, [&]() {
    const long N = ctx->N;
    double s = 0;
    for (long i = 0; i < 2 * N; ++i) {
      if (!std::isfinite(rhs[i])) throw std::invalid_argument("rhs must be finite");
      s += rhs[i] * rhs[i];
    }
    const double old_norm = ctx->rhs_norm;
    ctx->rhs_norm = std::sqrt(s);
    const double inv = ctx->rhs_norm > 0 ? 1.0 / ctx->rhs_norm : 0.0;
    ctx->stage.resize(2 * N);
    for (long i = 0; i < 2 * N; ++i) ctx->stage[i] = rhs[i] * inv;
    HIPCHK(hipMemcpyAsync(ctx->d_b.p, ctx->stage.data(), (size_t)N * 16, hipMemcpyHostToDevice, ctx->stream));
    // keep a previous device solution's physical value under the new normalisation
    if (ctx->have_solution && old_norm > 0 && ctx->rhs_norm > 0 && old_norm != ctx->rhs_norm)
      hipLaunchKernelGGL(scale_complex_kernel, dim3(ctx->red_blocks), dim3(256), 0, ctx->stream, ctx->d_x.p,
                         old_norm / ctx->rhs_norm, ctx->N);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->rhs_norm > 0) ctx->sol_scale = ctx->rhs_norm / ctx->f2->sA;
    ctx->f2->have_rhs = true;
    return WDG_OK;
  });
}

}  // extern "C"

namespace {

// Device copies of one contact's mode data (wdg2_set_contact_rhs).
struct ResUpload {
  DevBuf<double> grid, vals, energy, mx, mz, lift;
  DevBuf<double2> coef, achar, res;
  DevBuf<int> mask;
};

template <typename T>
void upload_vec(wdg_ctx* c, DevBuf<T>& b, const T* v, size_t n) {
  b.alloc(n);
  if (n) HIPCHK(hipMemcpyAsync(b.p, v, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
}

ResContact res_contact(wdg_ctx* c, const wdg2_contact_modes& m, ResUpload& u, int face, int nyd, int nrx, int nl) {
  if (m.n_modes < 0 || m.n_modes > kResMaxModes)
    throw std::invalid_argument("n_modes must be in 0.." + std::to_string(kResMaxModes));
  if (m.kind != WDG2_MODES_BOX && m.kind != WDG2_MODES_FOURIER) throw std::invalid_argument("unknown mode kind");
  if (!m.mask || !m.char_matrix || !m.lift) throw std::invalid_argument("contact mask, characteristic matrix and lift are required");
  ResContact rc{};
  rc.n_modes = m.n_modes;
  rc.kind = m.kind;
  rc.n_grid = m.n_grid;
  rc.face = face;
  rc.sign = face ? -1.0 : 1.0;
  rc.ef = m.fermi_level;
  std::vector<int> mask(m.mask, m.mask + nyd);
  upload_vec(c, u.mask, mask.data(), mask.size());
  upload_vec(c, u.achar, (const double2*)m.char_matrix, (size_t)nrx * nrx);
  upload_vec(c, u.lift, m.lift, (size_t)nl);
  rc.mask = u.mask.p;
  rc.achar = u.achar.p;
  if (m.n_modes > 0) {
    if (!m.grid || !m.values || !m.energy || !m.mass_x || !m.mass_z || m.n_grid < 2)
      throw std::invalid_argument("mode grid, values, energies and masses are required");
    if (!std::isfinite(m.fermi_level)) throw std::invalid_argument("the Fermi level must be finite");
    for (int i = 0; i < m.n_modes; ++i)
      if (!(m.mass_x[i] > 0) || !(m.mass_z[i] > 0) || !std::isfinite(m.energy[i]))
        throw std::invalid_argument("mode masses must be positive and energies finite");
    upload_vec(c, u.grid, m.grid, (size_t)m.n_grid);
    if (m.kind == WDG2_MODES_BOX) {
      for (int i = 1; i < m.n_grid; ++i)
        if (!(m.grid[i] > m.grid[i - 1])) throw std::invalid_argument("the mode grid must be strictly ascending");
      upload_vec(c, u.vals, m.values, (size_t)m.n_modes * m.n_grid);
      rc.vals = u.vals.p;
    } else {
      if (!(m.width > 0)) throw std::invalid_argument("the periodic cell width must be positive");
      upload_vec(c, u.coef, (const double2*)m.values, (size_t)m.n_modes * m.n_grid);
      rc.coef = u.coef.p;
      rc.y0 = m.y0;
      rc.inv_sqrt_w = 1.0 / std::sqrt(m.width);
    }
    upload_vec(c, u.energy, m.energy, (size_t)m.n_modes);
    upload_vec(c, u.mx, m.mass_x, (size_t)m.n_modes);
    upload_vec(c, u.mz, m.mass_z, (size_t)m.n_modes);
    rc.grid = u.grid.p;
    rc.energy = u.energy.p;
    rc.mass_x = u.mx.p;
    rc.mass_z = u.mz.p;
  }
  return rc;
}

}  // namespace

extern "C" {

int wdg2_set_contact_rhs(wdg_ctx* ctx, const wdg2_reservoir_grid* gr, const wdg2_contact_modes* source,
                         const wdg2_contact_modes* drain) {
  if (!ctx || !gr || !source || !drain) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not a full-2D context");
  return guarded(ctx, [&]() {
    F2Data& f = *ctx->f2;
    const wdg2_desc& d = f.d;
    if (d.ne_x < 2) throw std::invalid_argument("the inflow faces need at least two X elements");
    if (gr->n_kx < 2 || !gr->kx || !(gr->dkx > 0)) throw std::invalid_argument("k_x grid missing");
    if (gr->n_kz < 0 || (gr->n_kz > 0 && (!gr->kz || gr->n_kz < 2))) throw std::invalid_argument("k_z grid invalid");
    if (!gr->rho_x || !gr->rho_y || !gr->y_nodes) throw std::invalid_argument("relative cells and Y nodes required");
    if (!(gr->temp > 0) || !std::isfinite(gr->deg_factor) || !std::isfinite(gr->scale_x))
      throw std::invalid_argument("temperature must be positive, degeneracy and scale finite");
    const int nrx = d.n_rho_x, nry = d.n_rho_y, nyd = f.nyd, nxd = f.nxd, nl = d.nl_x;
    static thread_local ResUpload up[2];
    DevBuf<double> kx, kz, rx, ry, yn, prof;
    upload_vec(ctx, kx, gr->kx, (size_t)gr->n_kx);
    if (gr->n_kz > 0) upload_vec(ctx, kz, gr->kz, (size_t)gr->n_kz);
    upload_vec(ctx, rx, gr->rho_x, (size_t)nrx);
    upload_vec(ctx, ry, gr->rho_y, (size_t)nry);
    upload_vec(ctx, yn, gr->y_nodes, (size_t)nyd);
    const ResContact c0 = res_contact(ctx, *source, up[0], 0, nyd, nrx, nl);
    const ResContact c1 = res_contact(ctx, *drain, up[1], 1, nyd, nrx, nl);
    ResGrid g{};
    g.nkx = gr->n_kx; g.nkz = gr->n_kz; g.nrx = nrx; g.nry = nry; g.nyd = nyd; g.nxd = nxd; g.nl = nl;
    g.kx = kx.p; g.kz = gr->n_kz > 0 ? kz.p : nullptr; g.rx = rx.p; g.ry = ry.p; g.ynodes = yn.p;
    g.dkx = gr->dkx; g.temp = gr->temp; g.deg = gr->deg_factor; g.sx = gr->scale_x;
    prof.alloc((size_t)2 * kResMaxModes * nrx);
    const int mmax = std::max(1, std::max(c0.n_modes, c1.n_modes));
    hipLaunchKernelGGL(res_profile_kernel, dim3(mmax, 2), dim3(256), (size_t)gr->n_kx * sizeof(double), ctx->stream,
                       c0, c1, g, prof.p);
    const long N = ctx->N;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(ctx->red_blocks), dim3(256), 0, ctx->stream, ctx->d_b.p, N);
    const bool want0 = source->reservoir_out != nullptr, want1 = drain->reservoir_out != nullptr;
    const size_t nres = (size_t)nrx * nry * nyd;
    if (want0) up[0].res.alloc(nres);
    if (want1) up[1].res.alloc(nres);
    const int threads = std::min(256, ((nrx + 63) / 64) * 64);
    hipLaunchKernelGGL(res_rhs_kernel, dim3(nyd, nry, 2), dim3(threads), (size_t)(kResMaxModes + nrx) * sizeof(double2),
                       ctx->stream, c0, c1, g, prof.p, up[0].lift.p, up[1].lift.p, ctx->d_b.p,
                       want0 ? up[0].res.p : nullptr, want1 ? up[1].res.p : nullptr);
    HIPCHK(hipGetLastError());
    // normalise as wdg2_set_rhs does (b_hat = b / ||b||)
    const int nb = ctx->red_blocks;
    hipLaunchKernelGGL(norm2_partials_kernel, dim3(nb), dim3(256), 0, ctx->stream, ctx->d_b.p, N,
                       (double*)ctx->d_part.p);
    std::vector<double> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), ctx->d_part.p, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (want0)
      HIPCHK(hipMemcpyAsync(source->reservoir_out, up[0].res.p, nres * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (want1)
      HIPCHK(hipMemcpyAsync(drain->reservoir_out, up[1].res.p, nres * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double s2 = 0;
    for (double v : part) s2 += v;
    if (!std::isfinite(s2)) throw std::runtime_error("the contact right-hand side is not finite");
    const double old_norm = ctx->rhs_norm;
    ctx->rhs_norm = std::sqrt(s2);
    if (ctx->rhs_norm > 0)
      hipLaunchKernelGGL(scale_complex_kernel, dim3(nb), dim3(256), 0, ctx->stream, ctx->d_b.p, 1.0 / ctx->rhs_norm, N);
    if (ctx->have_solution && old_norm > 0 && ctx->rhs_norm > 0 && old_norm != ctx->rhs_norm)
      hipLaunchKernelGGL(scale_complex_kernel, dim3(nb), dim3(256), 0, ctx->stream, ctx->d_x.p,
                         old_norm / ctx->rhs_norm, N);
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->rhs_norm > 0) ctx->sol_scale = ctx->rhs_norm / f.sA;
    f.have_rhs = true;
    return WDG_OK;
  });
}

int wdg2_rowabs(wdg_ctx* ctx, double* out) {
  if (!ctx || !out) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not a full-2D context");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg2_set_potential has not been called");
  return guarded(ctx, [&]() {
    f2_diag_into(ctx, F2_ROWABS, ctx->d_y.p);
    std::vector<double2> h(ctx->N);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->d_y.p, (size_t)ctx->N * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (long i = 0; i < ctx->N; ++i) out[i] = h[i].x * ctx->f2->sA;
    return WDG_OK;
  });
}

int wdg2_set_precond(wdg_ctx* ctx, int32_t kind, int32_t* used) {
  if (!ctx) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not a full-2D context");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg2_set_potential has not been called");
  return guarded(ctx, [&]() {
    int u = 0;
    f2_build_precond(ctx, kind, 1e-12, true, &u);
    if (used) *used = u;
    return WDG_OK;
  });
}

int wdg2_set_line_precond(wdg_ctx* ctx, const wdg2_line_factors* lf) {
  if (!ctx || !lf) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not a full-2D context");
  if (!ctx->have_potential) return fail(ctx, WDG_ERR_STATE, "wdg2_set_potential has not been called");
  return guarded(ctx, [&]() {
    f2_set_line_impl(ctx, *lf);
    return WDG_OK;
  });
}

int wdg2_set_dv_mode(wdg_ctx* ctx, int32_t mode) {
  if (!ctx) return fail(nullptr, WDG_ERR_ARG, "null argument");
  if (!ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not a full-2D context");
  if (mode != WDG2_DV_INTERP && mode != WDG2_DV_TABLE) return fail(ctx, WDG_ERR_ARG, "unknown Delta V mode");
  return guarded(ctx, [&]() {
    if (ctx->f2->dv_mode == mode) return WDG_OK;
    ctx->f2->dv_mode = mode;
    if (ctx->have_potential) f2_configure_dv(ctx);   // else at wdg2_set_potential
    return WDG_OK;
  });
}

int wdg2_observables(wdg_ctx* ctx, const double* rho, const int32_t* idx, const double* w, const int32_t* counts,
                     double* out) {
  if (!ctx || !idx || !w || !counts || !out) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not a full-2D context");
  return guarded(ctx, [&]() {
    F2Data& f = *ctx->f2;
    F2ObsW ow{};
    for (int k = 0; k < 3; ++k) {
      if (counts[k] < 0 || counts[k] > 8) throw std::invalid_argument("counts must be in 0..8");
      ow.cnt[k] = counts[k];
      for (int j = 0; j < counts[k]; ++j) {
        if (idx[8 * k + j] < 0 || idx[8 * k + j] >= f.nrel) throw std::invalid_argument("weight index out of range");
        ow.idx[8 * k + j] = idx[8 * k + j];
        ow.w[8 * k + j] = w[8 * k + j];
      }
    }
    const double2* src;
    double scale;
    if (rho) {
      ctx->d_u64.alloc(ctx->N);
      HIPCHK(hipMemcpyAsync(ctx->d_u64.p, rho, (size_t)ctx->N * 16, hipMemcpyHostToDevice, ctx->stream));
      src = ctx->d_u64.p;
      scale = 1.0;
    } else {
      if (!ctx->have_solution) throw std::logic_error("no solution available: call wdg_solve first");
      src = ctx->d_x.p;
      scale = ctx->sol_scale;
    }
    ctx->d_obs.alloc((size_t)3 * f.ncen);
    hipLaunchKernelGGL(f2_observables_kernel, dim3((f.ncen + 255) / 256), dim3(256), 0, ctx->stream, src, f.ncen,
                       f.nrel, ow, scale, ctx->d_obs.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, ctx->d_obs.p, (size_t)3 * f.ncen * 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WDG_OK;
  });
}


int wdg2_poisson_setup(wdg_ctx* ctx, const wdg2_poisson_desc* d) {
  if (!ctx || !d) return fail(ctx, WDG_ERR_ARG, "null argument");
  if (!ctx->f2) return fail(ctx, WDG_ERR_UNSUPPORTED, "not a full-2D context");
  return guarded(ctx, [&]() {
    F2Data& f = *ctx->f2;
    const int nx = d->nx, nu = d->nu;
    if (nx < 2 || nu < 1 || d->n_entries < 1 || !d->a_row || !d->a_col || !d->a_val || !d->node_type ||
        !d->net_doping || !d->proj_x || !d->proj_y || !d->volume || !(d->temp > 0))
      throw std::invalid_argument("Poisson description: sizes, arrays and temperature required");
    if (d->x_dof != f.nxd || d->y_dof != f.nyd)
      throw std::invalid_argument("Poisson description: x_dof, y_dof must match the context's centre DOFs");
    if (d->n_weights < 1 || d->n_weights > kP2MaxWeights || !d->w_idx || !d->w_val)
      throw std::invalid_argument("Poisson description: 1..8 density weights required");
    auto p = std::make_shared<P2Data>();
    P2Data& q = *p;
    q.nx = nx;
    q.nu = nu;
    q.n = nx * nu;
    q.b = nu;
    q.xdof = d->x_dof;
    q.ydof = d->y_dof;
    q.ncen = f.ncen;
    q.nrel = f.nrel;
    q.temp = d->temp;
    q.w.cnt = d->n_weights;
    for (int j = 0; j < d->n_weights; ++j) {
      if (d->w_idx[j] < 0 || d->w_idx[j] >= f.nrel) throw std::invalid_argument("density weight index out of range");
      q.w.idx[j] = d->w_idx[j];
      q.w.w[j] = d->w_val[j];
    }
    const int n = q.n, b = q.b, W = 2 * b + 1;
    auto korder = [&](int u) { return (u / nx) + nu * (u % nx); };   // u = ix + nx iy -> k = iy + nu ix
    std::vector<double> band((size_t)n * W, 0.0);
    for (int e = 0; e < d->n_entries; ++e) {
      const int r = d->a_row[e], c = d->a_col[e];
      if (r < 0 || r >= n || c < 0 || c >= n) throw std::invalid_argument("Poisson operator index out of range");
      const int kr = korder(r), kc = korder(c);
      if (std::abs(kr - kc) > b)
        throw std::invalid_argument("Poisson operator couples nodes beyond neighbouring material columns");
      band[(size_t)kr * W + (kc - kr + b)] += d->a_val[e];
    }
    std::vector<int> type(n);
    std::vector<double> net(n);
    for (int u = 0; u < n; ++u) {
      const int t = d->node_type[u];
      if (t != 1 && t != 2) throw std::invalid_argument("Poisson node types must be 1 (free) or 2 (Dirichlet)");
      type[korder(u)] = t;
      net[korder(u)] = d->net_doping[u];
    }
    // nonzero column range of each projection row
    auto ranges = [](const double* m, int rows, int cols) {
      std::vector<int> r(2 * rows);
      for (int i = 0; i < rows; ++i) {
        int a = cols, z = 0;
        for (int j = 0; j < cols; ++j)
          if (m[(size_t)i * cols + j] != 0.0) {
            a = std::min(a, j);
            z = j + 1;
          }
        r[2 * i] = std::min(a, z);
        r[2 * i + 1] = z;
      }
      return r;
    };
    for (int u = 0; u < n; ++u)
      if (!(d->volume[u] > 0)) throw std::invalid_argument("Poisson dual volumes must be positive");
    auto up = [&](DevBuf<double>& buf, const double* v, size_t cnt) {
      buf.alloc(cnt);
      HIPCHK(hipMemcpy(buf.p, v, cnt * sizeof(double), hipMemcpyHostToDevice));
    };
    auto upi = [&](DevBuf<int>& buf, const std::vector<int>& v) {
      buf.alloc(v.size());
      HIPCHK(hipMemcpy(buf.p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
    };
    up(q.A, band.data(), band.size());
    up(q.net, net.data(), n);
    upi(q.type, type);
    up(q.px, d->proj_x, (size_t)nx * q.xdof);
    up(q.py, d->proj_y, (size_t)nu * q.ydof);
    upi(q.pxr, ranges(d->proj_x, nx, q.xdof));
    upi(q.pyr, ranges(d->proj_y, nu, q.ydof));
    up(q.vol, d->volume, n);
    q.J.alloc((size_t)n * W);
    for (DevBuf<double>* buf : {&q.nmat, &q.v, &q.f, &q.chg, &q.phi_in, &q.bc, &q.phi_out, &q.cand}) buf->alloc(n);
    q.ndg.alloc(q.ncen);
    q.info.alloc(4);
    ctx->p2 = p;
    return WDG_OK;
  });
}

int wdg2_poisson_step(wdg_ctx* const* ctxs, int32_t n, const double* phi_in, const double* bc, int32_t newton,
                      double alpha, double* phi_out, double* cand_out, double* n_out, wdg2_poisson_info* info) {
  if (!ctxs || n < 1 || !ctxs[0] || !phi_in || !bc || !phi_out || !info) return fail(nullptr, WDG_ERR_ARG, "null argument");
  wdg_ctx* c0 = ctxs[0];
  if (!c0->p2) return fail(c0, WDG_ERR_STATE, "wdg2_poisson_setup has not been called on the first context");
  if (!std::isfinite(alpha) || !(alpha > 0) || alpha > 1) return fail(c0, WDG_ERR_ARG, "alpha must lie in (0, 1]");
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i] || !ctxs[i]->f2) return batch_fail(ctxs[i], i, WDG_ERR_ARG, "every context must be a full-2D context");
    if (!ctxs[i]->have_solution) return batch_fail(ctxs[i], i, WDG_ERR_STATE, "no solution on the context yet");
    if (ctxs[i]->f2->ncen != c0->p2->ncen || ctxs[i]->f2->nrel != c0->p2->nrel)
      return batch_fail(ctxs[i], i, WDG_ERR_ARG, "the contexts must share the Poisson context's grid");
  }
  return guarded(c0, [&]() {
    P2Data& q = *c0->p2;
    const int m = q.n;
    for (int u = 0; u < m; ++u)
      if (!std::isfinite(phi_in[u]) || !std::isfinite(bc[u])) throw std::invalid_argument("phi_in and bc must be finite");
    // the summed density of the contexts' solutions (their streams drained first)
    for (int i = 0; i < n; ++i) {
      HIPCHK(hipStreamSynchronize(ctxs[i]->stream));
      hipLaunchKernelGGL(p2_density_kernel, dim3((q.ncen + 255) / 256), dim3(256), 0, c0->stream, ctxs[i]->d_x.p,
                         q.ncen, q.nrel, q.w, ctxs[i]->sol_scale, i > 0 ? 1 : 0, q.ndg.p);
    }
    hipLaunchKernelGGL(p2_project_kernel, dim3((m + 255) / 256), dim3(256), 0, c0->stream, q.ndg.p, q.nx, q.nu,
                       q.xdof, q.ydof, q.px.p, q.pxr.p, q.py.p, q.pyr.p, q.vol.p, q.nmat.p);
    HIPCHK(hipMemcpyAsync(q.phi_in.p, phi_in, m * sizeof(double), hipMemcpyHostToDevice, c0->stream));
    HIPCHK(hipMemcpyAsync(q.bc.p, bc, m * sizeof(double), hipMemcpyHostToDevice, c0->stream));
    P2Args a{q.nx, q.nu, q.n, q.b, q.A.p, q.type.p, q.net.p, q.nmat.p, q.phi_in.p, q.bc.p, q.J.p, q.v.p, q.f.p,
             q.chg.p, q.phi_out.p, q.cand.p, q.info.p,
             newton ? 1.602176634e-19 * 1.602176634e-19 / (1.380649e-23 * q.temp) : 0.0, alpha};
    hipLaunchKernelGGL((p2_newton_kernel<256>), dim3(1), dim3(256), 0, c0->stream, a);
    HIPCHK(hipGetLastError());
    double h[4];
    HIPCHK(hipMemcpyAsync(phi_out, q.phi_out.p, m * sizeof(double), hipMemcpyDeviceToHost, c0->stream));
    if (cand_out) HIPCHK(hipMemcpyAsync(cand_out, q.cand.p, m * sizeof(double), hipMemcpyDeviceToHost, c0->stream));
    if (n_out) HIPCHK(hipMemcpyAsync(n_out, q.nmat.p, m * sizeof(double), hipMemcpyDeviceToHost, c0->stream));
    HIPCHK(hipMemcpyAsync(h, q.info.p, sizeof h, hipMemcpyDeviceToHost, c0->stream));
    HIPCHK(hipStreamSynchronize(c0->stream));
    info->physical_relres_before = h[0];
    info->physical_relres_after = h[1];
    info->linear_relres = h[2];
    info->unmixed_step = h[3];
    if (!std::isfinite(h[2]) || !std::isfinite(h[3])) throw StatusError(WDG_ERR_NUMERIC, "non-finite Poisson update");
    return WDG_OK;
  });
}

}  // extern "C"
