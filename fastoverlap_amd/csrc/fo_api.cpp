// Host runtime and C ABI of libfastoverlap_hip (see include/fastoverlap_hip.h).
//
// A context owns one device, a default stream, the byte budget and the
// device-resident tables (Wigner-d and twiddles per grid side, twiddles per
// FFT side), built once and shared read-only by every batch.  A batch owns
// one device allocation holding all of its inputs, intermediates and
// outputs, and its own stream, so several batches run concurrently.
#include "fastoverlap_hip.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "fo_common.h"
#include "fo_host.h"
#include "fo_internal.h"
#include "fo_refine.h"

namespace fo {
// launchers (fo_spherical.hip / fo_periodic.hip)
void launch_wigner_table(hipStream_t, int, int, double*, float*);
void launch_twiddles(hipStream_t, int, double, double2*, float2*);
void launch_sph_prep(hipStream_t, int, const SphPair*, const double*, const double*, double*,
                     double*, double*, int);
void launch_sph_ylm(hipStream_t, int, const SphPair*, const int*, const double*, const double*,
                    const double2*, const double*, double2*, double2*, double*, double*);
void launch_sph_radial(hipStream_t, int, int, long long, const SphPair*, const SphGroup*,
                       const double*, const double*, const double*, double*, int);
void launch_harm_coeffs(hipStream_t, int, int, int, int, const SphPair*, const SphGroup*,
                        const int*, const double*, const double*, const double*, int, double, int,
                        const double*, const double2*, const double2*, double*, int, double2*,
                        float2*);
void launch_coef_norm2(hipStream_t, int, const SphPair*, const double2*, double*);
int harm_block_lx(int L, int nmax, int basis);
void launch_sph_contract(hipStream_t, int, int, int, const SphPair*, const SphGroup*,
                         const double*, const double*, const double2*, const double2*,
                         const double*, const double*, double2*, float2*);
size_t isoft_plane_smem(int, int, int);
int isoft_plane_use_fft(int, int, size_t);
void launch_isoft_plane(hipStream_t, int, int, int, int, long long, const SphPair*, const float*,
                        const float2*, const float2*, const float2*, int, int, PlaneCand*, float*,
                        double*, int, const float4*, int);
void launch_beta_gemm(hipStream_t, int, int, int, long long, int, const SphPair*, const float*,
                      const float2*, float4*, int, double*);
void launch_sph_select(hipStream_t, int, int, long long, int, const SphPair*, const double*,
                       const double2*, const double2*, const double2*, int, const PlaneCand*,
                       double2*, long long, const double*, SphPeak*, double, int);
void launch_sph_rotopt(hipStream_t, int, int, int, const SphPair*, const double2*, const double2*,
                       int, const SphPeak*, const double*, int, SphOpt*);
void launch_rec_table(hipStream_t, int, double*, double*, double2*, double*);
void launch_sph_finish(hipStream_t, int, int, const SphPair*, const SphGroup*, const int*,
                       const double*, const double*, const SphOpt*, int, double*, int*, int*,
                       double*, SphResult*);
void launch_coef_to_plane(hipStream_t, int, int, int, const SphPair*, const double2*, float2*);
void launch_soft_forward(hipStream_t, int, long long, const double*, const double2*, const double2*,
                         double*, double2*, double2*);
size_t soft_forward_smem(int);
void launch_sph_mirror(hipStream_t, int, int, const SphPair*, double2*);
void launch_sph_energy(hipStream_t, int, const double2*, double, double, double, const double*,
                       const double*, int, double*);
void launch_wigner_D(hipStream_t, int, double, double, double, double2*, double2*);
void launch_ylm_angles(hipStream_t, int, int, const double*, const double*, double2*);
size_t per_sf_smem(int);
void launch_per_sf(hipStream_t, int, int, const PerPair*, const PerGroup*, const int*,
                   const double*, const double*, double2*, double2*, double2*);
void launch_per_product(hipStream_t, int, int, const PerPair*, const double2*, const double2*,
                        double2*);
size_t per_plane_smem(int, int);
void launch_per_plane(hipStream_t, int, int, int, const PerPair*, const double2*, const float2*,
                      PlaneCand*, float*, float2*);
void launch_per_product_indexed(hipStream_t, int, int, const PerPair*, const double2*, const int*,
                                const int*, double2*);
void launch_find_peaks(hipStream_t, int, const float*, long long, float*, int, int, int, int, int,
                       double*, int*);
int find_peaks_max_window();
void launch_per_select(hipStream_t, int, int, int, const PerPair*, const double2*, const double2*,
                       const PlaneCand*, double2*, long long, PerPeak*);
void launch_per_refine(hipStream_t, int, int, const PerPair*, const PerGroup*, const int*,
                       const double*, const double*, const PerPeak*, int, int*, int*, int*,
                       double*, double*, PerResult*);
void launch_lap_align(hipStream_t, int, const double*, const double*, int, const int*, const int*,
                      const int*, int, const double*, int, int*, int*, double*);
}  // namespace fo

using namespace fo;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define FO_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return set_err(FO_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define FO_CHECK_LAUNCH()                                                         \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess)                                                         \
      return set_err(FO_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

static bool all_finite(const double* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}
static constexpr int kMaxL = 63;
static constexpr int kMaxGroup = 4096;
static constexpr int kHarmMaxN = 40;  // harmonic basis radial degree

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct SphTable {
  int n = 0, Lt = 0;
  long long Kt = 0;         // entries per beta of tab64 (compact layout, K(Lt))
  long long Kp = 0;         // entries per beta of tab32 (plane layout, rows(Lt) * n)
  double* tab64 = nullptr;
  float* tab32 = nullptr;
  double2* tw64 = nullptr;
  float2* tw32 = nullptr;
};
// Recurrence coefficients of the unnormalised Wigner d for every (l, m1, m2)
// with l <= kRecL, shared by all batches of a context.
static constexpr int kRecL = 63;
struct RecTable {
  double* rec = nullptr;    // 3 * K(kRecL)
  double* seedu = nullptr;  // (2 kRecL + 1)^2
  double2* leg = nullptr;   // Legendre recurrence coefficients, q(kRecL, kRecL) + 1
  double* cm = nullptr;     // P_m^m prefactors, kRecL + 1
};

struct PerTw {
  int F = 0;
  double2* tw64 = nullptr;
  float2* tw32 = nullptr;
};

struct fo_context {
  int device = 0;
  size_t max_lds = 65536;  // bytes of LDS one workgroup may use on this device
  hipStream_t stream = nullptr;
  unsigned long long budget = 0;
  unsigned long long used = 0;
  std::mutex mu;       // tables + accounting
  std::mutex call_mu;  // serialises the single-call helpers
  std::map<int, SphTable> sph;
  std::map<int, PerTw> per;
  // launch shapes chosen by the autotuner, keyed by (grid side, largest
  // degree, pair-count bucket, plane path)
  std::map<std::tuple<int, int, int, int>, std::array<int, 3>> tuned;
  RecTable rect;
  std::vector<void*> owned;  // table allocations
};

static int ctx_alloc(fo_context* c, size_t bytes, void** p) {
  {
    std::lock_guard<std::mutex> g(c->mu);
    if (c->budget && c->used + bytes > c->budget)
      return set_err(FO_ERR_NOMEM, "allocation of " + std::to_string(bytes) +
                                       " bytes exceeds the context budget");
    c->used += bytes;
  }
  hipError_t e = hipMalloc(p, bytes ? bytes : 256);
  if (e != hipSuccess) {
    std::lock_guard<std::mutex> g(c->mu);
    c->used -= bytes;
    return set_err(FO_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return FO_OK;
}
static void ctx_free(fo_context* c, void* p, size_t bytes) {
  if (!p) return;
  (void)hipFree(p);
  std::lock_guard<std::mutex> g(c->mu);
  c->used -= bytes;
}

// build = false: only look the table up (FO_ERR_STATE when it was never
// built), so that a caller that must not allocate never reaches hipMalloc.
static int get_sph_table(fo_context* c, int n, const SphTable** out, bool build = true) {
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->sph.find(n);
  if (it != c->sph.end()) {
    *out = &it->second;
    return FO_OK;
  }
  if (!build)
    return set_err(FO_ERR_STATE, "no Wigner table for grid side " + std::to_string(n) +
                                     " in this context (fo_sph_prepare builds it)");
  SphTable t;
  t.n = n;
  t.Lt = std::min(n / 2 - 1, kMaxL);
  t.Kt = kcount(t.Lt);
  t.Kp = plane_rows(t.Lt) * n;
  const size_t b64 = sizeof(double) * (size_t)t.Kt * n, b32 = sizeof(float) * (size_t)t.Kp * n;
  const size_t bt = (sizeof(double2) + sizeof(float2)) * (size_t)n;
  const size_t total = b64 + b32 + bt + 1024;
  if (c->budget && c->used + total > c->budget)
    return set_err(FO_ERR_NOMEM, "Wigner table exceeds the context budget");
  void* base = nullptr;
  hipError_t e = hipMalloc(&base, total);
  if (e != hipSuccess) return set_err(FO_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->used += total;
  c->owned.push_back(base);
  char* p = (char*)base;
  t.tab64 = (double*)p;
  p += (b64 + 255) / 256 * 256;
  t.tab32 = (float*)p;
  p += (b32 + 255) / 256 * 256;
  t.tw64 = (double2*)p;
  p += (sizeof(double2) * n + 255) / 256 * 256;
  t.tw32 = (float2*)p;
  FO_HIP(hipMemsetAsync(t.tab32, 0, b32, c->stream));  // entries |m2| > l stay zero
  launch_wigner_table(c->stream, n, t.Lt, t.tab64, t.tab32);
  launch_twiddles(c->stream, n, 1.0, t.tw64, t.tw32);
  FO_CHECK_LAUNCH();
  FO_HIP(hipStreamSynchronize(c->stream));
  c->sph[n] = t;
  *out = &c->sph[n];
  return FO_OK;
}

static int get_rec_table(fo_context* c, const RecTable** out, bool build = true) {
  std::lock_guard<std::mutex> g(c->mu);
  if (!c->rect.rec && !build)
    return set_err(FO_ERR_STATE, "recurrence tables not built (fo_sph_prepare builds them)");
  if (!c->rect.rec) {
    const size_t b1 = sizeof(double) * 3 * (size_t)kcount(kRecL);
    const size_t b2 = sizeof(double) * (size_t)(2 * kRecL + 1) * (2 * kRecL + 1);
    const size_t b3 = sizeof(double2) * (size_t)((kRecL + 1) * (kRecL + 2) / 2);
    const size_t b4 = sizeof(double) * (size_t)(kRecL + 1);
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t total = up(b1) + up(b2) + up(b3) + up(b4);
    void* base = nullptr;
    hipError_t e = hipMalloc(&base, total);
    if (e != hipSuccess) return set_err(FO_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->used += total;
    c->owned.push_back(base);
    char* cb = (char*)base;
    c->rect.rec = (double*)cb;
    c->rect.seedu = (double*)(cb + up(b1));
    c->rect.leg = (double2*)(cb + up(b1) + up(b2));
    c->rect.cm = (double*)(cb + up(b1) + up(b2) + up(b3));
    launch_rec_table(c->stream, kRecL, c->rect.rec, c->rect.seedu, c->rect.leg, c->rect.cm);
    FO_CHECK_LAUNCH();
    FO_HIP(hipStreamSynchronize(c->stream));
  }
  *out = &c->rect;
  return FO_OK;
}

static int get_per_tw(fo_context* c, int F, const PerTw** out) {
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->per.find(F);
  if (it != c->per.end()) {
    *out = &it->second;
    return FO_OK;
  }
  PerTw t;
  t.F = F;
  const size_t total = (sizeof(double2) + sizeof(float2)) * (size_t)F + 512;
  void* base = nullptr;
  hipError_t e = hipMalloc(&base, total);
  if (e != hipSuccess) return set_err(FO_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->used += total;
  c->owned.push_back(base);
  t.tw64 = (double2*)base;
  t.tw32 = (float2*)((char*)base + (sizeof(double2) * F + 255) / 256 * 256);
  launch_twiddles(c->stream, F, -1.0, t.tw64, t.tw32);
  FO_CHECK_LAUNCH();
  FO_HIP(hipStreamSynchronize(c->stream));
  c->per[F] = t;
  *out = &c->per[F];
  return FO_OK;
}

// Device buffers of the multi-peak search, allocated on a batch's first
// peak search (and grown when more peaks are asked for): one working copy of
// every kept grid, the fitted parameters and the fit counts.
struct PeakScratch {
  void* base = nullptr;
  size_t bytes = 0;
  float* work = nullptr;
  double* fit = nullptr;
  int* count = nullptr;
  int npeaks = 0;
};

static int peak_scratch(fo_context* c, PeakScratch& ps, long long ngrids, long long gsize,
                        int npeaks) {
  if (ps.base && ps.npeaks >= npeaks) return FO_OK;
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t bw = up(sizeof(float) * (size_t)ngrids * gsize);
  const size_t bf = up(sizeof(double) * 11 * (size_t)ngrids * npeaks);
  const size_t bc = up(sizeof(int) * (size_t)ngrids);
  if (ps.base) ctx_free(c, ps.base, ps.bytes);
  ps = PeakScratch{};
  int rc = ctx_alloc(c, bw + bf + bc, &ps.base);
  if (rc) return rc;
  ps.bytes = bw + bf + bc;
  ps.work = (float*)ps.base;
  ps.fit = (double*)((char*)ps.base + bw);
  ps.count = (int*)((char*)ps.base + bw + bf);
  ps.npeaks = npeaks;
  return FO_OK;
}

// Run the peak search over ngrids kept grids of gsize floats each and copy
// the results out (fit: ngrids x npeaks x 11, count: ngrids, residual: the
// working grids after the subtractions, optional).
static int run_peak_search(fo_context* c, hipStream_t s, PeakScratch& ps, const float* grids,
                           long long ngrids, int n0, int n1, int n2, int npeaks, int width,
                           double* fit, int* count, float* residual) {
  if (npeaks < 1 || npeaks > 64) return set_err(FO_ERR_INVALID, "npeaks must be in [1, 64]");
  if (width < 1 || width > find_peaks_max_window())
    return set_err(FO_ERR_UNSUPPORTED, "peak fit half width must be in [1, " +
                                           std::to_string(find_peaks_max_window()) + "]");
  const long long gsize = (long long)n0 * n1 * n2;
  int rc = peak_scratch(c, ps, ngrids, gsize, npeaks);
  if (rc) return rc;
  launch_find_peaks(s, (int)ngrids, grids, gsize, ps.work, n0, n1, n2, npeaks, width, ps.fit,
                    ps.count);
  FO_CHECK_LAUNCH();
  // the kernel leaves unfound peaks untouched: report zeros for them
  std::vector<double> hf((size_t)ngrids * ps.npeaks * 11);
  std::vector<int> hc((size_t)ngrids);
  FO_HIP(hipMemcpyAsync(hf.data(), ps.fit, sizeof(double) * hf.size(), hipMemcpyDeviceToHost, s));
  FO_HIP(hipMemcpyAsync(hc.data(), ps.count, sizeof(int) * hc.size(), hipMemcpyDeviceToHost, s));
  if (residual)
    FO_HIP(hipMemcpyAsync(residual, ps.work, sizeof(float) * (size_t)(ngrids * gsize),
                          hipMemcpyDeviceToHost, s));
  FO_HIP(hipStreamSynchronize(s));
  for (long long gi = 0; gi < ngrids; ++gi) {
    count[gi] = hc[gi];
    for (int k = 0; k < npeaks; ++k)
      for (int q = 0; q < 11; ++q)
        fit[(gi * npeaks + k) * 11 + q] =
            k < hc[gi] ? hf[((size_t)gi * npeaks + k) * 11 + q] : 0.0;
  }
  return FO_OK;
}

// bump layout inside one allocation
struct Layout {
  size_t off = 0;
  template <class T>
  size_t add(size_t count) {
    size_t o = off;
    off += (sizeof(T) * count + 255) / 256 * 256;
    return o;
  }
};

extern "C" {

int fo_version(void) { return 10000; }

const char* fo_last_error(void) { return g_err.c_str(); }

int fo_device_count(int* count) {
  if (!count) return set_err(FO_ERR_INVALID, "count is NULL");
  FO_HIP(hipGetDeviceCount(count));
  return FO_OK;
}

int fo_context_create(int device, unsigned long long hbm_budget, fo_context** out) {
  if (!out) return set_err(FO_ERR_INVALID, "out is NULL");
  int n = 0;
  FO_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return set_err(FO_ERR_INVALID, "no such device");
  FO_HIP(hipSetDevice(device));
  auto* c = new fo_context();
  c->device = device;
  c->budget = hbm_budget;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
      lds > 0)
    c->max_lds = (size_t)lds;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return set_err(FO_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  // device address of this device's allocation-free scratch (once per device)
  fo::noalloc_attach(device);
  *out = c;
  return FO_OK;
}

int fo_sph_prepare(fo_context* c, int grid_n) {
  if (!c) return set_err(FO_ERR_INVALID, "ctx is NULL");
  if (grid_n < 2 || grid_n > 256) return set_err(FO_ERR_UNSUPPORTED, "grid_n must be in [2, 256]");
  FO_HIP(hipSetDevice(c->device));
  const SphTable* t = nullptr;
  const RecTable* r = nullptr;
  int rc = get_sph_table(c, grid_n, &t);
  if (rc) return rc;
  return get_rec_table(c, &r);
}

int fo_context_destroy(fo_context* c) {
  if (!c) return FO_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (void* p : c->owned) (void)hipFree(p);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return FO_OK;
}

int fo_context_memory(fo_context* c, unsigned long long* used, unsigned long long* budget) {
  if (!c) return set_err(FO_ERR_INVALID, "ctx is NULL");
  std::lock_guard<std::mutex> g(c->mu);
  if (used) *used = c->used;
  if (budget) *budget = c->budget;
  return FO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// spherical batch
// ---------------------------------------------------------------------------
// Pairs whose harmonics and radial kernels are built and contracted
// together: the coefficient stage runs chunk by chunk through one reused
// Y / K buffer sized for the largest chunk, so those intermediates stay in
// the device's last-level cache instead of streaming through HBM.
struct CoefChunk {
  int p0 = 0, p1 = 0;   // pairs [p0, p1)
  int Lmax = 0;         // max L over the chunk
  long long maxjk = 0;  // max sum of group size^2 over the chunk
};

struct fo_sph_batch {
  fo_context* ctx = nullptr;
  hipStream_t stream = nullptr;
  const SphTable* tab = nullptr;
  const RecTable* rect = nullptr;
  SphLaunch lc{};
  int flags = 0;
  int maxgroup = 1;
  int fft_chunk = 0;  // 1: two-pass register transform in the plane kernel, 0: direct DFT
  long long Ntot = 0, Ktot = 0, C32tot = 0;
  // the last enqueued coefficient stage left the rows m1 < 0 of I64A unset
  bool mirror_pending = false;
  std::vector<SphPair> hp;
  std::vector<SphGroup> hg;
  std::vector<int> hgath;
  size_t bytes = 0;
  void* base = nullptr;
  // device pointers
  SphPair* pd = nullptr;
  SphGroup* grp = nullptr;
  int* gath = nullptr;
  double *pos1 = nullptr, *pos2 = nullptr, *X1c = nullptr, *X2c = nullptr, *X2R = nullptr,
         *X2out = nullptr, *scale = nullptr, *R = nullptr;
  double2 *Y1 = nullptr, *Y2 = nullptr;         // Y_l^m (m >= 0) of the two structures
  double* rad = nullptr;  // |x| of every grouped atom: structure 1, then structure 2
  double2 *I64A = nullptr, *I64B = nullptr, *scr = nullptr;
  float2 *I32A = nullptr, *I32B = nullptr;
  PlaneCand* cand = nullptr;
  double* abound = nullptr;  // per (pair, orientation): sum of weighted |C^l| (error bound)
  SphPeak* peaks = nullptr;
  SphOpt* opt = nullptr;
  SphResult* res = nullptr;
  int *perm = nullptr, *rowsol = nullptr;
  float* grid = nullptr;
  long long scr_stride = 0;
  // graphs per stage mask
  std::map<int, hipGraphExec_t> graphs;
  std::vector<CoefChunk> chunks;
  // timing
  bool timing = false;
  hipEvent_t ev[9] = {};
  std::vector<hipEvent_t> cev;  // 3 per chunk + 1: ylm | radial | contraction boundaries
  bool ev_made = false;
  bool cev_valid = false;       // the last timed run recorded cev
  // layout (byte offsets inside base) and ownership
  struct {
    size_t pd, grp, gath, pos1, pos2, X1c, X2c, X2R, X2out, scale, R, Y1, Y2, rad, I64A, I32A,
        I64B, I32B, cand, abound, peaks, opt, res, perm, rowsol, scr, grid, S, gtab;
  } off{};
  int gsum = 0;
  bool own_mem = false, own_stream = false;
  PeakScratch pk;  // multi-peak search buffers (allocated on first use)
  // FO_SPH_BETA_GEMM: even / odd-degree beta sums from k_beta_gemm
  // (split-BF16x3 on the matrix cores), [pair][beta][m1 <= Lmax][entry]
  float4* S = nullptr;
  // launch shapes (arithmetic-neutral): plane pairs per XCD chunk, a
  // reserved entry (kept for the interface, no effect), radial workgroups
  // per pair; 0 = default.  autotune: pick them on the first full run.
  std::array<int, 3> tune{{0, 15, 0}};
  bool autotune = false, tuned = false;
  // harmonic-oscillator basis (harm_nmax >= 0): radial degree, length scale,
  // basis polynomial table g[l][n][k] on the device, most groups of a pair
  int harm_nmax = -1;
  int harm_basis = 0;  // 0: SphericalHarmonicAlign's basis, 1: HARMONICCOEFFS'
  double harm_r0 = 1.0;
  double* gtab = nullptr;
  std::vector<double> hgtab;  // host copy (kept until the upload has run)
  int maxgroups = 1;
};

static int sph_enqueue(fo_sph_batch* b, int stages) {
  const SphLaunch& L = b->lc;
  hipStream_t s = b->stream;
  const int P = L.P;
  const bool tm = b->timing;
  auto mark = [&](int i) {
    if (tm) (void)hipEventRecord(b->ev[i], s);
  };
  mark(0);
  if (stages & FO_STAGE_PREP)
    launch_sph_prep(s, P, b->pd, b->pos1, b->pos2, b->X1c, b->X2c, b->scale, L.center);
  mark(1);
  if (stages & FO_STAGE_COEFF) {
    auto cmark = [&](size_t i) {
      if (tm) (void)hipEventRecord(b->cev[i], s);
    };
    for (size_t c = 0; c < b->chunks.size(); ++c) {
      const CoefChunk& C = b->chunks[c];
      const int Pc = C.p1 - C.p0;
      const SphPair* pdc = b->pd + C.p0;
      const double* scc = b->scale + C.p0;
      cmark(3 * c);
      launch_sph_ylm(s, Pc, pdc, b->gath, b->X1c, b->X2c, b->rect->leg, b->rect->cm, b->Y1, b->Y2,
                     b->rad, b->rad + b->hgath.size());
      cmark(3 * c + 1);
      if (b->harm_nmax >= 0) {
        launch_harm_coeffs(s, Pc, C.Lmax, b->maxgroups, L.maxgath, pdc, b->grp, b->gath, b->X1c,
                           b->X2c, scc, b->harm_nmax, b->harm_r0, b->harm_basis, b->gtab, b->Y1,
                           b->Y2, b->R, L.n, b->I64A, b->I32A);
        cmark(3 * c + 2);
        continue;
      }
      launch_sph_radial(s, Pc, C.Lmax, C.maxjk, pdc, b->grp, b->rad, b->rad + b->hgath.size(), scc,
                        b->R, b->tune[2]);
      cmark(3 * c + 2);
      launch_sph_contract(s, Pc, L.n, C.Lmax, pdc, b->grp, b->rad, b->rad + b->hgath.size(),
                          b->Y1, b->Y2, b->R, scc, b->I64A, b->I32A);
    }
    cmark(3 * b->chunks.size());
  }
  if (tm) b->cev_valid = (stages & FO_STAGE_COEFF) != 0;
  mark(2);
  mark(3);
  mark(4);
  if (stages & FO_STAGE_GRID) {
    if (b->S)
      launch_beta_gemm(s, P, L.n, L.Lmax, b->tab->Kp, L.Lt, b->pd, b->tab->tab32, b->I32A, b->S,
                       L.norient, b->abound);
    launch_isoft_plane(s, P, L.n, L.Lmax, b->fft_chunk, b->tab->Kp, b->pd, b->tab->tab32,
                       b->tab->tw32, b->I32A, b->I32B, L.mode, L.norient, b->cand, b->grid,
                       b->abound, b->tune[0], b->S, L.Lmax + 1);
    mark(5);
    // split-BF16x3 products carry up to 3.1 * 2^-16 of each |d I| beyond FP32
    launch_sph_select(s, P, L.n, b->tab->Kt, L.norient, b->pd, b->tab->tab64, b->tab->tw64,
                      b->I64A, b->I64B, L.mode, b->cand, b->scr, b->scr_stride, b->abound,
                      b->peaks, b->S ? 3.1 / 65536.0 : 0.0,
                      (b->flags & FO_SPH_EXACT_PLANES) ? 1 : 0);
  } else {
    mark(5);
  }
  mark(6);
  if (stages & FO_STAGE_ROTOPT)
    launch_sph_rotopt(s, P, L.norient, L.Lmax, b->pd, b->I64A, b->I64B, L.mode, b->peaks,
                      b->rect->seedu, kRecL, b->opt);
  mark(7);
  if (stages & FO_STAGE_FINISH)
    launch_sph_finish(s, P, b->maxgroup, b->pd, b->grp, b->gath, b->X1c, b->X2c, b->opt,
                      L.norient, b->X2R, b->perm, b->rowsol, b->X2out, b->res);
  mark(8);
  FO_CHECK_LAUNCH();
  return FO_OK;
}

namespace fo {

// Coefficients of the harmonic-oscillator basis functions in powers of r
// (fastoverlap/utils.py:408-427): g[l][n][k] multiplies r^{l+2k} in
// N_nl r^l L_n^{l+1/2}(r^2), N_nl = (2 n! r0^{-2l-3} / Gamma(n+l+3/2))^{1/2},
// L_n^a(x) = sum_k (-1)^k [prod_{i=k+1..n} (a+i)] / ((n-k)! k!) x^k.
std::vector<double> harmonic_basis_table(int L, int nmax, double r0) {
  const int NN = nmax + 1;
  std::vector<double> g((size_t)(L + 1) * NN * NN, 0.0);
  for (int l = 0; l <= L; ++l)
    for (int n = 0; n <= nmax; ++n) {
      const double a = l + 0.5;
      const double norm = std::exp(0.5 * (std::log(2.0) + std::lgamma(n + 1.0) -
                                          (2.0 * l + 3.0) * std::log(r0) - std::lgamma(n + l + 1.5)));
      for (int k = 0; k <= n; ++k) {
        double v = 1.0;
        for (int i = k + 1; i <= n; ++i) v *= a + i;
        for (int i = 2; i <= n - k; ++i) v /= i;
        for (int i = 2; i <= k; ++i) v /= i;
        g[((size_t)l * NN + n) * NN + k] = (k & 1 ? -v : v) * norm;
      }
    }
  return g;
}

// Byte budget of one coefficient chunk (harmonics of both structures plus
// the radial kernels).  Unlimited by default (one chunk: the kernels need
// many pairs per launch to fill the device); FO_COEF_CHUNK_MB caps the
// batch footprint of these intermediates at the cost of throughput.
static size_t coef_chunk_bytes() {
  if (const char* e = getenv("FO_COEF_CHUNK_MB")) {
    const long v = atol(e);
    if (v > 0) return (size_t)v << 20;
  }
  return SIZE_MAX;
}

int sph_batch_plan(fo_context* c, const fo_sph_batch_desc* d, fo_sph_batch** out,
                   bool build_tables, int harm_nmax, double harm_r0, int harm_basis) {
  if (!c || !d || !out) return set_err(FO_ERR_INVALID, "NULL argument");
  if (d->npairs < 1) return set_err(FO_ERR_INVALID, "npairs must be >= 1");
  if (!d->natoms || !d->L || !d->ngroups)
    return set_err(FO_ERR_INVALID, "natoms, L and ngroups are required");
  const int P = d->npairs, n = d->grid_n;
  if (n < 2 || n > 256) return set_err(FO_ERR_UNSUPPORTED, "grid_n must be in [2, 256]");
  const bool general = d->flags & FO_SPH_GENERAL;
  if (harm_nmax > kHarmMaxN)
    return set_err(FO_ERR_UNSUPPORTED, "harmonic radial degree nmax must be <= " +
                                           std::to_string(kHarmMaxN));
  if (harm_nmax >= 0 && !(harm_r0 > 0))
    return set_err(FO_ERR_INVALID, "harmonic length scale must be > 0");
  if (harm_nmax >= 0 && general)
    return set_err(FO_ERR_UNSUPPORTED, "the harmonic basis builds its own coefficients");
  auto b = std::make_unique<fo_sph_batch>();
  b->ctx = c;
  b->flags = d->flags;
  b->harm_nmax = harm_nmax;
  b->harm_r0 = harm_r0;
  b->harm_basis = harm_basis ? 1 : 0;
  int Lmax = 0, maxatoms = 0, maxgath = 0;
  long long maxjk = 0, Ntot = 0, Ktot = 0, C32tot = 0, radtot = 0, ylmtot = 0;
  long long radmax = 0, ylmmax = 0;  // largest chunk
  const size_t chunk_budget = coef_chunk_bytes();
  int gsum = 0, gatot = 0;
  for (int p = 0; p < P; ++p) gsum += d->ngroups[p];
  for (int q = 0; q < gsum; ++q) {
    if (!d->group_sizes || d->group_sizes[q] < 0) return set_err(FO_ERR_INVALID, "bad group size");
    gatot += d->group_sizes[q];
  }
  if (gatot > 0 && !d->group_atoms) return set_err(FO_ERR_INVALID, "group_atoms is NULL");
  b->hp.resize(P);
  b->hg.resize(gsum);
  b->hgath.resize(std::max(gatot, 1));
  int gq = 0, ga = 0;
  for (int p = 0; p < P; ++p) {
    const int N = d->natoms[p], Lp = d->L[p];
    if (N < 1) return set_err(FO_ERR_INVALID, "pair " + std::to_string(p) + ": natoms < 1");
    if (Lp < 0 || Lp > kMaxL)
      return set_err(FO_ERR_UNSUPPORTED, "pair " + std::to_string(p) + ": Jmax must be in [0, 63]");
    if (n < 2 * (Lp + 1))
      return set_err(FO_ERR_INVALID, "pair " + std::to_string(p) + ": grid_n < 2(Jmax+1)");
    SphPair& sp = b->hp[p];
    sp.natoms = N;
    sp.atom_off = (int)Ntot;
    sp.L = Lp;
    sp.ngroups = d->ngroups[p];
    sp.group_off = gq;
    sp.gath_off = ga;
    sp.scale = d->scale ? d->scale[p] : 0.0;
    sp.pad0 = 0;
    std::vector<char> seen(N, 0);
    long long jk = 0;
    int ng = 0;
    for (int q = 0; q < sp.ngroups; ++q) {
      const int sz = d->group_sizes[gq + q];
      if (sz > kMaxGroup)
        return set_err(FO_ERR_UNSUPPORTED, "permutation group larger than 4096 atoms");
      SphGroup& G = b->hg[gq + q];
      G.size = sz;
      G.gath_off = ng;
      G.jk_off = jk;
      G.rad_off = (long long)rad_slots(Lp) * jk;
      for (int i = 0; i < sz; ++i) {
        const int a = d->group_atoms[ga + ng + i];
        if (a < 0 || a >= N) return set_err(FO_ERR_INVALID, "group atom index out of range");
        if (seen[a]) return set_err(FO_ERR_INVALID, "atom listed in two permutation groups");
        seen[a] = 1;
        b->hgath[ga + ng + i] = a;
      }
      jk += (long long)sz * sz;
      ng += sz;
      b->maxgroup = std::max(b->maxgroup, sz);
    }
    sp.ngath = ng;
    sp.coef_off = Ktot;
    sp.c32_off = C32tot;
    Ktot += kcount(Lp);
    C32tot += plane_rows(Lp) * n;
    // coefficient chunks: K and Y offsets are relative to the pair's chunk
    // radial block: K_l[j][k] of every group, or for the harmonic basis
    // D[j][l][n] of both structures then c[group][q(l,m)][n] of both
    const long long nn = harm_nmax + 1;
    const long long lx = harm_nmax >= 0 ? harm_block_lx(Lp, harm_nmax, b->harm_basis) : 0;
    const long long rs = harm_nmax >= 0
                             ? 2 * ng * (lx + 1) * nn + 4 * (long long)sp.ngroups * plane_rows(Lp) * nn
                             : (long long)rad_slots(Lp) * jk;
    const long long ys = (long long)(Lp + 1) * (Lp + 2) / 2 * ng;
    b->maxgroups = std::max(b->maxgroups, sp.ngroups);
    const size_t cb = sizeof(double) * (size_t)(radtot + rs) + 2 * sizeof(double2) * (size_t)(ylmtot + ys);
    if (b->chunks.empty() || (cb > chunk_budget && radtot + ylmtot > 0)) {
      b->chunks.push_back(CoefChunk{p, p, 0, 0});
      radtot = ylmtot = 0;
    }
    CoefChunk& ch = b->chunks.back();
    ch.p1 = p + 1;
    ch.Lmax = std::max(ch.Lmax, Lp);
    ch.maxjk = std::max(ch.maxjk, jk);
    sp.rad_off = radtot;
    sp.ylm_off = ylmtot;
    radtot += rs;
    ylmtot += ys;
    radmax = std::max(radmax, radtot);
    ylmmax = std::max(ylmmax, ylmtot);
    Lmax = std::max(Lmax, Lp);
    maxatoms = std::max(maxatoms, N);
    maxgath = std::max(maxgath, ng);
    maxjk = std::max(maxjk, jk);
    Ntot += N;
    gq += sp.ngroups;
    ga += ng;
  }
  if (isoft_plane_smem(n, Lmax, 0) > c->max_lds)
    return set_err(FO_ERR_UNSUPPORTED, "grid_n and Jmax too large for one beta plane in LDS (" +
                                           std::to_string(isoft_plane_smem(n, Lmax, 0)) + " > " +
                                           std::to_string(c->max_lds) + " bytes)");
  b->fft_chunk = (d->flags & FO_SPH_DFT) ? 0 : isoft_plane_use_fft(n, Lmax, c->max_lds);
  const bool beta_gemm = (d->flags & FO_SPH_BETA_GEMM) != 0;
  if (beta_gemm && (general || !b->fft_chunk))
    return set_err(FO_ERR_UNSUPPORTED, "FO_SPH_BETA_GEMM needs the alignment mode and a "
                                       "power-of-two grid side in [16, 256]");
  if (lap_lds_bytes(b->maxgroup) > c->max_lds)
    return set_err(FO_ERR_UNSUPPORTED, "permutation group too large for the LDS assignment solver");
  int rc = get_sph_table(c, n, &b->tab, build_tables);
  if (rc) return rc;
  rc = get_rec_table(c, &b->rect, build_tables);
  if (rc) return rc;
  if (Lmax > b->tab->Lt) return set_err(FO_ERR_UNSUPPORTED, "Jmax exceeds the table degree");
  SphLaunch& lc = b->lc;
  lc.P = P;
  lc.n = n;
  lc.Lmax = Lmax;
  lc.Lt = b->tab->Lt;
  lc.maxgath = maxgath;
  lc.maxatoms = maxatoms;
  lc.maxjk = maxjk;
  lc.norient = (d->flags & FO_SPH_NO_INVERT) ? 1 : 2;
  lc.mode = general ? 1 : 0;
  if (general) lc.norient = 2;
  lc.keep_grid = (d->flags & FO_SPH_KEEP_GRID) ? 1 : 0;
  // the tuned shapes belong to the Gaussian-overlap coefficient kernels
  b->autotune = (d->flags & FO_SPH_AUTOTUNE) != 0 && harm_nmax < 0;
  lc.center = (d->flags & FO_SPH_NO_CENTER) ? 0 : 1;
  b->Ntot = Ntot;
  b->Ktot = Ktot;
  b->C32tot = C32tot;
  const int no = lc.norient;
  b->scr_stride = (long long)(Lmax + 1) * (2 * Lmax + 1) + (long long)(Lmax + 1) * n;
  b->gsum = gsum;
  Layout lay;
  auto& o = b->off;
  o.pd = lay.add<SphPair>(P);
  o.grp = lay.add<SphGroup>(std::max(gsum, 1));
  o.gath = lay.add<int>(b->hgath.size());
  o.pos1 = lay.add<double>(3 * Ntot);
  o.pos2 = lay.add<double>(3 * Ntot);
  o.X1c = lay.add<double>(3 * Ntot);
  o.X2c = lay.add<double>(3 * Ntot);
  o.X2R = lay.add<double>(3 * Ntot);
  o.X2out = lay.add<double>(3 * Ntot);
  o.scale = lay.add<double>(P);
  o.R = lay.add<double>(std::max(radmax, 1LL));
  o.Y1 = lay.add<double2>(std::max(ylmmax, 1LL));
  o.Y2 = lay.add<double2>(std::max(ylmmax, 1LL));
  o.rad = lay.add<double>(2 * b->hgath.size());
  o.I64A = lay.add<double2>(Ktot);
  o.I32A = lay.add<float2>(C32tot);
  o.I64B = general ? lay.add<double2>(Ktot) : 0;
  o.I32B = general ? lay.add<float2>(C32tot) : 0;
  o.cand = lay.add<PlaneCand>((size_t)P * no * n);
  o.abound = lay.add<double>((size_t)P * no);
  o.peaks = lay.add<SphPeak>((size_t)P * no);
  o.opt = lay.add<SphOpt>((size_t)P * no);
  o.res = lay.add<SphResult>(P);
  o.perm = lay.add<int>(Ntot);
  o.rowsol = lay.add<int>(Ntot + 1);
  o.scr = lay.add<double2>((size_t)P * no * b->scr_stride);
  o.grid = lc.keep_grid ? lay.add<float>((size_t)P * no * n * n * n) : 0;
  o.S = beta_gemm ? lay.add<float4>((size_t)P * n * (Lmax + 1) * n) : 0;
  o.gtab = harm_nmax >= 0 && !b->harm_basis ? lay.add<double>((size_t)(Lmax + 1) * (harm_nmax + 1) * (harm_nmax + 1)) : 0;
  b->bytes = lay.off;
  *out = b.release();
  return FO_OK;
}

size_t sph_batch_bytes(const fo_sph_batch* b) { return b->bytes; }

void sph_batch_outputs(const fo_sph_batch* b, SphOutputs* o) {
  o->res = b->res;
  o->X2 = b->X2out;
  o->perm = b->perm;
  o->natoms = b->Ntot;
  o->npairs = b->lc.P;
}

int sph_batch_bind(fo_sph_batch* b, void* mem, hipStream_t stream) {
  fo_context* c = b->ctx;
  if (mem) {
    b->base = mem;
    b->own_mem = false;
  } else {
    int rc = ctx_alloc(c, b->bytes, &b->base);
    if (rc) return rc;
    b->own_mem = true;
  }
  const bool general = b->lc.mode == 1;
  const auto& o = b->off;
  char* B = (char*)b->base;
  b->pd = (SphPair*)(B + o.pd);
  b->grp = (SphGroup*)(B + o.grp);
  b->gath = (int*)(B + o.gath);
  b->pos1 = (double*)(B + o.pos1);
  b->pos2 = (double*)(B + o.pos2);
  b->X1c = (double*)(B + o.X1c);
  b->X2c = (double*)(B + o.X2c);
  b->X2R = (double*)(B + o.X2R);
  b->X2out = (double*)(B + o.X2out);
  b->scale = (double*)(B + o.scale);
  b->R = (double*)(B + o.R);
  b->Y1 = (double2*)(B + o.Y1);
  b->Y2 = (double2*)(B + o.Y2);
  b->rad = (double*)(B + o.rad);
  b->I64A = (double2*)(B + o.I64A);
  b->I32A = (float2*)(B + o.I32A);
  b->I64B = general ? (double2*)(B + o.I64B) : nullptr;
  b->I32B = general ? (float2*)(B + o.I32B) : nullptr;
  b->cand = (PlaneCand*)(B + o.cand);
  b->abound = (double*)(B + o.abound);
  b->peaks = (SphPeak*)(B + o.peaks);
  b->opt = (SphOpt*)(B + o.opt);
  b->res = (SphResult*)(B + o.res);
  b->perm = (int*)(B + o.perm);
  b->rowsol = (int*)(B + o.rowsol);
  b->scr = (double2*)(B + o.scr);
  b->grid = b->lc.keep_grid ? (float*)(B + o.grid) : nullptr;
  b->S = o.S ? (float4*)(B + o.S) : nullptr;
  b->gtab = b->harm_nmax >= 0 && !b->harm_basis ? (double*)(B + o.gtab) : nullptr;
  if (stream) {
    b->stream = stream;
    b->own_stream = false;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return set_err(FO_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    b->own_stream = true;
  }
  FO_HIP(hipMemcpyAsync(b->pd, b->hp.data(), sizeof(SphPair) * b->lc.P, hipMemcpyHostToDevice,
                        b->stream));
  // plane-layout entries |m2| > l are never written by the kernels: zero once
  FO_HIP(hipMemsetAsync(b->I32A, 0, sizeof(float2) * (size_t)b->C32tot, b->stream));
  if (general)
    FO_HIP(hipMemsetAsync(b->I32B, 0, sizeof(float2) * (size_t)b->C32tot, b->stream));
  if (b->gsum)
    FO_HIP(hipMemcpyAsync(b->grp, b->hg.data(), sizeof(SphGroup) * b->gsum, hipMemcpyHostToDevice,
                          b->stream));
  FO_HIP(hipMemcpyAsync(b->gath, b->hgath.data(), sizeof(int) * b->hgath.size(),
                        hipMemcpyHostToDevice, b->stream));
  if (b->gtab) {
    b->hgtab = harmonic_basis_table(b->lc.Lmax, b->harm_nmax, b->harm_r0);
    FO_HIP(hipMemcpyAsync(b->gtab, b->hgtab.data(), sizeof(double) * b->hgtab.size(),
                          hipMemcpyHostToDevice, b->stream));
  }
  return FO_OK;
}

}  // namespace fo

extern "C" {

int fo_sph_batch_create(fo_context* c, const fo_sph_batch_desc* d, fo_sph_batch** out) {
  if (!out) return set_err(FO_ERR_INVALID, "out is NULL");
  fo_sph_batch* b = nullptr;
  int rc = sph_batch_plan(c, d, &b);
  if (rc) return rc;
  rc = sph_batch_bind(b, nullptr, nullptr);
  if (rc == FO_OK && hipStreamSynchronize(b->stream) != hipSuccess)
    rc = set_err(FO_ERR_HIP, "descriptor upload failed");
  if (rc) {
    fo_sph_batch_destroy(b);
    return rc;
  }
  *out = b;
  return FO_OK;
}

int fo_sph_batch_create_harmonic(fo_context* c, const fo_sph_batch_desc* d, int nmax,
                                 double harmscale, int basis, fo_sph_batch** out) {
  if (!out) return set_err(FO_ERR_INVALID, "out is NULL");
  if (nmax < 0) return set_err(FO_ERR_INVALID, "nmax must be >= 0");
  if (basis != 0 && basis != 1) return set_err(FO_ERR_INVALID, "basis must be 0 or 1");
  fo_sph_batch* b = nullptr;
  int rc = sph_batch_plan(c, d, &b, true, nmax, harmscale, basis);
  if (rc) return rc;
  rc = sph_batch_bind(b, nullptr, nullptr);
  if (rc == FO_OK && hipStreamSynchronize(b->stream) != hipSuccess)
    rc = set_err(FO_ERR_HIP, "descriptor upload failed");
  if (rc) {
    fo_sph_batch_destroy(b);
    return rc;
  }
  *out = b;
  return FO_OK;
}

int fo_sph_batch_harmonic_coeffs(fo_sph_batch* b, int pair, double* c1, double* c2) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  if (b->harm_nmax < 0) return set_err(FO_ERR_STATE, "not a harmonic-basis batch");
  if (pair < 0 || pair >= b->lc.P) return set_err(FO_ERR_INVALID, "pair out of range");
  if (b->chunks.size() != 1)
    return set_err(FO_ERR_STATE, "coefficients are kept only for a single-chunk batch");
  const SphPair& d = b->hp[pair];
  const size_t nn = b->harm_nmax + 1;
  const size_t one = (size_t)d.ngroups * plane_rows(d.L) * nn;
  const size_t lx = harm_block_lx(d.L, b->harm_nmax, b->harm_basis);
  const double* base = b->R + d.rad_off + 2 * (size_t)d.ngath * (lx + 1) * nn;
  FO_HIP(hipSetDevice(b->ctx->device));
  if (c1) FO_HIP(hipMemcpyAsync(c1, base, sizeof(double2) * one, hipMemcpyDeviceToHost, b->stream));
  if (c2)
    FO_HIP(hipMemcpyAsync(c2, base + 2 * one, sizeof(double2) * one, hipMemcpyDeviceToHost,
                          b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_sph_batch_coef_norm2(fo_sph_batch* b, double* norm2) {
  if (!b || !norm2) return set_err(FO_ERR_INVALID, "NULL argument");
  if (b->lc.mode != 0) return set_err(FO_ERR_UNSUPPORTED, "one coefficient set per pair only");
  FO_HIP(hipSetDevice(b->ctx->device));
  void* buf = nullptr;
  int rc = ctx_alloc(b->ctx, sizeof(double) * b->lc.P, &buf);
  if (rc) return rc;
  launch_coef_norm2(b->stream, b->lc.P, b->pd, b->I64A, (double*)buf);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(norm2, buf, sizeof(double) * b->lc.P, hipMemcpyDeviceToHost, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  ctx_free(b->ctx, buf, sizeof(double) * b->lc.P);
  if (e != hipSuccess) return set_err(FO_ERR_HIP, hipGetErrorString(e));
  return FO_OK;
}

int fo_sph_batch_destroy(fo_sph_batch* b) {
  if (!b) return FO_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  if (b->ev_made) {
    for (auto& e : b->ev) (void)hipEventDestroy(e);
    for (auto& e : b->cev) (void)hipEventDestroy(e);
  }
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  if (b->own_mem) ctx_free(b->ctx, b->base, b->bytes);
  if (b->pk.base) ctx_free(b->ctx, b->pk.base, b->pk.bytes);
  delete b;
  return FO_OK;
}

int fo_sph_batch_footprint(const fo_sph_batch* b, unsigned long long* bytes) {
  if (!b || !bytes) return set_err(FO_ERR_INVALID, "NULL argument");
  *bytes = b->bytes;
  return FO_OK;
}

int fo_sph_batch_norient(const fo_sph_batch* b) { return b ? b->lc.norient : 0; }

int fo_sph_batch_set_positions(fo_sph_batch* b, const double* p1, const double* p2) {
  if (!b || !p1 || !p2) return set_err(FO_ERR_INVALID, "NULL argument");
  if (!all_finite(p1, 3 * b->Ntot) || !all_finite(p2, 3 * b->Ntot))
    return set_err(FO_ERR_INVALID, "positions must be finite");
  FO_HIP(hipMemcpyAsync(b->pos1, p1, sizeof(double) * 3 * b->Ntot, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipMemcpyAsync(b->pos2, p2, sizeof(double) * 3 * b->Ntot, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_sph_batch_set_positions_device(fo_sph_batch* b, const void* p1, const void* p2) {
  if (!b || !p1 || !p2) return set_err(FO_ERR_INVALID, "NULL argument");
  FO_HIP(hipMemcpyAsync(b->pos1, p1, sizeof(double) * 3 * b->Ntot, hipMemcpyDeviceToDevice, b->stream));
  FO_HIP(hipMemcpyAsync(b->pos2, p2, sizeof(double) * 3 * b->Ntot, hipMemcpyDeviceToDevice, b->stream));
  return FO_OK;
}

int fo_sph_batch_set_coeffs(fo_sph_batch* b, const double* A, const double* Bc) {
  if (!b || !A) return set_err(FO_ERR_INVALID, "NULL argument");
  FO_HIP(hipMemcpyAsync(b->I64A, A, sizeof(double2) * b->Ktot, hipMemcpyHostToDevice, b->stream));
  b->mirror_pending = false;
  launch_coef_to_plane(b->stream, b->lc.P, b->lc.n, b->lc.Lmax, b->pd, b->I64A, b->I32A);
  if (b->lc.mode == 1) {
    if (!Bc) return set_err(FO_ERR_INVALID, "general mode needs the second coefficient set");
    FO_HIP(hipMemcpyAsync(b->I64B, Bc, sizeof(double2) * b->Ktot, hipMemcpyHostToDevice, b->stream));
    launch_coef_to_plane(b->stream, b->lc.P, b->lc.n, b->lc.Lmax, b->pd, b->I64B, b->I32B);
  }
  FO_CHECK_LAUNCH();
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_sph_batch_set_start(fo_sph_batch* b, const double* R0) {
  if (!b || !R0) return set_err(FO_ERR_INVALID, "NULL argument");
  const int no = b->lc.norient;
  std::vector<SphPeak> pk((size_t)b->lc.P * no);
  for (size_t i = 0; i < pk.size(); ++i) {
    memset(&pk[i], 0, sizeof(SphPeak));
    for (int k = 0; k < 3; ++k) pk[i].R0[k] = R0[3 * i + k];
  }
  FO_HIP(hipMemcpyAsync(b->peaks, pk.data(), sizeof(SphPeak) * pk.size(), hipMemcpyHostToDevice,
                        b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_sph_batch_set_rotation(fo_sph_batch* b, const double* R) {
  if (!b || !R) return set_err(FO_ERR_INVALID, "NULL argument");
  const int no = b->lc.norient;
  std::vector<SphOpt> op((size_t)b->lc.P * no);
  for (int p = 0; p < b->lc.P; ++p)
    for (int o = 0; o < no; ++o) {
      SphOpt& x = op[(size_t)p * no + o];
      memset(&x, 0, sizeof(SphOpt));
      for (int k = 0; k < 3; ++k) x.R[k] = R[3 * p + k];
      x.f = o == 0 ? 0.0 : HUGE_VAL;
    }
  FO_HIP(hipMemcpyAsync(b->opt, op.data(), sizeof(SphOpt) * op.size(), hipMemcpyHostToDevice,
                        b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

// The project's autotuner.  On the first full run of an FO_SPH_AUTOTUNE
// batch, each launch shape is chosen in turn (coordinate descent, the others
// held at their current choice) by timing its kernel over the batch's own
// data with HIP events on the batch stream: every candidate runs once to
// warm up and three times timed, the fastest mean wins.  Every candidate
// computes the same results (only the workgroup-to-work mapping and the
// tile shapes differ, not any reduction order), and the stages rewrite their
// outputs from unchanged inputs, so tuning leaves the batch's data as one
// eager run would.  The choice is cached in the context per (grid side,
// largest degree, pair-count bucket, plane path) and reused by later batches.
static int sph_autotune(fo_sph_batch* b) {
  fo_context* c = b->ctx;
  const SphLaunch& L = b->lc;
  int bucket = 1;
  while (bucket < L.P) bucket <<= 1;
  const auto key = std::make_tuple(L.n, L.Lmax, bucket, b->fft_chunk);
  {
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->tuned.find(key);
    if (it != c->tuned.end()) {
      b->tune = it->second;
      b->tuned = true;
      return FO_OK;
    }
  }
  hipStream_t s = b->stream;
  // inputs of every stage (and a warm start)
  int rc = sph_enqueue(b, FO_STAGE_PREP | FO_STAGE_COEFF | FO_STAGE_GRID);
  if (rc) return rc;
  hipEvent_t e0, e1;
  FO_HIP(hipEventCreate(&e0));
  FO_HIP(hipEventCreate(&e1));
  auto time_stage = [&](int knob, float* ms) -> int {
    auto launch = [&]() {
      if (knob == 0) {
        launch_isoft_plane(s, L.P, L.n, L.Lmax, b->fft_chunk, b->tab->Kp, b->pd, b->tab->tab32,
                           b->tab->tw32, b->I32A, b->I32B, L.mode, L.norient, b->cand, b->grid,
                           b->abound, b->tune[0], b->S, L.Lmax + 1);
      } else {
        for (const CoefChunk& C : b->chunks)
          launch_sph_radial(s, C.p1 - C.p0, C.Lmax, C.maxjk, b->pd + C.p0, b->grp, b->rad,
                            b->rad + b->hgath.size(), b->scale + C.p0, b->R, b->tune[2]);
      }
    };
    launch();
    FO_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 3; ++r) launch();
    FO_HIP(hipEventRecord(e1, s));
    FO_CHECK_LAUNCH();
    FO_HIP(hipEventSynchronize(e1));
    FO_HIP(hipEventElapsedTime(ms, e0, e1));
    return FO_OK;
  };
  // knob 1 is reserved (it chose the contraction's tile class below degree
  // 16, where one fused kernel now runs).  The first candidate of each knob
  // is the default shape; another replaces the best so far only when its
  // stage runs at least kTuneMargin faster, so timing noise (a few percent
  // between launches) cannot trade the default for an equal or slower
  // shape.
  const std::vector<int> cand[3] = {{16, 1, 2, 4, 8, 32}, {}, {1024, 64, 256, 4096}};
  constexpr float kTuneMargin = 0.03f;
  for (int knob = 0; knob < 3; ++knob) {
    if (cand[knob].empty() || (knob == 0 && !(L.P > 8))) continue;  // one chunk either way
    int best = b->tune[knob];
    float best_ms = 0.f;
    bool first = true;
    for (int v : cand[knob]) {
      b->tune[knob] = v;
      float ms = 0.f;
      rc = time_stage(knob, &ms);
      if (rc) break;
      if (first || ms < best_ms * (1.f - kTuneMargin)) {
        best = v;
        best_ms = ms;
        first = false;
      }
    }
    b->tune[knob] = best;
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  // the data were rewritten with other shapes: the coefficient rows m1 < 0
  // are refreshed on read-out as after any coefficient stage
  b->mirror_pending = true;
  b->tuned = true;
  std::lock_guard<std::mutex> g(c->mu);
  c->tuned[key] = b->tune;
  return FO_OK;
}

int fo_sph_batch_run(fo_sph_batch* b, int stages, int use_graph) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  FO_HIP(hipSetDevice(b->ctx->device));
  const int full = FO_STAGE_PREP | FO_STAGE_COEFF | FO_STAGE_GRID;
  if (b->autotune && !b->tuned && (stages & full) == full) {
    int rc = sph_autotune(b);
    if (rc) return rc;
  }
  if (stages & FO_STAGE_COEFF) b->mirror_pending = true;
  if (use_graph && !b->timing) {
    auto it = b->graphs.find(stages);
    if (it == b->graphs.end()) {
      hipGraph_t g;
      FO_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = sph_enqueue(b, stages);
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) return rc;
      if (e != hipSuccess) return set_err(FO_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
      hipGraphExec_t ex;
      FO_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
      it = b->graphs.emplace(stages, ex).first;
    }
    FO_HIP(hipGraphLaunch(it->second, b->stream));
    return FO_OK;
  }
  return sph_enqueue(b, stages);
}

int fo_sph_batch_sync(fo_sph_batch* b) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_sph_batch_results(fo_sph_batch* b, double* dist, double* X1, double* X2, int* perm,
                         double* R, double* rmat, int* inverted) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  const int P = b->lc.P;
  std::vector<SphResult> r(P);
  FO_HIP(hipMemcpyAsync(r.data(), b->res, sizeof(SphResult) * P, hipMemcpyDeviceToHost, b->stream));
  if (X1) FO_HIP(hipMemcpyAsync(X1, b->X1c, sizeof(double) * 3 * b->Ntot, hipMemcpyDeviceToHost, b->stream));
  if (X2) FO_HIP(hipMemcpyAsync(X2, b->X2out, sizeof(double) * 3 * b->Ntot, hipMemcpyDeviceToHost, b->stream));
  if (perm) FO_HIP(hipMemcpyAsync(perm, b->perm, sizeof(int) * b->Ntot, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  for (int p = 0; p < P; ++p) {
    if (dist) dist[p] = r[p].dist;
    if (R)
      for (int k = 0; k < 3; ++k) R[3 * p + k] = r[p].R[k];
    if (rmat)
      for (int k = 0; k < 9; ++k) rmat[9 * p + k] = r[p].rmat[k];
    if (inverted) inverted[p] = r[p].inverted;
  }
  return FO_OK;
}

int fo_sph_batch_peaks(fo_sph_batch* b, int* ind, double* fun, double* peak_value, double* margin,
                       double* delta, int* ncand, int* noverflow) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  const size_t m = (size_t)b->lc.P * b->lc.norient;
  std::vector<SphPeak> pk(m);
  std::vector<SphOpt> op(m);
  FO_HIP(hipMemcpyAsync(pk.data(), b->peaks, sizeof(SphPeak) * m, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipMemcpyAsync(op.data(), b->opt, sizeof(SphOpt) * m, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < m; ++i) {
    if (ind)
      for (int k = 0; k < 3; ++k) ind[3 * i + k] = pk[i].ind[k];
    if (fun) fun[i] = op[i].f;
    if (peak_value) peak_value[i] = pk[i].value;
    if (margin) margin[i] = pk[i].margin;
    if (delta) delta[i] = pk[i].delta;
    if (ncand) ncand[i] = pk[i].ncand;
    if (noverflow) noverflow[i] = pk[i].noverflow;
  }
  return FO_OK;
}

int fo_sph_batch_opt(fo_sph_batch* b, double* R, double* fun, double* gnorm, int* iters,
                     int* evals) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  const size_t m = (size_t)b->lc.P * b->lc.norient;
  std::vector<SphOpt> op(m);
  FO_HIP(hipMemcpyAsync(op.data(), b->opt, sizeof(SphOpt) * m, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < m; ++i) {
    if (R)
      for (int k = 0; k < 3; ++k) R[3 * i + k] = op[i].R[k];
    if (fun) fun[i] = op[i].f;
    if (gnorm) gnorm[i] = op[i].gnorm;
    if (iters) iters[i] = op[i].iters;
    if (evals) evals[i] = op[i].evals;
  }
  return FO_OK;
}

int fo_sph_batch_scales(fo_sph_batch* b, double* scale) {
  if (!b || !scale) return set_err(FO_ERR_INVALID, "NULL argument");
  FO_HIP(hipMemcpyAsync(scale, b->scale, sizeof(double) * b->lc.P, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_sph_batch_coeffs(fo_sph_batch* b, double* coef) {
  if (!b || !coef) return set_err(FO_ERR_INVALID, "NULL argument");
  if (b->mirror_pending) {
    launch_sph_mirror(b->stream, b->lc.P, b->lc.Lmax, b->pd, b->I64A);
    FO_CHECK_LAUNCH();
    b->mirror_pending = false;
  }
  FO_HIP(hipMemcpyAsync(coef, b->I64A, sizeof(double2) * b->Ktot, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_sph_batch_grid(fo_sph_batch* b, int pair, float* grid) {
  if (!b || !grid) return set_err(FO_ERR_INVALID, "NULL argument");
  if (!b->grid) return set_err(FO_ERR_STATE, "batch created without FO_SPH_KEEP_GRID");
  if (pair < 0 || pair >= b->lc.P) return set_err(FO_ERR_INVALID, "pair out of range");
  const size_t n3 = (size_t)b->lc.n * b->lc.n * b->lc.n * b->lc.norient;
  FO_HIP(hipMemcpyAsync(grid, b->grid + n3 * pair, sizeof(float) * n3, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_sph_batch_find_peaks(fo_sph_batch* b, int npeaks, int width, double* fit, int* count,
                            float* residual) {
  if (!b || !fit || !count) return set_err(FO_ERR_INVALID, "NULL argument");
  if (!b->grid) return set_err(FO_ERR_STATE, "batch created without FO_SPH_KEEP_GRID");
  FO_HIP(hipSetDevice(b->ctx->device));
  const int n = b->lc.n;
  return run_peak_search(b->ctx, b->stream, b->pk, b->grid, (long long)b->lc.P * b->lc.norient,
                         n, n, n, npeaks, width, fit, count, residual);
}

int fo_sph_batch_tuning(fo_sph_batch* b, int* params, int set) {
  if (!b || !params) return set_err(FO_ERR_INVALID, "NULL argument");
  if (set) {
    if (params[0] < 0 || params[0] > 64 || params[1] < -1 || params[1] > 15 || params[2] < 0)
      return set_err(FO_ERR_INVALID, "launch shape out of range");
    for (int k = 0; k < 3; ++k) b->tune[k] = params[k];
    b->tuned = true;
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
  } else {
    for (int k = 0; k < 3; ++k) params[k] = b->tune[k];
  }
  return FO_OK;
}

int fo_sph_batch_precision(fo_sph_batch* b, double* val32, double* val64, double* delta) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  const size_t m = (size_t)b->lc.P * b->lc.norient;
  std::vector<SphPeak> pk(m);
  FO_HIP(hipMemcpyAsync(pk.data(), b->peaks, sizeof(SphPeak) * m, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < m; ++i) {
    if (val32) val32[i] = pk[i].val32;
    if (val64) val64[i] = pk[i].value;
    if (delta) delta[i] = pk[i].delta;
  }
  return FO_OK;
}

int fo_sph_batch_start(fo_sph_batch* b, double* R0) {
  if (!b || !R0) return set_err(FO_ERR_INVALID, "NULL argument");
  const size_t m = (size_t)b->lc.P * b->lc.norient;
  std::vector<SphPeak> pk(m);
  FO_HIP(hipMemcpyAsync(pk.data(), b->peaks, sizeof(SphPeak) * m, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < m; ++i)
    for (int k = 0; k < 3; ++k) R0[3 * i + k] = pk[i].R0[k];
  return FO_OK;
}

int fo_sph_batch_timing(fo_sph_batch* b, int enable, float* ms8) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  if (enable >= 0) {
    if (enable && !b->ev_made) {
      for (auto& e : b->ev) FO_HIP(hipEventCreate(&e));
      b->cev.resize(3 * b->chunks.size() + 1);
      for (auto& e : b->cev) FO_HIP(hipEventCreate(&e));
      b->ev_made = true;
    }
    b->timing = enable != 0;
  }
  if (ms8 && b->ev_made) {
    FO_HIP(hipStreamSynchronize(b->stream));
    for (int i = 0; i < 8; ++i) FO_HIP(hipEventElapsedTime(&ms8[i], b->ev[i], b->ev[i + 1]));
    // coefficient stage (between events 1 and 4): harmonics / radial /
    // contraction, each summed over the chunks
    float t3[3] = {0.f, 0.f, 0.f};
    if (b->cev_valid)
      for (size_t i = 0; i + 1 < b->cev.size(); ++i) {
        float ms = 0.f;
        FO_HIP(hipEventElapsedTime(&ms, b->cev[i], b->cev[i + 1]));
        t3[i % 3] += ms;
      }
    for (int k = 0; k < 3; ++k) ms8[1 + k] = t3[k];
  }
  return FO_OK;
}

// ---------------------------------------------------------------------------
// single-call helpers
// ---------------------------------------------------------------------------
struct DevScratch {
  fo_context* c;
  void* p = nullptr;
  size_t n = 0;
  explicit DevScratch(fo_context* cc) : c(cc) {}
  ~DevScratch() { ctx_free(c, p, n); }
  int alloc(size_t bytes) {
    n = bytes;
    return ctx_alloc(c, bytes, &p);
  }
};

int fo_sph_harm(fo_context* c, int L, int npts, const double* theta, const double* phi, double* Y) {
  if (!c || !theta || !phi || !Y || L < 0 || npts < 0) return set_err(FO_ERR_INVALID, "bad argument");
  if (npts == 0) return FO_OK;
  std::lock_guard<std::mutex> g(c->call_mu);
  FO_HIP(hipSetDevice(c->device));
  const size_t W = 2 * L + 1, ny = (size_t)(L + 1) * W * npts;
  DevScratch s(c);
  int rc = s.alloc(sizeof(double) * 2 * npts + sizeof(double2) * ny + 512);
  if (rc) return rc;
  double* dth = (double*)s.p;
  double* dph = dth + npts;
  double2* dY = (double2*)((char*)s.p + ((sizeof(double) * 2 * npts + 255) / 256 * 256));
  FO_HIP(hipMemcpyAsync(dth, theta, sizeof(double) * npts, hipMemcpyHostToDevice, c->stream));
  FO_HIP(hipMemcpyAsync(dph, phi, sizeof(double) * npts, hipMemcpyHostToDevice, c->stream));
  launch_ylm_angles(c->stream, L, npts, dth, dph, dY);
  FO_CHECK_LAUNCH();
  FO_HIP(hipMemcpyAsync(Y, dY, sizeof(double2) * ny, hipMemcpyDeviceToHost, c->stream));
  FO_HIP(hipStreamSynchronize(c->stream));
  return FO_OK;
}

static void dense_to_compact_conj(int L, const double* X, std::vector<double>& out) {
  const int W = 2 * L + 1;
  out.assign(2 * (size_t)kcount(L), 0.0);
  for (int l = 0; l <= L; ++l)
    for (int m1 = -l; m1 <= l; ++m1)
      for (int m2 = -l; m2 <= l; ++m2) {
        const size_t dd = ((size_t)l * W + (m1 + W) % W) * W + (m2 + W) % W;
        const size_t k = (size_t)kidx(l, m1, m2);
        out[2 * k] = X[2 * dd];
        out[2 * k + 1] = -X[2 * dd + 1];
      }
}

int fo_sph_energy(fo_context* c, int L, const double* X, const double* rot, double* f, double* grad) {
  if (!c || !X || !rot || L < 0 || L > kMaxL) return set_err(FO_ERR_INVALID, "bad argument");
  const RecTable* rt = nullptr;
  int rc0 = get_rec_table(c, &rt);
  if (rc0) return rc0;
  std::lock_guard<std::mutex> g(c->call_mu);
  FO_HIP(hipSetDevice(c->device));
  std::vector<double> h;
  dense_to_compact_conj(L, X, h);
  DevScratch s(c);
  int rc = s.alloc(sizeof(double) * h.size() + 512);
  if (rc) return rc;
  double2* dI = (double2*)s.p;
  double* dout = (double*)((char*)s.p + (sizeof(double) * h.size() + 255) / 256 * 256);
  FO_HIP(hipMemcpyAsync(dI, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream));
  launch_sph_energy(c->stream, L, dI, rot[0], rot[1], rot[2], rt->rec, rt->seedu, kRecL, dout);
  FO_CHECK_LAUNCH();
  double o[10];
  FO_HIP(hipMemcpyAsync(o, dout, sizeof(o), hipMemcpyDeviceToHost, c->stream));
  FO_HIP(hipStreamSynchronize(c->stream));
  if (f) *f = o[0];
  if (grad) {
    grad[0] = o[1];
    grad[1] = o[2];
    grad[2] = o[3];
  }
  return FO_OK;
}

int fo_sph_wigner_D(fo_context* c, int L, const double* rot, double* D, double* grad) {
  if (!c || !rot || !D || L < 0 || L > kMaxL) return set_err(FO_ERR_INVALID, "bad argument");
  std::lock_guard<std::mutex> g(c->call_mu);
  FO_HIP(hipSetDevice(c->device));
  const size_t W = 2 * L + 1, m = (size_t)(L + 1) * W * W;
  DevScratch s(c);
  int rc = s.alloc(sizeof(double2) * 4 * m);
  if (rc) return rc;
  double2* dD = (double2*)s.p;
  double2* dG = dD + m;
  launch_wigner_D(c->stream, L, rot[0], rot[1], rot[2], dD, dG);
  FO_CHECK_LAUNCH();
  FO_HIP(hipMemcpyAsync(D, dD, sizeof(double2) * m, hipMemcpyDeviceToHost, c->stream));
  if (grad) FO_HIP(hipMemcpyAsync(grad, dG, sizeof(double2) * 3 * m, hipMemcpyDeviceToHost, c->stream));
  FO_HIP(hipStreamSynchronize(c->stream));
  return FO_OK;
}

int fo_sph_wigner_table(fo_context* c, int bw, double* Ds) {
  if (!c || !Ds || bw < 1 || bw - 1 > kMaxL) return set_err(FO_ERR_INVALID, "bad argument");
  const SphTable* t = nullptr;
  int rc = get_sph_table(c, 2 * bw, &t);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(c->call_mu);
  const int n = 2 * bw, J = bw - 1, W = 2 * bw - 1;
  std::vector<double> h((size_t)t->Kt * n);
  FO_HIP(hipMemcpyAsync(h.data(), t->tab64, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  FO_HIP(hipStreamSynchronize(c->stream));
  std::fill(Ds, Ds + (size_t)bw * W * W * n, 0.0);
  for (int l = 0; l <= J; ++l)
    for (int m1 = -l; m1 <= l; ++m1)
      for (int m2 = -l; m2 <= l; ++m2)
        for (int k = 0; k < n; ++k)
          Ds[(((size_t)l * W + (m1 + W) % W) * W + (m2 + W) % W) * n + k] =
              h[(size_t)k * t->Kt + kidx(l, m1, m2)];
  return FO_OK;
}

int fo_soft_forward(fo_context* c, int bw, const double* data, double* flmm) {
  if (!c || !data || !flmm || bw < 1 || bw - 1 > kMaxL)
    return set_err(FO_ERR_INVALID, "bad argument");
  const int n = 2 * bw, L = bw - 1, W = 2 * L + 1;
  if (soft_forward_smem(n) > c->max_lds)
    return set_err(FO_ERR_UNSUPPORTED, "bandwidth too large for the forward transform");
  const SphTable* t = nullptr;
  int rc = get_sph_table(c, n, &t);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(c->call_mu);
  FO_HIP(hipSetDevice(c->device));
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t n3 = (size_t)n * n * n, K = (size_t)kcount(L);
  const size_t bg = up(sizeof(double2) * n3), bs = up(sizeof(double2) * (size_t)W * n * W);
  const size_t bk = up(sizeof(double2) * K), bw8 = up(sizeof(double) * n);
  void* base = nullptr;
  rc = ctx_alloc(c, bg + bs + bk + bw8, &base);
  if (rc) return rc;
  char* p = (char*)base;
  double2* dg = (double2*)p;
  double2* S2 = (double2*)(p + bg);
  double2* df = (double2*)(p + bg + bs);
  double* w = (double*)(p + bg + bs + bk);
  hipError_t e = hipMemcpyAsync(dg, data, sizeof(double2) * n3, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    launch_soft_forward(c->stream, n, t->Kt, t->tab64, t->tw64, dg, w, S2, df);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(flmm, df, sizeof(double2) * K, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  ctx_free(c, base, bg + bs + bk + bw8);
  if (e != hipSuccess) return set_err(FO_ERR_HIP, std::string("forward SOFT: ") + hipGetErrorString(e));
  return FO_OK;
}

int fo_lap_align(fo_context* c, int N, const double* X1, const double* X2, int ngroups,
                 const int* gsizes, const int* gatoms, int metric, const double* box, int* perm,
                 double* dist) {
  if (!c || !X1 || !X2 || N < 1 || ngroups < 0) return set_err(FO_ERR_INVALID, "bad argument");
  if (metric == 1 && !box) return set_err(FO_ERR_INVALID, "periodic metric needs box");
  std::lock_guard<std::mutex> g(c->call_mu);
  FO_HIP(hipSetDevice(c->device));
  std::vector<int> goff(ngroups + 1, 0);
  int gmax = 1;
  for (int q = 0; q < ngroups; ++q) {
    goff[q + 1] = goff[q] + gsizes[q];
    gmax = std::max(gmax, gsizes[q]);
  }
  if (gmax > kMaxGroup || lap_lds_bytes(gmax) > c->max_lds)
    return set_err(FO_ERR_UNSUPPORTED, "permutation group too large");
  if (!all_finite(X1, 3 * (size_t)N) || !all_finite(X2, 3 * (size_t)N))
    return set_err(FO_ERR_INVALID, "positions must be finite");
  const int ga = goff[ngroups];
  for (int i = 0; i < ga; ++i)
    if (gatoms[i] < 0 || gatoms[i] >= N) return set_err(FO_ERR_INVALID, "group atom out of range");
  Layout lay;
  size_t o1 = lay.add<double>(3 * N), o2 = lay.add<double>(3 * N), og = lay.add<int>(std::max(ngroups, 1)),
         oo = lay.add<int>(ngroups + 1), oa = lay.add<int>(std::max(ga, 1)), op = lay.add<int>(N),
         orr = lay.add<int>(N + 1), od = lay.add<double>(1);
  DevScratch s(c);
  int rc = s.alloc(lay.off);
  if (rc) return rc;
  char* B = (char*)s.p;
  FO_HIP(hipMemcpyAsync(B + o1, X1, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
  FO_HIP(hipMemcpyAsync(B + o2, X2, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
  if (ngroups) FO_HIP(hipMemcpyAsync(B + og, gsizes, sizeof(int) * ngroups, hipMemcpyHostToDevice, c->stream));
  FO_HIP(hipMemcpyAsync(B + oo, goff.data(), sizeof(int) * (ngroups + 1), hipMemcpyHostToDevice, c->stream));
  if (ga) FO_HIP(hipMemcpyAsync(B + oa, gatoms, sizeof(int) * ga, hipMemcpyHostToDevice, c->stream));
  launch_lap_align(c->stream, N, (double*)(B + o1), (double*)(B + o2), ngroups, (int*)(B + og),
                   (int*)(B + oo), (int*)(B + oa), metric, box, gmax, (int*)(B + op),
                   (int*)(B + orr), (double*)(B + od));
  FO_CHECK_LAUNCH();
  if (perm) FO_HIP(hipMemcpyAsync(perm, B + op, sizeof(int) * N, hipMemcpyDeviceToHost, c->stream));
  double dd = 0.0;
  FO_HIP(hipMemcpyAsync(&dd, B + od, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FO_HIP(hipStreamSynchronize(c->stream));
  if (dist) *dist = dd;
  return FO_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// periodic batch
// ---------------------------------------------------------------------------
struct fo_per_batch {
  fo_context* ctx = nullptr;
  hipStream_t stream = nullptr;
  const PerTw* tw = nullptr;
  int P = 0, nk = 0, F = 0, niter = 10, flags = 0, maxgroup = 1, gsum = 0;
  long long Ntot = 0, cube = 0;
  std::vector<PerPair> hp;
  std::vector<PerGroup> hg;
  std::vector<int> hatoms;
  size_t bytes = 0;
  void* base = nullptr;
  PerPair* pd = nullptr;
  PerGroup* grp = nullptr;
  int* atoms = nullptr;
  double *pos1 = nullptr, *pos2 = nullptr, *X1out = nullptr, *X2out = nullptr;
  double2 *Cprod = nullptr, *C1 = nullptr, *C2 = nullptr, *scr = nullptr;
  PlaneCand* cand = nullptr;
  PerPeak* peaks = nullptr;
  PerResult* res = nullptr;
  int *perm = nullptr, *saveperm = nullptr, *rowsol = nullptr;
  float* grid = nullptr;
  long long scr_stride = 0;
  std::map<int, hipGraphExec_t> graphs;
  bool timing = false;
  hipEvent_t ev[5] = {};
  bool ev_made = false;
  struct {
    size_t pd, grp, at, p1, p2, x1, x2, cp, c1, c2, cand, pk, res, perm, sp, rs, scr, grid, cgrid;
  } off{};
  bool keepc = false;
  bool own_mem = false, own_stream = false;
  float2* cgrid = nullptr;   // complex f (FO_PER_KEEP_FFT)
  // per-structure coefficients shared by the pairs (set_coeffs_indexed)
  void* csbuf = nullptr;
  size_t csbytes = 0;
  double2* Cs = nullptr;
  int *idx1 = nullptr, *idx2 = nullptr;
  PeakScratch pk;
};

static int per_enqueue(fo_per_batch* b, int stages) {
  hipStream_t s = b->stream;
  const bool tm = b->timing;
  auto mark = [&](int i) {
    if (tm) (void)hipEventRecord(b->ev[i], s);
  };
  mark(0);
  if (stages & FO_PSTAGE_SF) {
    if ((b->flags & FO_PER_INPUT_COEFFS) && b->Cs)
      launch_per_product_indexed(s, b->P, b->nk, b->pd, b->Cs, b->idx1, b->idx2, b->Cprod);
    else if (b->flags & FO_PER_INPUT_COEFFS)
      launch_per_product(s, b->P, b->nk, b->pd, b->C1, b->C2, b->Cprod);
    else
      launch_per_sf(s, b->P, b->nk, b->pd, b->grp, b->atoms, b->pos1, b->pos2, b->Cprod,
                    (b->flags & FO_PER_KEEP_COEFFS) ? b->C1 : nullptr,
                    (b->flags & FO_PER_KEEP_COEFFS) ? b->C2 : nullptr);
  }
  mark(1);
  if (stages & FO_PSTAGE_GRID) {
    launch_per_plane(s, b->P, b->nk, b->F, b->pd, b->Cprod, b->tw->tw32, b->cand, b->grid,
                     b->cgrid);
    mark(2);
    launch_per_select(s, b->P, b->nk, b->F, b->pd, b->Cprod, b->tw->tw64, b->cand, b->scr,
                      b->scr_stride, b->peaks);
  } else {
    mark(2);
  }
  mark(3);
  if (stages & FO_PSTAGE_REFINE)
    launch_per_refine(s, b->P, b->maxgroup, b->pd, b->grp, b->atoms, b->pos1, b->pos2, b->peaks,
                      b->niter, b->perm, b->saveperm, b->rowsol, b->X1out, b->X2out, b->res);
  mark(4);
  FO_CHECK_LAUNCH();
  return FO_OK;
}

namespace fo {

int per_batch_plan(fo_context* c, const fo_per_batch_desc* d, fo_per_batch** out) {
  if (!c || !d || !out) return set_err(FO_ERR_INVALID, "NULL argument");
  if (d->npairs < 1) return set_err(FO_ERR_INVALID, "npairs must be >= 1");
  if (d->nk < 0 || d->nk > 31) return set_err(FO_ERR_UNSUPPORTED, "nk must be in [0, 31]");
  if (d->F < 2 * d->nk + 1 || d->F > 512) return set_err(FO_ERR_UNSUPPORTED, "F must be in [2nk+1, 512]");
  if (!d->natoms || !d->box || !d->scale || !d->ngroups)
    return set_err(FO_ERR_INVALID, "natoms, box, scale and ngroups are required");
  auto b = std::make_unique<fo_per_batch>();
  b->ctx = c;
  b->P = d->npairs;
  b->nk = d->nk;
  b->F = d->F;
  b->niter = d->niter;
  b->flags = d->flags;
  const int NB = 2 * d->nk + 1;
  b->cube = (long long)NB * NB * NB;
  int gsum = 0, gatot = 0;
  for (int p = 0; p < b->P; ++p) gsum += d->ngroups[p];
  for (int q = 0; q < gsum; ++q) gatot += d->group_sizes[q];
  b->gsum = gsum;
  b->hp.resize(b->P);
  b->hg.resize(std::max(gsum, 1));
  b->hatoms.resize(std::max(gatot, 1));
  long long Ntot = 0;
  int gq = 0, ga = 0;
  for (int p = 0; p < b->P; ++p) {
    PerPair& pp = b->hp[p];
    const int N = d->natoms[p];
    if (N < 1) return set_err(FO_ERR_INVALID, "natoms < 1");
    pp.natoms = N;
    pp.atom_off = (int)Ntot;
    pp.ngroups = d->ngroups[p];
    pp.group_off = gq;
    for (int k = 0; k < 3; ++k) {
      pp.box[k] = d->box[3 * p + k];
      if (!(pp.box[k] > 0)) return set_err(FO_ERR_INVALID, "box lengths must be > 0");
    }
    pp.scale = d->scale[p];
    if (!(pp.scale > 0)) return set_err(FO_ERR_INVALID, "scale must be > 0");
    pp.c_off = (long long)p * b->cube;
    std::vector<char> seen(N, 0);
    for (int q = 0; q < pp.ngroups; ++q) {
      const int sz = d->group_sizes[gq + q];
      if (sz > kMaxGroup) return set_err(FO_ERR_UNSUPPORTED, "permutation group too large");
      b->hg[gq + q].size = sz;
      b->hg[gq + q].atom_list_off = ga;
      for (int i = 0; i < sz; ++i) {
        const int a = d->group_atoms[ga + i];
        if (a < 0 || a >= N) return set_err(FO_ERR_INVALID, "group atom index out of range");
        if (seen[a]) return set_err(FO_ERR_INVALID, "atom listed in two permutation groups");
        seen[a] = 1;
        b->hatoms[ga + i] = a;
      }
      ga += sz;
      b->maxgroup = std::max(b->maxgroup, sz);
    }
    gq += pp.ngroups;
    Ntot += N;
  }
  b->Ntot = Ntot;
  if (per_plane_smem(b->nk, b->F) > c->max_lds || per_sf_smem(b->nk) > c->max_lds)
    return set_err(FO_ERR_UNSUPPORTED, "nk / F too large for LDS");
  if (lap_lds_bytes(b->maxgroup) > c->max_lds)
    return set_err(FO_ERR_UNSUPPORTED, "permutation group too large for the LDS assignment solver");
  int rc = get_per_tw(c, b->F, &b->tw);
  if (rc) return rc;
  b->scr_stride = (long long)NB * NB + (long long)NB * b->F;
  b->keepc = (b->flags & (FO_PER_KEEP_COEFFS | FO_PER_INPUT_COEFFS)) != 0;
  Layout lay;
  auto& o = b->off;
  o.pd = lay.add<PerPair>(b->P);
  o.grp = lay.add<PerGroup>(b->hg.size());
  o.at = lay.add<int>(b->hatoms.size());
  o.p1 = lay.add<double>(3 * Ntot);
  o.p2 = lay.add<double>(3 * Ntot);
  o.x1 = lay.add<double>(3 * Ntot);
  o.x2 = lay.add<double>(3 * Ntot);
  o.cp = lay.add<double2>((size_t)b->P * b->cube);
  o.c1 = b->keepc ? lay.add<double2>((size_t)std::max(gsum, 1) * b->cube) : 0;
  o.c2 = b->keepc ? lay.add<double2>((size_t)std::max(gsum, 1) * b->cube) : 0;
  o.cand = lay.add<PlaneCand>((size_t)b->P * b->F);
  o.pk = lay.add<PerPeak>(b->P);
  o.res = lay.add<PerResult>(b->P);
  o.perm = lay.add<int>(Ntot);
  o.sp = lay.add<int>(Ntot);
  o.rs = lay.add<int>(Ntot + 1);
  o.scr = lay.add<double2>((size_t)b->P * b->scr_stride);
  o.grid = (b->flags & FO_PER_KEEP_GRID) ? lay.add<float>((size_t)b->P * b->F * b->F * b->F) : 0;
  o.cgrid = (b->flags & FO_PER_KEEP_FFT) ? lay.add<float2>((size_t)b->P * b->F * b->F * b->F) : 0;
  b->bytes = lay.off;
  *out = b.release();
  return FO_OK;
}

size_t per_batch_bytes(const fo_per_batch* b) { return b->bytes; }

int per_batch_bind(fo_per_batch* b, void* mem, hipStream_t stream) {
  fo_context* c = b->ctx;
  if (mem) {
    b->base = mem;
    b->own_mem = false;
  } else {
    int rc = ctx_alloc(c, b->bytes, &b->base);
    if (rc) return rc;
    b->own_mem = true;
  }
  const auto& o = b->off;
  char* B = (char*)b->base;
  b->pd = (PerPair*)(B + o.pd);
  b->grp = (PerGroup*)(B + o.grp);
  b->atoms = (int*)(B + o.at);
  b->pos1 = (double*)(B + o.p1);
  b->pos2 = (double*)(B + o.p2);
  b->X1out = (double*)(B + o.x1);
  b->X2out = (double*)(B + o.x2);
  b->Cprod = (double2*)(B + o.cp);
  b->C1 = b->keepc ? (double2*)(B + o.c1) : nullptr;
  b->C2 = b->keepc ? (double2*)(B + o.c2) : nullptr;
  b->cand = (PlaneCand*)(B + o.cand);
  b->peaks = (PerPeak*)(B + o.pk);
  b->res = (PerResult*)(B + o.res);
  b->perm = (int*)(B + o.perm);
  b->saveperm = (int*)(B + o.sp);
  b->rowsol = (int*)(B + o.rs);
  b->scr = (double2*)(B + o.scr);
  b->grid = (b->flags & FO_PER_KEEP_GRID) ? (float*)(B + o.grid) : nullptr;
  b->cgrid = (b->flags & FO_PER_KEEP_FFT) ? (float2*)(B + o.cgrid) : nullptr;
  if (stream) {
    b->stream = stream;
    b->own_stream = false;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return set_err(FO_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    b->own_stream = true;
  }
  FO_HIP(hipMemcpyAsync(b->pd, b->hp.data(), sizeof(PerPair) * b->P, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipMemcpyAsync(b->grp, b->hg.data(), sizeof(PerGroup) * b->hg.size(), hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipMemcpyAsync(b->atoms, b->hatoms.data(), sizeof(int) * b->hatoms.size(), hipMemcpyHostToDevice, b->stream));
  return FO_OK;
}

int ctx_alloc_bytes(fo_context* c, size_t bytes, void** p) { return ctx_alloc(c, bytes, p); }
void ctx_free_bytes(fo_context* c, void* p, size_t bytes) { ctx_free(c, p, bytes); }
int ctx_device(const fo_context* c) { return c->device; }
hipStream_t ctx_stream(const fo_context* c) { return c->stream; }
int set_error(int code, const char* msg) { return set_err(code, msg); }

}  // namespace fo

extern "C" {

int fo_per_batch_create(fo_context* c, const fo_per_batch_desc* d, fo_per_batch** out) {
  if (!out) return set_err(FO_ERR_INVALID, "out is NULL");
  fo_per_batch* b = nullptr;
  int rc = per_batch_plan(c, d, &b);
  if (rc) return rc;
  rc = per_batch_bind(b, nullptr, nullptr);
  if (rc == FO_OK && hipStreamSynchronize(b->stream) != hipSuccess)
    rc = set_err(FO_ERR_HIP, "descriptor upload failed");
  if (rc) {
    fo_per_batch_destroy(b);
    return rc;
  }
  *out = b;
  return FO_OK;
}

int fo_per_batch_destroy(fo_per_batch* b) {
  if (!b) return FO_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  if (b->ev_made)
    for (auto& e : b->ev) (void)hipEventDestroy(e);
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  if (b->own_mem) ctx_free(b->ctx, b->base, b->bytes);
  if (b->csbuf) ctx_free(b->ctx, b->csbuf, b->csbytes);
  if (b->pk.base) ctx_free(b->ctx, b->pk.base, b->pk.bytes);
  delete b;
  return FO_OK;
}

int fo_per_batch_footprint(const fo_per_batch* b, unsigned long long* bytes) {
  if (!b || !bytes) return set_err(FO_ERR_INVALID, "NULL argument");
  *bytes = b->bytes;
  return FO_OK;
}

int fo_per_batch_set_positions(fo_per_batch* b, const double* p1, const double* p2) {
  if (!b || !p1 || !p2) return set_err(FO_ERR_INVALID, "NULL argument");
  if (!all_finite(p1, 3 * b->Ntot) || !all_finite(p2, 3 * b->Ntot))
    return set_err(FO_ERR_INVALID, "positions must be finite");
  FO_HIP(hipMemcpyAsync(b->pos1, p1, sizeof(double) * 3 * b->Ntot, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipMemcpyAsync(b->pos2, p2, sizeof(double) * 3 * b->Ntot, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_per_batch_set_positions_device(fo_per_batch* b, const void* p1, const void* p2) {
  if (!b || !p1 || !p2) return set_err(FO_ERR_INVALID, "NULL argument");
  FO_HIP(hipMemcpyAsync(b->pos1, p1, sizeof(double) * 3 * b->Ntot, hipMemcpyDeviceToDevice, b->stream));
  FO_HIP(hipMemcpyAsync(b->pos2, p2, sizeof(double) * 3 * b->Ntot, hipMemcpyDeviceToDevice, b->stream));
  return FO_OK;
}

int fo_per_batch_set_coeffs(fo_per_batch* b, const double* C1, const double* C2) {
  if (!b || !C1 || !C2) return set_err(FO_ERR_INVALID, "NULL argument");
  if (!b->C1) return set_err(FO_ERR_STATE, "batch created without FO_PER_INPUT_COEFFS");
  if (b->Cs) {
    // back to each pair's own coefficients (a captured graph read the shared ones)
    b->Cs = nullptr;
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
  }
  const size_t m = (size_t)b->gsum * b->cube;
  FO_HIP(hipMemcpyAsync(b->C1, C1, sizeof(double2) * m, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipMemcpyAsync(b->C2, C2, sizeof(double2) * m, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_per_batch_set_disp(fo_per_batch* b, const double* disp) {
  if (!b || !disp) return set_err(FO_ERR_INVALID, "NULL argument");
  std::vector<PerPeak> pk(b->P);
  for (int p = 0; p < b->P; ++p) {
    memset(&pk[p], 0, sizeof(PerPeak));
    for (int k = 0; k < 3; ++k) pk[p].disp[k] = disp[3 * p + k];
  }
  FO_HIP(hipMemcpyAsync(b->peaks, pk.data(), sizeof(PerPeak) * b->P, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_per_batch_run(fo_per_batch* b, int stages, int use_graph) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  FO_HIP(hipSetDevice(b->ctx->device));
  if (use_graph && !b->timing) {
    auto it = b->graphs.find(stages);
    if (it == b->graphs.end()) {
      hipGraph_t g;
      FO_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = per_enqueue(b, stages);
      hipError_t e = hipStreamEndCapture(b->stream, &g);
      if (rc) return rc;
      if (e != hipSuccess) return set_err(FO_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
      hipGraphExec_t ex;
      FO_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
      it = b->graphs.emplace(stages, ex).first;
    }
    FO_HIP(hipGraphLaunch(it->second, b->stream));
    return FO_OK;
  }
  return per_enqueue(b, stages);
}

int fo_per_batch_sync(fo_per_batch* b) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_per_batch_results(fo_per_batch* b, double* dist, double* X1, double* X2, int* perm,
                         double* disp) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  std::vector<PerResult> r(b->P);
  FO_HIP(hipMemcpyAsync(r.data(), b->res, sizeof(PerResult) * b->P, hipMemcpyDeviceToHost, b->stream));
  if (X1) FO_HIP(hipMemcpyAsync(X1, b->X1out, sizeof(double) * 3 * b->Ntot, hipMemcpyDeviceToHost, b->stream));
  if (X2) FO_HIP(hipMemcpyAsync(X2, b->X2out, sizeof(double) * 3 * b->Ntot, hipMemcpyDeviceToHost, b->stream));
  if (perm) FO_HIP(hipMemcpyAsync(perm, b->perm, sizeof(int) * b->Ntot, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  for (int p = 0; p < b->P; ++p) {
    if (dist) dist[p] = r[p].dist;
    if (disp)
      for (int k = 0; k < 3; ++k) disp[3 * p + k] = r[p].disp[k];
  }
  return FO_OK;
}

int fo_per_batch_peaks(fo_per_batch* b, int* ind, double* disp0, double* peak_value, double* margin,
                       double* delta, int* ncand, int* noverflow) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  std::vector<PerPeak> pk(b->P);
  FO_HIP(hipMemcpyAsync(pk.data(), b->peaks, sizeof(PerPeak) * b->P, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  for (int p = 0; p < b->P; ++p) {
    if (ind)
      for (int k = 0; k < 3; ++k) ind[3 * p + k] = pk[p].ind[k];
    if (disp0)
      for (int k = 0; k < 3; ++k) disp0[3 * p + k] = pk[p].disp[k];
    if (peak_value) peak_value[p] = pk[p].value;
    if (margin) margin[p] = pk[p].margin;
    if (delta) delta[p] = pk[p].delta;
    if (ncand) ncand[p] = pk[p].ncand;
    if (noverflow) noverflow[p] = pk[p].noverflow;
  }
  return FO_OK;
}

int fo_per_batch_coeffs(fo_per_batch* b, double* C1, double* C2) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  if (!b->C1) return set_err(FO_ERR_STATE, "batch created without FO_PER_KEEP_COEFFS");
  const size_t m = (size_t)b->gsum * b->cube;
  if (C1) FO_HIP(hipMemcpyAsync(C1, b->C1, sizeof(double2) * m, hipMemcpyDeviceToHost, b->stream));
  if (C2) FO_HIP(hipMemcpyAsync(C2, b->C2, sizeof(double2) * m, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_per_batch_product(fo_per_batch* b, double* C) {
  if (!b || !C) return set_err(FO_ERR_INVALID, "NULL argument");
  FO_HIP(hipMemcpyAsync(C, b->Cprod, sizeof(double2) * b->P * b->cube, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_per_batch_grid(fo_per_batch* b, int pair, float* g) {
  if (!b || !g) return set_err(FO_ERR_INVALID, "NULL argument");
  if (!b->grid) return set_err(FO_ERR_STATE, "batch created without FO_PER_KEEP_GRID");
  if (pair < 0 || pair >= b->P) return set_err(FO_ERR_INVALID, "pair out of range");
  const size_t f3 = (size_t)b->F * b->F * b->F;
  FO_HIP(hipMemcpyAsync(g, b->grid + f3 * pair, sizeof(float) * f3, hipMemcpyDeviceToHost, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_per_batch_fft(fo_per_batch* b, int pair, float* f) {
  if (!b || !f) return set_err(FO_ERR_INVALID, "NULL argument");
  if (!b->cgrid) return set_err(FO_ERR_STATE, "batch created without FO_PER_KEEP_FFT");
  if (pair < 0 || pair >= b->P) return set_err(FO_ERR_INVALID, "pair out of range");
  const size_t f3 = (size_t)b->F * b->F * b->F;
  FO_HIP(hipMemcpyAsync(f, b->cgrid + f3 * pair, sizeof(float2) * f3, hipMemcpyDeviceToHost,
                        b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  return FO_OK;
}

int fo_per_batch_set_coeffs_indexed(fo_per_batch* b, const double* Cs, int nstruct,
                                    const int* idx1, const int* idx2) {
  if (!b || !Cs || !idx1 || !idx2) return set_err(FO_ERR_INVALID, "NULL argument");
  if (!(b->flags & FO_PER_INPUT_COEFFS))
    return set_err(FO_ERR_STATE, "batch created without FO_PER_INPUT_COEFFS");
  if (nstruct < 1) return set_err(FO_ERR_INVALID, "nstruct must be >= 1");
  const int ng = b->hp[0].ngroups;
  for (int p = 0; p < b->P; ++p) {
    if (b->hp[p].ngroups != ng)
      return set_err(FO_ERR_INVALID, "shared coefficients need the same groups in every pair");
    if (idx1[p] < 0 || idx1[p] >= nstruct || idx2[p] < 0 || idx2[p] >= nstruct)
      return set_err(FO_ERR_INVALID, "structure index out of range");
  }
  FO_HIP(hipSetDevice(b->ctx->device));
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t bc = up(sizeof(double2) * (size_t)nstruct * ng * b->cube);
  const size_t bi = up(sizeof(int) * (size_t)b->P);
  if (b->csbuf) ctx_free(b->ctx, b->csbuf, b->csbytes);
  b->csbuf = nullptr;
  b->Cs = nullptr;
  int rc = ctx_alloc(b->ctx, bc + 2 * bi, &b->csbuf);
  if (rc) return rc;
  b->csbytes = bc + 2 * bi;
  b->Cs = (double2*)b->csbuf;
  b->idx1 = (int*)((char*)b->csbuf + bc);
  b->idx2 = (int*)((char*)b->csbuf + bc + bi);
  FO_HIP(hipMemcpyAsync(b->Cs, Cs, sizeof(double2) * (size_t)nstruct * ng * b->cube,
                        hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipMemcpyAsync(b->idx1, idx1, sizeof(int) * b->P, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipMemcpyAsync(b->idx2, idx2, sizeof(int) * b->P, hipMemcpyHostToDevice, b->stream));
  FO_HIP(hipStreamSynchronize(b->stream));
  // a captured launch sequence would still read the pairs' own coefficients
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
  return FO_OK;
}

int fo_per_batch_find_peaks(fo_per_batch* b, int npeaks, int width, double* fit, int* count,
                            float* residual) {
  if (!b || !fit || !count) return set_err(FO_ERR_INVALID, "NULL argument");
  if (!b->grid) return set_err(FO_ERR_STATE, "batch created without FO_PER_KEEP_GRID");
  FO_HIP(hipSetDevice(b->ctx->device));
  return run_peak_search(b->ctx, b->stream, b->pk, b->grid, b->P, b->F, b->F, b->F, npeaks,
                         width, fit, count, residual);
}

int fo_per_batch_timing(fo_per_batch* b, int enable, float* ms4) {
  if (!b) return set_err(FO_ERR_INVALID, "NULL batch");
  if (enable >= 0) {
    if (enable && !b->ev_made) {
      for (auto& e : b->ev) FO_HIP(hipEventCreate(&e));
      b->ev_made = true;
    }
    b->timing = enable != 0;
  }
  if (ms4 && b->ev_made) {
    FO_HIP(hipStreamSynchronize(b->stream));
    for (int i = 0; i < 4; ++i) FO_HIP(hipEventElapsedTime(&ms4[i], b->ev[i], b->ev[i + 1]));
  }
  return FO_OK;
}

}  // extern "C"
