// Multi-peak search on 3D overlap grids (spherical SO(3) grids and periodic
// translation grids) for gfx950: fastoverlap/utils.py:347-396 (findPeaks /
// fitPeak / _gaussian), used by fastoverlap/sphericalAlignment.py:196-235
// (findRotations, malign) and fastoverlap/periodicAlignment.py:442-456
// (findDisps with npeaks > 1).
//
// One workgroup per grid.  From f = g - min(g), repeated npeaks times:
//   1. argmax of f (first occurrence in C order on ties);
//   2. least-squares fit of  A exp(-sum_{i<=j} s_ij d_i d_j) + mu,
//      d = x - x0, on the (2w+1)^3 window around it (indices wrapped
//      periodically), from A = f(argmax), mu = 0, s = identity, x0 = 0;
//   3. record the peak at argmax + x0, subtract the fitted function (with its
//      background mu, unwrapped indices) from the whole grid;
// stopping at the first fit that fails.  The fit is a Levenberg-Marquardt
// iteration in FP64 with the analytic Jacobian, Marquardt's diagonal scaling
// and Nielsen's gain-ratio damping update (damping starts at 1e-3 of the
// largest diagonal entry); the grid itself is the FP32 grid of the plane
// kernels.
#include "fo_common.h"
#include "fo_internal.h"

namespace fo {

constexpr int kNPar = 11;                              // A, mu, 6 x s_ij, 3 x x0
constexpr int kNSum = kNPar * (kNPar + 1) / 2 + kNPar + 1;  // J^T J (66), J^T r (11), r.r
constexpr int kMaxWin = 3;                             // window half width
constexpr int kPtsPerThread = 2;                       // (2*3+1)^3 = 343 <= 2 * 256
constexpr int kMaxIter = 200;

// model value and Jacobian row at offset (x, y, z) from the window centre
FO_DEV double gauss_model(const double* p, const double d[3], double* J) {
  const double dx = d[0] - p[8], dy = d[1] - p[9], dz = d[2] - p[10];
  const double q = p[2] * dx * dx + p[3] * dx * dy + p[4] * dx * dz + p[5] * dy * dy +
                   p[6] * dy * dz + p[7] * dz * dz;
  const double e = exp(-q);
  const double Ae = p[0] * e;
  if (J) {
    J[0] = e;
    J[1] = 1.0;
    J[2] = -Ae * dx * dx;
    J[3] = -Ae * dx * dy;
    J[4] = -Ae * dx * dz;
    J[5] = -Ae * dy * dy;
    J[6] = -Ae * dy * dz;
    J[7] = -Ae * dz * dz;
    // dq/dx0_m = -dq/dd_m
    J[8] = Ae * (2.0 * p[2] * dx + p[3] * dy + p[4] * dz);
    J[9] = Ae * (p[3] * dx + 2.0 * p[5] * dy + p[6] * dz);
    J[10] = Ae * (p[4] * dx + p[6] * dy + 2.0 * p[7] * dz);
  }
  return Ae + p[1];
}

// Solve (H + lam diag(H)) x = -g by Cholesky (thread 0); false if not SPD.
FO_DEV bool damped_solve(const double* H, const double* g, double lam, double* x) {
  double M[kNPar][kNPar];
  for (int i = 0; i < kNPar; ++i)
    for (int j = 0; j <= i; ++j) {
      double v = H[i * kNPar + j];
      if (i == j) v += lam * fmax(H[i * kNPar + i], 1e-300);
      M[i][j] = v;
    }
  for (int j = 0; j < kNPar; ++j) {
    double s = M[j][j];
    for (int k = 0; k < j; ++k) s -= M[j][k] * M[j][k];
    if (!(s > 0.0)) return false;
    const double r = sqrt(s);
    M[j][j] = r;
    for (int i = j + 1; i < kNPar; ++i) {
      double t = M[i][j];
      for (int k = 0; k < j; ++k) t -= M[i][k] * M[j][k];
      M[i][j] = t / r;
    }
  }
  double y[kNPar];
  for (int i = 0; i < kNPar; ++i) {
    double t = -g[i];
    for (int k = 0; k < i; ++k) t -= M[i][k] * y[k];
    y[i] = t / M[i][i];
  }
  for (int i = kNPar - 1; i >= 0; --i) {
    double t = y[i];
    for (int k = i + 1; k < kNPar; ++k) t -= M[k][i] * x[k];
    x[i] = t / M[i][i];
  }
  for (int i = 0; i < kNPar; ++i)
    if (!isfinite(x[i])) return false;
  return true;
}

__global__ void __launch_bounds__(256) k_find_peaks(
    const float* __restrict__ grids, long long grid_stride, float* __restrict__ work, int n0,
    int n1, int n2, int npeaks, int win, double* __restrict__ out_fit,
    int* __restrict__ out_count) {
  __shared__ double s_part[4][kNSum];
  __shared__ double s_sum[kNSum];
  __shared__ double s_H[kNPar * kNPar];
  __shared__ double s_p[kNPar], s_pt[kNPar], s_g[kNPar];
  __shared__ float s_rv[16];
  __shared__ int s_ri[16];
  __shared__ int s_state;      // 0 iterate, 1 converged, 2 failed
  __shared__ int s_use_trial;  // evaluate at the trial parameters
  __shared__ double s_lam, s_nu, s_cost, s_pred;
  __shared__ int s_fev;  // function evaluations, counted as a difference-Jacobian solver would
  __shared__ int s_ind[3];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const long long g = blockIdx.x;
  const long long tot = (long long)n0 * n1 * n2;
  const float* src = grids + g * grid_stride;
  float* f = work + g * tot;
  // f = g - min(g)
  float mn = INFINITY;
  for (long long i = t; i < tot; i += blockDim.x) mn = fminf(mn, src[i]);
  {
    for (int off = 32; off > 0; off >>= 1) mn = fminf(mn, __shfl_xor(mn, off, 64));
    if (lane == 0) s_rv[wid] = mn;
    __syncthreads();
    mn = fminf(fminf(s_rv[0], s_rv[1]), fminf(s_rv[2], s_rv[3]));
    __syncthreads();
  }
  for (long long i = t; i < tot; i += blockDim.x) f[i] = src[i] - mn;
  __syncthreads();
  const int W = 2 * win + 1, K = W * W * W;
  int found = 0;
  for (int it = 0; it < npeaks; ++it) {
    // 1. argmax (first occurrence)
    float bv = -INFINITY;
    int bi = 0x7fffffff;
    for (long long i = t; i < tot; i += blockDim.x) argmax_merge(bv, bi, f[i], (int)i);
    block_argmax(bv, bi, s_rv, s_ri);
    if (t == 0) {
      s_ind[0] = bi / (n1 * n2);
      s_ind[1] = (bi / n2) % n1;
      s_ind[2] = bi % n2;
    }
    __syncthreads();
    // 2. the window's points owned by this thread (offsets and data)
    double dx[kPtsPerThread][3], yv[kPtsPerThread];
    int own = 0;
    for (int k = t; k < K && own < kPtsPerThread; k += blockDim.x, ++own) {
      const int a = k / (W * W) - win, b = (k / W) % W - win, c = k % W - win;
      const int ia = ((s_ind[0] + a) % n0 + n0) % n0;
      const int ib = ((s_ind[1] + b) % n1 + n1) % n1;
      const int ic = ((s_ind[2] + c) % n2 + n2) % n2;
      dx[own][0] = a;
      dx[own][1] = b;
      dx[own][2] = c;
      yv[own] = f[((long long)ia * n1 + ib) * n2 + ic];
    }
    if (t == 0) {
      const double p0[kNPar] = {(double)bv, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0};
      for (int i = 0; i < kNPar; ++i) s_p[i] = s_pt[i] = p0[i];
      s_state = 0;
      s_use_trial = 0;
      s_lam = -1.0;  // set from the first normal equations
      s_nu = 2.0;
      s_fev = 0;
    }
    __syncthreads();
    // Levenberg-Marquardt: each pass evaluates the residuals at s_p (full
    // normal equations) or at the trial point s_pt (cost only)
    for (int pass = 0; pass < 4 * kMaxIter; ++pass) {
      const bool trial = s_use_trial != 0;
      const double* pp = trial ? s_pt : s_p;
      double acc[kNSum];
      for (int i = 0; i < kNSum; ++i) acc[i] = 0.0;
      for (int q = 0; q < own; ++q) {
        double J[kNPar];
        const double r = gauss_model(pp, dx[q], trial ? nullptr : J) - yv[q];
        acc[kNSum - 1] += r * r;
        if (!trial) {
          int h = 0;
          for (int i = 0; i < kNPar; ++i)
            for (int j = 0; j <= i; ++j) acc[h++] += J[i] * J[j];
          for (int i = 0; i < kNPar; ++i) acc[h++] += J[i] * r;
        }
      }
      const int nsum = trial ? 1 : kNSum;
      for (int i = 0; i < nsum; ++i) {
        const int slot = trial ? kNSum - 1 : i;
        const double v = wave_sum(acc[slot]);
        if (lane == 0) s_part[wid][slot] = v;
      }
      __syncthreads();
      if (t < kNSum && (!trial || t == kNSum - 1))
        s_sum[t] = s_part[0][t] + s_part[1][t] + s_part[2][t] + s_part[3][t];
      __syncthreads();
      if (t == 0) {
        const double cost = s_sum[kNSum - 1];
        // the reference's solver evaluates the residuals once per trial and
        // kNPar times per forward-difference Jacobian, and gives up at
        // 200 (kNPar + 1) evaluations
        s_fev += trial ? 1 : (s_fev == 0 ? 1 + kNPar : kNPar);
        double hmax = 1e-300;
        if (!trial) {
          // fresh normal equations at s_p: accepted point
          s_cost = cost;
          int h = 0;
          for (int i = 0; i < kNPar; ++i)
            for (int j = 0; j <= i; ++j) {
              s_H[i * kNPar + j] = s_sum[h];
              s_H[j * kNPar + i] = s_sum[h];
              ++h;
            }
          for (int i = 0; i < kNPar; ++i) hmax = fmax(hmax, s_H[i * kNPar + i]);
          if (s_lam < 0.0) s_lam = 1e-3 * hmax;  // first point of this fit
          if (!isfinite(cost)) s_state = 2;
        } else {
          for (int i = 0; i < kNPar; ++i) hmax = fmax(hmax, s_H[i * kNPar + i]);
          // gain ratio: actual over predicted reduction of the cost
          const double rho = s_pred > 0.0 ? (s_cost - cost) / s_pred : -1.0;
          if (isfinite(cost) && rho > 0.0) {
            const double rel = (s_cost - cost) / fmax(s_cost, 1e-300);
            const double prel = s_pred / fmax(s_cost, 1e-300);
            double dn = 0.0, pn = 0.0;
            for (int i = 0; i < kNPar; ++i) {
              dn += (s_pt[i] - s_p[i]) * (s_pt[i] - s_p[i]);
              pn += s_p[i] * s_p[i];
            }
            for (int i = 0; i < kNPar; ++i) s_p[i] = s_pt[i];
            // Nielsen's damping update
            const double t = 2.0 * rho - 1.0;
            s_lam *= fmax(1.0 / 3.0, 1.0 - t * t * t);
            s_nu = 2.0;
            // tolerances of the reference's least-squares fit (1.49012e-8):
            // actual and predicted relative reductions, or the relative step
            if ((rel <= 1.49012e-8 && prel <= 1.49012e-8) ||
                sqrt(dn) <= 1.49012e-8 * (sqrt(pn) + 1.49012e-8))
              s_state = 1;
            s_use_trial = 0;  // next pass: normal equations at the new point
          } else {
            s_lam *= s_nu;
            s_nu *= 2.0;
            if (s_lam > 1e16 * fmax(1.0, hmax)) s_state = 1;  // no descent left
          }
        }
        if (s_state == 0 && (!trial || s_use_trial)) {
          // propose a step from the current normal equations
          // J^T r of the accepted point (kept across rejected trials)
          if (!trial)
            for (int i = 0; i < kNPar; ++i) s_g[i] = s_sum[kNPar * (kNPar + 1) / 2 + i];
          double x[kNPar];
          const double lim = 1e16 * fmax(1.0, hmax);
          while (!damped_solve(s_H, s_g, s_lam, x) && s_lam < lim) s_lam *= 10.0;
          if (s_lam >= lim) {
            s_state = 1;
          } else {
            // predicted reduction of the cost: x . (lam D x - g), D = diag(H)
            double pr = 0.0;
            for (int i = 0; i < kNPar; ++i) {
              s_pt[i] = s_p[i] + x[i];
              pr += x[i] * (s_lam * fmax(s_H[i * kNPar + i], 1e-300) * x[i] - s_g[i]);
            }
            s_pred = pr;
            s_use_trial = 1;
          }
        }
        if (s_state == 0 && (s_fev >= 200 * (kNPar + 1) || pass == 4 * kMaxIter - 1))
          s_state = 2;
      }
      __syncthreads();
      if (s_state != 0) break;
    }
    if (s_state != 1) break;  // a failed fit ends the search
    // 3. record; subtract the fitted function from the whole grid
    const double c0 = s_ind[0] + s_p[8], c1 = s_ind[1] + s_p[9], c2 = s_ind[2] + s_p[10];
    if (t == 0) {
      // A, mu, s00, s01, s02, s11, s12, s22, then the centre in grid indices
      double* o = out_fit + (g * npeaks + it) * kNPar;
      for (int i = 0; i < 8; ++i) o[i] = s_p[i];
      o[8] = c0;
      o[9] = c1;
      o[10] = c2;
    }
    const double A = s_p[0], mu = s_p[1];
    const double s00 = s_p[2], s01 = s_p[3], s02 = s_p[4], s11 = s_p[5], s12 = s_p[6],
                 s22 = s_p[7];
    for (long long i = t; i < tot; i += blockDim.x) {
      const double ex = (double)(i / ((long long)n1 * n2)) - c0;
      const double ey = (double)((i / n2) % n1) - c1;
      const double ez = (double)(i % n2) - c2;
      const double q = s00 * ex * ex + s01 * ex * ey + s02 * ex * ez + s11 * ey * ey +
                       s12 * ey * ez + s22 * ez * ez;
      f[i] = (float)((double)f[i] - (A * exp(-q) + mu));
    }
    ++found;
    __syncthreads();
  }
  if (t == 0) out_count[g] = found;
}

void launch_find_peaks(hipStream_t s, int ngrids, const float* grids, long long grid_stride,
                       float* work, int n0, int n1, int n2, int npeaks, int win,
                       double* out_fit, int* out_count) {
  hipLaunchKernelGGL(k_find_peaks, dim3(ngrids), dim3(256), 0, s, grids, grid_stride, work, n0,
                     n1, n2, npeaks, win, out_fit, out_count);
}

int find_peaks_max_window() { return kMaxWin; }

}  // namespace fo
