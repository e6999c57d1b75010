// FP64 refinement building blocks run by one workgroup per problem:
//  * exact linear assignment (shortest augmenting path with potentials,
//    the method family of Jonker & Volgenant 1987 that the reference cites at
//    f90/alignutils.f90:482-580; the Python path solves the same LAP with
//    munkres, fastoverlap/utils.py:34-46);
//  * Kearsley (1989) quaternion superposition (fastoverlap/utils.py:169-275);
//  * Euler-angle rotation matrix (fastoverlap/utils.py:447-460).
#pragma once
#include "fo_common.h"

namespace fo {

constexpr double kLapInf = 1e300;

// LDS layout of one LAP problem of size <= nmax (bytes = lap_lds_bytes(nmax)).
struct LapShared {
  double* u;
  double* v;
  double* minv;
  int* p;
  int* way;
  int* used;
  int* ra;  // column assigned to row i by the initial column reduction (0: free)
};
FO_HD size_t lap_lds_bytes(int nmax) {
  return (size_t)(nmax + 1) * (3 * sizeof(double) + 4 * sizeof(int));
}
FO_DEV LapShared lap_carve(unsigned char* base, int nmax) {
  LapShared s;
  s.u = (double*)base;
  s.v = s.u + (nmax + 1);
  s.minv = s.v + (nmax + 1);
  s.p = (int*)(s.minv + (nmax + 1));
  s.way = s.p + (nmax + 1);
  s.used = s.way + (nmax + 1);
  s.ra = s.used + (nmax + 1);
  return s;
}

// Wavefront argmin of (value, index) pairs, ties to the smaller index:
// butterflies inside rows of 16 lanes by DPP (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror), then the four row
// results through readlane.  Every lane must be active.
template <int CTRL>
FO_DEV void argmin_dpp_step(double& v, int& i) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, true);
  const double ov = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  const int oi = __builtin_amdgcn_update_dpp(0, i, CTRL, 0xf, 0xf, true);
  argmin_merge(v, i, ov, oi);
}
FO_DEV double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
FO_DEV void wave_argmin_dpp(double& v, int& i) {
  argmin_dpp_step<0xB1>(v, i);
  argmin_dpp_step<0x4E>(v, i);
  argmin_dpp_step<0x141>(v, i);
  argmin_dpp_step<0x140>(v, i);
  double rv = readlane_d(v, 0);
  int ri = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
  for (int r = 16; r < 64; r += 16) argmin_merge(rv, ri, readlane_d(v, r), __builtin_amdgcn_readlane(i, r));
  v = rv;
  i = ri;
}

// Minimum-cost assignment of an n x n cost matrix given by cost(i, j)
// (0-based).  rowsol[i] = column assigned to row i.  Ties in the column
// search are broken toward the smaller column index, so the result does not
// depend on the launch width.  Every thread of the block must call this;
// a one-wavefront block reduces without LDS or barrier.
// Returns false (rowsol = identity) if a non-finite cost stops the search.
template <class CostF>
FO_DEV bool lap_solve(int n, CostF cost, LapShared s, int* rowsol, double* red_v,
                      int* red_i) {
  const int t = threadIdx.x, bd = blockDim.x;
  // column reduction (Jonker & Volgenant 1987): v[j] = min_i c(i, j), and
  // column j goes to its minimising row (ties: smallest row) unless an
  // earlier column took that row.  The duals (u = 0, v) are feasible and
  // tight on these pairs, so the search below only runs for the free rows.
  for (int j = t; j <= n; j += bd) {
    double mn = 0.0;
    int mi = 0;
    if (j > 0) {
      mn = kLapInf;
      for (int i = 1; i <= n; ++i) {
        const double c = cost(i - 1, j - 1);
        if (c < mn) {
          mn = c;
          mi = i;
        }
      }
      if (!(mn < kLapInf)) mn = 0.0;  // non-finite column: left to the search
    }
    s.u[j] = 0.0;
    s.v[j] = mn;
    s.p[j] = 0;
    s.way[j] = mi;  // candidate row, read once below
    s.ra[j] = 0;
  }
  __syncthreads();
  if (t == 0)
    for (int j = 1; j <= n; ++j) {
      const int i = s.way[j];
      if (i > 0 && !s.ra[i]) {
        s.ra[i] = j;
        s.p[j] = i;
      }
    }
  __syncthreads();
  for (int i = 1; i <= n; ++i) {
    if (s.ra[i]) continue;  // uniform: every thread reads the same LDS word
    for (int j = t; j <= n; j += bd) {
      s.minv[j] = kLapInf;
      s.used[j] = 0;
    }
    if (t == 0) s.p[0] = i;
    __syncthreads();
    int j0 = 0;
    for (;;) {
      if (t == 0) s.used[j0] = 1;
      __syncthreads();
      const int i0 = s.p[j0];
      const double ui0 = s.u[i0];
      double best = kLapInf;
      int bj = 0x7fffffff;
      for (int j = 1 + t; j <= n; j += bd) {
        if (!s.used[j]) {
          double cur = cost(i0 - 1, j - 1) - ui0 - s.v[j];
          double mv = s.minv[j];
          if (cur < mv) {
            mv = cur;
            s.minv[j] = cur;
            s.way[j] = j0;
          }
          if (mv < best || (mv == best && j < bj)) {
            best = mv;
            bj = j;
          }
        }
      }
      if (bd == 64)
        wave_argmin_dpp(best, bj);
      else
        block_argmin(best, bj, red_v, red_i);
      if (bj == 0x7fffffff || !(best < kLapInf)) {
        // no finite reduced cost left: the costs are not finite numbers
        for (int k = t; k < n; k += bd) rowsol[k] = k;
        __syncthreads();
        return false;
      }
      const double delta = best;
      for (int j = t; j <= n; j += bd) {
        if (s.used[j]) {
          s.u[s.p[j]] += delta;
          s.v[j] -= delta;
        } else {
          s.minv[j] -= delta;
        }
      }
      __syncthreads();
      j0 = bj;
      if (s.p[j0] == 0) break;
    }
    if (t == 0) {
      do {
        int j1 = s.way[j0];
        s.p[j0] = s.p[j1];
        j0 = j1;
      } while (j0);
    }
    __syncthreads();
  }
  for (int j = 1 + t; j <= n; j += bd) rowsol[s.p[j] - 1] = j - 1;
  __syncthreads();
  return true;
}

// Cyclic Jacobi eigen-decomposition of a symmetric 4x4 matrix (FP64).
FO_DEV void jacobi4(double A[4][4], double V[4][4], double w[4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int p = 0; p < 4; ++p) {
      dia += A[p][p] * A[p][p];
      for (int q = p + 1; q < 4; ++q) off += A[p][q] * A[p][q];
    }
    if (off == 0.0 || off < 1e-36 * dia) break;
    for (int p = 0; p < 4; ++p) {
      for (int q = p + 1; q < 4; ++q) {
        double apq = A[p][q];
        if (apq == 0.0) continue;
        double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
        double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(tt * tt + 1.0), sn = tt * c;
        for (int k = 0; k < 4; ++k) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - sn * akq;
          A[k][q] = sn * akp + c * akq;
        }
        for (int k = 0; k < 4; ++k) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - sn * aqk;
          A[q][k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < 4; ++k) {
          double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - sn * vkq;
          V[k][q] = sn * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < 4; ++i) w[i] = A[i][i];
}

// Quaternion -> rotation matrix (fastoverlap/utils.py:255-275), row-major.
FO_DEV void q2mx(const double qin[4], double R[9]) {
  double nq = sqrt(qin[0] * qin[0] + qin[1] * qin[1] + qin[2] * qin[2] + qin[3] * qin[3]);
  double Q[4] = {qin[0] / nq, qin[1] / nq, qin[2] / nq, qin[3] / nq};
  R[0] = 2.0 * (0.5 - Q[2] * Q[2] - Q[3] * Q[3]);
  R[4] = 2.0 * (0.5 - Q[1] * Q[1] - Q[3] * Q[3]);
  R[8] = 2.0 * (0.5 - Q[1] * Q[1] - Q[2] * Q[2]);
  R[1] = 2.0 * (Q[1] * Q[2] - Q[0] * Q[3]);
  R[3] = 2.0 * (Q[1] * Q[2] + Q[0] * Q[3]);
  R[2] = 2.0 * (Q[1] * Q[3] + Q[0] * Q[2]);
  R[6] = 2.0 * (Q[1] * Q[3] - Q[0] * Q[2]);
  R[5] = 2.0 * (Q[2] * Q[3] - Q[0] * Q[1]);
  R[7] = 2.0 * (Q[2] * Q[3] + Q[0] * Q[1]);
}

// Kearsley QMAT accumulated over one atom pair (x1 - com1, x2 - com2).
FO_DEV void qmat_accum(double q[10], double x1, double y1, double z1, double x2,
                       double y2, double z2) {
  double XM = x1 - x2, YM = y1 - y2, ZM = z1 - z2;
  double XP = x1 + x2, YP = y1 + y2, ZP = z1 + z2;
  q[0] += XM * XM + YM * YM + ZM * ZM;
  q[1] += -YP * ZM + YM * ZP;
  q[2] += -XM * ZP + XP * ZM;
  q[3] += -XP * YM + XM * YP;
  q[4] += YP * YP + ZP * ZP + XM * XM;
  q[5] += XM * YM - XP * YP;
  q[6] += XM * ZM - XP * ZP;
  q[7] += XP * XP + ZP * ZP + YM * YM;
  q[8] += YM * ZM - YP * ZP;
  q[9] += XP * XP + YP * YP + ZM * ZM;
}

// Solve the 4x4 problem: dist = sqrt(min eigenvalue) with the reference's
// clamp of small negative eigenvalues (fastoverlap/utils.py:236-248), and
// the rotation MR (row-major).
FO_DEV void kearsley_solve(const double q[10], double* dist, double MR[9]) {
  double A[4][4] = {{q[0], q[1], q[2], q[3]},
                    {q[1], q[4], q[5], q[6]},
                    {q[2], q[5], q[7], q[8]},
                    {q[3], q[6], q[8], q[9]}};
  double V[4][4], w[4];
  jacobi4(A, V, w);
  int im = 0;
  for (int i = 1; i < 4; ++i)
    if (w[i] < w[im]) im = i;
  double e = w[im];
  if (e < 0.0) e = (fabs(e) < 1e-6) ? 0.0 : -e;
  *dist = sqrt(e);
  double qv[4] = {V[0][im], V[1][im], V[2][im], V[3][im]};
  q2mx(qv, MR);
}

// EulerM(a, b, y) = My . Mb . Ma  (fastoverlap/utils.py:447-460), row-major.
FO_HD void euler_matrix(double a, double b, double y, double M[9]) {
  double sa = sin(a), ca = cos(a), sb = sin(b), cb = cos(b), sy = sin(y), cy = cos(y);
  // Mb . Ma
  double T[9] = {cb * ca, -cb * sa, -sb, sa, ca, 0.0, sb * ca, -sb * sa, cb};
  // My . T
  M[0] = cy * T[0] - sy * T[3];
  M[1] = cy * T[1] - sy * T[4];
  M[2] = cy * T[2] - sy * T[5];
  M[3] = sy * T[0] + cy * T[3];
  M[4] = sy * T[1] + cy * T[4];
  M[5] = sy * T[2] + cy * T[5];
  M[6] = T[6];
  M[7] = T[7];
  M[8] = T[8];
}

}  // namespace fo
