// ---------------------------------------------------------------------------
// Harmonic-oscillator basis coefficients (SphericalHarmonicAlign,
// fastoverlap/sphericalAlignment.py:276-391).  Per structure and atom j the
// smeared density's projection on the basis r^l L_n^{l+1/2}(r^2) e^{-r^2/2r0^2}
// (the reference's polynomial in r, fastoverlap/utils.py:414-427):
//   D[j][l][n] = sum_{k<=n} g[n][l][k] J_k(l, r_j),
//   J_k(l, r) = int_0^inf e^{-(x^2+r^2)/2s^2} e^{-x^2/2r0^2} i_l(x r/s^2) x^{l+2k} x^2 dx
// in closed form (sphericalAlignment.py:287-300, s = l + 2k there).  With
// u = (s^-2 + r0^-2)^-1, b = l + 3/2 and z = r^2 u / (2 s^4), Kummer's
// transformation M(a, b, z) = e^z M(b - a, b, -z) turns the reference's
// hyp1f1(b + k, b, z) e^{-r^2/2s^2} into the terminating, all-positive
//   (b)_k M(-k, b, -z) = sum_i C(k, i) z^i (b + i)_{k-i}
// times e^{-r^2 / 2(r0^2 + s^2)}, so
//   J_k = (2 pi u)^{3/2} (r u / s^2)^l (2u)^k (b)_k M(-k, b, -z) e^{-r^2/2(r0^2+s^2)}
// (no overflow-prone e^z, no cancellation).  Then per group
//   c[q(l,m)][n] = sum_{j in group} D[j][l][n] conj(Y_l^m(j))   (m >= 0,
// sphericalAlignment.py:353-355) and
//   C^l_{m1 m2} = sum_groups sum_n conj(c1[l,m1][n]) c2[l,m2][n]  (:363-368)
// with c[l,-m] = (-1)^m conj(c[l,m]).
// Basis 1 is the Fortran module's (HARMONICNL, f90/fastclusters.f90:492-540):
// the normalised oscillator functions
//   g_nl(r) = N (r/r0)^l e^{-r^2/2r0^2} L_n^{l+1/2}(r^2/r0^2),
// whose integrals 4 pi int g_nl e^{-(r^2+r_j^2)/2s^2} i_l(r r_j/s^2) r^2 dr
// follow a two-index recurrence over (n, l <= L + 2 nmax).
// Block of a pair inside the coefficient chunk (doubles): D1, D2
// [ngath][LX+1][nmax+1] (LX = L, or L + 2 nmax for basis 1), then c1, c2
// [ngroups][q(L)][nmax+1] complex.
// ---------------------------------------------------------------------------
#include "fo_common.h"
#include "fo_internal.h"

namespace fo {

// per (chunk pair, structure): D[j][l][n] for the pair's gathered atoms
__global__ void __launch_bounds__(256) k_harm_radial(
    const SphPair* __restrict__ pd, const int* __restrict__ gath, const double* __restrict__ X1c,
    const double* __restrict__ X2c, const double* __restrict__ scale_eff, int nmax, double r0,
    const double* __restrict__ gtab, double* __restrict__ R) {
  const SphPair d = pd[blockIdx.x];
  const int s = blockIdx.y, L = d.L, Ng = d.ngath, NN = nmax + 1;
  const double* X = s ? X2c : X1c;
  double* D = R + d.rad_off + (long long)s * Ng * (L + 1) * NN;
  const double sg = scale_eff[blockIdx.x];
  const double s2 = sg * sg, r02 = r0 * r0;
  const double u = 1.0 / (1.0 / s2 + 1.0 / r02);
  const double pre = pow(2.0 * kPi * u, 1.5);
  for (int w = threadIdx.x; w < Ng * (L + 1); w += blockDim.x) {
    const int g = w / (L + 1), l = w % (L + 1);
    const int a = gath[d.gath_off + g];
    const double* x = X + 3LL * (d.atom_off + a);
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2], r = sqrt(r2);
    const double b = l + 1.5, z = 0.5 * r2 * u / (s2 * s2);
    double* Dg = D + ((long long)g * (L + 1) + l) * NN;
    const double* gl = gtab + (long long)l * NN * NN;  // [l][n][k], k <= n
    for (int n = 0; n <= nmax; ++n) Dg[n] = 0.0;
    // (2 pi u)^{3/2} (r u / s^2)^l e^{-r^2 / 2(r0^2 + s^2)} (2u)^k
    double tk = pre * ipow(r * u / s2, l) * exp(-0.5 * r2 / (r02 + s2));
    double bk = 1.0;  // (b)_k
    for (int k = 0; k <= nmax; ++k) {
      // (b)_k M(-k, b, -z) = sum_i C(k,i) z^i (b+i)_{k-i}, all terms positive:
      // term_0 = (b)_k, term_{i+1} = term_i (k-i)/(i+1) z/(b+i)
      double term = bk, sum = bk;
      for (int i = 0; i < k; ++i) {
        term *= (double)(k - i) / (i + 1) * z / (b + i);
        sum += term;
      }
      const double J = tk * sum;
      // D[n] = sum_{k <= n} g[n][k] J_k, accumulated in ascending k
      for (int n = k; n <= nmax; ++n) Dg[n] += gl[n * NN + k] * J;
      tk *= 2.0 * u;
      bk *= b + k;
    }
  }
}

// Basis 1, one thread per (chunk pair, structure, gathered atom): the
// recurrence of HARMONICNL over rows n = 0..nmax, degrees J <= LX - 2n
// (row 1: J <= LX - 2), into D[j][J][n].
__global__ void __launch_bounds__(64) k_harm_radial_nl(
    const SphPair* __restrict__ pd, const int* __restrict__ gath, const double* __restrict__ X1c,
    const double* __restrict__ X2c, const double* __restrict__ scale_eff, int nmax, double r0,
    double* __restrict__ R) {
  const SphPair d = pd[blockIdx.x];
  const int s = blockIdx.y, g = blockIdx.z * blockDim.x + threadIdx.x;
  const int L = d.L, Ng = d.ngath, NN = nmax + 1, LX = L + 2 * nmax;
  if (g >= Ng) return;
  const double* X = s ? X2c : X1c;
  double* D = R + d.rad_off + ((long long)s * Ng + g) * (LX + 1) * NN;
  const double sg = scale_eff[blockIdx.x], s2 = sg * sg;
  const int a = gath[d.gath_off + g];
  const double* x = X + 3LL * (d.atom_off + a);
  const double rj = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  auto at = [&](int n, int J) -> double& { return D[(long long)J * NN + n]; };
  const double rs = 1.0 / (r0 * r0 + s2);
  at(0, 0) = sqrt(2.0 * sqrt(kPi) * ipow(r0 * rs, 3)) * s2 * sg * exp(-0.5 * rj * rj * rs) * 4.0 * kPi;
  const double f0 = sqrt(2.0) * r0 * rj * rs;
  for (int J = 1; J <= LX; ++J) at(0, J) = f0 / sqrt(1.0 + 2.0 * J) * at(0, J - 1);
  const double q = s2 / rj / r0;  // the reference divides by r_j (no limit at the centre)
  for (int n = 1; n <= nmax; ++n) {
    const double sn = sqrt((double)n);
    const int Jmax = n == 1 ? LX - 2 : LX - 2 * n;
    for (int J = 0; J <= Jmax; ++J) {
      double v = sqrt(n + J + 0.5) * at(n - 1, J) - (2.0 * J + 3.0) * q * at(n - 1, J + 1) -
                 sqrt(n + J + 1.5) * at(n - 1, J + 2);
      if (n >= 2) v += sqrt(n - 1.0) * at(n - 2, J + 2);
      at(n, J) = v / sn;
    }
  }
}

// per (chunk pair, structure, group): c[q(l,m)][n] = sum_j D[j][l][n] conj(Y[q][j])
__global__ void __launch_bounds__(256) k_harm_coef(const SphPair* __restrict__ pd,
                                                   const SphGroup* __restrict__ grp, int nmax,
                                                   const double2* __restrict__ Y1,
                                                   const double2* __restrict__ Y2,
                                                   double* __restrict__ R, int basis) {
  const SphPair d = pd[blockIdx.x];
  const int s = blockIdx.y, gi = blockIdx.z;
  if (gi >= d.ngroups) return;
  const int L = d.L, Ng = d.ngath, NN = nmax + 1;
  const int LX = basis ? L + 2 * nmax : L;
  const long long nq = plane_rows(L);
  const SphGroup G = grp[d.group_off + gi];
  const double2* Y = (s ? Y2 : Y1) + d.ylm_off;
  const double* D = R + d.rad_off + (long long)s * Ng * (LX + 1) * NN;
  double2* c = reinterpret_cast<double2*>(R + d.rad_off + 2LL * Ng * (LX + 1) * NN) +
               ((long long)s * d.ngroups + gi) * nq * NN;
  for (long long e = threadIdx.x; e < nq * NN; e += blockDim.x) {
    const int q = (int)(e / NN), n = (int)(e % NN);
    int l = 0;
    while (qlm(l + 1, 0) <= q) ++l;
    double cx = 0.0, cy = 0.0;
    for (int i = 0; i < G.size; ++i) {
      const int j = G.gath_off + i;
      const double dv = D[((long long)j * (LX + 1) + l) * NN + n];
      const double2 y = Y[(long long)q * Ng + j];
      cx += dv * y.x;
      cy -= dv * y.y;
    }
    c[e] = make_double2(cx, cy);
  }
}

// per (chunk pair, degree l): C^l_{m1 m2} = sum_groups sum_n conj(c1[l,m1][n]) c2[l,m2][n],
// rows m1 >= 0, written to the compact FP64 and the FP32 plane layouts
__global__ void __launch_bounds__(256) k_harm_contract(const SphPair* __restrict__ pd, int nmax,
                                                       int basis, const double* __restrict__ R,
                                                       int n, double2* __restrict__ I64,
                                                       float2* __restrict__ I32) {
  const SphPair d = pd[blockIdx.x];
  const int l = blockIdx.y, L = d.L;
  if (l > L) return;
  const int NN = nmax + 1, W = 2 * l + 1, LX = basis ? L + 2 * nmax : L;
  const long long nq = plane_rows(L);
  const double2* c1 =
      reinterpret_cast<const double2*>(R + d.rad_off + 2LL * d.ngath * (LX + 1) * NN);
  const double2* c2 = c1 + (long long)d.ngroups * nq * NN;
  for (int e = threadIdx.x; e < (l + 1) * W; e += blockDim.x) {
    const int m1 = e / W, m2 = e % W - l;
    const int am2 = m2 < 0 ? -m2 : m2;
    // c[l,-m] = (-1)^m conj(c[l,m])
    const double s2r = (m2 < 0 && (am2 & 1)) ? -1.0 : 1.0, s2i = m2 < 0 ? -s2r : 1.0;
    double vx = 0.0, vy = 0.0;
    for (int g = 0; g < d.ngroups; ++g) {
      const double2* a = c1 + ((long long)g * nq + qlm(l, m1)) * NN;
      const double2* b = c2 + ((long long)g * nq + qlm(l, am2)) * NN;
      for (int k = 0; k < NN; ++k) {
        const double2 x = a[k], y0 = b[k];
        const double yx = s2r * y0.x, yy = s2i * y0.y;
        // conj(x) y
        vx += x.x * yx + x.y * yy;
        vy += x.x * yy - x.y * yx;
      }
    }
    I64[d.coef_off + kidx(l, m1, m2)] = make_double2(vx, vy);
    I32[d.c32_off + (long long)qlm(l, m1) * n + plane_pos(plane_k(m2, n), n)] =
        make_float2((float)vx, (float)vy);
  }
}

// sum over the full coefficient set of |C^l_{m1 m2}|^2 from the rows
// m1 >= 0 (|C_{-m1,-m2}| = |C_{m1,m2}|), one workgroup per pair
__global__ void __launch_bounds__(256) k_coef_norm2(const SphPair* __restrict__ pd,
                                                    const double2* __restrict__ I64,
                                                    double* __restrict__ out) {
  const SphPair d = pd[blockIdx.x];
  const int L = d.L;
  double v[1] = {0.0};
  for (int l = 0; l <= L; ++l) {
    const int W = 2 * l + 1;
    for (int e = threadIdx.x; e < (l + 1) * W; e += blockDim.x) {
      const int m1 = e / W, m2 = e % W - l;
      const double2 c = I64[d.coef_off + kidx(l, m1, m2)];
      v[0] += (m1 ? 2.0 : 1.0) * (c.x * c.x + c.y * c.y);
    }
  }
  __shared__ double sc[8];
  block_sum_n<1>(v, sc);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

int harm_block_lx(int L, int nmax, int basis) { return basis ? L + 2 * nmax : L; }

void launch_harm_coeffs(hipStream_t s, int P, int Lmax, int maxgroups, int maxgath,
                        const SphPair* pd, const SphGroup* grp, const int* gath, const double* X1c,
                        const double* X2c, const double* scale_eff, int nmax, double r0, int basis,
                        const double* gtab, const double2* Y1, const double2* Y2, double* R,
                        int n, double2* I64, float2* I32) {
  if (P <= 0) return;
  if (basis)
    hipLaunchKernelGGL(k_harm_radial_nl, dim3(P, 2, (maxgath + 63) / 64), dim3(64), 0, s, pd,
                       gath, X1c, X2c, scale_eff, nmax, r0, R);
  else
    hipLaunchKernelGGL(k_harm_radial, dim3(P, 2), dim3(256), 0, s, pd, gath, X1c, X2c,
                       scale_eff, nmax, r0, gtab, R);
  hipLaunchKernelGGL(k_harm_coef, dim3(P, 2, maxgroups), dim3(256), 0, s, pd, grp, nmax, Y1, Y2,
                     R, basis);
  hipLaunchKernelGGL(k_harm_contract, dim3(P, Lmax + 1), dim3(256), 0, s, pd, nmax, basis, R, n,
                     I64, I32);
}

void launch_coef_norm2(hipStream_t s, int P, const SphPair* pd, const double2* I64, double* out) {
  if (P > 0) hipLaunchKernelGGL(k_coef_norm2, dim3(P), dim3(256), 0, s, pd, I64, out);
}

}  // namespace fo
