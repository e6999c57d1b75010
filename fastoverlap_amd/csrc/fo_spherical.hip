// Spherical (cluster) overlap-alignment kernels for gfx950.
//
// Pipeline per batch of pairs sharing one SO(3) grid side n = 2B:
//   k_wigner_table   FP64 Wigner-d table on the beta grid (cached per n)
//   k_sph_prep       centroids, automatic kernel width
//   k_sph_ylm        Y_lm of every grouped atom (FP64)
//   k_sph_radial     scaled i_l(x) * Gaussian radial kernel (FP64)
//   k_sph_contract   C^l = Y1 . K_l . Y2^H per degree (FP64)
//   k_isoft_fft      inverse SOFT of both orientations per beta plane (FP32),
//                    two-pass register transforms (power-of-two n); k_isoft_plane
//                    is the direct-DFT form for other n;
//                    top candidates per plane
//   k_sph_select     FP64 re-evaluation of every candidate within the FP32
//                    error bound of the maximum -> exact argmax, parabolic
//                    offsets, grid Euler angles
//   k_sph_rotopt     FP64 Newton maximisation of Re sum conj(I) D(R)
//   k_sph_finish     orientation choice, LAP, quaternion fit (FP64)
#include "fo_common.h"
#include "fo_internal.h"
#include "fo_refine.h"
#include <algorithm>

namespace fo {

// ---------------------------------------------------------------------------
// Wigner-d table: tab[k * Kt + kidx(l,m1,m2)] = sqrt((2l+1)/2) d^l_{m1m2}(b_k)
// with b_k = pi (2k+1) / (2n)  (fastoverlap/soft.py:55, :73-96).  Evaluated
// with the three-term recurrence in the degree of Kostelec & Rockmore, "FFTs
// on the rotation group" (J. Fourier Anal. Appl. 14, 2008), seeded at
// l0 = max(|m1|,|m2|) by the closed form.
// ---------------------------------------------------------------------------
// Coefficients of  dn^{j+1} = B (cos b - C) dn^j - A dn^{j-1}  for the
// normalised dn^j = sqrt((2j+1)/2) d^j.  With the "edge" factor
// e_j = sqrt((j^2 - m1^2)(j^2 - m2^2)), the unnormalised recurrence reads
//   e_{j+1} d^{j+1} = (j+1)(2j+1) (cos b - m1 m2 / (j(j+1))) d^j
//                     - (j+1)/j e_j d^{j-1};
// rescaling each d^j by sqrt((2j+1)/2) gives
//   B = (j+1) sqrt((2j+1)(2j+3)) / e_{j+1},
//   A = (j+1)/j sqrt((2j+3)/(2j-1)) e_j / e_{j+1},   C = m1 m2 / (j(j+1)),
// and A = C = 0 at j = 0 (only m1 = m2 = 0 reaches it).
FO_DEV void recurrence_terms(int l, int m1, int m2, double& A, double& B, double& C) {
  const double j = (double)l, a2 = (double)m1 * m1, b2 = (double)m2 * m2;
  const double jp = j + 1.0;
  const double inv_e_next = 1.0 / sqrt((jp * jp - a2) * (jp * jp - b2));
  B = jp * sqrt((2.0 * j + 1.0) * (2.0 * j + 3.0)) * inv_e_next;
  if (l == 0) {
    A = 0.0;
    C = 0.0;
    return;
  }
  const double e_cur = sqrt((j * j - a2) * (j * j - b2));
  A = (jp / j) * sqrt((2.0 * j + 3.0) / (2.0 * j - 1.0)) * e_cur * inv_e_next;
  C = (double)m1 * m2 / (j * jp);
}

// log of sqrt(binomial(mu+nu, mu)) and the sign xi of the edge element
FO_DEV double seed_log_factor(int mu, int nu) {
  return 0.5 * (lgamma((double)(mu + nu + 1)) - lgamma((double)(mu + 1)) -
                lgamma((double)(nu + 1)));
}
FO_DEV double seed_xi(int m1, int m2) {
  return (m2 < m1 && ((m1 - m2) & 1)) ? -1.0 : 1.0;
}

// tab64: compact layout, [k][kidx(l,m1,m2)]; tab32: plane layout (rows
// m1 >= 0 only, zero entries pre-cleared by the caller), [k][row][pos].
__global__ void k_wigner_table(int n, int Lt, double* __restrict__ tab64,
                               float* __restrict__ tab32) {
  const int W = 2 * Lt + 1;
  const long long total = (long long)n * W * W;
  const long long Kt = kcount(Lt);
  const long long Kp = plane_rows(Lt) * n;
  for (long long w = blockIdx.x * (long long)blockDim.x + threadIdx.x; w < total;
       w += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(w / (W * W));
    const int rem = (int)(w % (W * W));
    const int m1 = rem / W - Lt, m2 = rem % W - Lt;
    const double beta = kPi * (2.0 * k + 1.0) / (2.0 * n);
    const double c2 = cos(0.5 * beta), s2 = sin(0.5 * beta), cb = cos(beta);
    const int l0 = max(abs(m1), abs(m2));
    const int mu = abs(m1 - m2), nu = abs(m1 + m2);
    const double seed = sqrt((2.0 * l0 + 1.0) / 2.0) * seed_xi(m1, m2) *
                        exp(seed_log_factor(mu, nu) + mu * log(s2) + nu * log(c2));
    double prev = 0.0, cur = seed;
    double* row64 = tab64 + (long long)k * Kt;
    float* row32 = tab32 + (long long)k * Kp + plane_pos(plane_k(m2, n), n);
    row64[kidx(l0, m1, m2)] = cur;
    if (m1 >= 0) row32[(long long)qlm(l0, m1) * n] = (float)cur;
    for (int l = l0; l < Lt; ++l) {
      double A, B, C;
      recurrence_terms(l, m1, m2, A, B, C);
      const double nxt = B * (cb - C) * cur - A * prev;
      prev = cur;
      cur = nxt;
      row64[kidx(l + 1, m1, m2)] = cur;
      if (m1 >= 0) row32[(long long)qlm(l + 1, m1) * n] = (float)cur;
    }
  }
}

// twiddles e^{+2 pi i k / n} (inverse DFT sign of fastoverlap/soft.py:124-125)
__global__ void k_twiddles(int n, double sign, double2* tw64, float2* tw32) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double s, c;
    sincospi(sign * 2.0 * (double)k / (double)n, &s, &c);
    tw64[k] = make_double2(c, s);
    tw32[k] = make_float2((float)c, (float)s);
  }
}

// ---------------------------------------------------------------------------
// Prep: centroids (fastoverlap/sphericalAlignment.py:129-137) and automatic
// kernel width (mean nearest-neighbour distance / 3 averaged over the two
// structures, :35-41 and :145-147).
// ---------------------------------------------------------------------------
__global__ void k_sph_prep(const SphPair* __restrict__ pd, const double* __restrict__ pos1,
                           const double* __restrict__ pos2, double* __restrict__ X1c,
                           double* __restrict__ X2c, double* __restrict__ scale_eff,
                           int center) {
  __shared__ double red[16];
  const SphPair d = pd[blockIdx.x];
  const int N = d.natoms, t = threadIdx.x, bd = blockDim.x;
  const double* p1 = pos1 + 3LL * d.atom_off;
  const double* p2 = pos2 + 3LL * d.atom_off;
  double c[6] = {0, 0, 0, 0, 0, 0};
  for (int a = t; a < N; a += bd)
    for (int q = 0; q < 3; ++q) {
      c[q] += p1[3 * a + q];
      c[3 + q] += p2[3 * a + q];
    }
  for (int q = 0; q < 6; ++q) c[q] = block_sum(c[q], red) / (double)N;
  if (!center)
    for (int q = 0; q < 6; ++q) c[q] = 0.0;
  double* x1 = X1c + 3LL * d.atom_off;
  double* x2 = X2c + 3LL * d.atom_off;
  for (int a = t; a < N; a += bd)
    for (int q = 0; q < 3; ++q) {
      x1[3 * a + q] = p1[3 * a + q] - c[q];
      x2[3 * a + q] = p2[3 * a + q] - c[3 + q];
    }
  if (d.scale > 0.0) {
    if (t == 0) scale_eff[blockIdx.x] = d.scale;
    return;
  }
  double s1 = 0.0, s2 = 0.0;
  for (int a = t; a < N; a += bd) {
    double m1 = 1e300, m2 = 1e300;
    for (int b = 0; b < N; ++b) {
      if (b == a) continue;
      double dx = p1[3 * a] - p1[3 * b], dy = p1[3 * a + 1] - p1[3 * b + 1],
             dz = p1[3 * a + 2] - p1[3 * b + 2];
      m1 = fmin(m1, dx * dx + dy * dy + dz * dz);
      dx = p2[3 * a] - p2[3 * b];
      dy = p2[3 * a + 1] - p2[3 * b + 1];
      dz = p2[3 * a + 2] - p2[3 * b + 2];
      m2 = fmin(m2, dx * dx + dy * dy + dz * dz);
    }
    s1 += sqrt(m1);
    s2 += sqrt(m2);
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  if (t == 0) scale_eff[blockIdx.x] = (s1 / N + s2 / N) / 6.0;
}

// ---------------------------------------------------------------------------
// Spherical harmonics of every grouped atom (fastoverlap/utils.py:296-302,
// fastoverlap/sphericalAlignment.py:57-65), FP64, for m >= 0 only:
//   Y[q(l,m)][g] = P_l^m(cos theta) e^{i m phi},  q(l,m) = l(l+1)/2 + m,
// with P the normalised associated Legendre function (Condon-Shortley
// phase); Y_l^{-m} = (-1)^m conj(Y_l^m) is formed where it is read.  One
// thread per (atom, m): P_m^m = c_m sin^m(theta), then the three-term
// recurrence in l with tabulated coefficients (k_leg_table).  An atom at the
// centre takes theta = phi = 0 (only l = 0 survives there).
// ---------------------------------------------------------------------------
// XCD-aware decode of a 1-D launch of 8 ceil(P/8) per workgroups over
// (pair, item): workgroup w serves pair (w % 8) + 8 (w / (8 per)) and item
// (w / 8) % per.  Workgroups are dealt to the 8 XCDs round-robin, so all
// items of a pair run on one XCD and share its L2.  False for the padding
// workgroups (pair >= P).
FO_DEV bool xcd_decode(int P, int per, int& p, int& item) {
  const int w = blockIdx.x;
  p = (w & 7) + 8 * (w / (8 * per));
  item = (w >> 3) % per;
  return p < P;
}
static inline unsigned xcd_grid(int P, int per) { return (unsigned)(8 * ((P + 7) / 8) * per); }

__global__ void k_sph_ylm(const SphPair* __restrict__ pd, const int* __restrict__ gath,
                          const double* __restrict__ X1c, const double* __restrict__ X2c,
                          const double2* __restrict__ leg, const double* __restrict__ cm,
                          double2* __restrict__ Y1, double2* __restrict__ Y2,
                          double* __restrict__ rad1, double* __restrict__ rad2) {
  const SphPair d = pd[blockIdx.x];
  const double* X = blockIdx.y ? X2c : X1c;
  double2* Yp = (blockIdx.y ? Y2 : Y1) + d.ylm_off;
  double* rad = (blockIdx.y ? rad2 : rad1) + d.gath_off;
  const int L = d.L, Ng = d.ngath;
  for (int w = threadIdx.x; w < Ng * (L + 1); w += blockDim.x) {
    const int m = w / Ng, g = w % Ng;
    const int a = gath[d.gath_off + g];
    const double x = X[3LL * (d.atom_off + a)], y = X[3LL * (d.atom_off + a) + 1],
                 z = X[3LL * (d.atom_off + a) + 2];
    const double rxy = sqrt(x * x + y * y), r = sqrt(x * x + y * y + z * z);
    if (m == 0) rad[g] = r;  // |x| per grouped atom, for the fused contraction
    double ct = 1.0, st = 0.0, phi = 0.0;
    if (r > 0.0) {
      ct = z / r;
      st = rxy / r;
      phi = atan2(y, x);
    }
    double sp, cp;
    sincos((double)m * phi, &sp, &cp);
    double p2 = 0.0, p1 = cm[m] * ipow(st, m);
    Yp[(long long)qlm(m, m) * Ng + g] = make_double2(p1 * cp, p1 * sp);
    for (int l = m + 1; l <= L; ++l) {
      const double2 ab = leg[qlm(l, m)];
      const double pl = ab.x * (ct * p1 - ab.y * p2);
      Yp[(long long)qlm(l, m) * Ng + g] = make_double2(pl * cp, pl * sp);
      p2 = p1;
      p1 = pl;
    }
  }
}

// Legendre recurrence coefficients P_l^m = a (cos(theta) P_{l-1}^m - b P_{l-2}^m):
// leg[q(l,m)] = (a, b) with a = sqrt((4l^2-1)/(l^2-m^2)),
// b = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)) for l >= m+2, (sqrt(2m+3), 0) for
// l = m+1; cm[m] = (-1)^m sqrt(1/(4 pi)) prod_{k<=m} sqrt((2k+1)/(2k)).
__global__ void k_leg_table(int Lt, double2* __restrict__ leg, double* __restrict__ cm) {
  const int nq = qlm(Lt, Lt) + 1;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nq; e += gridDim.x * blockDim.x) {
    int l = 0;
    while (qlm(l + 1, 0) <= e) ++l;
    const int m = e - qlm(l, 0);
    double2 v = make_double2(0.0, 0.0);
    if (l == m + 1) {
      v.x = sqrt(2.0 * m + 3.0);
    } else if (l >= m + 2) {
      const double dl = l, dm = m;
      v.x = sqrt((4.0 * dl * dl - 1.0) / (dl * dl - dm * dm));
      v.y = sqrt(((dl - 1.0) * (dl - 1.0) - dm * dm) / (4.0 * (dl - 1.0) * (dl - 1.0) - 1.0));
    }
    leg[e] = v;
  }
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m <= Lt; m += gridDim.x * blockDim.x) {
    double c = sqrt(1.0 / (4.0 * kPi));
    for (int k = 1; k <= m; ++k) c *= -sqrt((2.0 * k + 1.0) / (2.0 * k));
    cm[m] = c;
  }
}

// ---------------------------------------------------------------------------
// Radial kernel il[l,j,k] = i_l(x) exp(-(r1^2 + r2^2)/4s^2), x = r1 r2/(2s^2)
// (fastoverlap/sphericalAlignment.py:269-272), evaluated overflow-free as
// [i_l(x) e^{-x}] exp(-(r1-r2)^2/4s^2) in FP64.  i_l(x) e^{-x}: power series
// for x < 1; otherwise Miller's backward recurrence started at
// l_s = ceil(sqrt(T^2 + 40x)) + 10 for the highest stored degree T (where
// the minimal-solution contamination is below 1e-17) and normalised by
// i_0(x) e^{-x} = -expm1(-2x)/(2x).  The first pass finds the normalisation
// at l = 0; the second resumes at T and stores the slot degrees as it
// passes them (rad_slot_deg).
// Layout: R[rad_off + grp.rad_off + (slot*Ns + j)*Ns + k].
// ---------------------------------------------------------------------------
FO_DEV double radial_series(int l, double x, double exg) {
  // i_l(x) e^{-x} g = x^l/(2l+1)!! sum_k (x^2/2)^k / (k! (2l+3)...(2l+2k+1)) e^{-x} g
  const double h = 0.5 * x * x;
  double lead = 1.0;
  for (int i = 1; i <= l; ++i) lead *= x / (2.0 * i + 1.0);
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 40; ++k) {
    term *= h / (k * (2.0 * l + 2.0 * k + 1.0));
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  return lead * sum * exg;
}
// The values are stored non-temporally: 1.7 GB per C2 step that the
// contraction reads back from HBM anyway, so they need not displace L2 lines
// (0.93 -> 0.84 ms; the contraction's time unchanged).
FO_DEV void radial_store(int L, double x, double gauss, double* out, long long stride) {
  const int ns = rad_slots(L);
  if (x == 0.0) {
    for (int s = 0; s < ns; ++s)
      __builtin_nontemporal_store(rad_slot_deg(s, L) == 0 ? gauss : 0.0, &out[s * stride]);
    return;
  }
  if (x < 1.0) {
    const double exg = exp(-x) * gauss;
    for (int s = 0; s < ns; ++s)
      __builtin_nontemporal_store(radial_series(rad_slot_deg(s, L), x, exg), &out[s * stride]);
    return;
  }
  // i_0(x) e^{-x} times the Gaussian fixes the normalisation of Miller's
  // downward recurrence.  Pass 1 runs the recurrence to l = 0 without
  // storing, keeping its state at the top slot degree; pass 2 resumes from
  // that state (the identical values) and stores f_l * scal at the slot
  // degrees (no array of values held in registers, so more wavefronts per
  // SIMD overlap their stores).
  const int top = rad_slot_deg(ns - 1, L), bot = rad_slot_deg(0, L);
  const double i0 = -expm1(-2.0 * x) / (2.0 * x);
  const int ls = (int)ceil(sqrt((double)top * top + 40.0 * x)) + 10;
  const double ix = 1.0 / x;
  double fp1 = 0.0, f = 1.0;
#pragma unroll 2
  for (int l = ls; l > top; --l) {
    const double fm = fp1 + (2.0 * l + 1.0) * ix * f;
    fp1 = f;
    f = fm;
  }
  const double ft = f, ft1 = fp1;  // f_top, f_{top+1}
#pragma unroll 2
  for (int l = top; l > 0; --l) {
    const double fm = fp1 + (2.0 * l + 1.0) * ix * f;
    fp1 = f;
    f = fm;
  }
  const double scal = i0 / f * gauss;
  f = ft;
  fp1 = ft1;
  // degrees top .. bot (uniform over the workgroup: L is)
  int s = ns - 1;
  for (int l = top; l >= bot; --l) {
    while (s >= 0 && rad_slot_deg(s, L) == l) __builtin_nontemporal_store(f * scal, &out[(s--) * stride]);
    const double fm = fp1 + (2.0 * l + 1.0) * ix * f;
    fp1 = f;
    f = fm;
  }
}
__global__ void __launch_bounds__(256, 5) k_sph_radial(const SphPair* __restrict__ pd, const SphGroup* __restrict__ grp,
                             const double* __restrict__ rad1, const double* __restrict__ rad2,
                             const double* __restrict__ scale_eff, double* __restrict__ R) {
  // the pair's first kRadGroups group records are staged in LDS (one load
  // latency for the whole workgroup instead of a dependent chain per entry)
  constexpr int kRadGroups = 16;
  __shared__ SphGroup sgrp[kRadGroups];
  const SphPair d = pd[blockIdx.y];
  if (d.ngroups < 1) return;  // whole workgroup: no (j, k) entries
  const double s = scale_eff[blockIdx.y];
  const double i2s2 = 1.0 / (2.0 * s * s), i4s2 = 1.0 / (4.0 * s * s);
  const SphGroup* gsrc = grp + d.group_off;
  if ((int)threadIdx.x < min(d.ngroups, kRadGroups)) sgrp[threadIdx.x] = gsrc[threadIdx.x];
  __syncthreads();
  const SphGroup* g = d.ngroups <= kRadGroups ? sgrp : gsrc;
  const SphGroup last = g[d.ngroups - 1];
  const long long total = last.jk_off + (long long)last.size * last.size;
  for (long long w = blockIdx.x * (long long)blockDim.x + threadIdx.x; w < total;
       w += (long long)gridDim.x * blockDim.x) {
    int q = 0;
    while (q + 1 < d.ngroups && w >= g[q + 1].jk_off) ++q;
    const int Ns = g[q].size;
    const int jk = (int)(w - g[q].jk_off);  // < Ns^2 <= 2^24
    const int j = jk / Ns, k = jk - j * Ns;
    // |x| of the two atoms, as the harmonics kernel recorded them
    const double r1 = rad1[d.gath_off + g[q].gath_off + j];
    const double r2 = rad2[d.gath_off + g[q].gath_off + k];
    const double x = r1 * r2 * i2s2;
    const double gauss = exp(-(r1 - r2) * (r1 - r2) * i4s2);
    radial_store(d.L, x, gauss, R + d.rad_off + g[q].rad_off + (long long)j * Ns + k,
                 (long long)Ns * Ns);
  }
}

// ---------------------------------------------------------------------------
// Contraction I^l_{m1m2} = 4 pi^{5/2} s^3 sum_s sum_{j,k in s}
//   Y1[l,m1,j] K_l[j,k] conj(Y2[l,m2,k])      (sphericalAlignment.py:273,:175)
// Only rows m1 >= 0 are contracted: Y_l^{-m} = (-1)^m conj(Y_l^m) and K_l
// is real, so I^l_{-m1,-m2} = (-1)^{m1+m2} conj(I^l_{m1m2}) and the m1 < 0
// rows are stored from the computed ones.
// ---------------------------------------------------------------------------
// m1 rows per workgroup: 16 for the column-tile class of degrees <= 15
// (at most 16 rows), 32 above
template <int NCT>
constexpr int contract_rows() {
  return NCT <= 2 ? 16 : 32;
}

// ---------------------------------------------------------------------------
// MFMA form of the contraction (v_mfma_f64_16x16x4f64).  One workgroup per
// (pair, l, block of MB rows m1 >= 0).  Per species group and tile of kCT
// atoms k of structure 2:
//   stage Y2[m][k-tile] (built from P2, E2) in LDS
//   A[rr][k] = sum_j Y1'[rr][j] K_l[j][k]: real rows rr = 2r + {Re, Im} of
//              row m1a + r; j in chunks of kJC staged in LDS one chunk ahead
//              (cooperative coalesced loads, double buffer); wave w computes
//              rows 16w..16w+15 of both 16-wide column tiles
//   A -> LDS (complex [row][k], aliasing the chunk buffers)
//   C[r][m] += sum_k A[r][k] conj(Y2[m][k]) as four real products
//              (Cr += Ar Y2r + Ai Y2i, Ci += Ai Y2r - Ar Y2i); complex 16x16
//              tiles dealt to the waves round-robin.
// MFMA lane layout: A operand lane -> (row lane%16, k lane/16), B operand
// lane -> (k lane/16, col lane%16), result register i of a lane -> (row
// 4 i + lane/16, col lane%16).  The reduction order of every output is
// fixed, so results do not depend on the batch a pair is in.
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));

FO_DEV v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int kCT = 32;  // atoms of structure 2 per tile
constexpr int kJC = 16;  // atoms of structure 1 per staged chunk

// LDS of one contraction workgroup; NCT = 16-wide column tiles of Y2.  The
// chunk buffers of the first product and the A / Y2 tiles of the second are
// never live together.
template <int NCT, int MB = contract_rows<NCT>()>
struct ContractLds {
  union {
    struct {
      double y1[2][2 * MB][kJC + 1];   // [buffer][real row][j]
      double kc[2][kJC][kCT + 1];      // [buffer][j][k]
    } a;
    struct {
      double2 ac[MB][kCT + 1];         // A tile [row][k]
      double2 ys[kCT][16 * NCT + 1];   // Y2 tile [k][m]
    } b;
  } u;
};

template <int NCT>
FO_DEV void contract_mfma_body(const SphPair& d, int l, int b, const SphGroup* __restrict__ grp,
                               const double2* __restrict__ Y1, const double2* __restrict__ Y2,
                               const double* __restrict__ R, double sc,
                               double2* __restrict__ I64, float2* __restrict__ I32, int n,
                               ContractLds<NCT>& S) {
  constexpr int MB = contract_rows<NCT>();  // m1 rows per workgroup
  constexpr int MP = 16 * NCT;            // padded columns of Y2
  constexpr int NY = kCT * MP / 256;      // staged Y2 entries per thread
  const int M = 2 * l + 1;
  const int m1a = b * MB;
  const int rows = min(MB, l + 1 - m1a);
  const int nrt = (rows + 15) >> 4;       // complex row tiles of C
  const int nrtA = (2 * rows + 15) >> 4;  // real row tiles of A
  const int nct = (M + 15) >> 4;          // column tiles of C
  // the wave index is read into a scalar register so that the tile
  // conditions below are uniform branches (MFMA needs all lanes active)
  const int t = threadIdx.x, w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int ln = t & 63, col = ln & 15, kq = ln >> 4;
  // first product: wave w owns A tiles (row tile (w >> 1) + 2 qa, column
  // tile w & 1), qa < 2.  Second product: wave w computes the real (w even)
  // or imaginary (w odd) part of complex C tiles (w >> 1) + 2 qb, qb < NCT.
  const int actA = w & 1, part = w & 1;
  v4d cacc[NCT];
#pragma unroll
  for (int q = 0; q < NCT; ++q) cacc[q] = v4d{0.0, 0.0, 0.0, 0.0};
  const int Ng = d.ngath;
  const double2* Y1l = Y1 + d.ylm_off;
  const double2* Y2l = Y2 + d.ylm_off;
  for (int g = 0; g < d.ngroups; ++g) {
    const SphGroup G = grp[d.group_off + g];
    const int Ns = G.size, go = G.gath_off;
    // degree l >= 16 is stored in radial slot l - 13
    const double* Rl = R + d.rad_off + G.rad_off + (long long)(l - 13) * Ns * Ns;
    // staging roles of this thread:
    //   Y1 chunk:  j = j0 + t % kJC, complex rows t / kJC + kYR i, i < kNY1
    //              (each gives the real rows 2r and 2r + 1)
    //   K chunk:   k = k0 + (t & 31), j = j0 + (t >> 5) + 8 i, i < kNK
    //   Y2 tile:   k = k0 + (t & 31), m = (t >> 5) + 8 i, i < NY
    constexpr int kYR = 256 / kJC;          // complex rows per Y1 staging round
    constexpr int kNY1 = MB * kJC / 256;    // Y1 entries staged per thread
    constexpr int kNK = kJC / 8;            // K entries staged per thread
    const int yj = t % kJC, yr0 = t / kJC;
    const int kk_t = t & 31, kj0 = t >> 5;
    const double2* Y1g = Y1l + go;
    const double2* Y2g = Y2l + go;
    for (int k0 = 0; k0 < Ns; k0 += kCT) {
      const int kt = k0 + kk_t;
      const bool kok = kt < Ns;
      double2 sy[kNY1];
      double sk[kNK];
      double2 yv[NY];
#define FO_LOAD_CHUNK(J0)                                                           \
  {                                                                                 \
    const int j_ = (J0) + yj;                                                       \
    _Pragma("unroll") for (int i = 0; i < kNY1; ++i) {                              \
      const int r = yr0 + kYR * i;                                                  \
      sy[i] = (r < rows && j_ < Ns) ? Y1g[qlm(l, m1a + r) * Ng + j_]                \
                                    : make_double2(0.0, 0.0);                       \
    }                                                                               \
    _Pragma("unroll") for (int i = 0; i < kNK; ++i) {                               \
      const int j = (J0) + kj0 + 8 * i;                                             \
      sk[i] = (j < Ns && kok) ? Rl[j * Ns + kt] : 0.0;                              \
    }                                                                               \
  }
#define FO_STORE_CHUNK(BUF)                                                         \
  {                                                                                 \
    _Pragma("unroll") for (int i = 0; i < kNY1; ++i) {                              \
      S.u.a.y1[BUF][2 * (yr0 + kYR * i)][yj] = sy[i].x;                             \
      S.u.a.y1[BUF][2 * (yr0 + kYR * i) + 1][yj] = sy[i].y;                         \
    }                                                                               \
    _Pragma("unroll") for (int i = 0; i < kNK; ++i) S.u.a.kc[BUF][kj0 + 8 * i][kk_t] = sk[i]; \
  }
// Y2 row m2 = m - l: Y_l^{m2} = P2 E2 for m2 >= 0, (-1)^|m2| conj(P2 E2) for m2 < 0.
// Every lane loads (entries outside the tile read element 0 of the group's
// block and are zeroed afterwards), so the loads issue back to back.
#define FO_LOAD_Y2                                                                  \
  {                                                                                 \
    _Pragma("unroll") for (int i = 0; i < NY; ++i) {                                \
      const int m = kj0 + 8 * i;                                                    \
      const bool ok = m < M && kok;                                                 \
      const int am = abs(m - l);                                                    \
      yv[i] = Y2g[ok ? qlm(l, am) * Ng + kt : 0];                                   \
    }                                                                               \
    _Pragma("unroll") for (int i = 0; i < NY; ++i) {                                \
      const int m = kj0 + 8 * i, m2 = m - l;                                        \
      const bool ok = m < M && kok;                                                 \
      const double sgn = (m2 < 0 && (m2 & 1)) ? -1.0 : 1.0;                         \
      yv[i] = ok ? make_double2(sgn * yv[i].x, (m2 < 0 ? -sgn : 1.0) * yv[i].y)     \
                 : make_double2(0.0, 0.0);                                          \
    }                                                                               \
  }
      FO_LOAD_CHUNK(0);
      __syncthreads();  // previous tile's readers of the union are done
      FO_STORE_CHUNK(0);
      __syncthreads();
      v4d a0 = v4d{0.0, 0.0, 0.0, 0.0}, a1 = a0;
      const int nch = (Ns + kJC - 1) / kJC;
      for (int c = 0; c < nch; ++c) {
        const int buf = c & 1;
        if (c + 1 < nch) FO_LOAD_CHUNK((c + 1) * kJC);
        // A tiles of this wave: rows 16 rt, columns 16 actA.  The row-tile
        // conditions are wave-uniform and outside the k loop, so every LDS
        // operand of the chunk is read before the first MFMA waits on one.
        {
          const int rt0 = w >> 1, rt1 = rt0 + 2;
          double bv[kJC / 4], ya[kJC / 4], yb[kJC / 4];
          if (rt0 < nrtA) {
#pragma unroll
            for (int s4 = 0; s4 < kJC / 4; ++s4) {
              const int jj = 4 * s4 + kq;
              bv[s4] = S.u.a.kc[buf][jj][16 * actA + col];
              ya[s4] = S.u.a.y1[buf][16 * rt0 + col][jj];
            }
            if (rt1 < nrtA) {
#pragma unroll
              for (int s4 = 0; s4 < kJC / 4; ++s4)
                yb[s4] = S.u.a.y1[buf][16 * rt1 + col][4 * s4 + kq];
#pragma unroll
              for (int s4 = 0; s4 < kJC / 4; ++s4) {
                a0 = mfma64(ya[s4], bv[s4], a0);
                a1 = mfma64(yb[s4], bv[s4], a1);
              }
            } else {
#pragma unroll
              for (int s4 = 0; s4 < kJC / 4; ++s4) a0 = mfma64(ya[s4], bv[s4], a0);
            }
          }
        }
        if (c + 1 < nch) FO_STORE_CHUNK(buf ^ 1);
        __syncthreads();
      }
      FO_LOAD_Y2;
#undef FO_LOAD_CHUNK
#undef FO_STORE_CHUNK
#undef FO_LOAD_Y2
      // A and Y2 tiles -> LDS (the chunk buffers are free after the barrier)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int rt = (w >> 1) + 2 * h;
        if (rt < nrtA) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int r2 = 16 * rt + 4 * i + kq;
            double* e0 = (double*)&S.u.b.ac[r2 >> 1][16 * actA + col];
            e0[r2 & 1] = h == 0 ? a0[i] : a1[i];
          }
        }
      }
#pragma unroll
      for (int i = 0; i < NY; ++i) S.u.b.ys[kk_t][kj0 + 8 * i] = yv[i];
      __syncthreads();
      // second product, k-steps outer so consecutive MFMAs update
      // different accumulators.  Every k-step reads the operands of all NCT
      // tile slots first (row clamped into the block, so the reads stay in
      // the A tile), then issues the MFMAs of this wave's tiles, ctile =
      // (w >> 1) + 2 qb below nrt * nct (wave-uniform).
      {
        const int ntl = nrt * nct;
#pragma unroll 2
        for (int s8 = 0; s8 < kCT / 4; ++s8) {
          const int kk = 4 * s8 + kq;
          double2 av[NCT], yw[NCT];
#pragma unroll
          for (int qb = 0; qb < NCT; ++qb) {
            const int ctile = (w >> 1) + 2 * qb;
            const int rt = min(ctile / nct, nrt - 1), ct = ctile % nct;
            const int r = 16 * rt + col;
            av[qb] = r < rows ? S.u.b.ac[r][kk] : make_double2(0.0, 0.0);
            yw[qb] = S.u.b.ys[kk][16 * ct + col];
          }
#pragma unroll
          for (int qb = 0; qb < NCT; ++qb) {
            if ((w >> 1) + 2 * qb >= ntl) continue;  // wave-uniform
            // real part: Ar Y2r + Ai Y2i; imaginary part: Ai Y2r - Ar Y2i
            const double a0 = part ? av[qb].y : av[qb].x, a1 = part ? -av[qb].x : av[qb].y;
            cacc[qb] = mfma64(a0, yw[qb].x, cacc[qb]);
            cacc[qb] = mfma64(a1, yw[qb].y, cacc[qb]);
          }
        }
      }
    }
  }
  // C tiles: real parts in waves 0, 2, imaginary parts in waves 1, 3;
  // exchange through LDS so one lane writes both parts of an entry
  __syncthreads();
  double2 (*cb)[kCT + 1] = S.u.b.ac;  // reuse as [MB][kCT + 1] staging per column tile pass
  const double norm = 4.0 * pow(kPi, 2.5) * sc * sc * sc;
  for (int ct = 0; ct < nct; ++ct) {
    // tiles (rt, ct) for rt < nrt: ctile = rt * nct + ct
#pragma unroll
    for (int qb = 0; qb < NCT; ++qb) {
      const int ctile = (w >> 1) + 2 * qb;
      const int rt = ctile / nct, cc = ctile % nct;
      if (rt >= nrt || cc != ct) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 16 * rt + 4 * i + kq;
        double* e = (double*)&cb[r][col];
        e[part] = cacc[qb][i];
      }
    }
    __syncthreads();
    for (int e = t; e < MB * 16; e += 256) {
      const int r = e >> 4, mc = e & 15, m = 16 * ct + mc;
      if (r < rows && m < M) {
        const int m1 = m1a + r, m2 = m - l;
        const long long id = d.coef_off + kidx(l, m1, m2);
        const double2 cv = cb[r][mc];
        const double2 v = make_double2(norm * cv.x, norm * cv.y);
        I64[id] = v;
        I32[d.c32_off + (long long)qlm(l, m1) * n + plane_pos(plane_k(m2, n), n)] =
            make_float2((float)v.x, (float)v.y);
      }
    }
    __syncthreads();
  }
  // Only rows m1 >= 0 are written: the selection and the refinement read
  // nothing else, the m1 < 0 rows of I64 are filled by k_sph_mirror when the
  // coefficients are read out, and the plane-layout entries |m2| > l stay
  // at the zeros the batch was bound with.
}

// NCTMAX: column tiles of the largest degree in the launch (2l+1 <= 16 NCTMAX);
// the launch covers the degrees l_lo..l_hi of every pair.
template <int NCTMAX>
__global__ void __launch_bounds__(256, NCTMAX <= 2 ? 5 : 3) k_sph_contract_mfma(
    const SphPair* __restrict__ pd, const SphGroup* __restrict__ grp,
    const double2* __restrict__ Y1, const double2* __restrict__ Y2,
    const double* __restrict__ R, const double* __restrict__ scale_eff,
    double2* __restrict__ I64, float2* __restrict__ I32, int n, int P, int nbl, int l_lo,
    int l_hi) {
  __shared__ ContractLds<NCTMAX> S;
  int pp, b;
  if (!xcd_decode(P, nbl, pp, b)) return;
  const SphPair d = pd[pp];
  const int L = min(d.L, l_hi);
  int l = l_lo;
  for (; l <= L; ++l) {
    constexpr int MB = contract_rows<NCTMAX>();
    const int nb = (l + 1 + MB - 1) / MB;
    if (b < nb) break;
    b -= nb;
  }
  if (l > L) return;
  const double sc = scale_eff[pp];
  contract_mfma_body<NCTMAX>(d, l, b, grp, Y1, Y2, R, sc, I64, I32, n, S);
}

// Raw buffer access (gfx9 descriptor): a per-lane byte offset plus a
// wavefront-uniform one (scalar register); offsets at or beyond `bytes`
// read zero.
FO_DEV __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}
FO_DEV float4 load_f4(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const u4 w = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return make_float4(__uint_as_float(w.x), __uint_as_float(w.y), __uint_as_float(w.z),
                     __uint_as_float(w.w));
}

FO_DEV double load_d(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  const u2 w = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  return __hiloint2double((int)w.y, (int)w.x);
}
FO_DEV double2 load_d2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const u4 w = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return make_double2(__hiloint2double((int)w.y, (int)w.x), __hiloint2double((int)w.w, (int)w.z));
}

// ---------------------------------------------------------------------------
// Degrees 0..15: radial factors and contraction in one pass.  The radial
// kernel stored K at the top degree h of each block of eight (block 0:
// 0..min(7, L), block 1: 8..min(15, L)) and at h + 1 (slots 2b, 2b + 1).
// One workgroup per (pair, four degrees): quarter s of block b covers
// degrees h - 4s down to max(8b, h - 4s - 3).  It starts from the stored
// pair and descends with the radial kernel's recurrence
// K_{l-1} = K_{l+1} + (2l+1)/x K_l (i_l is the minimal solution, stable
// downwards), so the K_l of degrees <= 15 never pass through HBM.  Entries
// with x < 2^-40, where the stored values of degree 16 can be subnormal
// (x = 0: an atom at the centre), take the power series instead.
// Per species group and tile of KT = 64 atoms k of structure 2:
//   j in chunks of kJL: the Y1 rows of the four degrees are staged in LDS
//   (the next chunk is loaded into registers meanwhile); wave w owns the
//   column tile of atoms k0 + 16 w .. + 15 and every row tile of A (one for
//   block 0, whose 2(l+1) <= 16 real rows fit one, two for block 1), so
//   each lane steps the recurrence of the four (j, k) entries it supplies
//   as the MFMA B operand once for all row tiles, and A_l += Y1'_l K_l
//   accumulates per degree in registers.
//   Then two degrees at a time: their A_l tiles and Y2_l rows -> LDS,
//   C^l += A_l conj(Y2_l)^T as in k_sph_contract_mfma, one real or
//   imaginary part of one column tile per wave (block 0: even degrees on
//   waves 0-1, odd degrees on waves 2-3).
// Each output's reduction order is fixed: j in order inside A, k in order
// inside C, groups in order.
// ---------------------------------------------------------------------------
constexpr double kTinyX = 0x1p-40;
constexpr int kJL = 16;  // atoms j per staged chunk
static_assert(kJL == 16, "the chunk staging maps thread t to atom t & 15 of a chunk");
constexpr int kNS4 = kJL / 4;  // MFMA k-steps per chunk
// 1/(2l+1)!!, l = 0..15
__constant__ double kInvDfact[16] = {
    1.0, 1.0 / 3, 1.0 / 15, 1.0 / 105, 1.0 / 945, 1.0 / 10395, 1.0 / 135135, 1.0 / 2027025,
    1.0 / 34459425, 1.0 / 654729075, 1.0 / 13749310575.0, 1.0 / 316234143225.0,
    1.0 / 7905853580625.0, 1.0 / 213458046676875.0, 1.0 / 6190283353629375.0,
    1.0 / 191898783962510625.0};

template <int BLK>
struct LowCfg {
  static constexpr int RTW = BLK ? 2 : 1;     // real-row tiles of A per degree (per wave)
  static constexpr int KT = 64;               // atoms k per tile: one column tile per wave
  static constexpr int NQ = BLK ? 58 : 26;    // complex Y rows of four degrees (12..15, 4..7)
  static constexpr int AR = BLK ? 16 : 8;     // complex rows m1 <= l
  static constexpr int NMC = BLK ? 32 : 16;   // columns m = m2 + l (2l + 1 < NMC)
  static constexpr int NC = BLK ? 4 : 2;      // C tile parts held per wave
  static constexpr int NYI = (NQ * kJL + 255) / 256;  // staged Y1 entries per thread
  static constexpr int NM = BLK ? 16 : 8;     // rows m2 = 0..l of Y2 (l < NM)
  static constexpr int NYV = KT * NM / 256;           // staged Y2 entries per thread
};

template <int BLK>
struct LowLds {
  using Cf = LowCfg<BLK>;
  union {
    double y1[2 * Cf::NQ][kJL + 1];     // Y1 rows of a j chunk, [real row][j]
    double2 ac[2][Cf::AR][Cf::KT + 1];  // A_l tiles of two degrees [degree][m1][k]
  } u;
  double2 ys[2][Cf::KT][Cf::NM + 1];    // Y2_l rows m2 >= 0 of two degrees [degree][k][m2]
  double rj[kJL];                       // |x1_j| of the chunk's atoms (-1: none)
  double irj[kJL];                      // 1/|x1_j| (0: none)
};

// A copy of v that the compiler cannot see through: lane-dependent index
// arithmetic derived from it stays where it is used instead of being hoisted
// out of the loops (where it would hold registers for the whole loop nest).
FO_DEV int opaque_lane(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <int BLK>
__global__ void __launch_bounds__(256, 2) k_sph_contract_low(
    const SphPair* __restrict__ pd, const SphGroup* __restrict__ grp,
    const double* __restrict__ rad1, const double* __restrict__ rad2,
    const double2* __restrict__ Y1, const double2* __restrict__ Y2,
    const double* __restrict__ R, const double* __restrict__ scale_eff,
    double2* __restrict__ I64, float2* __restrict__ I32, int n, int P) {
  using Cf = LowCfg<BLK>;
  constexpr int KT = Cf::KT;
  constexpr double kBig = 0x1p600;  // stands for 1/r of an atom at the centre
  __shared__ LowLds<BLK> S;
  int pp, qs;
  if (!xcd_decode(P, 2, pp, qs)) return;
  const SphPair d = pd[pp];
  const int lo = 8 * BLK;
  if (lo > d.L) return;
  const int h = min(lo + 7, d.L), top = h - 4 * qs;
  if (top < lo) return;
  const int bot = max(lo, top - 3), nd = top - bot + 1;
  const double sc = scale_eff[pp];
  const double i2s2 = 1.0 / (2.0 * sc * sc), i4s2 = 1.0 / (4.0 * sc * sc), s2x2 = 2.0 * sc * sc;
  const int t = threadIdx.x, w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int ln = t & 63, col = ln & 15, kq = ln >> 4;
  const int ct = w;                             // A column tile of this wave (all row tiles)
  const int part = w & 1;                       // C part: 0 real, 1 imaginary
  const int ct2 = BLK ? (w >> 1) : 0;           // C column tile
  const int q0 = qlm(bot, 0);
  // Every workgroup steps four degrees top - dd; those below bot (dd >= nd,
  // only when L < 15) are computed into accumulators that are never
  // stored, so the degree loops carry no branches.  LDS offset of this
  // lane's A operand rows per degree: real row 16 rt + col = 2 m1 + (Re, Im)
  // of row tile rt (rows m1 > l read row l, degrees below bot read bot:
  // never stored)
  int arow[4][Cf::RTW];
#pragma unroll
  for (int dd = 0; dd < 4; ++dd) {
    const int l = max(top - dd, bot);
#pragma unroll
    for (int rt = 0; rt < Cf::RTW; ++rt)
      arow[dd][rt] = (2 * (qlm(l, min((16 * rt + col) >> 1, l)) - q0) + (col & 1)) * (kJL + 1);
  }
  v4d cacc[Cf::NC];
#pragma unroll
  for (int q = 0; q < Cf::NC; ++q) cacc[q] = v4d{0.0, 0.0, 0.0, 0.0};
  const int Ng = d.ngath;
  // Y1 of the pair (reads past its end return zero); this thread stages
  // row (t >> 4) + 16 i of the quarter and atom t & 15 of each chunk
  const __amdgpu_buffer_rsrc_t ry1 = buf_rsrc(Y1 + d.ylm_off, (int)(plane_rows(d.L) * Ng * 16));
  const int vy1 = ((t >> 4) * Ng + (t & 15)) * 16;
  for (int g = 0; g < d.ngroups; ++g) {
    const SphGroup G = grp[d.group_off + g];
    const int Ns = G.size, go = G.gath_off;
    const long long ss = (long long)Ns * Ns;
    // stored K_h and K_{h+1} of the group ([j][k]; reads past the end: zero)
    const __amdgpu_buffer_rsrc_t rq0 = buf_rsrc(R + d.rad_off + G.rad_off + 2 * BLK * ss, (int)(ss * 8));
    const __amdgpu_buffer_rsrc_t rq1 =
        buf_rsrc(R + d.rad_off + G.rad_off + (2 * BLK + 1) * ss, (int)(ss * 8));
    const double2* Y2g = Y2 + d.ylm_off + go;
    const double* ra1 = rad1 + d.gath_off + go;  // |x| of the group's atoms
    const double* ra2 = rad2 + d.gath_off + go;
    for (int k0 = 0; k0 < Ns; k0 += KT) {
      const int kk = k0 + 16 * ct + col;  // this lane's B operand column
      const bool kok = kk < Ns;
      const double r2 = kok ? ra2[kk] : 0.0;
      // 1/x = (1/r1) (1/(r2 i2s2)): the second factor per lane, 0 for
      // columns beyond the group (their entries stay zero)
      const double ir2 = !kok ? 0.0 : (r2 > 0.0 ? s2x2 / r2 : kBig);
      // byte offset of entry (4 s4 + kq, kk) of a chunk, past the end for
      // columns beyond the group
      const int vq = kok ? (kq * Ns + kk) * 8 : 0x7ff00000;
      v4d acc[4][Cf::RTW];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int rt = 0; rt < Cf::RTW; ++rt) acc[q][rt] = v4d{0.0, 0.0, 0.0, 0.0};
      // registers of the next chunk: Y1 entries, the stored K_h, K_{h+1}
      // of this lane's entries, one radius (-1: no atom)
      double2 yb[Cf::NYI];
      double kh[kNS4], kh1[kNS4], rjv = -1.0;
#define FO_LOW_LOAD_CHUNK(J0)                                                               \
  {                                                                                         \
    const int j0_ = (J0);                                                                   \
    _Pragma("unroll") for (int i = 0; i < Cf::NYI; ++i)                                     \
      yb[i] = load_d2(ry1, vy1 + ((q0 + 16 * i) * Ng + go + j0_) * 16, 0);                  \
    _Pragma("unroll") for (int s4 = 0; s4 < kNS4; ++s4) {                                   \
      const int o_ = vq + (j0_ + 4 * s4) * Ns * 8;                                          \
      kh[s4] = load_d(rq0, o_, 0);                                                          \
      kh1[s4] = load_d(rq1, o_, 0);                                                         \
    }                                                                                       \
    if (t < kJL) rjv = j0_ + t < Ns ? ra1[j0_ + t] : -1.0;                                  \
  }
      const int nch = (Ns + kJL - 1) / kJL;
      FO_LOW_LOAD_CHUNK(0);
      for (int c = 0; c < nch; ++c) {
        const int j0 = c * kJL;
        __syncthreads();  // readers of the previous chunk (or tile) are done
#pragma unroll
        for (int i = 0; i < Cf::NYI; ++i) {
          const int qq = (t >> 4) + 16 * i;
          if (qq < Cf::NQ) {
            S.u.y1[2 * qq][t & 15] = yb[i].x;
            S.u.y1[2 * qq + 1][t & 15] = yb[i].y;
          }
        }
        if (t < kJL) {
          S.rj[t] = rjv;
          S.irj[t] = rjv > 0.0 ? 1.0 / rjv : (rjv == 0.0 ? kBig : 0.0);
        }
        double cur[kNS4], up[kNS4];
#pragma unroll
        for (int s4 = 0; s4 < kNS4; ++s4) {
          cur[s4] = kh[s4];
          up[s4] = kh1[s4];
        }
        __syncthreads();
        if (c + 1 < nch) FO_LOW_LOAD_CHUNK(j0 + kJL)
        // entries (j0 + 4 s4 + kq, kk): cx = 1/x (0 for empty entries);
        // for x < 2^-40 the series, which is x^l/(2l+1)!! e^{-x} g to FP64
        // precision there (its next term is below 2^-81 relative, and
        // e^{-x} = 1 - x): cx = e^{-x} g, and the recurrence values are
        // replaced at every degree
        double cx[kNS4];
        int tiny = 0;
#pragma unroll
        for (int s4 = 0; s4 < kNS4; ++s4) {
          cx[s4] = S.irj[4 * s4 + kq] * ir2;
          if (cx[s4] > 0x1p40) {
            tiny |= 1 << s4;
            const double r1 = S.rj[4 * s4 + kq], x = r1 * r2 * i2s2;
            cx[s4] = (1.0 - x) * exp(-(r1 - r2) * (r1 - r2) * i4s2);
          }
        }
        // degrees above this quarter: recurrence only
        for (int l = h; l > top; --l) {
          const double cl = 2.0 * l + 1.0;
#pragma unroll
          for (int s4 = 0; s4 < kNS4; ++s4) {
            const double nx = up[s4] + cl * cx[s4] * cur[s4];
            up[s4] = cur[s4];
            cur[s4] = nx;
          }
        }
#pragma unroll
        for (int dd = 0; dd < 4; ++dd) {
          const int l = top - dd;
          if (tiny) {
            const int la = max(l, 0);
#pragma unroll
            for (int s4 = 0; s4 < kNS4; ++s4)
              if (tiny & (1 << s4))
                cur[s4] = ipow(S.rj[4 * s4 + kq] * r2 * i2s2, la) * cx[s4] * kInvDfact[la];
          }
          double ya[Cf::RTW][kNS4];
#pragma unroll
          for (int rt = 0; rt < Cf::RTW; ++rt)
#pragma unroll
            for (int s4 = 0; s4 < kNS4; ++s4) ya[rt][s4] = (&S.u.y1[0][0])[arow[dd][rt] + 4 * s4 + kq];
#pragma unroll
          for (int s4 = 0; s4 < kNS4; ++s4)
#pragma unroll
            for (int rt = 0; rt < Cf::RTW; ++rt) acc[dd][rt] = mfma64(ya[rt][s4], cur[s4], acc[dd][rt]);
          const double cl = 2.0 * l + 1.0;
#pragma unroll
          for (int s4 = 0; s4 < kNS4; ++s4) {
            const double nx = up[s4] + cl * cx[s4] * cur[s4];
            up[s4] = cur[s4];
            cur[s4] = nx;
          }
        }
      }
      // second product, two degrees at a time: their Y2_l rows m2 = 0..l
      // (entry k = k0 + e % KT, m2 = e / KT, coalesced in k; the columns
      // m2 < 0 are formed when read, Y_l^{-m} = (-1)^m conj(Y_l^m)) and
      // their A_l tiles -> LDS (the A tiles aliasing the chunk buffer), then
      // C^l += A_l conj(Y2_l)^T
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int tl = opaque_lane(t), cl_ = tl & 15, kq_ = (tl >> 4) & 3;
        double2 yv[2][Cf::NYV];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int dd = 2 * hh + q, l_ = top - dd;
#pragma unroll
          for (int i = 0; i < Cf::NYV; ++i) {
            const int e = tl + 256 * i, k = k0 + e % KT, m2 = e / KT;
            yv[q][i] = (dd < nd && m2 <= l_ && k < Ns) ? Y2g[(long long)qlm(l_, m2) * Ng + k]
                                                        : make_double2(0.0, 0.0);
          }
        }
        __syncthreads();  // readers of the last chunk / the previous pair of degrees are done
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int dd = 2 * hh + q, l = top - dd;
#pragma unroll
          for (int rt = 0; rt < Cf::RTW; ++rt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int r = 16 * rt + 4 * i + kq_, m1 = r >> 1;
              if (m1 <= l) ((double*)&S.u.ac[q][m1][16 * ct + cl_])[r & 1] = acc[dd][rt][i];
            }
#pragma unroll
          for (int i = 0; i < Cf::NYV; ++i) {
            const int e = tl + 256 * i;
            S.ys[q][e % KT][e / KT] = yv[q][i];  // rows m2 > l hold zeros
          }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int dd = 2 * hh + q, l = top - dd;
          if (BLK || (w >> 1) == (dd & 1)) {  // uniform
            v4d& cq = cacc[BLK ? dd : (dd >> 1)];
            // column m = 16 ct2 + col of C: m2 = m - l, read as row |m2|
            // (times -1 for odd negative m2, conjugated for negative m2)
            const int m2 = 16 * ct2 + cl_ - l, am2 = min(abs(m2), Cf::NM - 1);
            const double sr = (m2 < 0 && (m2 & 1)) ? -1.0 : 1.0;
            const double si = m2 < 0 ? -sr : sr;
            const bool rok = cl_ <= l;
            const int ar = min(cl_, Cf::AR - 1);
#pragma unroll 4
            for (int s = 0; s < KT / 4; ++s) {
              const int kk2 = 4 * s + kq_;
              const double2 av = rok ? S.u.ac[q][ar][kk2] : make_double2(0.0, 0.0);
              const double2 yr = S.ys[q][kk2][am2];
              const double a0 = part ? av.y : av.x, a1 = part ? -av.x : av.y;
              cq = mfma64(a0, sr * yr.x, cq);
              cq = mfma64(a1, si * yr.y, cq);
            }
          }
        }
      }
#undef FO_LOW_LOAD_CHUNK
    }
  }
  // C^l of the four degrees: both parts of an entry through LDS, then one
  // lane writes it (FP64 compact and FP32 plane layouts)
  const double norm = 4.0 * pow(kPi, 2.5) * sc * sc * sc;
#pragma unroll
  for (int dd = 0; dd < 4; ++dd) {
    if (dd < nd) {
      const int l = top - dd, M = 2 * l + 1, tl = opaque_lane(t);
      __syncthreads();
      if (BLK || (w >> 1) == (dd & 1)) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = 4 * i + ((tl >> 4) & 3);
          if (r < Cf::AR)
            ((double*)&S.u.ac[0][r][16 * ct2 + (tl & 15)])[part] = cacc[BLK ? dd : (dd >> 1)][i];
        }
      }
      __syncthreads();
      for (int e = tl; e < Cf::AR * Cf::NMC; e += 256) {
        const int r = e / Cf::NMC, mc = e % Cf::NMC;
        if (r <= l && mc < M) {
          const int m2 = mc - l;
          const double2 cv = S.u.ac[0][r][mc];
          const double2 v = make_double2(norm * cv.x, norm * cv.y);
          I64[d.coef_off + kidx(l, r, m2)] = v;
          I32[d.c32_off + (long long)qlm(l, r) * n + plane_pos(plane_k(m2, n), n)] =
              make_float2((float)v.x, (float)v.y);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Inverse SOFT of two Hermitian coefficient sets, one beta plane per
// workgroup (fastoverlap/soft.py:115-125).  In alignment mode the second set
// is the inverted orientation, B^l = (-1)^l A^l, so one beta contraction
// serves both: with E/O the even/odd-l partial sums, S_A = E + O and
// S_B = E - O.  Hermitian symmetry S[-m1,-m2] = conj S[m1,m2] makes the
// grid real, so only m1 >= 0 rows are contracted and transformed:
//   T[m1][g]  = sum_{m2} S[m1][m2] e^{i m2 g}         (gamma pass)
//   out[a][g] = Re T[0][g] + 2 sum_{m1>0} Re(T[m1][g] e^{i m1 a})   (alpha pass)
// Each plane keeps its kCandPerPlane largest values and an upper bound on
// the rest, for the FP64 selection kernel.
// ---------------------------------------------------------------------------
// Inverse DFT of R = 2, 4, 8 or 16 values in registers (radix-2
// decimation in time, natural order in and out), twiddles e^{+2 pi i k/16}.
template <int R>
FO_DEV constexpr int bitrev(int i) {
  int r = 0;
  for (int b = 1; b < R; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

template <int R>
FO_DEV void dft_reg(float2 (&v)[R]) {
  constexpr float C16[16] = {1.0f,
                             0.92387953251128674f,
                             0.70710678118654752f,
                             0.38268343236508977f,
                             0.0f,
                             -0.38268343236508977f,
                             -0.70710678118654752f,
                             -0.92387953251128674f,
                             -1.0f,
                             -0.92387953251128674f,
                             -0.70710678118654752f,
                             -0.38268343236508977f,
                             0.0f,
                             0.38268343236508977f,
                             0.70710678118654752f,
                             0.92387953251128674f};
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int j = bitrev<R>(i);
    if (j > i) {
      const float2 tmp = v[i];
      v[i] = v[j];
      v[j] = tmp;
    }
  }
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
#pragma unroll
    for (int i = 0; i < R; i += len) {
#pragma unroll
      for (int k = 0; k < len / 2; ++k) {
        const int ti = k * (16 / len);
        // a e^{+2 pi i ti/16} (cos = C16[ti], sin = C16[(ti + 12) % 16]);
        // the factors 1, i and the eighth turns are written out, as the
        // compiler may not drop products with 0 (x * 0 can be -0 or NaN)
        const float2 a = v[i + k + len / 2];
        constexpr float c8 = 0.70710678118654752f;
        float2 tt;
        if (ti == 0)
          tt = a;
        else if (ti == 4)
          tt = make_float2(-a.y, a.x);
        else if (ti == 2)
          tt = make_float2(c8 * (a.x - a.y), c8 * (a.x + a.y));
        else if (ti == 6)
          tt = make_float2(-c8 * (a.x + a.y), c8 * (a.x - a.y));
        else
          tt = cmul(a, make_float2(C16[ti], C16[(ti + 12) & 15]));
        const float2 u = v[i + k];
        v[i + k] = make_float2(u.x + tt.x, u.y + tt.y);
        v[i + k + len / 2] = make_float2(u.x - tt.x, u.y - tt.y);
      }
    }
  }
}

// Per-thread record of the largest grid value and a bound on the others,
// per orientation (named fields: no runtime-indexed private arrays).
struct CandOne {
  float bv = -INFINITY, dm = -INFINITY;
  int bi = 0x7fffffff;
  FO_DEV void add(int flat, float val) {
    if (val > bv) {
      dm = fmaxf(dm, bv);
      bv = val;
      bi = flat;
    } else {
      dm = fmaxf(dm, val);
    }
  }
};
struct CandTrack {
  CandOne c0, c1;
  FO_DEV void add(int o, int flat, float val) {
    if (o == 0)
      c0.add(flat, val);
    else
      c1.add(flat, val);
  }
  FO_DEV const CandOne& get(int o) const { return o == 0 ? c0 : c1; }
};

// Ordering of LDS accesses between lanes of one wavefront (no s_barrier).
FO_DEV void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP row shifts (within rows of 16 lanes): row_shl:S gives lane i the value
// of lane i + S, row_shr:S that of lane i - S.
template <int CTRL>
FO_DEV float dpp_mov(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, true));
}

// Records (largest value v, its flat index i, largest of the other values
// s): larger value first, ties to the smaller index; the loser's largest
// value joins the others (so s counts a repeated maximum).
FO_DEV void top2_merge(float& v, int& i, float& s, float ov, int oi, float os) {
  const bool take = ov > v || (ov == v && oi < i);
  s = fmaxf(fmaxf(s, os), take ? v : ov);
  v = take ? ov : v;
  i = take ? oi : i;
}
// Plane candidates of a 256-thread workgroup: wavefront w fills entry w of
// the plane's record with its largest value and the largest of its other
// values (no LDS, no barrier).  One butterfly of (v, i, s) records per
// orientation inside each row of 16 lanes by DPP (quad_perm [1,0,3,2],
// quad_perm [2,3,0,1], row_half_mirror, row_mirror), the orientations
// interleaved, then the four row results through readlane.  Every thread
// must call it.
template <int CTRL>
FO_DEV void top2_dpp_step(float (&v)[2], int (&i)[2], float (&s)[2]) {
  float ov[2], os[2];
  int oi[2];
#pragma unroll
  for (int o = 0; o < 2; ++o) {
    ov[o] = dpp_mov<CTRL>(v[o]);
    oi[o] = __builtin_amdgcn_update_dpp(0, i[o], CTRL, 0xf, 0xf, true);
    os[o] = dpp_mov<CTRL>(s[o]);
  }
#pragma unroll
  for (int o = 0; o < 2; ++o) top2_merge(v[o], i[o], s[o], ov[o], oi[o], os[o]);
}
FO_DEV void plane_candidates(const CandTrack& ct, int norient, PlaneCand* out, int ostride) {
  const int wl = threadIdx.x & 63, w = threadIdx.x >> 6;
  float v[2] = {ct.c0.bv, ct.c1.bv}, s[2] = {ct.c0.dm, ct.c1.dm};
  int i[2] = {ct.c0.bi, ct.c1.bi};
  top2_dpp_step<0xB1>(v, i, s);
  top2_dpp_step<0x4E>(v, i, s);
  top2_dpp_step<0x141>(v, i, s);
  top2_dpp_step<0x140>(v, i, s);
#pragma unroll
  for (int o = 0; o < 2; ++o) {
    float rv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[o]), 0));
    float rs = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s[o]), 0));
    int ri = __builtin_amdgcn_readlane(i[o], 0);
#pragma unroll
    for (int r = 16; r < 64; r += 16)
      top2_merge(rv, ri, rs, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[o]), r)),
                 __builtin_amdgcn_readlane(i[o], r),
                 __int_as_float(__builtin_amdgcn_readlane(__float_as_int(s[o]), r)));
    if (o < norient && wl == 0) {
      PlaneCand* dst = out + (long long)o * ostride;
      dst->val[w] = rv;
      dst->idx[w] = ri;
      dst->rest[w] = rs;
    }
  }
}

// One stage of the lane transpose: lanes i and i ^ S (S < 16, groups
// aligned within a row of 16 lanes) swap the registers whose index bit S
// differs from their own lane bit.
template <int R, int S>
FO_DEV void transpose_stage(float2 (&v)[R], int lane) {
  const bool hi = (lane & S) != 0;
#pragma unroll
  for (int q = 0; q < R / 2; ++q) {
    const int r0 = (q / S) * 2 * S + (q % S), r1 = r0 + S;
    const float2 A = v[r0], B = v[r1];
    // hi lanes receive B of lane i - S, lo lanes A of lane i + S
    const float2 fromlo = make_float2(dpp_mov<0x110 + S>(B.x), dpp_mov<0x110 + S>(B.y));
    const float2 fromhi = make_float2(dpp_mov<0x100 + S>(A.x), dpp_mov<0x100 + S>(A.y));
    v[r0] = hi ? fromlo : A;
    v[r1] = hi ? B : fromhi;
  }
}

// Transpose of R x R values held by R consecutive lanes (lane i, register r)
// -> (lane r, register i): log2 R stages.
template <int R>
FO_DEV void lane_transpose(float2 (&v)[R], int lane) {
  static_assert(R == 2 || R == 4 || R == 8 || R == 16, "radix");
  transpose_stage<R, 1>(v, lane);
  if constexpr (R >= 4) transpose_stage<R, 2>(v, lane);
  if constexpr (R >= 8) transpose_stage<R, 4>(v, lane);
  if constexpr (R >= 16) transpose_stage<R, 8>(v, lane);
}

// Second half of a two-pass transform of length n = N1 N2: lane j1 < N1
// holds the twiddled first-pass outputs v[k2] (k2 < N2); afterwards lane
// k2 < N2 holds the N1 inputs of its length-N1 DFT in u[j1].  Square
// shapes exchange in registers; others through the group's LDS row xr
// (index i stored at i + i / N1, so both access patterns are
// bank-conflict free), ordered within the wavefront.
template <int N1, int N2>
FO_DEV void two_pass_exchange(float2 (&v)[N2], float2 (&u)[N1], int lane, float2* xr) {
  if constexpr (N1 == N2) {
    lane_transpose<N1>(v, lane);
#pragma unroll
    for (int j = 0; j < N1; ++j) u[j] = v[j];
  } else {
    if (lane < N1)
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) {
        const int i = lane + N1 * k2;
        xr[i + i / N1] = v[k2];
      }
    wave_sync();
    if (lane < N2)
#pragma unroll
      for (int j1 = 0; j1 < N1; ++j1) {
        const int i = j1 + N1 * lane;
        u[j1] = xr[i + i / N1];
      }
    wave_sync();
  }
}

template <int N1, int N2>
constexpr int fft_tpf() {
  return N1 > N2 ? N1 : N2;
}

// ---------------------------------------------------------------------------
// Power-of-two grids, n = N1 N2 (N1, N2 <= 16, radix-2 register DFTs).  One
// workgroup (256 threads) per (pair, beta plane), groups of TPF = max(N1,N2)
// lanes, one block barrier between the passes and one for the candidates.
//  gamma pass: group <-> row m1 >= 0 (both orientations).  Lane j1 < N1
//    owns the entries m2 = k (k <= L) or k - n (k >= n - L) of
//    k = j1 + N1 j2, contracts them over l in an l-outer loop (N2
//    independent loads per step), then the row transform: DFT of length N2
//    in registers, twiddle e^{2 pi i j1 k2 / n}, lane exchange, DFT of
//    length N1; output k2 + N2 k1 is stored at T[o][g][m1] (odd stride LS).
//  alpha pass: group <-> column g, the same transform of
//    X[m1] = T_A[m1][g] + i T_B[m1][g] (m1 >= 0) with its Hermitian
//    extension X[-m1] = conj T_A + i conj T_B; Re / Im of the result are
//    the grid values of the two orientations.
// Workgroup w of the 1-D launch serves pair (w % 8) + 8 (w / 8n) and plane
// (w / 8) % n, so all planes of a pair run on one XCD and share its L2.
// ---------------------------------------------------------------------------
// A = sum_l sqrt((2l+1)/2) sum_{m1,m2} |I^l_{m1 m2}| for the selection's
// error bound, from the per-thread partial sums (FP32 magnitudes of the FP32
// coefficients, summed in FP64).  The margin 1 + 2^-16 covers the FP32
// rounding of the coefficients and of the magnitudes, so A stays an upper
// bound of the FP64 sum.  Alignment mode: one set serves both orientations.
// Every thread of the workgroup must call it.
FO_DEV void plane_abound(const double (&asum)[2], bool two_sets, int norient, double* out) {
  __shared__ double sc[8];
  double v[2] = {asum[0], asum[1]};
  block_sum_n<2>(v, sc);
  if (threadIdx.x == 0)
    for (int o = 0; o < norient; ++o) out[o] = (two_sets && o == 1 ? v[1] : v[0]) * (1.0 + 1.0 / 65536.0);
}

// Degree segments of the plane kernel's beta contraction.  Lane k1 owns
// entries k = k1 + N1 k2 (k2 < N2); entry k2 of every lane is zero for all
// degrees below the smallest |m2| of its N1 values.  Segment s covers the
// degrees s N1 .. (s + 1) N1 - 1 (the last one everything above), and its
// mask lists the k2 that are nonzero in some lane at some degree of it.
template <int N1, int N2>
FO_HD constexpr int k2_min_abs_m2(int k2) {
  constexpr int n = N1 * N2;
  const int lo = N1 * k2, hi = lo + N1 - 1;
  return lo < n / 2 ? lo : n - hi;
}
template <int N1, int N2>
FO_HD constexpr unsigned seg_mask(int s) {
  if (s >= N2 / 2 - 1) return (1u << N2) - 1u;
  unsigned m = 0;
  for (int k2 = 0; k2 < N2; ++k2)
    if (k2_min_abs_m2<N1, N2>(k2) <= (s + 1) * N1 - 1) m |= 1u << k2;
  return m;
}
template <int S, int NS, class F>
FO_DEV void static_for(F&& f) {
  if constexpr (S < NS) {
    f(std::integral_constant<int, S>{});
    static_for<S + 1, NS>(f);
  }
}

// FROMS: the beta contraction was done by k_beta_gemm; its even / odd-degree
// sums are read from S (rows m1 <= Srows - 1) instead of contracted here.
template <int N1, int N2, int MODE, bool FROMS>
__global__ void __launch_bounds__(256) k_isoft_fft(
    const SphPair* __restrict__ pd, int P, int C, long long Kp, const float* __restrict__ tab32,
    const float2* __restrict__ twg, const float2* __restrict__ IA,
    const float2* __restrict__ IB, int norient, PlaneCand* __restrict__ cand,
    float* __restrict__ grid, double* __restrict__ abound, const float4* __restrict__ Sg,
    int Srows) {
  constexpr int n = N1 * N2;
  constexpr int TPF = fft_tpf<N1, N2>();
  static_assert(TPF == N1 && N2 % 4 == 0, "lanes own N2 contiguous plane-layout entries");
  constexpr int NG = 256 / TPF;           // lane groups per workgroup
  constexpr int GPW = 64 / TPF;           // lane groups (rows) per wavefront
  constexpr int XS = n + n / N1;          // padded exchange row (non-square)
  extern __shared__ __align__(16) unsigned char smem[];
  // workgroups are dealt to the 8 XCDs round-robin; each XCD walks its
  // pairs in chunks of C, beta-major inside a chunk, so the coefficient
  // blocks of C pairs and a few beta slices of the table share its L2
  const int wg = blockIdx.x;
  const int sx = wg >> 3, chunk = sx / (C * n), rr = sx - chunk * C * n;
  const int ib = rr / C, p = (wg & 7) + 8 * (chunk * C + rr % C);
  if (p >= P) return;  // whole workgroup: padding of P to a multiple of 8 C
  const SphPair d = pd[p];
  const int L = d.L, t = threadIdx.x;
  const int LS = (L + 1) | 1;
  float2* T = (float2*)smem;                              // [2][n][LS]
  float2* X = T + 2 * n * LS;                             // [NG][XS] (non-square)
  const int grp = t / TPF, lane = t % TPF;
  const int wrow = __builtin_amdgcn_readfirstlane(t >> 6) * GPW;  // first row of the wavefront
  float2* xr = X + grp * XS;
  float2 tw[N2];
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) tw[k2] = twg[(lane * k2) % n];
  const int no = norient;
  // plane-layout operands of this pair and beta: rows q(l, m1) of n entries,
  // lane k1's entries k1 + N1*k2 contiguous at N2*k1
  const int rbytes = (int)plane_rows(L) * n;
  const __amdgpu_buffer_rsrc_t rtb = buf_rsrc(tab32 + (long long)ib * Kp, rbytes * 4);
  const __amdgpu_buffer_rsrc_t ria = buf_rsrc(IA + d.c32_off, rbytes * 8);
  const __amdgpu_buffer_rsrc_t rib =
      buf_rsrc(MODE != 0 ? (const void*)(IB + d.c32_off) : (const void*)(IA + d.c32_off), rbytes * 8);
  // the beta = 0 workgroup also sums the coefficient magnitudes for the
  // selection's error bound (each entry is loaded once per workgroup)
  const bool wantA = ib == 0;
  double asum[2] = {0.0, 0.0};

  // ---- gamma pass (rows m1 in rounds of NG)
  for (int m1_0 = 0; m1_0 <= L; m1_0 += NG) {
    const int m1 = m1_0 + grp;
    const bool row = m1 <= L;
    const int m1w = m1_0 + wrow;  // wavefront-uniform
    // a wavefront whose rows all lie above L has nothing to transform (the
    // pass has no block-wide barrier, and the exchange is wavefront-local)
    if (m1w > L) continue;
    // sums over degrees l = m1w, m1w + 2, ... (A) and m1w + 1, m1w + 3, ... (B):
    // the even / odd-l partial sums (alignment mode) in some order, or, in
    // general mode, the two coefficient sets (A, B) over all degrees
    float2 e[N2], od[N2];
    if constexpr (FROMS) {
      const float4* src = Sg + (((long long)p * n + ib) * Srows + (row ? m1 : 0)) * n + lane * N2;
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) {
        const float4 v = row ? src[j2] : make_float4(0.f, 0.f, 0.f, 0.f);
        e[j2] = make_float2(v.x, v.y);
        od[j2] = make_float2(v.z, v.w);
      }
    } else {
    float2 a[N2], bq[N2];
#pragma unroll
    for (int j2 = 0; j2 < N2; ++j2) a[j2] = bq[j2] = make_float2(0.f, 0.f);
    const int vo = (m1 * n + lane * N2) * 4;  // lane's entries in a degree block, table bytes
    // one degree: s += d^l * I^l over the lane's N2 entries, of which only
    // those in MASK can be nonzero (the others are skipped: table and
    // coefficients both hold exact zeros there)
    auto step = [&](auto MC, int l, float2(&sa)[N2], float2(&sb)[N2]) {
      constexpr unsigned MASK = decltype(MC)::value;
      const int so = qlm(l, 0) * n * 4;  // degree block, uniform
      float dv[N2];
      float2 ca[N2];
#pragma unroll
      for (int i = 0; i < N2 / 4; ++i) {
        if (!((MASK >> (4 * i)) & 15u)) continue;
        const float4 w = load_f4(rtb, vo + 16 * i, so);
        dv[4 * i] = w.x;
        dv[4 * i + 1] = w.y;
        dv[4 * i + 2] = w.z;
        dv[4 * i + 3] = w.w;
      }
#pragma unroll
      for (int i = 0; i < N2 / 2; ++i) {
        if (!((MASK >> (2 * i)) & 3u)) continue;
        const float4 w = load_f4(ria, 2 * vo + 16 * i, 2 * so);
        ca[2 * i] = make_float2(w.x, w.y);
        ca[2 * i + 1] = make_float2(w.z, w.w);
      }
      // weight sqrt((2l+1)/2), doubled for m1 > 0 (the m1 < 0 rows)
      const float wl = wantA ? sqrtf(l + 0.5f) * (m1 == 0 ? 1.f : 2.f) : 0.f;
      if (wantA)
#pragma unroll
        for (int j2 = 0; j2 < N2; ++j2)
          if ((MASK >> j2) & 1u) asum[0] += wl * sqrtf(ca[j2].x * ca[j2].x + ca[j2].y * ca[j2].y);
      if constexpr (MODE != 0) {
        float2 cb[N2];
#pragma unroll
        for (int i = 0; i < N2 / 2; ++i) {
          if (!((MASK >> (2 * i)) & 3u)) continue;
          const float4 w = load_f4(rib, 2 * vo + 16 * i, 2 * so);
          cb[2 * i] = make_float2(w.x, w.y);
          cb[2 * i + 1] = make_float2(w.z, w.w);
        }
        if (wantA)
#pragma unroll
          for (int j2 = 0; j2 < N2; ++j2)
            if ((MASK >> j2) & 1u) asum[1] += wl * sqrtf(cb[j2].x * cb[j2].x + cb[j2].y * cb[j2].y);
#pragma unroll
        for (int j2 = 0; j2 < N2; ++j2) {
          if (!((MASK >> j2) & 1u)) continue;
          sa[j2].x += dv[j2] * ca[j2].x;
          sa[j2].y += dv[j2] * ca[j2].y;
          sb[j2].x += dv[j2] * cb[j2].x;
          sb[j2].y += dv[j2] * cb[j2].y;
        }
      } else {
#pragma unroll
        for (int j2 = 0; j2 < N2; ++j2) {
          if (!((MASK >> j2) & 1u)) continue;
          sa[j2].x += dv[j2] * ca[j2].x;
          sa[j2].y += dv[j2] * ca[j2].y;
        }
      }
    };
    // the wavefront's rows start at degree m1w .. m1w + GPW - 1: row m1
    // takes part from l = m1 on; from l = m1w + GPW every row does.  Rows
    // beyond L load nothing, so every access stays inside rows(L) * n.
    // Degrees run in segments of N1 with their masks; m1w is a multiple of
    // GPW, so when GPW <= N1 the first GPW degrees lie in one segment and
    // the degree pairs (l, l + 1) of the alignment mode never straddle one.
    constexpr int NSEG = N2 / 2;
    if (row) {
      const int lp = min(m1w + GPW, L + 1);
      int l = m1w;
      // rows joining: the segment of m1w, or every entry if GPW > N1
      auto head = [&](auto MC) {
        if constexpr (MODE == 0) {
          for (; l < lp; l += 2) {
            if (m1 <= l) step(MC, l, a, a);
            if (l + 1 <= L && m1 <= l + 1) step(MC, l + 1, bq, bq);
          }
        } else {
          for (; l < lp; ++l)
            if (m1 <= l) step(MC, l, a, bq);
        }
      };
      if constexpr (GPW <= N1) {
        const int s0 = min(m1w / N1, NSEG - 1);  // wavefront-uniform
        static_for<0, NSEG>([&](auto S) {
          if (S.value == s0) head(std::integral_constant<unsigned, seg_mask<N1, N2>(S.value)>{});
        });
      } else {
        head(std::integral_constant<unsigned, (1u << N2) - 1u>{});
      }
      // every row: segment by segment
      static_for<0, NSEG>([&](auto S) {
        constexpr unsigned M = seg_mask<N1, N2>(S.value);
        const int send = S.value == NSEG - 1 ? L : min(L, (S.value + 1) * N1 - 1);
        if constexpr (MODE == 0) {
          for (; l <= send; l += 2) {
            step(std::integral_constant<unsigned, M>{}, l, a, a);
            if (l + 1 <= L) step(std::integral_constant<unsigned, M>{}, l + 1, bq, bq);
          }
        } else {
          for (; l <= send; ++l) step(std::integral_constant<unsigned, M>{}, l, a, bq);
        }
      });
    }
    // alignment mode: e = even-l sums, od = odd-l sums
    const bool swap = MODE == 0 && (m1w & 1);
#pragma unroll
    for (int j2 = 0; j2 < N2; ++j2) {
      e[j2] = swap ? bq[j2] : a[j2];
      od[j2] = swap ? a[j2] : bq[j2];
    }
    }  // FROMS
    for (int o = 0; o < no; ++o) {
      float2 v[N2], u[N1];
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) {
        if constexpr (MODE == 0)
          v[j2] = o == 0 ? make_float2(e[j2].x + od[j2].x, e[j2].y + od[j2].y)
                         : make_float2(e[j2].x - od[j2].x, e[j2].y - od[j2].y);
        else
          v[j2] = o == 0 ? e[j2] : od[j2];
      }
      dft_reg<N2>(v);
#pragma unroll
      for (int k2 = 1; k2 < N2; ++k2) v[k2] = cmul(v[k2], tw[k2]);  // tw[0] = 1
      two_pass_exchange<N1, N2>(v, u, lane, xr);
      dft_reg<N1>(u);
      if (row && lane < N2)
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) T[(o * n + lane + N2 * k1) * LS + m1] = u[k1];
    }
  }
  __syncthreads();
  // (FROMS: the error-bound sum comes from k_coef_abound)
  if (!FROMS && wantA) plane_abound(asum, MODE != 0, no, abound + (long long)p * no);

  // ---- alpha pass (columns g in rounds of NG)
  CandTrack ct;
  const float2* TA = T;
  const float2* TB = T + n * LS;
  for (int g0 = 0; g0 < n; g0 += NG) {
    const int g = g0 + grp;
    const bool col = g < n;
    float2 v[N2], u[N1];
    // X[m1] = T_A + i T_B (m1 > 0), conj T_A + i conj T_B (m1 < 0); T[0][g]
    // is real for a Hermitian set, so only its real parts are kept
    const int gr = (col ? g : 0) * LS;
#pragma unroll
    for (int j2 = 0; j2 < N2; ++j2) {
      const int k = lane + N1 * j2;
      const int m1 = k <= L ? k : k - n;
      const bool ok = col && lane < N1 && (k <= L || k >= n - L);
      const int am = ok ? abs(m1) : 0;
      const float s1 = ok ? (float)((m1 > 0) - (m1 < 0)) : 0.f;
      const float2 A = TA[gr + am];
      const float2 B = no == 2 ? TB[gr + am] : make_float2(0.f, 0.f);
      const float x = A.x - s1 * B.y, y = s1 * A.y + B.x;
      v[j2] = ok ? make_float2(x, y) : make_float2(0.f, 0.f);
    }
    dft_reg<N2>(v);
#pragma unroll
    for (int k2 = 1; k2 < N2; ++k2) v[k2] = cmul(v[k2], tw[k2]);  // tw[0] = 1
    two_pass_exchange<N1, N2>(v, u, lane, xr);
    dft_reg<N1>(u);
    if (col && lane < N2) {
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) {
        const int a = lane + N2 * k1;
        const int flat = (a * n + ib) * n + g;
        if (grid) grid[((long long)p * no) * n * n * n + flat] = u[k1].x;
        ct.add(0, flat, u[k1].x);
        if (no == 2) {
          if (grid) grid[((long long)p * no + 1) * n * n * n + flat] = u[k1].y;
          ct.add(1, flat, u[k1].y);
        }
      }
    }
  }
  plane_candidates(ct, no, cand + ((long long)p * no) * n + ib, n);
}

// Direct-DFT planes (any n): the beta contraction per (m1 >= 0, m2) entry,
// then the gamma and alpha passes as direct DFTs through LDS.
__global__ void __launch_bounds__(256) k_isoft_plane(
    const SphPair* __restrict__ pd, int n, long long Kp, const float* __restrict__ tab32,
    const float2* __restrict__ twg, const float2* __restrict__ IA,
    const float2* __restrict__ IB, int mode, int norient, PlaneCand* __restrict__ cand,
    float* __restrict__ grid, double* __restrict__ abound) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int ib = blockIdx.x, p = blockIdx.y;
  const SphPair d = pd[p];
  const int L = d.L, W = 2 * L + 1, t = threadIdx.x, bd = blockDim.x;
  float2* tw = (float2*)smem;                 // [n]
  float2* S = tw + n;                          // [2][L+1][W]
  float2* T = S + 2 * (L + 1) * W;             // [2][L+1][n]
  for (int k = t; k < n; k += bd) tw[k] = twg[k];
  // plane-layout operands: entry (l, m1, m2) at row q(l, m1), position pos
  const float* tb = tab32 + (long long)ib * Kp;
  const float2* ia = IA + d.c32_off;
  const float2* ibb = IB + d.c32_off;
  // beta contraction (the beta = 0 workgroup also sums the coefficient
  // magnitudes for the selection's error bound)
  const bool wantA = ib == 0;
  double asum[2] = {0.0, 0.0};
  for (int e = t; e < (L + 1) * W; e += bd) {
    const int m1 = e / W, m2 = e % W - L;
    const int l0 = max(m1, abs(m2));
    const int pos = plane_pos(plane_k(m2, n), n);
    const float wm = m1 == 0 ? 1.f : 2.f;
    float2 s0 = make_float2(0.f, 0.f), s1 = make_float2(0.f, 0.f);
    if (mode == 0) {
      for (int l = l0; l <= L; ++l) {
        const long long id = (long long)qlm(l, m1) * n + pos;
        const float dv = tb[id];
        const float2 c = ia[id];
        if (wantA) asum[0] += wm * sqrtf(l + 0.5f) * sqrtf(c.x * c.x + c.y * c.y);
        if (l & 1) {
          s1.x += dv * c.x;
          s1.y += dv * c.y;
        } else {
          s0.x += dv * c.x;
          s0.y += dv * c.y;
        }
      }
      S[e] = make_float2(s0.x + s1.x, s0.y + s1.y);
      S[(L + 1) * W + e] = make_float2(s0.x - s1.x, s0.y - s1.y);
    } else {
      for (int l = l0; l <= L; ++l) {
        const long long id = (long long)qlm(l, m1) * n + pos;
        const float dv = tb[id];
        const float2 ca = ia[id], cb = ibb[id];
        if (wantA) {
          asum[0] += wm * sqrtf(l + 0.5f) * sqrtf(ca.x * ca.x + ca.y * ca.y);
          asum[1] += wm * sqrtf(l + 0.5f) * sqrtf(cb.x * cb.x + cb.y * cb.y);
        }
        s0.x += dv * ca.x;
        s0.y += dv * ca.y;
        s1.x += dv * cb.x;
        s1.y += dv * cb.y;
      }
      S[e] = s0;
      S[(L + 1) * W + e] = s1;
    }
  }
  __syncthreads();
  if (wantA) plane_abound(asum, mode != 0, norient, abound + (long long)p * norient);
  CandTrack ct;
  auto consider = [&](int o, int a, int g, float val) {
    const int flat = (a * n + ib) * n + g;
    if (grid) grid[((long long)p * norient + o) * n * n * n + flat] = val;
    ct.add(o, flat, val);
  };
  // gamma pass: T[o][m1][g] = sum_{m2} S[o][m1][m2] e^{i m2 g}
  for (int e = t; e < norient * (L + 1) * n; e += bd) {
    const int o = e / ((L + 1) * n);
    const int rem = e % ((L + 1) * n);
    const int m1 = rem / n, g = rem % n;
    const float2* Sr = S + (o * (L + 1) + m1) * W;
    float2 acc = make_float2(0.f, 0.f);
    int ph = (n - (L % n) * g % n) % n;  // (-L * g) mod n
    for (int m2i = 0; m2i < W; ++m2i) {
      const float2 w = tw[ph];
      const float2 sv = Sr[m2i];
      acc.x += sv.x * w.x - sv.y * w.y;
      acc.y += sv.x * w.y + sv.y * w.x;
      ph += g;
      if (ph >= n) ph -= n;
    }
    T[(o * (L + 1) + m1) * n + g] = acc;
  }
  __syncthreads();
  // alpha pass: out[a][g] = Re T[0][g] + 2 sum_{m1>0} Re(T[m1][g] e^{i m1 a})
  for (int e = t; e < n * n; e += bd) {
    const int a = e / n, g = e % n;
    for (int o = 0; o < norient; ++o) {
      const float2* To = T + o * (L + 1) * n;
      float acc = 0.f;
      int ph = 0;
      for (int m1 = 1; m1 <= L; ++m1) {
        ph += a;
        if (ph >= n) ph -= n;
        const float2 w = tw[ph];
        const float2 v = To[m1 * n + g];
        acc += v.x * w.x - v.y * w.y;
      }
      consider(o, a, g, To[g].x + 2.f * acc);
    }
  }
  plane_candidates(ct, norient, cand + ((long long)p * norient) * n + ib, n);
}

// ---------------------------------------------------------------------------
// FP64 evaluation of the overlap grid at grid points (workgroup cooperative):
//   g(a,b,c) = sum_{m1>=0,m2} w_m1 Re( S64[m1,m2](b) e^{i(m1 alpha + m2 gamma)} )
// with S64 = sum_l sgn^l dtilde64 I64 (alignment mode) or the B set.
// ---------------------------------------------------------------------------
struct GridCtx {
  int n, L;
  long long Kt;
  const double* tab64;
  const double2* tw64;
  const double2* I;   // pair's FP64 coefficients (A set)
  const double2* IB;  // B set (general mode) or nullptr
  double sgn;         // +1 / -1 (alignment mode)
};

// FP64 grid values at up to kEvalBatch points whose beta indices take at most
// kEvalBeta distinct values, in one sweep over the coefficients (each
// coefficient is read once for all points).  Point q: alpha pa[q], gamma
// pg[q], beta slot pb[q] into betas[].  Each point's value is reduced with
// the tree of block_sum.
constexpr int kEvalBatch = 8;
constexpr int kEvalBeta = 4;
constexpr int kSelUnroll = 3;  // degrees per batch of operand loads

// Entry e of the (m1 >= 0, m2) entries of degree <= L in shell order: shell
// l0 = max(m1, |m2|) holds 4 l0 + 1 entries from e = l0 (2 l0 - 1) on
// (m1 = l0, m2 = -l0..l0, then m1 < l0 with m2 = +-l0).  The lanes of a
// wavefront then start their degree recurrences at nearly the same l.
FO_DEV void shell_entry(int e, int& l0, int& m1, int& m2) {
  l0 = (int)((1.0 + sqrt(1.0 + 8.0 * e)) * 0.25);
  while (l0 * (2 * l0 - 1) > e) --l0;
  while ((l0 + 1) * (2 * l0 + 1) <= e) ++l0;
  const int r = e - l0 * (2 * l0 - 1);
  if (r <= 2 * l0) {
    m1 = l0;
    m2 = r - l0;
  } else {
    m1 = (r - 2 * l0 - 1) >> 1;
    m2 = ((r - 2 * l0 - 1) & 1) ? -l0 : l0;
  }
}
// compact index of degree l + 1 from that of degree l (same m1, m2) once the
// entry exists (l >= l0); held at the last existing degree otherwise
FO_DEV int next_kidx(int id, int l, int l0, int m1, int L) {
  return (l >= l0 && l < L) ? id + (2 * l + 1) * (2 * l + 1) + 2 * m1 + 4 * l + 4 : id;
}

// The entries are swept in blocks of 64 (one wavefront) in shell order, so
// the first blocks carry the longest degree loops; a wavefront that finishes
// its block takes the next untaken one (LDS counter ctr), which evens out the
// wavefronts' loop counts.  Each block's sums are reduced over its wavefront
// into part[blk][q] (LDS, (L+1)(2L+1)/64 + 1 rows) and the blocks are added
// in block order, so the result does not depend on which wavefront took
// which block.
FO_DEV void grid_eval64_batch(const GridCtx& c, int np, const int* pa, const int* pg,
                              const int* pb, int nbeta, const int* betas, double* out,
                              double* part, int* ctr) {
  const int L = c.L, W = 2 * L + 1, n = c.n;
  const double2* I = c.IB ? c.IB : c.I;
  const double sg = c.IB ? 1.0 : c.sgn;
  const double* tb[kEvalBeta];
#pragma unroll
  for (int b = 0; b < kEvalBeta; ++b) tb[b] = c.tab64 + (long long)betas[b < nbeta ? b : 0] * c.Kt;
  // entries (m1 >= 0, m2) in shell order (shell_entry), so the degree loop of
  // a wavefront is uniform and the operands of two degrees are loaded at once
  const int ne = (L + 1) * W, nblk = (ne + 63) >> 6;
  const int lane = threadIdx.x & 63, nwv = blockDim.x >> 6;
  if (threadIdx.x == 0) *ctr = nwv;
  __syncthreads();
  for (int blk = threadIdx.x >> 6; blk < nblk;) {  // wavefront-uniform
    const int e = (blk << 6) + lane;
    const bool act = e < ne;
    int l0 = L + 1, m1 = 0, m2 = 0;
    if (act) shell_entry(e, l0, m1, m2);
    double sx[kEvalBeta], sy[kEvalBeta];
#pragma unroll
    for (int b = 0; b < kEvalBeta; ++b) sx[b] = sy[b] = 0.0;
    // The column entries of a shell (m1 < l0, m2 = +-l0) read the table at
    // the transposed entry, d_{m1 m2} = (-1)^{m1-m2} d_{m2 m1}: row m2, so
    // the lanes of a wavefront read consecutive table words (the rows of
    // the compact layout are contiguous, its columns 2l+1 words apart).
    const bool col = m1 < l0;
    const int tm1 = col ? m2 : m1, tm2 = col ? m1 : m2;
    double sl = ((l0 & 1) ? sg : 1.0) * ((col && ((m1 - m2) & 1)) ? -1.0 : 1.0);
    const int lw = __builtin_amdgcn_readfirstlane(l0);  // smallest l0 of the wavefront
    int id = act ? (int)kidx(l0, m1, m2) : 0;
    int td = act ? (int)kidx(l0, tm1, tm2) : 0;
    for (int l = lw; l <= L; l += kSelUnroll) {  // wavefront-uniform
      int ids[kSelUnroll], tds[kSelUnroll];
      ids[0] = id;
      tds[0] = td;
#pragma unroll
      for (int q = 1; q < kSelUnroll; ++q) {
        ids[q] = next_kidx(ids[q - 1], l + q - 1, l0, m1, L);
        tds[q] = next_kidx(tds[q - 1], l + q - 1, l0, tm1, L);
      }
      double2 Iv[kSelUnroll];
      double dv[kSelUnroll][kEvalBeta];
#pragma unroll
      for (int q = 0; q < kSelUnroll; ++q) {
        Iv[q] = I[ids[q]];
#pragma unroll
        for (int b = 0; b < kEvalBeta; ++b) dv[q][b] = b < nbeta ? tb[b][tds[q]] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kSelUnroll; ++q) {
        if (l + q <= L && l + q >= l0) {
#pragma unroll
          for (int b = 0; b < kEvalBeta; ++b) {
            const double d = dv[q][b] * sl;
            sx[b] += d * Iv[q].x;
            sy[b] += d * Iv[q].y;
          }
          sl *= sg;
        }
      }
      id = next_kidx(ids[kSelUnroll - 1], l + kSelUnroll - 1, l0, m1, L);
      td = next_kidx(tds[kSelUnroll - 1], l + kSelUnroll - 1, l0, tm1, L);
    }
    const double wm = m1 == 0 ? 1.0 : 2.0;
    double acc[kEvalBatch];
#pragma unroll
    for (int q = 0; q < kEvalBatch; ++q) {
      acc[q] = 0.0;
      if (q < np) {
        const int b = pb[q];
        double x = sx[0], y = sy[0];
#pragma unroll
        for (int bb = 1; bb < kEvalBeta; ++bb)
          if (b == bb) {
            x = sx[bb];
            y = sy[bb];
          }
        int ph = (m1 * pa[q] + m2 * pg[q]) % n;  // |m1 pa + m2 pg| < 2 L n
        if (ph < 0) ph += n;
        const double2 w = c.tw64[ph];
        acc[q] = wm * (x * w.x - y * w.y);
      }
    }
#pragma unroll
    for (int q = 0; q < kEvalBatch; ++q) {
      const double v = wave_sum(acc[q]);
      if (lane == 0) part[blk * kEvalBatch + q] = v;
    }
    int nb = 0;
    if (lane == 0) nb = atomicAdd(ctr, 1);
    blk = __builtin_amdgcn_readfirstlane(nb);
  }
  __syncthreads();
  if (threadIdx.x < np) {
    double v = 0.0;
    for (int b = 0; b < nblk; ++b) v += part[b * kEvalBatch + threadIdx.x];
    out[threadIdx.x] = v;
  }
}

// Full FP64 evaluation of one beta plane; returns its argmax (ties -> lower
// flat index).  Three stages, each thread-parallel:
//   S[m1][m2] = sum_l s_l d^l_{m1 m2}(beta) I^l_{m1 m2}   (entries in shell
//     order, two degrees of loads in flight, as in grid_eval64_batch),
//   T[m1][g]  = sum_m2 S[m1][m2] e^{i m2 g 2pi/n}          (thread per (m1, g)),
//   v(a, g)   = Re T[0][g] + 2 sum_{m1>=1} Re(T[m1][g] e^{i m1 a 2pi/n})
//     (thread per (g, kPlaneA values of a): each T load serves them all).
// S lives in the LDS buffer sS when (L+1)(2L+1) <= kSelS (L <= 31), else
// in the global scratch Sg; T in the global scratch Tg[(L+1)n]; the n
// twiddles are staged in the LDS buffer stw (n <= 256).  Phases advance by
// integer steps mod n, so every twiddle is a table entry.
constexpr int kSelS = 2048;
constexpr int kPlaneA = 8;
FO_DEV void plane_eval64(const GridCtx& c, int ib, double2* Sg, double2* Tg, double2* sS,
                         double2* stw, double* red, int* redi, double& bestv, int& besti) {
  const int L = c.L, W = 2 * L + 1, n = c.n, t = threadIdx.x, nt = blockDim.x;
  const double* tb = c.tab64 + (long long)ib * c.Kt;
  const double2* I = c.IB ? c.IB : c.I;
  const double sg = c.IB ? 1.0 : c.sgn;
  const bool sl_lds = (L + 1) * W <= kSelS;
  for (int k = t; k < n; k += nt) stw[k] = c.tw64[k];
  // (1) beta sums
  const int ne = (L + 1) * W;
  for (int e0 = 0; e0 < ne; e0 += nt) {
    const int e = e0 + t;
    const bool act = e < ne;
    int l0 = L + 1, m1 = 0, m2 = 0;
    if (act) shell_entry(e, l0, m1, m2);
    const int lw = __builtin_amdgcn_readfirstlane(l0);  // smallest l0 of the wavefront
    double sx = 0.0, sy = 0.0, sl = (l0 & 1) ? sg : 1.0;
    int id = act ? (int)kidx(l0, m1, m2) : 0;
    for (int l = lw; l <= L; l += 2) {  // wavefront-uniform
      int ids[2];
      ids[0] = id;
      ids[1] = next_kidx(id, l, l0, m1, L);
      double2 Iv[2];
      double dv[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        Iv[q] = I[ids[q]];
        dv[q] = tb[ids[q]];
      }
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (l + q <= L && l + q >= l0) {
          const double d = dv[q] * sl;
          sx += d * Iv[q].x;
          sy += d * Iv[q].y;
          sl *= sg;
        }
      }
      id = next_kidx(ids[1], l + 1, l0, m1, L);
    }
    if (act) (sl_lds ? sS : Sg)[m1 * W + m2 + L] = make_double2(sx, sy);
  }
  __threadfence_block();
  __syncthreads();
  // (2) gamma sums
  auto gamma = [&](const double2* S) {
    for (int e = t; e < (L + 1) * n; e += nt) {
      const int m1 = e / n, g = e % n;
      int ph = (L * g) % n;
      ph = ph ? n - ph : 0;  // (-L g) mod n
      const double2* row = S + m1 * W;
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll 4
      for (int m2i = 0; m2i < W; ++m2i) {
        acc = cadd(acc, cmul(row[m2i], stw[ph]));
        ph += g;
        if (ph >= n) ph -= n;
      }
      Tg[e] = acc;
    }
  };
  if (sl_lds)
    gamma(sS);
  else
    gamma(Sg);
  __threadfence_block();
  __syncthreads();
  // (3) alpha sums and the plane's argmax
  double bv = -INFINITY;
  int bi = 0x7fffffff;
  const int nch = (n + kPlaneA - 1) / kPlaneA;
  for (int it = t; it < n * nch; it += nt) {
    const int g = it % n, a0 = (it / n) * kPlaneA;
    double acc[kPlaneA];
    int ph[kPlaneA];
    const double t0 = Tg[g].x;
#pragma unroll
    for (int j = 0; j < kPlaneA; ++j) {
      acc[j] = t0;
      ph[j] = 0;
    }
#pragma unroll 2
    for (int m1 = 1; m1 <= L; ++m1) {
      const double2 v = Tg[m1 * n + g];
#pragma unroll
      for (int j = 0; j < kPlaneA; ++j) {
        ph[j] += a0 + j < n ? a0 + j : 0;  // (values of a >= n are not used)
        if (ph[j] >= n) ph[j] -= n;
        const double2 w = stw[ph[j]];
        acc[j] += 2.0 * (v.x * w.x - v.y * w.y);
      }
    }
#pragma unroll
    for (int j = 0; j < kPlaneA; ++j)
      if (a0 + j < n) argmax_merge(bv, bi, acc[j], ((a0 + j) * n + ib) * n + g);
  }
  block_argmax(bv, bi, red, redi);
  bestv = bv;
  besti = bi;
  __syncthreads();
}

// Selection: exact argmax of the grid of one (pair, orientation).
// delta bounds the FP32 grid error: |g| <= A = sum_l sqrt((2l+1)/2) sum|I^l|,
// and every FP32 stage (coefficient and table rounding, <= L+1 products per
// beta sum, the two DFT passes of <= 2L+1 terms with rounded twiddles) adds
// at most a few units of 2^-24 A each; delta = (8(L+1) + 32) 2^-24 A covers
// their sum.  Every grid point within delta of the FP32 maximum is
// re-evaluated in FP64; planes whose unlisted values could be within delta
// are re-evaluated in full (all_planes: every plane, the exhaustive check).

__global__ void __launch_bounds__(256) k_sph_select(
    const SphPair* __restrict__ pd, int n, long long Kt, const double* __restrict__ tab64,
    const double2* __restrict__ tw64, const double2* __restrict__ IA,
    const double2* __restrict__ IB, int mode, int norient, const PlaneCand* __restrict__ cand,
    double2* __restrict__ scratch, long long scratch_stride, const double* __restrict__ abound,
    SphPeak* __restrict__ peaks, int P, double extra_rel, int all_planes) {
  __shared__ double red[16 * kEvalBatch];
  __shared__ int redi[16];
  __shared__ int s_cnt, s_ovf;
  __shared__ int s_cand[kMaxCand];
  __shared__ int s_ovfl[256];
  __shared__ int s_pa[kEvalBatch], s_pg[kEvalBatch], s_pb[kEvalBatch], s_beta[kEvalBeta];
  __shared__ int s_np, s_nb, s_blk;
  // s_planeS also holds the block sums of grid_eval64_batch (L <= 63)
  static_assert(2 * kSelS >= kEvalBatch * ((64 * 127 + 63) / 64), "block sums");
  __shared__ double s_val[kEvalBatch];
  __shared__ double2 s_planeS[kSelS], s_planeTw[256];  // full-plane evaluation only
  int p, o;
  if (!xcd_decode(P, norient, p, o)) return;
  const SphPair d = pd[p];
  const int L = d.L, t = threadIdx.x;
  GridCtx c;
  c.n = n;
  c.L = L;
  c.Kt = Kt;
  c.tab64 = tab64;
  c.tw64 = tw64;
  c.I = IA + d.coef_off;
  c.IB = (mode == 1 && o == 1) ? IB + d.coef_off : nullptr;
  c.sgn = (mode == 0 && o == 1) ? -1.0 : 1.0;
  const PlaneCand* pc = cand + ((long long)p * norient + o) * n;
  // FP32 maximum over planes
  float mv = -INFINITY;
  int mi = 0x7fffffff;
  for (int b = t; b < n; b += blockDim.x)
    for (int r = 0; r < kCandPerPlane; ++r) argmax_merge(mv, mi, pc[b].val[r], pc[b].idx[r]);
  {
    __shared__ float rv[16];
    __shared__ int ri[16];
    block_argmax(mv, mi, rv, ri);
  }
  // error bound: A = sum_l sqrt((2l+1)/2) sum |I^l| >= every |grid value|,
  // from the plane kernel (Hermitian sets: |I[-m1,-m2]| = |I[m1,m2]|)
  const double A = abound[(long long)p * norient + o];
  // extra_rel: error of the beta contraction beyond FP32 rounding (the
  // split-BF16x3 products: 3.1 * 2^-16 of each |d I|)
  const double delta = ((8.0 * (L + 1) + 32.0) * 5.9604644775390625e-08 + extra_rel) * A;
  // |g32 - g64| <= delta everywhere, so the FP64 maximum has an FP32 value
  // of at least (FP32 maximum) - 2 delta
  const double thr = (double)mv - 2.0 * delta;
  if (t == 0) {
    s_cnt = 0;
    s_ovf = 0;
  }
  __syncthreads();
  for (int b = t; b < n; b += blockDim.x) {
    float rest = pc[b].rest[0];
    for (int r = 1; r < kCandPerPlane; ++r) rest = fmaxf(rest, pc[b].rest[r]);
    if (all_planes || (double)rest >= thr) {
      int q = atomicAdd(&s_ovf, 1);
      s_ovfl[q] = b;
    } else {
      for (int r = 0; r < kCandPerPlane; ++r)
        if ((double)pc[b].val[r] >= thr) {
          int q = atomicAdd(&s_cnt, 1);
          if (q < kMaxCand) s_cand[q] = pc[b].idx[r];
        }
    }
  }
  __syncthreads();
  const int ncand = min(s_cnt, kMaxCand), novf = s_ovf;
  const bool cand_overflow = s_cnt > kMaxCand;
  // sort candidate list (thread 0; small) for a launch-independent order
  if (t == 0) {
    for (int i = 1; i < ncand; ++i) {
      int v = s_cand[i], j = i - 1;
      while (j >= 0 && s_cand[j] > v) {
        s_cand[j + 1] = s_cand[j];
        --j;
      }
      s_cand[j + 1] = v;
    }
    for (int i = 1; i < novf; ++i) {
      int v = s_ovfl[i], j = i - 1;
      while (j >= 0 && s_ovfl[j] > v) {
        s_ovfl[j + 1] = s_ovfl[j];
        --j;
      }
      s_ovfl[j + 1] = v;
    }
  }
  __syncthreads();
  double best = -INFINITY, second = -INFINITY;
  int besti = 0x7fffffff;
  double2* Sg = scratch + ((long long)p * norient + o) * scratch_stride;
  double2* Tg = Sg + (L + 1) * (2 * L + 1);
  // one listed candidate (the common case): it is the argmax, and it is
  // evaluated in the same sweep as its six neighbours below
  const bool merged = !cand_overflow && novf == 0 && ncand == 1;
  if (merged) {
    besti = s_cand[0];
  } else if (cand_overflow) {
    // too many near-maximal points to list: evaluate every plane in FP64
    for (int b = 0; b < n; ++b) {
      double v;
      int i;
      plane_eval64(c, b, Sg, Tg, s_planeS, s_planeTw, red, redi, v, i);
      if (v > best || (v == best && i < besti)) {
        second = best;
        best = v;
        besti = i;
      } else if (v > second) {
        second = v;
      }
    }
  } else {
    // listed candidates in passes of <= kEvalBatch points with <= kEvalBeta
    // distinct beta planes (s_cand is sorted by flat index)
    for (int q0 = 0; q0 < ncand;) {
      if (t == 0) {
        int np = 0, nb = 0, q = q0;
        for (; q < ncand && np < kEvalBatch; ++q) {
          const int f = s_cand[q];
          const int ib = (f / n) % n;
          int slot = -1;
          for (int b = 0; b < nb; ++b)
            if (s_beta[b] == ib) slot = b;
          if (slot < 0) {
            if (nb == kEvalBeta) break;
            slot = nb;
            s_beta[nb++] = ib;
          }
          s_pa[np] = f / (n * n);
          s_pg[np] = f % n;
          s_pb[np] = slot;
          ++np;
        }
        s_np = np;
        s_nb = nb;
      }
      __syncthreads();
      const int np = s_np;
      grid_eval64_batch(c, np, s_pa, s_pg, s_pb, s_nb, s_beta, s_val, (double*)s_planeS, &s_blk);
      __syncthreads();
      for (int q = 0; q < np; ++q) {
        const int f = s_cand[q0 + q];
        const double v = s_val[q];
        if (v > best || (v == best && f < besti)) {
          second = best;
          best = v;
          besti = f;
        } else if (v > second) {
          second = v;
        }
      }
      q0 += np;
      __syncthreads();
    }
    for (int q = 0; q < novf; ++q) {
      double v;
      int i;
      plane_eval64(c, s_ovfl[q], Sg, Tg, s_planeS, s_planeTw, red, redi, v, i);
      if (v > best || (v == best && i < besti)) {
        second = best;
        best = v;
        besti = i;
      } else if (v > second) {
        second = v;
      }
    }
  }
  // parabolic offsets on |g| with periodic wrap (fastoverlap/utils.py:319-338);
  // the six neighbours (after the argmax itself when merged) in one sweep
  const int ind[3] = {besti / (n * n), (besti / n) % n, besti % n};
  const int q0 = merged ? 1 : 0;
  if (t == 0) {
    const int bm = (ind[1] + n - 1) % n, bp = (ind[1] + 1) % n;
    s_beta[0] = ind[1];
    s_beta[1] = bp;
    s_beta[2] = bm;
    s_nb = 3;
    if (merged) {
      s_pa[0] = ind[0];
      s_pg[0] = ind[2];
      s_pb[0] = 0;
    }
    for (int ax = 0; ax < 3; ++ax)
      for (int sd = 0; sd < 2; ++sd) {
        const int q = q0 + 2 * ax + sd;  // sd 0: +1, 1: -1
        const int dlt = sd == 0 ? 1 : n - 1;
        s_pa[q] = ax == 0 ? (ind[0] + dlt) % n : ind[0];
        s_pg[q] = ax == 2 ? (ind[2] + dlt) % n : ind[2];
        s_pb[q] = ax == 1 ? (sd == 0 ? 1 : 2) : 0;
      }
  }
  __syncthreads();
  grid_eval64_batch(c, q0 + 6, s_pa, s_pg, s_pb, s_nb, s_beta, s_val, (double*)s_planeS,
                    &s_blk);
  __syncthreads();
  if (merged) best = s_val[0];
  double off[3];
  for (int ax = 0; ax < 3; ++ax) {
    const double y1 = fabs(s_val[q0 + 2 * ax]);
    const double y3 = fabs(s_val[q0 + 2 * ax + 1]);
    const double y2 = fabs(best);
    off[ax] = (y3 - y1) / (2.0 * (2.0 * y2 - y1 - y3));
  }
  if (t == 0) {
    SphPeak pk;
    for (int ax = 0; ax < 3; ++ax) pk.ind[ax] = ind[ax];
    pk.ncand = cand_overflow ? -1 : ncand;
    pk.noverflow = cand_overflow ? n : novf;
    pk.pad = 0;
    pk.value = best;
    pk.val32 = mv;
    pk.R0[0] = 2.0 * kPi / n * (ind[0] - off[0]);
    pk.R0[1] = kPi / n * (ind[1] - off[1]) + 0.5 * kPi / n;
    pk.R0[2] = 2.0 * kPi / n * (ind[2] - off[2]);
    pk.margin = best - second;
    pk.delta = delta;
    peaks[(long long)p * norient + o] = pk;
  }
}

// ---------------------------------------------------------------------------
// Rotation refinement: minimise f(R) = -Re sum conj(I^l_{m1m2}) D^l_{m1m2}(R),
// D = e^{-i m1 a} d^l_{m1m2}(b) e^{-i m2 y} with the unnormalised d
// (fastoverlap/sphericalAlignment.py:67-113).  The reference runs scipy
// L-BFGS-B from the grid angle; here a damped Newton iteration with the
// exact gradient and Hessian (d, d', d'' by the differentiated recurrence)
// converges to the same local optimum to full FP64 precision.
//
// The recurrence coefficients of the unnormalised d,
//   d^{l+1} = Bu_l (cos b - C_l) d^l - Au_l d^{l-1},
//   Bu_l = B_l sqrt((2l+1)/(2l+3)),  Au_l = A_l sqrt((2l-1)/(2l+3)),
// and the edge prefactor xi sqrt(binom(mu+nu, mu)) are precomputed once per
// context (k_rec_table), so the inner loop has no FP64 square root or
// division.
// ---------------------------------------------------------------------------
__global__ void k_rec_table(int Lt, double* __restrict__ rec, double* __restrict__ seedu) {
  const int W = 2 * Lt + 1;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < W * W; e += gridDim.x * blockDim.x) {
    const int m1 = e / W - Lt, m2 = e % W - Lt;
    const int l0 = max(abs(m1), abs(m2));
    const int mu = abs(m1 - m2), nu = abs(m1 + m2);
    seedu[e] = seed_xi(m1, m2) * exp(seed_log_factor(mu, nu));
    for (int l = l0; l <= Lt; ++l) {
      double A = 0.0, B = 0.0, C = 0.0;
      if (l < Lt) recurrence_terms(l, m1, m2, A, B, C);
      double* r = rec + 3 * kidx(l, m1, m2);
      // A_0 = 0 (no d^{-1} term); the square root is undefined at l = 0
      r[0] = l > 0 ? A * sqrt((2.0 * l - 1.0) / (2.0 * l + 3.0)) : 0.0;
      r[1] = B * sqrt((2.0 * l + 1.0) / (2.0 * l + 3.0));
      r[2] = C;
    }
  }
}

struct FGH {
  double v[10];  // f, fa, fb, fy, faa, fab, fay, fbb, fby, fyy
};

// Recurrence coefficients of the unnormalised d from two small tables in
// LDS instead of the per-entry table: with s_l(m) = sqrt(l^2 - m^2) and
// e_l = s_l(m1) s_l(m2),
//   Bu_l = (l+1)(2l+1) / e_{l+1},  Au_l = (l+1)/l e_l / e_{l+1},
//   C_l = m1 m2 / (l(l+1))        (Au_0 = C_0 = 0),
// the same values as k_rec_table's (to rounding).  st / ist hold s_l(m) and
// 1 / s_l(m) (0 for m >= l) for l, m <= Lp1 - 1, row stride Lp1; al / cl
// hold (l+1)/l and 1/(l(l+1)) (0 at l = 0).
struct RecLds {
  const double* st;
  const double* ist;
  const double* al;
  const double* cl;
  int Lp1;
};
FO_DEV size_t rec_lds_doubles(int Lp1) { return 2 * (size_t)Lp1 * Lp1 + 2 * (size_t)Lp1; }
FO_DEV RecLds build_rec_lds(double* t, int Lp1) {
  double *st = t, *ist = t + Lp1 * Lp1, *al = ist + Lp1 * Lp1, *cl = al + Lp1;
  for (int e = threadIdx.x; e < Lp1 * Lp1; e += blockDim.x) {
    const int l = e / Lp1, m = e % Lp1;
    const double v = m < l ? sqrt((double)(l * l - m * m)) : 0.0;
    st[e] = v;
    ist[e] = m < l ? 1.0 / v : 0.0;
  }
  for (int l = threadIdx.x; l < Lp1; l += blockDim.x) {
    al[l] = l > 0 ? (double)(l + 1) / l : 0.0;
    cl[l] = l > 0 ? 1.0 / ((double)l * (l + 1)) : 0.0;
  }
  return RecLds{st, ist, al, cl, Lp1};
}

// `ph`: LDS for 3L + 2 phase factors; `part`: LDS for (kFghBlocks + 1) * 10
// doubles; `ctr`: an LDS int.
constexpr int kEvalUnroll = 2;  // degrees per batch of operand loads
constexpr int kFghBlocks = (64 * 64 + 63 + 63) / 64;  // item blocks at L = 63
template <bool kLdsRec = false>
FO_DEV void eval_fgh(const double2* __restrict__ I, double sgn, int L, double a, double b,
                     double y, const double* __restrict__ rec, const double* __restrict__ seedu,
                     int Lt, FGH& out, double* part, double2* ph, int* ctr,
                     RecLds rl = RecLds{}) {
  const int W = 2 * L + 1, Wt = 2 * Lt + 1;
  const double cb = cos(b), sb = sin(b), c2 = cos(0.5 * b), s2 = sin(0.5 * b);
  // phase factors e^{-i m1 a} (m1 = 0..L) and e^{-i m2 y} (m2 = -L..L)
  for (int q = threadIdx.x; q < 3 * L + 2; q += blockDim.x) {
    double sp, cp;
    if (q <= L)
      sincos(-(q * a), &sp, &cp);
    else
      sincos(-((q - L - 1 - L) * y), &sp, &cp);
    ph[q] = make_double2(cp, sp);
  }
  if (threadIdx.x == 0) *ctr = blockDim.x >> 6;
  __syncthreads();
  // The sum runs over the entries (m1 >= 0, m2).  d and its b-derivatives
  // obey d_{m1,m2} = (-1)^{m1-m2} d_{m2,m1} = d_{-m2,-m1}, so one recurrence
  // serves two entries: item (l0, m2), |m2| <= l0 (the row m1 = l0 of shell
  // l0), also carries (m2, l0) for 0 <= m2 < l0 (sign (-1)^{l0-m2}) or
  // (-m2, -l0) for -l0 < m2 < 0; the entry (0, -l0), l0 >= 1, is an item of
  // its own.  Items in shell order: shell l0 >= 1 starts at item
  // l0^2 + l0 - 1 with its 2 l0 + 1 row items, then (0, -l0); so the lanes
  // of a wavefront start their recurrences at nearly the same l and the loop
  // over l is wavefront-uniform.  Blocks of 64 items (one wavefront): the
  // first blocks carry the longest loops, and a wavefront that finishes its
  // block takes the next untaken one (LDS counter ctr).  Each block's sums
  // are reduced over its wavefront into part[blk][10] (LDS, kFghBlocks
  // rows) and the blocks are added in block order, so the result does not
  // depend on which wavefront took which block.
  const int ne = (L + 1) * (L + 1) + L, nblk = (ne + 63) >> 6;
  const int lane = threadIdx.x & 63;
  for (int blk = threadIdx.x >> 6; blk < nblk;) {  // wavefront-uniform
    const int e = (blk << 6) + lane;
    const bool act = e < ne;
    int l0 = L + 1, m1 = 0, m2 = 0;
    bool row = false;  // an item of the row m1 = l0
    if (act && e == 0) {
      l0 = 0;
      row = true;
    } else if (act) {
      l0 = (int)((sqrt(4.0 * e + 5.0) - 1.0) * 0.5);
      while (l0 * l0 + l0 - 1 > e) --l0;
      while ((l0 + 1) * (l0 + 1) + l0 <= e) ++l0;
      const int r = e - (l0 * l0 + l0 - 1);
      row = r <= 2 * l0;
      m1 = row ? l0 : 0;
      m2 = row ? r - l0 : -l0;
    }
    const bool hasp = row && m2 > -l0 && m2 < l0;
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int pm1 = m2 >= 0 ? m2 : -m2, pm2 = m2 >= 0 ? l0 : -l0;
    const int mu = abs(m1 - m2), nu = abs(m1 + m2);
    // edge element g = s2^mu c2^nu and its b-derivatives
    const double fct = seedu[(m1 + Lt) * Wt + (m2 + Lt)];
    const double smu = ipow(s2, mu), cnu = ipow(c2, nu);
    const double smu1 = mu ? ipow(s2, mu - 1) : 0.0, cnu1 = nu ? ipow(c2, nu - 1) : 0.0;
    const double smu2 = mu > 1 ? ipow(s2, mu - 2) : 0.0, cnu2 = nu > 1 ? ipow(c2, nu - 2) : 0.0;
    const double g0 = smu * cnu;
    const double g1 = 0.5 * (mu * smu1 * cnu * c2 - nu * smu * s2 * cnu1);
    // g'' = 1/4 [mu(mu-1) s^{mu-2} c^{nu+2} - (mu(nu+1) + nu(mu+1)) s^mu c^nu
    //            + nu(nu-1) s^{mu+2} c^{nu-2}]
    const double g2 = 0.25 * (mu * (mu - 1) * smu2 * cnu * c2 * c2 -
                              (mu * (nu + 1.0) + nu * (mu + 1.0)) * smu * cnu +
                              nu * (nu - 1) * smu * s2 * s2 * cnu2);
    double c0x = 0, c0y = 0, c1x = 0, c1y = 0, c2x = 0, c2y = 0;
    double q0x = 0, q0y = 0, q1x = 0, q1y = 0, q2x = 0, q2y = 0;  // partner
    double d0 = fct * g0, d1 = fct * g1, dd2 = fct * g2;
    double p0 = 0, p1 = 0, p2 = 0;
    double sl = (l0 & 1) ? sgn : 1.0;
    auto step = [&](int l, const double2& Iv, const double2& Pv, double Au, double Bu, double C) {
      // conj(sgn^l I) * d
      const double ix = sl * Iv.x, iy = -sl * Iv.y;
      const double jx = sl * Pv.x, jy = -sl * Pv.y;
      c0x += d0 * ix;
      c0y += d0 * iy;
      c1x += d1 * ix;
      c1y += d1 * iy;
      c2x += dd2 * ix;
      c2y += dd2 * iy;
      q0x += d0 * jx;
      q0y += d0 * jy;
      q1x += d1 * jx;
      q1y += d1 * jy;
      q2x += dd2 * jx;
      q2y += dd2 * jy;
      sl *= sgn;
      if (l == L) return;
      const double bc = Bu * (cb - C);
      const double n0 = bc * d0 - Au * p0;
      const double n1 = bc * d1 - Bu * sb * d0 - Au * p1;
      const double n2 = bc * dd2 - Bu * (cb * d0 + 2.0 * sb * d1) - Au * p2;
      p0 = d0;
      p1 = d1;
      p2 = dd2;
      d0 = n0;
      d1 = n1;
      dd2 = n2;
    };
    // operands (coefficients of the item and its partner, recurrence terms
    // to l + 1) of kEvalUnroll degrees at a time, all loads issued before
    // the steps.  The loaded degree is max(l, l0) clamped to L (terms of
    // degree L are not used).
    auto next_id = [&](int id, int l) { return next_kidx(id, l, l0, m1, L); };
    // smallest l0 of the wavefront (the items (0, -l0) that follow the
    // shells can share a wavefront with shell L)
    int lmin = l0;
    for (int off = 32; off > 0; off >>= 1) lmin = min(lmin, __shfl_xor(lmin, off, 64));
    const int lw = __builtin_amdgcn_readfirstlane(lmin);
    int id = act ? (int)kidx(l0, m1, m2) : 0;
    int pid = hasp ? (int)kidx(l0, pm1, pm2) : 0;
    for (int l = lw; l <= L; l += kEvalUnroll) {  // wavefront-uniform
      int ids[kEvalUnroll], pids[kEvalUnroll];
      double2 Iv[kEvalUnroll], Pv[kEvalUnroll];
      double Au[kEvalUnroll], Bu[kEvalUnroll], C[kEvalUnroll];
      ids[0] = id;
      pids[0] = pid;
#pragma unroll
      for (int q = 1; q < kEvalUnroll; ++q) {
        ids[q] = next_id(ids[q - 1], l + q - 1);
        pids[q] = hasp ? next_kidx(pids[q - 1], l + q - 1, l0, pm1, L) : 0;
      }
#pragma unroll
      for (int q = 0; q < kEvalUnroll; ++q) {
        Iv[q] = I[ids[q]];
        Pv[q] = I[pids[q]];
        if constexpr (kLdsRec) {
          // degree lq = l + q (wavefront-uniform); lanes below their first
          // degree take harmless values (their steps are skipped)
          const int lq = min(l + q, L), am1 = m1, am2 = m2 < 0 ? -m2 : m2;
          const double* s0 = rl.st + lq * rl.Lp1;
          const double* i1 = rl.ist + (lq + 1) * rl.Lp1;
          const double ie = i1[am1] * i1[am2];  // 1 / e_{l+1}
          const double el = s0[am1] * s0[am2];  // e_l
          Bu[q] = (double)((lq + 1) * (2 * lq + 1)) * ie;
          Au[q] = rl.al[lq] * el * ie;
          C[q] = (double)(m1 * m2) * rl.cl[lq];
        } else {
          Au[q] = rec[3 * ids[q]];
          Bu[q] = rec[3 * ids[q] + 1];
          C[q] = rec[3 * ids[q] + 2];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the loads in degree order
      }
#pragma unroll
      for (int q = 0; q < kEvalUnroll; ++q)
        if (l + q <= L && l + q >= l0) step(l + q, Iv[q], Pv[q], Au[q], Bu[q], C[q]);
      id = next_id(ids[kEvalUnroll - 1], l + kEvalUnroll - 1);
      pid = hasp ? next_kidx(pids[kEvalUnroll - 1], l + kEvalUnroll - 1, l0, pm1, L) : 0;
    }
    // the item's entry (m1, m2) and its partner (pm1, pm2): weight 2 for
    // m1 > 0 (the rows m1 < 0 by Hermitian symmetry), the partner's sign
    const double psg = (m2 >= 0 && ((l0 - m2) & 1)) ? -1.0 : 1.0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int a1 = k ? pm1 : m1, a2 = k ? pm2 : m2;
      const double w = k ? (hasp ? (a1 == 0 ? psg : 2.0 * psg) : 0.0) : (a1 == 0 ? 1.0 : 2.0);
      const double x0 = k ? q0x : c0x, y0 = k ? q0y : c0y, x1 = k ? q1x : c1x;
      const double y1 = k ? q1y : c1y, x2 = k ? q2x : c2x, y2 = k ? q2y : c2y;
      const double2 ea = ph[a1], ey = ph[2 * L + 1 + a2];
      const double cp = ea.x * ey.x - ea.y * ey.y, sp = ea.x * ey.y + ea.y * ey.x;
      // Re(c E) and Re(-i c E) = Im(c E)
      const double r0 = x0 * cp - y0 * sp, i0 = x0 * sp + y0 * cp;
      const double r1 = x1 * cp - y1 * sp, i1 = x1 * sp + y1 * cp;
      const double r2 = x2 * cp - y2 * sp;
      const double M1 = a1, M2 = a2;
      acc[0] -= w * r0;
      acc[1] -= w * M1 * i0;          // d/da: Re(-i m1 c E) = m1 Im(cE)
      acc[2] -= w * r1;
      acc[3] -= w * M2 * i0;
      acc[4] -= w * (-M1 * M1 * r0);
      acc[5] -= w * M1 * i1;
      acc[6] -= w * (-M1 * M2 * r0);
      acc[7] -= w * r2;
      acc[8] -= w * M2 * i1;
      acc[9] -= w * (-M2 * M2 * r0);
    }
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      const double v = wave_sum(acc[q]);
      if (lane == 0) part[blk * 10 + q] = v;
    }
    int nb = 0;
    if (lane == 0) nb = atomicAdd(ctr, 1);
    blk = __builtin_amdgcn_readfirstlane(nb);
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    double v = 0.0;
    for (int b = 0; b < nblk; ++b) v += part[b * 10 + threadIdx.x];
    part[kFghBlocks * 10 + threadIdx.x] = v;
  }
  __syncthreads();
  // the sums are uniform over the workgroup: keep them in scalar registers
#pragma unroll
  for (int q = 0; q < 10; ++q) {
    const long long b = __double_as_longlong(part[kFghBlocks * 10 + q]);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    out.v[q] = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

FO_DEV bool chol3_solve(const double H[6], double lam, const double g[3], double x[3]) {
  // H = [haa hab hay; hab hbb hby; hay hby hyy] stored (aa,ab,ay,bb,by,yy)
  double a00 = H[0] + lam, a10 = H[1], a20 = H[2], a11 = H[3] + lam, a21 = H[4], a22 = H[5] + lam;
  if (!(a00 > 0)) return false;
  const double l00 = sqrt(a00), l10 = a10 / l00, l20 = a20 / l00;
  const double t11 = a11 - l10 * l10;
  if (!(t11 > 0)) return false;
  const double l11 = sqrt(t11), l21 = (a21 - l20 * l10) / l11;
  const double t22 = a22 - l20 * l20 - l21 * l21;
  if (!(t22 > 0)) return false;
  const double l22 = sqrt(t22);
  // solve L L^T x = -g
  const double z0 = -g[0] / l00, z1 = (-g[1] - l10 * z0) / l11, z2 = (-g[2] - l20 * z0 - l21 * z1) / l22;
  x[2] = z2 / l22;
  x[1] = (z1 - l21 * x[2]) / l11;
  x[0] = (z0 - l10 * x[1] - l20 * x[2]) / l00;
  return true;
}

// Damped Newton iteration on f(R) from the grid peak: one evaluation site
// (the start, then trial points with backtracking: the step is halved at
// most 8 times), at most 40 accepted steps.
struct NewtonState {
  double x[3], xt[3], dx[3], cur[10];
  double step, slope;
  int nev, it, ls, have, stop;
};

FO_DEV void newton_init(NewtonState& s, const double R0[3]) {
  for (int q = 0; q < 3; ++q) {
    s.x[q] = s.xt[q] = R0[q];
    s.dx[q] = 0.0;
  }
  s.step = 1.0;
  s.slope = 0.0;
  s.nev = s.it = s.ls = s.have = s.stop = 0;
}

// Consume the evaluation r at s.xt; sets s.stop or the next s.xt.
FO_DEV void newton_update(NewtonState& s, const FGH& r) {
  ++s.nev;
  if (s.have) {
    if (!(r.v[0] <= s.cur[0] + 1e-4 * s.step * s.slope)) {
      // reject: halve the step
      if (++s.ls >= 8) {
        s.stop = 1;
        return;
      }
      s.step *= 0.5;
      for (int q = 0; q < 3; ++q) s.xt[q] = s.x[q] + s.step * s.dx[q];
      return;
    }
    for (int q = 0; q < 3; ++q) s.x[q] = s.xt[q];
    ++s.it;
  }
  for (int q = 0; q < 10; ++q) s.cur[q] = r.v[q];
  s.have = 1;
  if (s.it >= 40) {
    s.stop = 1;
    return;
  }
  const double fscale = fabs(s.cur[0]) + 1e-300;
  const double g[3] = {s.cur[1], s.cur[2], s.cur[3]};
  const double gn = fmax(fabs(g[0]), fmax(fabs(g[1]), fabs(g[2])));
  if (gn <= 1e-12 * fscale) {
    s.stop = 1;
    return;
  }
  // Newton direction with Levenberg damping until positive definite
  const double H[6] = {s.cur[4], s.cur[5], s.cur[6], s.cur[7], s.cur[8], s.cur[9]};
  const double hs = fmax(fabs(H[0]), fmax(fabs(H[3]), fmax(fabs(H[5]), 1e-30)));
  double dx[3];
  bool ok = chol3_solve(H, 0.0, g, dx);
  const bool pure = ok;  // undamped Newton step (positive definite Hessian)
  for (int k = 0; !ok && k < 12; ++k) ok = chol3_solve(H, hs * pow(10.0, -8.0 + k), g, dx);
  if (!ok)
    for (int q = 0; q < 3; ++q) dx[q] = -g[q] / hs;
  for (int q = 0; q < 3; ++q) s.dx[q] = dx[q];
  s.slope = g[0] * dx[0] + g[1] * dx[1] + g[2] * dx[2];
  const double dxn = fmax(fabs(dx[0]), fmax(fabs(dx[1]), fabs(dx[2])));
  if (-0.5 * s.slope <= 1e-13 * fscale || dxn <= 1e-12) {
    s.stop = 1;
    return;
  }
  // an undamped Newton step this short lands within O(|dx|^2) (~1e-8 rad)
  // of the optimum: take it without evaluating there, with the objective of
  // the quadratic model (error O(|dx|^3), ~1e-12 of f).  The rotation only
  // seeds the assignment; the distance comes from the quaternion fit.
  if (pure && dxn <= 1e-4) {
    for (int q = 0; q < 3; ++q) s.x[q] += dx[q];
    s.cur[0] += 0.5 * s.slope;
    ++s.it;
    s.stop = 1;
    return;
  }
  s.step = 1.0;
  s.ls = 0;
  for (int q = 0; q < 3; ++q) s.xt[q] = s.x[q] + dx[q];
}

// Converged when the gradient, the Newton model decrease -g.dx/2 or the step
// reaches the level where rounding noise of f (a sum of ~K(L) terms)
// dominates: |g| <= 1e-12 |f|, decrease <= 1e-13 |f|, step <= 1e-12 rad.
// The line search halves at most 8 times.
__global__ void __launch_bounds__(256, 3) k_sph_rotopt(const SphPair* __restrict__ pd,
                                                    const double2* __restrict__ IA,
                                                    const double2* __restrict__ IB, int mode,
                                                    const SphPeak* __restrict__ peaks, int norient,
                                                    const double* __restrict__ seedu, int Lt,
                                                    int Lmax, SphOpt* __restrict__ opt, int P) {
  __shared__ double part[(kFghBlocks + 1) * 10];
  __shared__ double2 ph[3 * 64 + 2];
  __shared__ int s_blk;
  extern __shared__ __align__(16) double rtab[];  // rec_lds_doubles(Lmax + 2)
  int p, o;
  if (!xcd_decode(P, norient, p, o)) return;
  const SphPair d = pd[p];
  const int L = d.L;
  const RecLds rl = build_rec_lds(rtab, Lmax + 2);
  __syncthreads();
  const double sgn = (mode == 0 && o == 1) ? -1.0 : 1.0;
  const double2* I = ((mode == 1 && o == 1) ? IB : IA) + d.coef_off;
  // the iteration state lives in LDS, updated by thread 0 between
  // evaluations, so no register holds it across an evaluation
  __shared__ NewtonState ns;
  if (threadIdx.x == 0) {
    const SphPeak pk = peaks[(long long)p * norient + o];
    newton_init(ns, pk.R0);
  }
  __syncthreads();
  for (;;) {
    FGH r;
    eval_fgh<true>(I, sgn, L, ns.xt[0], ns.xt[1], ns.xt[2], nullptr, seedu, Lt, r, part, ph,
                   &s_blk, rl);
    if (threadIdx.x == 0) newton_update(ns, r);
    __syncthreads();
    if (ns.stop) break;
  }
  if (threadIdx.x == 0) {
    SphOpt r;
    r.R[0] = ns.x[0];
    r.R[1] = ns.x[1];
    r.R[2] = ns.x[2];
    r.f = ns.cur[0];
    // gradient at the last evaluated point
    r.gnorm = fmax(fabs(ns.cur[1]), fmax(fabs(ns.cur[2]), fabs(ns.cur[3])));
    r.iters = ns.it;
    r.evals = ns.nev;
    opt[(long long)p * norient + o] = r;
  }
}

// ---------------------------------------------------------------------------
// Finish: orientation choice (inverted wins on a strictly smaller objective,
// sphericalAlignment.py:184-187), X2 <- (+-X2) EulerM(R) (:124), per-group
// LAP on squared distances with rows = X1 (utils.py:49-56, :82-167), and the
// quaternion fit (utils.py:169-275): X2out = X2R[perm] MR^T (:126-127).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_sph_finish(
    const SphPair* __restrict__ pd, const SphGroup* __restrict__ grp, const int* __restrict__ gath,
    const double* __restrict__ X1c, const double* __restrict__ X2c, const SphOpt* __restrict__ opt,
    int norient, int lap_nmax, int pos_lds, double* __restrict__ X2R, int* __restrict__ perm,
    int* __restrict__ rowsol, double* __restrict__ X2out, SphResult* __restrict__ res) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[16];
  __shared__ int redi[16];
  const int p = blockIdx.x, t = threadIdx.x;
  const SphPair d = pd[p];
  const int N = d.natoms;
  int inv = 0;
  if (norient == 2 && opt[(long long)p * 2 + 1].f < opt[(long long)p * 2].f) inv = 1;
  const SphOpt ob = opt[(long long)p * norient + inv];
  double M[9];
  euler_matrix(ob.R[0], ob.R[1], ob.R[2], M);
  const double sg = inv ? -1.0 : 1.0;
  const double* x1 = X1c + 3LL * d.atom_off;
  const double* x2 = X2c + 3LL * d.atom_off;
  double* xr = X2R + 3LL * d.atom_off;
  int* pm = perm + d.atom_off;
  for (int a = t; a < N; a += blockDim.x) {
    const double u = sg * x2[3 * a], v = sg * x2[3 * a + 1], w = sg * x2[3 * a + 2];
    // row vector times M
    xr[3 * a] = u * M[0] + v * M[3] + w * M[6];
    xr[3 * a + 1] = u * M[1] + v * M[4] + w * M[7];
    xr[3 * a + 2] = u * M[2] + v * M[5] + w * M[8];
    pm[a] = a;
  }
  __threadfence_block();
  __syncthreads();
  LapShared ls = lap_carve(smem, lap_nmax);
  // the group's positions in gathered order, staged in LDS when they fit
  double* g1 = (double*)(smem + ((lap_lds_bytes(lap_nmax) + 15) & ~(size_t)15));
  double* g2 = g1 + 3 * lap_nmax;
  int* rs = rowsol + d.atom_off;
  for (int q = 0; q < d.ngroups; ++q) {
    const SphGroup G = grp[d.group_off + q];
    const int* ga = gath + d.gath_off + G.gath_off;
    if (pos_lds) {
      for (int e = t; e < 3 * G.size; e += blockDim.x) {
        const int i = e / 3, k = e - 3 * i;
        g1[e] = x1[3 * ga[i] + k];
        g2[e] = xr[3 * ga[i] + k];
      }
      __syncthreads();
    }
    auto cost = [&](int i, int j) {
      const double* A = pos_lds ? g1 + 3 * i : x1 + 3 * ga[i];
      const double* B = pos_lds ? g2 + 3 * j : xr + 3 * ga[j];
      const double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
      return dx * dx + dy * dy + dz * dz;
    };
    lap_solve(G.size, cost, ls, rs, red, redi);
    for (int i = t; i < G.size; i += blockDim.x) pm[ga[i]] = ga[rs[i]];
    __threadfence_block();
    __syncthreads();
  }
  // quaternion fit on all atoms, by the first wavefront (its sums do not
  // depend on the block size)
  __shared__ double sMR[9];
  __shared__ double sdist;
  if (t < 64) {
    double c[6] = {0, 0, 0, 0, 0, 0};
    for (int a = t; a < N; a += 64)
      for (int k = 0; k < 3; ++k) {
        c[k] += x1[3 * a + k];
        c[3 + k] += xr[3 * pm[a] + k];
      }
    for (int k = 0; k < 6; ++k) c[k] = wave_sum(c[k]) / (double)N;
    double q[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = t; a < N; a += 64) {
      const double* B = xr + 3 * pm[a];
      qmat_accum(q, x1[3 * a] - c[0], x1[3 * a + 1] - c[1], x1[3 * a + 2] - c[2], B[0] - c[3],
                 B[1] - c[4], B[2] - c[5]);
    }
    for (int k = 0; k < 10; ++k) q[k] = wave_sum(q[k]);
    if (t == 0) {
      double MR[9], dist;
      kearsley_solve(q, &dist, MR);
      for (int k = 0; k < 9; ++k) sMR[k] = MR[k];
      sdist = dist;
    }
  }
  __syncthreads();
  double* xo = X2out + 3LL * d.atom_off;
  for (int a = t; a < N; a += blockDim.x) {
    const double* B = xr + 3 * pm[a];
    for (int i = 0; i < 3; ++i) xo[3 * a + i] = B[0] * sMR[3 * i] + B[1] * sMR[3 * i + 1] + B[2] * sMR[3 * i + 2];
  }
  if (t == 0) {
    SphResult r;
    r.dist = sdist;
    // total map for row vectors: x -> sg * x M MR^T
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += M[3 * i + k] * sMR[3 * j + k];
        r.rmat[3 * i + j] = sg * s;
      }
    r.R[0] = ob.R[0];
    r.R[1] = ob.R[1];
    r.R[2] = ob.R[2];
    r.inverted = inv;
    r.status = 0;
    res[p] = r;
  }
}

}  // namespace fo

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace fo {

// ---------------------------------------------------------------------------
// The beta contraction as a GEMM on the matrix cores, split-BF16x3 with FP32
// accumulation.  For each row m1 >= 0 and plane entry (m2), over a tile of 16
// pairs and every beta:
//   S_par[beta][pair] = sum_{l = par (mod 2)} dn^l_{m1m2}(beta) C^l_{m1m2}[pair]
// i.e. (beta x l) . (l x pairs) per (m1, m2) -- the Wigner-d x coefficient
// product of fastoverlap/soft.py:119-123 batched over pairs.  par = even /
// odd degrees: the two orientations' sums are E + O and E - O.  Each FP32
// operand x is split x = xh + xl (xh = bf16(x), xl = bf16(x - xh)); a b is
// taken as ah bh + ah bl + al bh (three v_mfma_f32_16x16x16bf16, products
// exact in FP32), which leaves at most 3.1 * 2^-16 |a b| per term.
// Output S[pair][beta][m1][pos] = (E.re, E.im, O.re, O.im), rows m1 <= Lmax
// (the batch's largest degree), in the plane layout's entry order.
// MFMA lane layout (16x16x16): A lane -> (row lane%16, k 4(lane/16)..+3),
// B lane -> (k 4(lane/16)..+3, column lane%16), result register r of a lane
// -> (row 4(lane/16) + r, column lane%16).
// ---------------------------------------------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

FO_DEV unsigned short bf16_rne(float x) {
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
FO_DEV void split_bf16(float x, short& hi, short& lo) {
  const unsigned short h = bf16_rne(x);
  hi = (short)h;
  lo = (short)bf16_rne(x - __uint_as_float((unsigned)h << 16));
}
FO_DEV v4f mfma_bf16(v4s a, v4s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

constexpr int kGemmPos = 8;  // plane entries per wavefront

// one wavefront per (tile of 16 pairs, row m1, kGemmPos plane entries)
__global__ void __launch_bounds__(64) k_beta_gemm(const SphPair* __restrict__ pd, int P, int n,
                                                  int Lmax, long long Kp, int Lt,
                                                  const float* __restrict__ tab32,
                                                  const float2* __restrict__ IA,
                                                  float4* __restrict__ S) {
  const int lane = threadIdx.x, col = lane & 15, kq = lane >> 4;
  const int nblk = n / kGemmPos;
  int w = blockIdx.x;
  const int pb = w % nblk;
  w /= nblk;
  const int m1 = w % (Lmax + 1);
  const int tile = w / (Lmax + 1);
  // this lane's pair: B operand column and output column
  const int pj = tile * 16 + col;
  const bool pok = pj < P;
  const SphPair d = pd[pok ? pj : 0];
  const int Lp = pok ? d.L : -1;
  const int n1 = plane_n1(n), n2 = n / n1;
  const int nbt = n / 16;
  const int R = Lmax + 1;
  for (int q = 0; q < kGemmPos; ++q) {
    const int pos = kGemmPos * pb + q;
    const int k = (pos % n2) * n1 + pos / n2;  // inverse of plane_pos
    const int m2 = k < n / 2 ? k : k - n;
    const int l0 = max(m1, abs(m2));
    // coefficients of this lane's pair (B), split once per (parity, chunk)
    v4s bh[2][2][2], bl[2][2][2];  // [parity][chunk][re, im]
    const int nch = l0 <= Lmax ? (Lmax - l0) / 32 + 1 : 0;  // <= 2 for Lmax <= 63
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
      for (int kc = 0; kc < 2; ++kc) {
        const int lf = l0 + ((l0 + par) & 1) + 32 * kc;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int l = lf + 2 * (4 * kq + e);
          float2 c = make_float2(0.f, 0.f);
          if (kc < nch && l <= Lp) c = IA[d.c32_off + (long long)qlm(l, m1) * n + pos];
          short h, lo;
          split_bf16(c.x, h, lo);
          bh[par][kc][0][e] = h;
          bl[par][kc][0][e] = lo;
          split_bf16(c.y, h, lo);
          bh[par][kc][1][e] = h;
          bl[par][kc][1][e] = lo;
        }
      }
    for (int bt = 0; bt < nbt; ++bt) {
      v4f acc[2][2];
#pragma unroll
      for (int par = 0; par < 2; ++par) acc[par][0] = acc[par][1] = v4f{0.f, 0.f, 0.f, 0.f};
      const int beta = bt * 16 + col;  // A operand row of this lane
#pragma unroll
      for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int kc = 0; kc < 2; ++kc) {
          if (kc >= nch) continue;  // wavefront-uniform
          const int lf = l0 + ((l0 + par) & 1) + 32 * kc;
          v4s ah, al;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int l = lf + 2 * (4 * kq + e);
            float a = 0.f;
            if (l <= Lmax && l <= Lt)
              a = tab32[(long long)beta * Kp + (long long)qlm(l, m1) * n + pos];
            short h, lo;
            split_bf16(a, h, lo);
            ah[e] = h;
            al[e] = lo;
          }
#pragma unroll
          for (int ri = 0; ri < 2; ++ri) {
            acc[par][ri] = mfma_bf16(ah, bh[par][kc][ri], acc[par][ri]);
            acc[par][ri] = mfma_bf16(ah, bl[par][kc][ri], acc[par][ri]);
            acc[par][ri] = mfma_bf16(al, bh[par][kc][ri], acc[par][ri]);
          }
        }
      if (pok)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int b = bt * 16 + 4 * kq + r;
          S[(((long long)pj * n + b) * R + m1) * n + pos] =
              make_float4(acc[0][0][r], acc[0][1][r], acc[1][0][r], acc[1][1][r]);
        }
    }
  }
}

// error bound input of the selection for the GEMM arm: A = sum_l
// sqrt((2l+1)/2) sum |C^l| (rows m1 > 0 count twice), one workgroup per pair
__global__ void __launch_bounds__(256) k_coef_abound(const SphPair* __restrict__ pd, int n,
                                                     const float2* __restrict__ IA, int norient,
                                                     double* __restrict__ abound) {
  const SphPair d = pd[blockIdx.x];
  const long long tot = plane_rows(d.L) * n;
  double s = 0.0;
  for (long long e = threadIdx.x; e < tot; e += blockDim.x) {
    const int q = (int)(e / n);
    int l = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
    while (qlm(l + 1, 0) <= q) ++l;
    while (qlm(l, 0) > q) --l;
    const int m1 = q - qlm(l, 0);
    const float2 c = IA[d.c32_off + e];
    const float wl = sqrtf(l + 0.5f) * (m1 == 0 ? 1.f : 2.f);
    s += wl * sqrtf(c.x * c.x + c.y * c.y);
  }
  __shared__ double sc[8];
  double v[1] = {s};
  block_sum_n<1>(v, sc);
  if (threadIdx.x == 0)
    for (int o = 0; o < norient; ++o)
      abound[(long long)blockIdx.x * norient + o] = v[0] * (1.0 + 1.0 / 65536.0);
}

}  // namespace fo


namespace fo {

static inline int cdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

void launch_wigner_table(hipStream_t s, int n, int Lt, double* tab64, float* tab32) {
  const long long total = (long long)n * (2 * Lt + 1) * (2 * Lt + 1);
  const int blocks = (int)std::min<long long>(cdiv(total, 256), 65535);
  hipLaunchKernelGGL(k_wigner_table, dim3(blocks), dim3(256), 0, s, n, Lt, tab64, tab32);
}

void launch_twiddles(hipStream_t s, int n, double sign, double2* tw64, float2* tw32) {
  hipLaunchKernelGGL(k_twiddles, dim3(cdiv(n, 256)), dim3(256), 0, s, n, sign, tw64, tw32);
}

void launch_sph_prep(hipStream_t s, int P, const SphPair* pd, const double* pos1,
                     const double* pos2, double* X1c, double* X2c, double* scale_eff,
                     int center) {
  hipLaunchKernelGGL(k_sph_prep, dim3(P), dim3(256), 0, s, pd, pos1, pos2, X1c, X2c, scale_eff,
                     center);
}

void launch_sph_ylm(hipStream_t s, int P, const SphPair* pd, const int* gath, const double* X1c,
                    const double* X2c, const double2* leg, const double* cm, double2* Y1,
                    double2* Y2, double* rad1, double* rad2) {
  hipLaunchKernelGGL(k_sph_ylm, dim3(P, 2), dim3(256), 0, s, pd, gath, X1c, X2c, leg, cm, Y1, Y2,
                     rad1, rad2);
}

void launch_sph_radial(hipStream_t s, int P, int Lmax, long long maxjk, const SphPair* pd,
                       const SphGroup* grp, const double* rad1, const double* rad2,
                       const double* scale_eff, double* R, int bx_cap) {
  // workgroups per pair: one per 256 (j, k) entries up to bx_cap (the rest
  // is a grid-stride loop); every entry is computed the same way
  const int bx = std::max(1, std::min(cdiv(maxjk, 256), bx_cap > 0 ? bx_cap : 1024));
  (void)Lmax;  // one kernel for every degree
  hipLaunchKernelGGL(k_sph_radial, dim3(bx, P), dim3(256), 0, s, pd, grp, rad1, rad2, scale_eff,
                     R);
}

void launch_sph_contract(hipStream_t s, int P, int n, int Lmax, const SphPair* pd,
                         const SphGroup* grp, const double* rad1, const double* rad2,
                         const double2* Y1, const double2* Y2, const double* R,
                         const double* scale_eff, double2* I64, float2* I32) {
  // degrees 0..15: radial recurrence and contraction fused, one workgroup
  // per (pair, block of eight degrees); degrees >= 16: one launch per
  // column-tile class, workgroups per (pair, degree, block of 32 rows)
  const dim3 b(256), gl(xcd_grid(P, 2));
  hipLaunchKernelGGL(k_sph_contract_low<0>, gl, b, 0, s, pd, grp, rad1, rad2, Y1, Y2, R,
                     scale_eff, I64, I32, n, P);
  if (Lmax >= 8)
    hipLaunchKernelGGL(k_sph_contract_low<1>, gl, b, 0, s, pd, grp, rad1, rad2, Y1, Y2, R,
                       scale_eff, I64, I32, n, P);
  auto launch = [&](int K, int lo, int hi) {
    hi = std::min(hi, Lmax);
    if (lo > hi) return;
    int nb = 0;
    const int MB = K <= 2 ? contract_rows<2>() : contract_rows<4>();
    for (int l = lo; l <= hi; ++l) nb += (l + 1 + MB - 1) / MB;
    const dim3 g(xcd_grid(P, nb));
    if (K == 4)
      hipLaunchKernelGGL(k_sph_contract_mfma<4>, g, b, 0, s, pd, grp, Y1, Y2, R, scale_eff, I64,
                         I32, n, P, nb, lo, hi);
    else if (K == 6)
      hipLaunchKernelGGL(k_sph_contract_mfma<6>, g, b, 0, s, pd, grp, Y1, Y2, R, scale_eff, I64,
                         I32, n, P, nb, lo, hi);
    else
      hipLaunchKernelGGL(k_sph_contract_mfma<8>, g, b, 0, s, pd, grp, Y1, Y2, R, scale_eff, I64,
                         I32, n, P, nb, lo, hi);
  };
  // column-tile classes: 2l+1 <= 64 / 96 / 128 (the six-tile class holds two
  // waves per SIMD and two workgroups per CU, the eight-tile class one)
  launch(4, 16, 31);
  launch(6, 32, 47);
  launch(8, 48, 63);
}

// two-pass factorisation n = N1 N2 of the power-of-two grids (0: direct DFT)
static bool fft_shape(int n, int& n1, int& n2) {
  n1 = plane_n1(n);
  n2 = n1 ? n / n1 : 0;
  return n1 != 0;
}

size_t isoft_plane_smem(int n, int Lmax, int use_fft) {
  int n1, n2;
  if (use_fft && fft_shape(n, n1, n2)) {
    const size_t ls = (size_t)((Lmax + 1) | 1);
    const int tpf = std::max(n1, n2);
    const size_t x = n1 == n2 ? 0 : (size_t)(256 / tpf) * (n + n / n1);
    return sizeof(float2) * (2 * (size_t)n * ls + x);
  }
  return sizeof(float2) * ((size_t)n + 2 * (size_t)(Lmax + 1) * (2 * Lmax + 1) +
                           2 * (size_t)(Lmax + 1) * n);
}

// 1 if the two-pass transform applies to n and fits max_lds, else 0.
int isoft_plane_use_fft(int n, int Lmax, size_t max_lds) {
  int n1, n2;
  return fft_shape(n, n1, n2) && isoft_plane_smem(n, Lmax, 1) <= max_lds ? 1 : 0;
}

void launch_isoft_plane(hipStream_t s, int P, int n, int Lmax, int use_fft, long long Kp,
                        const SphPair* pd, const float* tab32, const float2* tw32,
                        const float2* IA, const float2* IB, int mode, int norient,
                        PlaneCand* cand, float* grid, double* abound, int chunk,
                        const float4* S, int Srows) {
  const size_t sm = isoft_plane_smem(n, Lmax, use_fft);
  int n1, n2;
  if (!use_fft || !fft_shape(n, n1, n2)) {
    hipLaunchKernelGGL(k_isoft_plane, dim3(n, P), dim3(256), sm, s, pd, n, Kp, tab32, tw32, IA,
                       IB, mode, norient, cand, grid, abound);
    return;
  }
  // pairs per chunk of one XCD (workgroups of a chunk walk its pairs
  // beta-major); 16 unless chosen by the autotuner
  const int C = std::max(1, std::min(chunk > 0 ? chunk : 16, (P + 7) / 8));
  const dim3 g((unsigned)(8 * ((P + 8 * C - 1) / (8 * C)) * C * n)), b(256);
#define FO_FFT(N1, N2)                                                                         \
  if (S)                                                                                        \
    hipLaunchKernelGGL((k_isoft_fft<N1, N2, 0, true>), g, b, sm, s, pd, P, C, Kp, tab32, tw32,  \
                       IA, IB, norient, cand, grid, abound, S, Srows);                          \
  else if (mode == 0)                                                                           \
    hipLaunchKernelGGL((k_isoft_fft<N1, N2, 0, false>), g, b, sm, s, pd, P, C, Kp, tab32, tw32, \
                       IA, IB, norient, cand, grid, abound, S, Srows);                          \
  else                                                                                          \
    hipLaunchKernelGGL((k_isoft_fft<N1, N2, 1, false>), g, b, sm, s, pd, P, C, Kp, tab32, tw32, \
                       IA, IB, norient, cand, grid, abound, S, Srows)
  switch (n) {
    case 16: FO_FFT(4, 4); break;
    case 32: FO_FFT(8, 4); break;
    case 64: FO_FFT(8, 8); break;
    case 128: FO_FFT(16, 8); break;
    default: FO_FFT(16, 16); break;
  }
#undef FO_FFT
}

void launch_beta_gemm(hipStream_t s, int P, int n, int Lmax, long long Kp, int Lt,
                      const SphPair* pd, const float* tab32, const float2* IA, float4* S,
                      int norient, double* abound) {
  const int tiles = (P + 15) / 16;
  const long long wgs = (long long)tiles * (Lmax + 1) * (n / kGemmPos);
  hipLaunchKernelGGL(k_beta_gemm, dim3((unsigned)wgs), dim3(64), 0, s, pd, P, n, Lmax, Kp, Lt,
                     tab32, IA, S);
  hipLaunchKernelGGL(k_coef_abound, dim3(P), dim3(256), 0, s, pd, n, IA, norient, abound);
}

void launch_sph_select(hipStream_t s, int P, int n, long long Kt, int norient, const SphPair* pd,
                       const double* tab64, const double2* tw64, const double2* IA,
                       const double2* IB, int mode, const PlaneCand* cand, double2* scratch,
                       long long scratch_stride, const double* abound, SphPeak* peaks,
                       double extra_rel, int all_planes) {
  hipLaunchKernelGGL(k_sph_select, dim3(xcd_grid(P, norient)), dim3(256), 0, s, pd, n, Kt, tab64,
                     tw64, IA, IB, mode, norient, cand, scratch, scratch_stride, abound, peaks, P,
                     extra_rel, all_planes);
}

void launch_sph_rotopt(hipStream_t s, int P, int norient, int Lmax, const SphPair* pd,
                       const double2* IA, const double2* IB, int mode, const SphPeak* peaks,
                       const double* seedu, int Lt, SphOpt* opt) {
  const size_t smem = sizeof(double) * (2 * (size_t)(Lmax + 2) * (Lmax + 2) + 2 * (Lmax + 2));
  hipLaunchKernelGGL(k_sph_rotopt, dim3(xcd_grid(P, norient)), dim3(256), smem, s, pd, IA, IB,
                     mode, peaks, norient, seedu, Lt, Lmax, opt, P);
}

void launch_sph_finish(hipStream_t s, int P, int lap_nmax, const SphPair* pd, const SphGroup* grp,
                       const int* gath, const double* X1c, const double* X2c, const SphOpt* opt,
                       int norient, double* X2R, int* perm, int* rowsol, double* X2out,
                       SphResult* res) {
  // one wavefront per pair: the assignment search is a chain of small
  // reductions, cheapest without cross-wave barriers, until the largest
  // group reaches 192 atoms: then four, whose column scans are shorter (the
  // assignment and the fit do not depend on the block size)
  // positions of the largest group in LDS beside the assignment arrays when
  // that stays within 64 KiB
  const size_t lap = (lap_lds_bytes(lap_nmax) + 15) & ~(size_t)15;
  const size_t pos = 6 * sizeof(double) * (size_t)lap_nmax;
  const int pos_lds = lap + pos <= 65536 ? 1 : 0;
  const int nthr = lap_nmax >= 192 ? 256 : 64;
  hipLaunchKernelGGL(k_sph_finish, dim3(P), dim3(nthr), lap + (pos_lds ? pos : 0), s, pd, grp,
                     gath, X1c, X2c, opt, norient, lap_nmax, pos_lds, X2R, perm, rowsol, X2out,
                     res);
}

}  // namespace fo

// ---------------------------------------------------------------------------
// Single-call helpers behind the drop-in class methods.
// ---------------------------------------------------------------------------
namespace fo {

// FP32 plane-layout copy of compact FP64 coefficients, one thread per
// (row q(l, m1), k) entry of pair blockIdx.y (zero where |m2| > l)
// Rows m1 < 0 of the compact C^l from the Hermitian identity
// I^l_{-m1,-m2} = (-1)^{m1+m2} conj(I^l_{m1,m2}) (the coefficient sets of
// real densities; fastoverlap/sphericalAlignment.py:269-273 builds them in
// full).  One thread per (l, m1 > 0, m2) entry of a pair.
__global__ void k_sph_mirror(const SphPair* __restrict__ pd, double2* __restrict__ I) {
  const SphPair d = pd[blockIdx.y];
  const int L = d.L;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  // entries (l, m1 > 0, m2) of degree l: l (2l + 1); prefix sum_{l' < l} = (l-1) l (4l+1) / 6
  if (e >= (long long)L * (L + 1) * (4 * L + 5) / 6) return;
  int l = (int)cbrt(1.5 * (double)e);
  while (l > 0 && (long long)(l - 1) * l * (4 * l + 1) / 6 > e) --l;
  while ((long long)l * (l + 1) * (4 * l + 5) / 6 <= e) ++l;
  const int r = (int)(e - (long long)(l - 1) * l * (4 * l + 1) / 6);
  const int m1 = 1 + r / (2 * l + 1), m2 = r % (2 * l + 1) - l;
  const double2 v = I[d.coef_off + kidx(l, m1, m2)];
  const double sg = ((m1 + m2) & 1) ? -1.0 : 1.0;
  I[d.coef_off + kidx(l, -m1, -m2)] = make_double2(sg * v.x, -sg * v.y);
}

__global__ void k_coef_to_plane(const SphPair* __restrict__ pd, int n,
                                const double2* __restrict__ a, float2* __restrict__ b) {
  const SphPair d = pd[blockIdx.y];
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= plane_rows(d.L) * n) return;
  const int row = (int)(e / n), k = (int)(e % n);
  int l = (int)((sqrt(8.0 * row + 1.0) - 1.0) * 0.5);
  while (qlm(l + 1, 0) <= row) ++l;
  while (qlm(l, 0) > row) --l;
  const int m1 = row - qlm(l, 0);
  const int m2 = k < n / 2 ? k : k - n;
  float2 v = make_float2(0.f, 0.f);
  if (abs(m2) <= l) {
    const double2 c = a[d.coef_off + kidx(l, m1, m2)];
    v = make_float2((float)c.x, (float)c.y);
  }
  b[d.c32_off + (long long)row * n + plane_pos(k, n)] = v;
}

// getEnergyGradient (fastoverlap/sphericalAlignment.py:98-100) for one
// coefficient set given as conj(X) in compact layout.
__global__ void __launch_bounds__(256) k_sph_energy(int L, const double2* __restrict__ Iconj,
                                                    double a, double b, double y,
                                                    const double* __restrict__ rec,
                                                    const double* __restrict__ seedu, int Lt,
                                                    double* __restrict__ out) {
  __shared__ double part[(kFghBlocks + 1) * 10];
  __shared__ double2 ph[3 * 64 + 2];
  __shared__ int s_blk;
  FGH r;
  eval_fgh(Iconj, 1.0, L, a, b, y, rec, seedu, Lt, r, part, ph, &s_blk);
  if (threadIdx.x == 0)
    for (int q = 0; q < 10; ++q) out[q] = r.v[q];
}

// calcWignerMatrices(rot) (fastoverlap/sphericalAlignment.py:67-96):
// D = e^{-i m1 a} d e^{-i m2 y} (unnormalised d) and its gradient, dense
// (l, m1 mod W, m2 mod W) complex.
__global__ void k_wigner_D(int L, double a, double b, double y, double2* __restrict__ D,
                           double2* __restrict__ G) {
  const int W = 2 * L + 1;
  const double cb = cos(b), sb = sin(b), c2 = cos(0.5 * b), s2 = sin(0.5 * b);
  const long long plane = (long long)(L + 1) * W * W;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < W * W; e += gridDim.x * blockDim.x) {
    const int m1 = e / W - L, m2 = e % W - L;
    const int l0 = max(abs(m1), abs(m2));
    const int mu = abs(m1 - m2), nu = abs(m1 + m2);
    const double fct = sqrt((2.0 * l0 + 1.0) * 0.5) * seed_xi(m1, m2) * exp(seed_log_factor(mu, nu));
    double d0 = fct * ipow(s2, mu) * ipow(c2, nu);
    double d1 = fct * 0.5 *
                ((mu ? mu * ipow(s2, mu - 1) * ipow(c2, nu + 1) : 0.0) -
                 (nu ? nu * ipow(s2, mu + 1) * ipow(c2, nu - 1) : 0.0));
    double p0 = 0.0, p1 = 0.0;
    double sa, ca, sy, cy;
    sincos(-(double)m1 * a, &sa, &ca);
    sincos(-(double)m2 * y, &sy, &cy);
    const double2 ph = cmul(make_double2(ca, sa), make_double2(cy, sy));
    for (int l = 0; l <= L; ++l) {
      const long long id = (long long)l * W * W + (long long)((m1 + W) % W) * W + (m2 + W) % W;
      if (l < l0) {
        D[id] = make_double2(0.0, 0.0);
        for (int k = 0; k < 3; ++k) G[k * plane + id] = make_double2(0.0, 0.0);
        continue;
      }
      const double un = sqrt(2.0 / (2.0 * l + 1.0));
      const double2 Dv = cscale(ph, un * d0);
      D[id] = Dv;
      G[id] = make_double2(m1 * Dv.y, -m1 * Dv.x);                // -i m1 D
      G[plane + id] = cscale(ph, un * d1);
      G[2 * plane + id] = make_double2(m2 * Dv.y, -m2 * Dv.x);    // -i m2 D
      if (l == L) break;
      double A, B, C;
      recurrence_terms(l, m1, m2, A, B, C);
      const double n0 = B * (cb - C) * d0 - A * p0;
      const double n1 = B * (-sb * d0 + (cb - C) * d1) - A * p1;
      p0 = d0;
      p1 = d1;
      d0 = n0;
      d1 = n1;
    }
  }
}

// sphHarm(theta, phi) (fastoverlap/sphericalAlignment.py:57-65):
// Y[(l*W + m mod W)*npts + j].
__global__ void k_ylm_angles(int L, int npts, const double* __restrict__ theta,
                             const double* __restrict__ phi, double2* __restrict__ Y) {
  const int W = 2 * L + 1;
  const double inv4pi = 1.0 / (4.0 * kPi);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < npts; j += gridDim.x * blockDim.x) {
    double st, ct, sp, cp;
    sincos(theta[j], &st, &ct);
    sincos(phi[j], &sp, &cp);
    for (int l = 0; l <= L; ++l)
      for (int m = 0; m < W; ++m) Y[((long long)l * W + m) * npts + j] = make_double2(0.0, 0.0);
    double pmm = sqrt(inv4pi);
    double2 em = make_double2(1.0, 0.0);
    for (int m = 0; m <= L; ++m) {
      if (m > 0) {
        pmm *= -sqrt((2.0 * m + 1.0) / (2.0 * m)) * st;
        em = cmul(em, make_double2(cp, sp));
      }
      double p2 = 0.0, p1 = pmm;
      for (int l = m; l <= L; ++l) {
        double pl;
        if (l == m) {
          pl = pmm;
        } else if (l == m + 1) {
          pl = sqrt(2.0 * m + 3.0) * ct * pmm;
        } else {
          const double aa = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
          const double bb = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) /
                                 (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
          pl = aa * (ct * p1 - bb * p2);
        }
        if (l > m) {
          p2 = p1;
          p1 = pl;
        }
        const double2 yp = cscale(em, pl);
        Y[((long long)l * W + m) * npts + j] = yp;
        if (m > 0) {
          const double sg = (m & 1) ? -1.0 : 1.0;
          Y[((long long)l * W + (W - m)) * npts + j] = make_double2(sg * yp.x, -sg * yp.y);
        }
      }
    }
  }
}

void launch_sph_mirror(hipStream_t s, int P, int Lmax, const SphPair* pd, double2* I) {
  const long long tot = (long long)Lmax * (Lmax + 1) * (4 * Lmax + 5) / 6;
  if (tot == 0) return;
  hipLaunchKernelGGL(k_sph_mirror, dim3(cdiv(tot, 256), P), dim3(256), 0, s, pd, I);
}

// ---------------------------------------------------------------------------
// Forward SOFT (fastoverlap/soft.py:98-113), FP64:
//   S2[m1][k][m2] = n^-2 sum_{j1,j2} g[j1][k][j2] e^{-2 pi i (m1 j1 + m2 j2)/n}
//   f^l_{m1m2}    = sum_k w_k dn^l_{m1m2}(b_k) S2[m1][k][m2]
// for |m1|, |m2| <= L = n/2 - 1; w_k are the quadrature weights of
// fastoverlap/soft.py:63-71.
// ---------------------------------------------------------------------------
__global__ void k_soft_weights(int bw, double* __restrict__ w) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 2 * bw; k += gridDim.x * blockDim.x) {
    const double a = (2.0 * k + 1.0) / (4.0 * bw);  // (2k+1) pi / (4 bw), in units of pi
    double s = 0.0;
    for (int j = 0; j < bw; ++j) s += 2.0 / (2.0 * j + 1.0) * sinpi((2.0 * j + 1.0) * a);
    w[k] = s * sinpi(a) / bw;
  }
}

// one workgroup per beta plane k: the 2D DFT of g[., k, .] at |m| <= L
__global__ void k_soft_fwd_plane(int n, const double2* __restrict__ g,
                                 const double2* __restrict__ tw, double2* __restrict__ S2) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = blockIdx.x, L = n / 2 - 1, W = 2 * L + 1, t = threadIdx.x;
  double2* T = (double2*)smem;  // [n][W]: sum over j2 for each j1, m2
  for (int e = t; e < n * W; e += blockDim.x) {
    const int j1 = e / W, m2 = e % W - L;
    double2 acc = make_double2(0.0, 0.0);
    for (int j2 = 0; j2 < n; ++j2) {
      // e^{-2 pi i m2 j2 / n} = conj(tw[(m2 j2) mod n])
      const int q = ((m2 * j2) % n + n) % n;
      acc = cadd(acc, cmul(g[((long long)j1 * n + k) * n + j2], cconj(tw[q])));
    }
    T[e] = acc;
  }
  __syncthreads();
  const double inv = 1.0 / ((double)n * n);
  for (int e = t; e < W * W; e += blockDim.x) {
    const int m1 = e / W - L, m2i = e % W;
    double2 acc = make_double2(0.0, 0.0);
    for (int j1 = 0; j1 < n; ++j1) {
      const int q = ((m1 * j1) % n + n) % n;
      acc = cadd(acc, cmul(T[j1 * W + m2i], cconj(tw[q])));
    }
    S2[((long long)(m1 + L) * n + k) * W + m2i] = cscale(acc, inv);
  }
}

// one thread per compact (l, m1, m2), l <= L
__global__ void k_soft_fwd_beta(int n, long long Kt, const double* __restrict__ tab64,
                                const double* __restrict__ w, const double2* __restrict__ S2,
                                double2* __restrict__ f) {
  const int L = n / 2 - 1, W = 2 * L + 1;
  const long long K = kcount(L);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < K;
       e += (long long)gridDim.x * blockDim.x) {
    int l = 0;
    while (kbase(l + 1) <= e) ++l;
    const int r = (int)(e - kbase(l));
    const int m1 = r / (2 * l + 1) - l, m2 = r % (2 * l + 1) - l;
    double2 acc = make_double2(0.0, 0.0);
    for (int k = 0; k < n; ++k)
      acc = cadd(acc, cscale(S2[((long long)(m1 + L) * n + k) * W + (m2 + L)],
                             w[k] * tab64[(long long)k * Kt + e]));
    f[e] = acc;
  }
}

void launch_soft_forward(hipStream_t s, int n, long long Kt, const double* tab64,
                         const double2* tw64, const double2* g, double* w, double2* S2,
                         double2* f) {
  const int L = n / 2 - 1, W = 2 * L + 1;
  hipLaunchKernelGGL(k_soft_weights, dim3(1), dim3(256), 0, s, n / 2, w);
  hipLaunchKernelGGL(k_soft_fwd_plane, dim3(n), dim3(256), sizeof(double2) * (size_t)n * W, s,
                     n, g, tw64, S2);
  const long long K = kcount(L);
  hipLaunchKernelGGL(k_soft_fwd_beta, dim3((unsigned)std::min<long long>((K + 255) / 256, 4096)),
                     dim3(256), 0, s, n, Kt, tab64, w, S2, f);
}

size_t soft_forward_smem(int n) { return sizeof(double2) * (size_t)n * (n - 1); }

void launch_coef_to_plane(hipStream_t s, int P, int n, int Lmax, const SphPair* pd,
                          const double2* a, float2* b) {
  const long long e = plane_rows(Lmax) * n;
  hipLaunchKernelGGL(k_coef_to_plane, dim3((unsigned)((e + 255) / 256), P), dim3(256), 0, s, pd,
                     n, a, b);
}
void launch_sph_energy(hipStream_t s, int L, const double2* Iconj, double a, double b, double y,
                       const double* rec, const double* seedu, int Lt, double* out) {
  hipLaunchKernelGGL(k_sph_energy, dim3(1), dim3(256), 0, s, L, Iconj, a, b, y, rec, seedu, Lt,
                     out);
}

void launch_rec_table(hipStream_t s, int Lt, double* rec, double* seedu, double2* leg,
                      double* cm) {
  const int W = 2 * Lt + 1;
  hipLaunchKernelGGL(k_rec_table, dim3((W * W + 127) / 128), dim3(128), 0, s, Lt, rec, seedu);
  hipLaunchKernelGGL(k_leg_table, dim3((qlm(Lt, Lt) + 128) / 128), dim3(128), 0, s, Lt, leg, cm);
}
void launch_wigner_D(hipStream_t s, int L, double a, double b, double y, double2* D, double2* G) {
  const int W = 2 * L + 1;
  hipLaunchKernelGGL(k_wigner_D, dim3((W * W + 127) / 128), dim3(128), 0, s, L, a, b, y, D, G);
}
void launch_ylm_angles(hipStream_t s, int L, int npts, const double* th, const double* ph,
                       double2* Y) {
  hipLaunchKernelGGL(k_ylm_angles, dim3((npts + 127) / 128), dim3(128), 0, s, L, npts, th, ph, Y);
}

}  // namespace fo
