// Periodic (bulk) overlap-alignment kernels for gfx950.
//
//   k_per_sf       structure factors C_s(k) = sum_j e^{-i k.r_j} for the
//                  (2n+1)^3 k-cube, separable per axis, and the weighted
//                  product sum_s C1_s conj(C2_s) e^{-|k|^2 s^2}   (FP64)
//   k_per_product  the same product from caller-supplied coefficients
//   k_per_plane    |forward 3D DFT| of the corner-padded product on an F^3
//                  grid, one x-plane per workgroup (pruned: only 2n+1 of F
//                  inputs per axis are non-zero), top candidates   (FP32)
//   k_per_select   FP64 re-evaluation of every candidate within the FP32
//                  error bound -> exact argmax, parabolic offsets, disp
//   k_per_refine   LAP + mean-shift iteration (FP64)
// (reference: fastoverlap/periodicAlignment.py:27-80, :384-460)
#include "fo_common.h"
#include "fo_internal.h"
#include "fo_refine.h"
#include <algorithm>

namespace fo {

constexpr int kSFChunk = 4;   // (b,c) entries per thread per chunk
constexpr int kSFTile = 32;   // atoms per LDS tile

FO_DEV double2 phase64(double kfrac_times2) {
  // e^{-i pi * kfrac_times2}
  double s, c;
  sincospi(kfrac_times2, &s, &c);
  return make_double2(c, -s);
}

// One workgroup per (x-index a of the k-cube, pair).
__global__ void __launch_bounds__(256) k_per_sf(
    const PerPair* __restrict__ pd, const PerGroup* __restrict__ grp,
    const int* __restrict__ atoms, const double* __restrict__ pos1,
    const double* __restrict__ pos2, int nk, double2* __restrict__ Cprod,
    double2* __restrict__ C1out, double2* __restrict__ C2out) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NB = 2 * nk + 1;
  const int a = blockIdx.x, p = blockIdx.y, t = threadIdx.x;
  const PerPair d = pd[p];
  double2* ex = (double2*)smem;          // [kSFTile]
  double2* ey = ex + kSFTile;            // [kSFTile][NB]
  double2* ez = ey + kSFTile * NB;       // [kSFTile][NB]
  const long long plane = (long long)NB * NB;
  const long long cube = plane * NB;
  for (int e0 = 0; e0 < NB * NB; e0 += kSFChunk * blockDim.x) {
    double2 prod[kSFChunk];
    for (int q = 0; q < kSFChunk; ++q) prod[q] = make_double2(0.0, 0.0);
    for (int g = 0; g < d.ngroups; ++g) {
      const PerGroup G = grp[d.group_off + g];
      double2 acc[2][kSFChunk];
      for (int s = 0; s < 2; ++s) {
        const double* pos = (s ? pos2 : pos1) + 3LL * d.atom_off;
        for (int q = 0; q < kSFChunk; ++q) acc[s][q] = make_double2(0.0, 0.0);
        for (int j0 = 0; j0 < G.size; j0 += kSFTile) {
          const int jt = min(kSFTile, G.size - j0);
          __syncthreads();
          for (int w = t; w < jt * (1 + 2 * NB); w += blockDim.x) {
            const int j = w / (1 + 2 * NB), r = w % (1 + 2 * NB);
            const int at = atoms[G.atom_list_off + j0 + j];
            const double* x = pos + 3LL * at;
            if (r == 0) {
              ex[j] = phase64(2.0 * (a - nk) * x[0] / d.box[0]);
            } else if (r <= NB) {
              ey[j * NB + r - 1] = phase64(2.0 * (r - 1 - nk) * x[1] / d.box[1]);
            } else {
              ez[j * NB + r - 1 - NB] = phase64(2.0 * (r - 1 - NB - nk) * x[2] / d.box[2]);
            }
          }
          __syncthreads();
          // e^{-i kx x} (fixed kx per workgroup) folded into the y phases once
          // per tile: the product the sum below used to form per entry
          for (int w = t; w < jt * NB; w += blockDim.x) ey[w] = cmul(ex[w / NB], ey[w]);
          __syncthreads();
          for (int q = 0; q < kSFChunk; ++q) {
            const int e = e0 + q * blockDim.x + t;
            if (e >= NB * NB) break;
            const int b = e / NB, c = e % NB;
            double2 sacc = acc[s][q];
            for (int j = 0; j < jt; ++j) sacc = cadd(sacc, cmul(ey[j * NB + b], ez[j * NB + c]));
            acc[s][q] = sacc;
          }
        }
      }
      for (int q = 0; q < kSFChunk; ++q) {
        const int e = e0 + q * blockDim.x + t;
        if (e >= NB * NB) break;
        prod[q] = cadd(prod[q], cmul(acc[0][q], cconj(acc[1][q])));
        if (C1out) {
          const long long o = (long long)(d.group_off + g) * cube + a * plane + e;
          C1out[o] = acc[0][q];
          C2out[o] = acc[1][q];
        }
      }
    }
    for (int q = 0; q < kSFChunk; ++q) {
      const int e = e0 + q * blockDim.x + t;
      if (e >= NB * NB) break;
      const int b = e / NB, c = e % NB;
      const double kx = 2.0 * kPi * (a - nk) / d.box[0];
      const double ky = 2.0 * kPi * (b - nk) / d.box[1];
      const double kz = 2.0 * kPi * (c - nk) / d.box[2];
      const double w = exp(-(kx * kx + ky * ky + kz * kz) * d.scale * d.scale);
      Cprod[d.c_off + a * plane + e] = cscale(prod[q], w);
    }
  }
}

// Weighted product from caller-supplied per-group coefficients
// (the Cs= path of fastoverlap/periodicAlignment.py:408-438).
__global__ void k_per_product(const PerPair* __restrict__ pd, int nk,
                              const double2* __restrict__ C1, const double2* __restrict__ C2,
                              double2* __restrict__ Cprod) {
  const int NB = 2 * nk + 1, p = blockIdx.y;
  const PerPair d = pd[p];
  const long long cube = (long long)NB * NB * NB;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < cube;
       e += (long long)gridDim.x * blockDim.x) {
    const int a = (int)(e / (NB * NB)), b = (int)((e / NB) % NB), c = (int)(e % NB);
    double2 s = make_double2(0.0, 0.0);
    for (int g = 0; g < d.ngroups; ++g) {
      const long long o = (long long)(d.group_off + g) * cube + e;
      s = cadd(s, cmul(C1[o], cconj(C2[o])));
    }
    const double kx = 2.0 * kPi * (a - nk) / d.box[0];
    const double ky = 2.0 * kPi * (b - nk) / d.box[1];
    const double kz = 2.0 * kPi * (c - nk) / d.box[2];
    Cprod[d.c_off + e] = cscale(s, exp(-(kx * kx + ky * ky + kz * kz) * d.scale * d.scale));
  }
}

// The same product from per-structure coefficients shared by many pairs:
// pair p multiplies structure idx1[p] with structure idx2[p] (alignGroup's
// coefficient reuse, fastoverlap/periodicAlignment.py:466-473).  Cs holds
// nstruct x ngroups cubes; every pair has the same ngroups.
__global__ void k_per_product_indexed(const PerPair* __restrict__ pd, int nk,
                                      const double2* __restrict__ Cs, const int* __restrict__ idx1,
                                      const int* __restrict__ idx2, double2* __restrict__ Cprod) {
  const int NB = 2 * nk + 1, p = blockIdx.y;
  const PerPair d = pd[p];
  const long long cube = (long long)NB * NB * NB;
  const double2* A = Cs + (long long)idx1[p] * d.ngroups * cube;
  const double2* B = Cs + (long long)idx2[p] * d.ngroups * cube;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < cube;
       e += (long long)gridDim.x * blockDim.x) {
    const int a = (int)(e / (NB * NB)), b = (int)((e / NB) % NB), c = (int)(e % NB);
    double2 s = make_double2(0.0, 0.0);
    for (int g = 0; g < d.ngroups; ++g)
      s = cadd(s, cmul(A[(long long)g * cube + e], cconj(B[(long long)g * cube + e])));
    const double kx = 2.0 * kPi * (a - nk) / d.box[0];
    const double ky = 2.0 * kPi * (b - nk) / d.box[1];
    const double kz = 2.0 * kPi * (c - nk) / d.box[2];
    Cprod[d.c_off + e] = cscale(s, exp(-(kx * kx + ky * ky + kz * kz) * d.scale * d.scale));
  }
}

// |f(x,y,z)|, f = sum_{a,b,c} C[a,b,c] e^{-2 pi i (a x + b y + c z)/F}, for
// one x-plane per workgroup (numpy fftn with C at the array corner,
// fastoverlap/periodicAlignment.py:439-440).
__global__ void __launch_bounds__(256) k_per_plane(
    const PerPair* __restrict__ pd, int nk, int F, const double2* __restrict__ Cprod,
    const float2* __restrict__ twg, PlaneCand* __restrict__ cand, float* __restrict__ grid,
    float2* __restrict__ cgrid) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NB = 2 * nk + 1;
  const int x = blockIdx.x, p = blockIdx.y, t = threadIdx.x, bd = blockDim.x;
  const PerPair d = pd[p];
  float2* tw = (float2*)smem;       // [F]
  float2* H = tw + F;               // [NB][NB]
  float2* G = H + NB * NB;          // [NB][F]
  float* rv = (float*)(G + NB * F);
  int* ri = (int*)(rv + 16);
  for (int k = t; k < F; k += bd) tw[k] = twg[k];
  __syncthreads();
  const double2* C = Cprod + d.c_off;
  for (int e = t; e < NB * NB; e += bd) {
    float2 acc = make_float2(0.f, 0.f);
    int ph = 0;
    for (int a = 0; a < NB; ++a) {
      const double2 cv = C[(long long)a * NB * NB + e];
      const float2 w = tw[ph];
      const float cx = (float)cv.x, cy = (float)cv.y;
      acc.x += cx * w.x - cy * w.y;
      acc.y += cx * w.y + cy * w.x;
      ph += x;
      if (ph >= F) ph -= F;
    }
    H[e] = acc;
  }
  __syncthreads();
  for (int e = t; e < NB * F; e += bd) {
    const int b = e / F, z = e % F;
    float2 acc = make_float2(0.f, 0.f);
    int ph = 0;
    for (int c = 0; c < NB; ++c) {
      const float2 h = H[b * NB + c];
      const float2 w = tw[ph];
      acc.x += h.x * w.x - h.y * w.y;
      acc.y += h.x * w.y + h.y * w.x;
      ph += z;
      if (ph >= F) ph -= F;
    }
    G[e] = acc;
  }
  __syncthreads();
  float bv = -INFINITY, dm = -INFINITY;
  int bi = 0x7fffffff;
  for (int e = t; e < F * F; e += bd) {
    const int y = e / F, z = e % F;
    float2 acc = make_float2(0.f, 0.f);
    int ph = 0;
    for (int b = 0; b < NB; ++b) {
      const float2 gv = G[b * F + z];
      const float2 w = tw[ph];
      acc.x += gv.x * w.x - gv.y * w.y;
      acc.y += gv.x * w.y + gv.y * w.x;
      ph += y;
      if (ph >= F) ph -= F;
    }
    const float val = sqrtf(acc.x * acc.x + acc.y * acc.y);
    const int flat = (x * F + y) * F + z;
    if (grid) grid[(long long)p * F * F * F + flat] = val;
    if (cgrid) cgrid[(long long)p * F * F * F + flat] = acc;
    if (val > bv) {
      dm = fmaxf(dm, bv);
      bv = val;
      bi = flat;
    } else {
      dm = fmaxf(dm, val);
    }
  }
  bool taken = false;
  PlaneCand out;
  float fifth = -INFINITY;
  for (int r = 0; r <= kCandPerPlane; ++r) {
    float v = taken ? -INFINITY : bv;
    int i = taken ? 0x7fffffff : bi;
    block_argmax(v, i, rv, ri);
    if (r < kCandPerPlane) {
      out.val[r] = v;
      out.idx[r] = i;
    } else {
      fifth = v;
    }
    if (!taken && i == bi && v == bv) taken = true;
  }
  int dummy = 0;
  block_argmax(dm, dummy, rv, ri);
  // one part: the block's top kCandPerPlane, the bound in rest[0]
  out.rest[0] = fmaxf(fifth, dm);
  for (int r = 1; r < kCandPerPlane; ++r) out.rest[r] = -INFINITY;
  if (t == 0) cand[(long long)p * F + x] = out;
}

struct PerGridCtx {
  int NB, F;
  const double2* C;
  const double2* tw64;
};

// FP64 |f| at one grid point (workgroup cooperative)
FO_DEV double per_eval64(const PerGridCtx& c, int x, int y, int z, double* red) {
  const int NB = c.NB, F = c.F;
  double sx = 0.0, sy = 0.0;
  for (int e = threadIdx.x; e < NB * NB * NB; e += blockDim.x) {
    const int a = e / (NB * NB), b = (e / NB) % NB, cc = e % NB;
    const int ph = (int)(((long long)a * x + (long long)b * y + (long long)cc * z) % F);
    const double2 v = cmul(c.C[e], c.tw64[ph]);
    sx += v.x;
    sy += v.y;
  }
  sx = block_sum(sx, red);
  sy = block_sum(sy, red);
  return sqrt(sx * sx + sy * sy);
}

FO_DEV void per_plane64(const PerGridCtx& c, int x, double2* Hg, double2* Gg, double* red,
                        int* redi, double& bestv, int& besti) {
  const int NB = c.NB, F = c.F;
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    double2 acc = make_double2(0.0, 0.0);
    for (int a = 0; a < NB; ++a)
      acc = cadd(acc, cmul(c.C[(long long)a * NB * NB + e], c.tw64[(int)(((long long)a * x) % F)]));
    Hg[e] = acc;
  }
  __threadfence_block();
  __syncthreads();
  for (int e = threadIdx.x; e < NB * F; e += blockDim.x) {
    const int b = e / F, z = e % F;
    double2 acc = make_double2(0.0, 0.0);
    for (int cc = 0; cc < NB; ++cc)
      acc = cadd(acc, cmul(Hg[b * NB + cc], c.tw64[(int)(((long long)cc * z) % F)]));
    Gg[e] = acc;
  }
  __threadfence_block();
  __syncthreads();
  double bv = -INFINITY;
  int bi = 0x7fffffff;
  for (int e = threadIdx.x; e < F * F; e += blockDim.x) {
    const int y = e / F, z = e % F;
    double2 acc = make_double2(0.0, 0.0);
    for (int b = 0; b < NB; ++b)
      acc = cadd(acc, cmul(Gg[b * F + z], c.tw64[(int)(((long long)b * y) % F)]));
    argmax_merge(bv, bi, sqrt(acc.x * acc.x + acc.y * acc.y), (x * F + y) * F + z);
  }
  block_argmax(bv, bi, red, redi);
  bestv = bv;
  besti = bi;
  __syncthreads();
}

// Error bound of the FP32 planes: |f| <= A = sum |C|; rounding of C, of the
// twiddles and of three DFT passes of NB terms each stays below
// (4 NB + 32) 2^-24 A.  Candidates within 2 delta of the FP32 maximum are
// re-evaluated in FP64.
__global__ void __launch_bounds__(256) k_per_select(
    const PerPair* __restrict__ pd, int nk, int F, const double2* __restrict__ Cprod,
    const double2* __restrict__ tw64, const PlaneCand* __restrict__ cand,
    double2* __restrict__ scratch, long long scratch_stride, PerPeak* __restrict__ peaks) {
  __shared__ double red[16];
  __shared__ int redi[16];
  __shared__ float rv[16];
  __shared__ int ri[16];
  __shared__ int s_cnt, s_ovf;
  __shared__ int s_cand[kMaxCand];
  __shared__ int s_ovfl[512];
  const int p = blockIdx.x, t = threadIdx.x;
  const PerPair d = pd[p];
  const int NB = 2 * nk + 1;
  PerGridCtx c;
  c.NB = NB;
  c.F = F;
  c.C = Cprod + d.c_off;
  c.tw64 = tw64;
  const PlaneCand* pc = cand + (long long)p * F;
  float mv = -INFINITY;
  int mi = 0x7fffffff;
  for (int b = t; b < F; b += blockDim.x)
    for (int r = 0; r < kCandPerPlane; ++r) argmax_merge(mv, mi, pc[b].val[r], pc[b].idx[r]);
  block_argmax(mv, mi, rv, ri);
  double A = 0.0;
  for (int e = t; e < NB * NB * NB; e += blockDim.x) A += hypot(c.C[e].x, c.C[e].y);
  A = block_sum(A, red);
  const double delta = (4.0 * NB + 32.0) * 5.9604644775390625e-08 * A;
  const double thr = (double)mv - 2.0 * delta;
  if (t == 0) {
    s_cnt = 0;
    s_ovf = 0;
  }
  __syncthreads();
  for (int b = t; b < F; b += blockDim.x) {
    float rest = pc[b].rest[0];
    for (int r = 1; r < kCandPerPlane; ++r) rest = fmaxf(rest, pc[b].rest[r]);
    if ((double)rest >= thr) {
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
  double2* Hg = scratch + (long long)p * scratch_stride;
  double2* Gg = Hg + NB * NB;
  auto take = [&](double v, int i) {
    if (v > best || (v == best && i < besti)) {
      second = best;
      best = v;
      besti = i;
    } else if (v > second) {
      second = v;
    }
  };
  if (cand_overflow) {
    for (int x = 0; x < F; ++x) {
      double v;
      int i;
      per_plane64(c, x, Hg, Gg, red, redi, v, i);
      take(v, i);
    }
  } else {
    for (int q = 0; q < ncand; ++q) {
      const int f = s_cand[q];
      take(per_eval64(c, f / (F * F), (f / F) % F, f % F, red), f);
    }
    for (int q = 0; q < novf; ++q) {
      double v;
      int i;
      per_plane64(c, s_ovfl[q], Hg, Gg, red, redi, v, i);
      take(v, i);
    }
  }
  const int ind[3] = {besti / (F * F), (besti / F) % F, besti % F};
  double off[3];
  for (int ax = 0; ax < 3; ++ax) {
    int ip[3] = {ind[0], ind[1], ind[2]}, im[3] = {ind[0], ind[1], ind[2]};
    ip[ax] = (ind[ax] + 1) % F;
    im[ax] = (ind[ax] + F - 1) % F;
    const double y1 = per_eval64(c, ip[0], ip[1], ip[2], red);
    const double y3 = per_eval64(c, im[0], im[1], im[2], red);
    off[ax] = (y3 - y1) / (2.0 * (2.0 * best - y1 - y3));
  }
  if (t == 0) {
    PerPeak pk;
    for (int ax = 0; ax < 3; ++ax) {
      pk.ind[ax] = ind[ax];
      pk.disp[ax] = (ind[ax] - off[ax]) * d.box[ax] / F;
    }
    pk.ncand = cand_overflow ? -1 : ncand;
    pk.noverflow = cand_overflow ? F : novf;
    pk.pad = 0;
    pk.value = best;
    pk.margin = best - second;
    pk.delta = delta;
    peaks[p] = pk;
  }
}

// ---------------------------------------------------------------------------
// Refinement (fastoverlap/periodicAlignment.py:27-80) with the intended LAP
// semantics: Euclidean minimum-image cost with rows = X1, X2[perm] ~ X1.
// ---------------------------------------------------------------------------
FO_DEV double pwrap(double v, double box) { return v - rint(v / box) * box; }

__global__ void __launch_bounds__(256) k_per_refine(
    const PerPair* __restrict__ pd, const PerGroup* __restrict__ grp,
    const int* __restrict__ atoms, const double* __restrict__ pos1,
    const double* __restrict__ pos2, const PerPeak* __restrict__ peaks, int niter, int lap_nmax,
    int* __restrict__ perm, int* __restrict__ saveperm, int* __restrict__ rowsol,
    double* __restrict__ X1out, double* __restrict__ X2out, PerResult* __restrict__ res) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[16];
  __shared__ int redi[16];
  __shared__ int s_same;
  const int p = blockIdx.x, t = threadIdx.x;
  const PerPair d = pd[p];
  const int N = d.natoms;
  const double* x = pos1 + 3LL * d.atom_off;
  const double* y = pos2 + 3LL * d.atom_off;
  int* pm = perm + d.atom_off;
  int* sp = saveperm + d.atom_off;
  int* rs = rowsol + d.atom_off;
  const double bx = d.box[0], by = d.box[1], bz = d.box[2];
  double disp[3] = {peaks[p].disp[0], peaks[p].disp[1], peaks[p].disp[2]};
  LapShared ls = lap_carve(smem, lap_nmax);

  auto hungarian = [&](int* out) {
    for (int a = t; a < N; a += blockDim.x) out[a] = a;
    __threadfence_block();
    __syncthreads();
    for (int g = 0; g < d.ngroups; ++g) {
      const PerGroup G = grp[d.group_off + g];
      const int* ga = atoms + G.atom_list_off;
      const double d0 = disp[0], d1 = disp[1], d2 = disp[2];
      auto cost = [&](int i, int j) {
        const double* A = x + 3 * ga[i];
        const double* B = y + 3 * ga[j];
        const double dx = pwrap(A[0] - (B[0] - d0), bx);
        const double dy = pwrap(A[1] - (B[1] - d1), by);
        const double dz = pwrap(A[2] - (B[2] - d2), bz);
        return sqrt(dx * dx + dy * dy + dz * dz);
      };
      lap_solve(G.size, cost, ls, rs, red, redi);
      for (int i = t; i < G.size; i += blockDim.x) out[ga[i]] = ga[rs[i]];
      __threadfence_block();
      __syncthreads();
    }
  };
  auto mean_shift = [&](const int* pr) {
    double s[3] = {0, 0, 0};
    for (int a = t; a < N; a += blockDim.x) {
      const double* B = y + 3 * pr[a];
      s[0] += pwrap(x[3 * a] - (B[0] - disp[0]), bx);
      s[1] += pwrap(x[3 * a + 1] - (B[1] - disp[1]), by);
      s[2] += pwrap(x[3 * a + 2] - (B[2] - disp[2]), bz);
    }
    for (int k = 0; k < 3; ++k) disp[k] -= block_sum(s[k], red) / (double)N;
  };

  hungarian(sp);
  int it = 0;
  for (; it < niter; ++it) {
    mean_shift(sp);
    hungarian(pm);
    if (t == 0) s_same = 1;
    __syncthreads();
    for (int a = t; a < N; a += blockDim.x)
      if (pm[a] != sp[a]) s_same = 0;
    __syncthreads();
    const bool same = s_same;
    __syncthreads();
    if (same) break;
    for (int a = t; a < N; a += blockDim.x) sp[a] = pm[a];
    __threadfence_block();
    __syncthreads();
  }
  if (niter == 0) {
    for (int a = t; a < N; a += blockDim.x) pm[a] = sp[a];
    __threadfence_block();
    __syncthreads();
  }
  mean_shift(pm);
  double* o1 = X1out + 3LL * d.atom_off;
  double* o2 = X2out + 3LL * d.atom_off;
  double ss = 0.0;
  for (int a = t; a < N; a += blockDim.x) {
    const double* B = y + 3 * pm[a];
    const double u0 = pwrap(x[3 * a], bx), u1 = pwrap(x[3 * a + 1], by), u2 = pwrap(x[3 * a + 2], bz);
    const double v0 = pwrap(B[0] - disp[0], bx), v1 = pwrap(B[1] - disp[1], by),
                 v2 = pwrap(B[2] - disp[2], bz);
    o1[3 * a] = u0;
    o1[3 * a + 1] = u1;
    o1[3 * a + 2] = u2;
    o2[3 * a] = v0;
    o2[3 * a + 1] = v1;
    o2[3 * a + 2] = v2;
    const double e0 = pwrap(u0 - v0, bx), e1 = pwrap(u1 - v1, by), e2 = pwrap(u2 - v2, bz);
    ss += e0 * e0 + e1 * e1 + e2 * e2;
  }
  ss = block_sum(ss, red);
  if (t == 0) {
    PerResult r;
    r.dist = sqrt(ss);
    r.disp[0] = disp[0];
    r.disp[1] = disp[1];
    r.disp[2] = disp[2];
    r.iters = it;
    r.status = 0;
    res[p] = r;
  }
}

}  // namespace fo

namespace fo {

static inline int pcdiv(long long a, long long b) { return (int)((a + b - 1) / b); }

size_t per_sf_smem(int nk) { return sizeof(double2) * (size_t)kSFTile * (1 + 2 * (2 * nk + 1)); }

void launch_per_sf(hipStream_t s, int P, int nk, const PerPair* pd, const PerGroup* grp,
                   const int* atoms, const double* pos1, const double* pos2, double2* Cprod,
                   double2* C1out, double2* C2out) {
  hipLaunchKernelGGL(k_per_sf, dim3(2 * nk + 1, P), dim3(256), per_sf_smem(nk), s, pd, grp, atoms,
                     pos1, pos2, nk, Cprod, C1out, C2out);
}

void launch_per_product(hipStream_t s, int P, int nk, const PerPair* pd, const double2* C1,
                        const double2* C2, double2* Cprod) {
  const long long cube = (long long)(2 * nk + 1) * (2 * nk + 1) * (2 * nk + 1);
  hipLaunchKernelGGL(k_per_product, dim3(std::min(pcdiv(cube, 256), 256), P), dim3(256), 0, s, pd,
                     nk, C1, C2, Cprod);
}

size_t per_plane_smem(int nk, int F) {
  const int NB = 2 * nk + 1;
  return sizeof(float2) * ((size_t)F + (size_t)NB * NB + (size_t)NB * F) + 16 * sizeof(float) +
         16 * sizeof(int);
}

void launch_per_plane(hipStream_t s, int P, int nk, int F, const PerPair* pd,
                      const double2* Cprod, const float2* tw32, PlaneCand* cand, float* grid,
                      float2* cgrid) {
  hipLaunchKernelGGL(k_per_plane, dim3(F, P), dim3(256), per_plane_smem(nk, F), s, pd, nk, F,
                     Cprod, tw32, cand, grid, cgrid);
}

void launch_per_product_indexed(hipStream_t s, int P, int nk, const PerPair* pd,
                                const double2* Cs, const int* idx1, const int* idx2,
                                double2* Cprod) {
  const long long cube = (long long)(2 * nk + 1) * (2 * nk + 1) * (2 * nk + 1);
  hipLaunchKernelGGL(k_per_product_indexed, dim3(std::min(pcdiv(cube, 256), 256), P), dim3(256),
                     0, s, pd, nk, Cs, idx1, idx2, Cprod);
}

void launch_per_select(hipStream_t s, int P, int nk, int F, const PerPair* pd,
                       const double2* Cprod, const double2* tw64, const PlaneCand* cand,
                       double2* scratch, long long scratch_stride, PerPeak* peaks) {
  hipLaunchKernelGGL(k_per_select, dim3(P), dim3(256), 0, s, pd, nk, F, Cprod, tw64, cand,
                     scratch, scratch_stride, peaks);
}

void launch_per_refine(hipStream_t s, int P, int lap_nmax, const PerPair* pd, const PerGroup* grp,
                       const int* atoms, const double* pos1, const double* pos2,
                       const PerPeak* peaks, int niter, int* perm, int* saveperm, int* rowsol,
                       double* X1out, double* X2out, PerResult* res) {
  hipLaunchKernelGGL(k_per_refine, dim3(P), dim3(256), lap_lds_bytes(lap_nmax), s, pd, grp, atoms,
                     pos1, pos2, peaks, niter, lap_nmax, perm, saveperm, rowsol, X1out, X2out,
                     res);
}

}  // namespace fo

// ---------------------------------------------------------------------------
// Stand-alone LAP of two point sets (Hungarian methods of the drop-in
// classes: fastoverlap/sphericalAlignment.py:102-108 with squared distances,
// fastoverlap/periodicAlignment.py:94-107 with minimum-image distances).
// ---------------------------------------------------------------------------
namespace fo {

__global__ void __launch_bounds__(256) k_lap_align(int N, const double* __restrict__ X1,
                                                   const double* __restrict__ X2, int ngroups,
                                                   const int* __restrict__ gsize,
                                                   const int* __restrict__ goff,
                                                   const int* __restrict__ gatoms, int metric,
                                                   double bx, double by, double bz, int lap_nmax,
                                                   int* __restrict__ perm,
                                                   int* __restrict__ rowsol,
                                                   double* __restrict__ dist) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[16];
  __shared__ int redi[16];
  const int t = threadIdx.x;
  LapShared ls = lap_carve(smem, lap_nmax);
  for (int a = t; a < N; a += blockDim.x) perm[a] = a;
  __threadfence_block();
  __syncthreads();
  for (int g = 0; g < ngroups; ++g) {
    const int* ga = gatoms + goff[g];
    auto cost = [&](int i, int j) {
      const double* A = X1 + 3 * ga[i];
      const double* B = X2 + 3 * ga[j];
      double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
      if (metric == 1) {
        dx = pwrap(dx, bx);
        dy = pwrap(dy, by);
        dz = pwrap(dz, bz);
        return sqrt(dx * dx + dy * dy + dz * dz);
      }
      return dx * dx + dy * dy + dz * dz;
    };
    lap_solve(gsize[g], cost, ls, rowsol, red, redi);
    for (int i = t; i < gsize[g]; i += blockDim.x) perm[ga[i]] = ga[rowsol[i]];
    __threadfence_block();
    __syncthreads();
  }
  double ss = 0.0;
  for (int a = t; a < N; a += blockDim.x) {
    const double* A = X1 + 3 * a;
    const double* B = X2 + 3 * perm[a];
    double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
    if (metric == 1) {
      dx = pwrap(dx, bx);
      dy = pwrap(dy, by);
      dz = pwrap(dz, bz);
    }
    ss += dx * dx + dy * dy + dz * dz;
  }
  ss = block_sum(ss, red);
  if (t == 0) *dist = sqrt(ss);
}

void launch_lap_align(hipStream_t s, int N, const double* X1, const double* X2, int ngroups,
                      const int* gsize, const int* goff, const int* gatoms, int metric,
                      const double* box, int lap_nmax, int* perm, int* rowsol, double* dist) {
  hipLaunchKernelGGL(k_lap_align, dim3(1), dim3(256), lap_lds_bytes(lap_nmax), s, N, X1, X2,
                     ngroups, gsize, goff, gatoms, metric, box ? box[0] : 1.0, box ? box[1] : 1.0,
                     box ? box[2] : 1.0, lap_nmax, perm, rowsol, dist);
}

}  // namespace fo
