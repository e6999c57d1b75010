// Shared device/host helpers for the fastoverlap MI355X (gfx950) kernels.
//
// Conventions (fixed by the reference's Python classes):
//  * C^l_{m1 m2} is stored "compact l-major": for l = 0..L, m1 = -l..l,
//    m2 = -l..l, at index kidx(l,m1,m2) = l(4l^2-1)/3 + (m1+l)(2l+1) + (m2+l).
//    The layout does not depend on L, so any prefix K(L) of a table built for
//    Lt >= L serves degree L.  (reference: the (l, m1, m2) enumeration order of
//    fastoverlap/sphericalAlignment.py:52-55 and fastoverlap/soft.py:57-59)
//  * SO(3) grid axes are (alpha, beta, gamma) = (axis 0, 1, 2) of an n^3 grid,
//    flat index (ia*n + ib)*n + ig (fastoverlap/soft.py:119-125).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FO_DEV __device__ __forceinline__
#define FO_HD __host__ __device__ __forceinline__

namespace fo {

constexpr double kPi = 3.141592653589793238462643383279502884;

FO_HD long long kcount(int L) {  // number of (l,m1,m2) with l <= L
  return (long long)(L + 1) * (2 * L + 1) * (2 * L + 3) / 3;
}
FO_HD long long kbase(int l) {  // first compact index of degree l
  return (long long)l * (2 * l - 1) * (2 * l + 1) / 3;
}
FO_HD long long kidx(int l, int m1, int m2) {
  return kbase(l) + (long long)(m1 + l) * (2 * l + 1) + (m2 + l);
}

// triangular index of (l, m), 0 <= m <= l
FO_HD int qlm(int l, int m) { return l * (l + 1) / 2 + m; }

// ---- stored radial slots ---------------------------------------------------
// The radial kernel stores, per (j, k) of a species group, the values of a
// few degrees ("slots", [slot][j][k]).  Degrees 0..15 are contracted in
// blocks of eight (0..7, 8..15), each from the two stored degrees at its top
// by the downward recurrence; degrees >= 16 one per slot.  Slot degrees, in
// increasing order:
//   L <= 7 :  L, L+1                    (block 0 = 0..L)
//   L <= 15:  7, 8, L, L+1              (block 1 = 8..L)
//   L >= 16:  7, 8, 15, 16, 17, ..., L  (degree l >= 16 in slot l - 13)
// Block b reads slots 2b (its top degree) and 2b + 1 (one above).
FO_HD int rad_slots(int L) { return L <= 7 ? 2 : (L <= 15 ? 4 : L - 12); }
FO_HD int rad_slot_deg(int s, int L) {
  if (L <= 7) return L + s;
  if (s < 2) return 7 + s;
  return L <= 15 ? L + s - 2 : s + 13;
}

// ---- plane layout (FP32 operands of the iSOFT plane kernels) --------------
// Rows (l, m1) with 0 <= m1 <= l at row l(l+1)/2 + m1 (l-major, so the rows
// of degrees <= L are a prefix), n entries per row indexed by the gamma
// transform input k (m2 = k for k < n/2, k - n above) and zero where |m2| > l.
// Inside a row, entry k sits at plane_pos(k, n): lane-major for the grid
// sides of the two-pass transform (n = N1*N2, lane k1 owns k = k1 + N1*k2 and
// its N2 entries are contiguous), in k order for other n.
FO_HD int plane_n1(int n) {  // N1 of the two-pass transform, 0 if none
  switch (n) {
    case 16: return 4;
    case 32: case 64: return 8;
    case 128: case 256: return 16;
    default: return 0;
  }
}
FO_HD long long plane_rows(int L) { return (long long)(L + 1) * (L + 2) / 2; }
FO_HD int plane_k(int m2, int n) { return m2 < 0 ? m2 + n : m2; }
FO_HD int plane_pos(int k, int n) {
  const int n1 = plane_n1(n);
  return n1 ? (k % n1) * (n / n1) + k / n1 : k;
}

// ---- complex helpers (double2 / float2 = re, im) -------------------------
FO_HD double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
FO_HD float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
FO_HD double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
FO_HD float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
FO_HD double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
FO_HD float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
FO_HD double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
FO_HD float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// integer power by squaring (exponent >= 0); valid for negative bases
FO_HD double ipow(double x, int e) {
  double r = 1.0;
  while (e > 0) {
    if (e & 1) r *= x;
    x *= x;
    e >>= 1;
  }
  return r;
}

// ---- wave / block reductions (wave64) -------------------------------------
template <typename T>
FO_DEV T wave_sum(T v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
template <typename T>
FO_DEV T wave_max(T v) {
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Deterministic block sum (fixed tree independent of launch size).
// `scratch` must hold blockDim.x/64 elements.  All threads receive the sum.
template <typename T>
FO_DEV T block_sum(T v, T* scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  T r = 0;
  for (int w = 0; w < nw; ++w) r += scratch[w];
  return r;
}

// Deterministic block sum of NV values at once (two barriers).  `scratch`
// must hold NV * blockDim.x/64 elements.  All threads receive the sums.
template <int NV, typename T>
FO_DEV void block_sum_n(T (&v)[NV], T* scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) scratch[wid * NV + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    T r = 0;
    for (int w = 0; w < nw; ++w) r += scratch[w * NV + k];
    v[k] = r;
  }
}

// Block argmax on (value, index): larger value wins, ties -> smaller index.
FO_DEV void argmax_merge(float& v, int& i, float ov, int oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}
FO_DEV void argmax_merge(double& v, int& i, double ov, int oi) {
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}
template <typename T>
FO_DEV void wave_argmax(T& v, int& i) {
  for (int off = 32; off > 0; off >>= 1) {
    T ov = __shfl_xor(v, off, 64);
    int oi = __shfl_xor(i, off, 64);
    argmax_merge(v, i, ov, oi);
  }
}
// scratch: two arrays of blockDim.x/64 entries.  All threads get the result.
template <typename T>
FO_DEV void block_argmax(T& v, int& i, T* sv, int* si) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  wave_argmax(v, i);
  __syncthreads();
  if (lane == 0) {
    sv[wid] = v;
    si[wid] = i;
  }
  __syncthreads();
  v = sv[0];
  i = si[0];
  for (int w = 1; w < nw; ++w) argmax_merge(v, i, sv[w], si[w]);
}

// Block argmin with ties -> smaller index.
FO_DEV void argmin_merge(double& v, int& i, double ov, int oi) {
  if (ov < v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}
FO_DEV void block_argmin(double& v, int& i, double* sv, int* si) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_xor(v, off, 64);
    int oi = __shfl_xor(i, off, 64);
    argmin_merge(v, i, ov, oi);
  }
  __syncthreads();
  if (lane == 0) {
    sv[wid] = v;
    si[wid] = i;
  }
  __syncthreads();
  v = sv[0];
  i = si[0];
  for (int w = 1; w < nw; ++w) argmin_merge(v, i, sv[w], si[w]);
}

}  // namespace fo
