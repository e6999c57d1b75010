// Internal (not exported) declarations shared by the kernel translation units
// and the host runtime of libfastoverlap_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fo {

// ---------------------------------------------------------------------------
// Spherical batch: one descriptor per pair (device copy is a plain array).
// ---------------------------------------------------------------------------
struct SphPair {
  int natoms;          // atoms in each structure of the pair
  int atom_off;        // first atom in the concatenated position arrays
  int L;               // Jmax of the pair
  int ngroups;         // permutation groups (species)
  int group_off;       // first group in the group arrays
  int ngath;           // atoms that belong to a group (sum of group sizes)
  int gath_off;        // first entry in the gathered-atom index list
  int pad0;
  long long coef_off;  // first complex entry of the pair's C^l (K(L) entries)
  long long rad_off;   // first double of the pair's radial block
  long long ylm_off;   // first complex entry of the pair's Y block Y[q(l,m)][g], m >= 0 (per structure)
  long long c32_off;   // first complex entry of the pair's FP32 C^l in the plane layout
  double scale;        // Gaussian width sigma (<= 0: automatic)
};

struct SphGroup {
  int size;            // atoms in the group
  int gath_off;        // offset of the group inside the pair's gathered list
  long long jk_off;    // prefix of size^2 over the pair's earlier groups
  long long rad_off;   // (L+1) * jk_off : offset inside the pair's radial block
};

// Per (pair, orientation, plane) candidate record written by the plane
// kernels: the plane's values are split into kCandPerPlane parts (the
// spherical kernels: one per wavefront) and entry r lists the largest value
// of part r with a bound on the part's other values.  Entries are unordered;
// parts with no values hold -inf.
constexpr int kCandPerPlane = 4;
constexpr int kMaxCand = 64;     // FP64-re-evaluated candidates listed per grid
struct PlaneCand {
  float val[kCandPerPlane];   // largest value of each part
  int idx[kCandPerPlane];     // its flat grid index (ia*n+ib)*n+ig
  float rest[kCandPerPlane];  // upper bound on every other value of the part
};

// Per (pair, orientation) peak record written by the selection kernel.
struct SphPeak {
  int ind[3];          // integer argmax (alpha, beta, gamma)
  int ncand;           // FP64-re-evaluated candidates
  int noverflow;       // planes re-evaluated in full in FP64
  int pad;
  double value;        // FP64 grid value at the argmax
  double R0[3];        // interpolated Euler angles (fastoverlap/soft.py:127-130)
  double margin;       // FP64 value gap to the runner-up candidate (or +inf)
  double delta;        // FP32 error bound used to collect candidates
  double val32;        // largest FP32 grid value (plane kernels)
};

struct SphOpt {
  double R[3];         // optimised Euler angles
  double f;            // objective -Re sum conj(I) D(R)
  double gnorm;        // gradient max-norm at the last evaluated point
  int iters;
  int evals;           // objective evaluations, line search included
};

struct SphResult {
  double dist;
  double rmat[9];      // total rotation applied to (+/-)X2, row-major
  double R[3];         // Euler angles of the chosen orientation
  int inverted;
  int status;
};

// ---------------------------------------------------------------------------
// Periodic batch
// ---------------------------------------------------------------------------
struct PerPair {
  int natoms;
  int atom_off;
  int ngroups;
  int group_off;
  double box[3];
  double scale;        // sigma (> 0, resolved on the host)
  long long c_off;     // first complex entry of the pair's (2n+1)^3 product
};

struct PerGroup {
  int size;
  int atom_list_off;   // offset of the group in the atom index list
};

struct PerPeak {
  int ind[3];
  int ncand;
  int noverflow;
  int pad;
  double value;
  double disp[3];      // initial displacement (find_max * box / F)
  double margin;
  double delta;
};

struct PerResult {
  double dist;
  double disp[3];
  int iters;
  int status;
};

// ---- kernel launchers (implemented in the .hip files) ---------------------
struct SphLaunch {
  int P;               // pairs
  int n;               // grid side 2B
  int Lmax;            // max L over the pairs
  int Lt;              // degree of the resident Wigner table
  int maxgath;         // max gathered atoms over pairs
  int maxatoms;        // max atoms over pairs
  long long maxjk;     // max sum of group size^2 over pairs
  int norient;         // 1 or 2
  int mode;            // 0 = alignment (B = (-1)^l A), 1 = general (A,B given)
  int keep_grid;       // write the FP32 grid
  int center;          // subtract centroids
};

}  // namespace fo
