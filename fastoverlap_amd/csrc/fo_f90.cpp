// Fortran-facing entry points, bound by the iso_c_binding modules in
// fastoverlap_amd/fortran/fastoverlap_hip.f90.  They mirror the argument
// meaning of the reference's Fortran ALIGN routines:
//   CLUSTERFASTOVERLAP::ALIGN (f90/fastclusters.f90:129-269)
//   BULKFASTOVERLAP::ALIGN    (f90/fastbulk.f90:279-334)
// with permutation groups in the 1-based SETPERM layout
// (f90/fastutils.f90:117-167).  Errors are returned as status codes (the
// reference STOPs); the numerical work is the batch pipeline of fo_api.cpp.
#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "fastoverlap_hip.h"
#include "fo_host.h"

namespace {

std::mutex g_mu;
fo_context* g_ctx = nullptr;

int default_ctx(fo_context** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx) {
    const char* dev = std::getenv("FASTOVERLAP_DEVICE");
    int rc = fo_context_create(dev ? std::atoi(dev) : 0, 0, &g_ctx);
    if (rc) return rc;
  }
  *out = g_ctx;
  return FO_OK;
}

int groups_from_f90(int natoms, int ngroups, const int* npermsize, const int* permgroup,
                    std::vector<int>& sizes, std::vector<int>& atoms) {
  if (ngroups <= 0) {
    sizes.assign(1, natoms);
    atoms.resize(natoms);
    for (int a = 0; a < natoms; ++a) atoms[a] = a;
    return FO_OK;
  }
  int tot = 0;
  for (int g = 0; g < ngroups; ++g) {
    sizes.push_back(npermsize[g]);
    tot += npermsize[g];
  }
  atoms.resize(tot);
  for (int i = 0; i < tot; ++i) {
    atoms[i] = permgroup[i] - 1;
    if (atoms[i] < 0 || atoms[i] >= natoms)
      return fo::set_error(FO_ERR_INVALID, "PERMGROUP entry out of range (1-based indices)");
  }
  return FO_OK;
}

// The batch of the last call of each kind, kept while the next call has
// the same shape (GMIN / OPTIM align one system size over and over): the
// call then uploads positions and replays the batch's captured launch
// graph.  The key lists every input that shapes the batch.  Calls are
// serialised (the reference's Fortran modules are not re-entrant either);
// the batches live until a different shape replaces them.
struct CachedBatch {
  std::vector<double> key;
  fo_sph_batch* sph = nullptr;
  fo_per_batch* per = nullptr;
};
std::mutex g_call_mu;
CachedBatch g_sph_cache, g_per_cache;

std::vector<double> shape_key(std::initializer_list<double> head, const std::vector<int>& sizes,
                              const std::vector<int>& atoms) {
  std::vector<double> k(head);
  k.push_back((double)sizes.size());
  for (int v : sizes) k.push_back(v);
  for (int v : atoms) k.push_back(v);
  return k;
}

int next_fast_len(int target) {
  if (target <= 6) return target;
  long long best = 1LL << 40;
  for (long long p5 = 1; p5 < 2LL * target; p5 *= 5)
    for (long long p35 = p5; p35 < 2LL * target; p35 *= 3) {
      long long N = p35;
      while (N < target) N *= 2;
      if (N < best) best = N;
    }
  return (int)best;
}

// NFSPACE of f90/fastbulk.f90:268-274 / :312-318
int nfspace_for(int nwave) {
  if (2 * nwave + 1 <= 200) return next_fast_len(4 * nwave + 3);
  int f = 1;
  while (f < 4 * nwave + 3) f *= 2;
  return f;
}


// The 48 operations of the cubic point group O_h as signed axis
// permutations, in the order of OHOPS (f90/alignutils.f90:830-987): six
// characters per operation, (sign, input axis) of output axes x, y, z, so
// "+y-x+z" maps (x, y, z) to (y, -x, z).  The order decides which of two
// equally good operations is reported (the first strictly better one is
// kept, f90/fastbulk.f90:458-481).
constexpr char kOhOps[] =
    "+x+y+z" "-x-y+z" "+y+z+x" "+y-x+z" "-x-y-z" "-y+z-x" "-y+x+z" "+x+y-z"
    "-y-z+x" "+y-z-x" "+z+x+y" "+z-x-y" "+z-y+x" "+z+y-x" "-y-z-x" "+y-z+x"
    "-z+x-y" "-z-x+y" "-z+y+x" "-z-y-x" "+y+z-x" "-y+z+x" "+x-y-z" "-x+y-z"
    "+x-z+y" "-x-z-y" "+x+z-y" "-x+z+y" "-z-x-y" "-z+x+y" "+z-x+y" "+z+x-y"
    "-y-x-z" "+y+x-z" "-x+y+z" "+x-y+z" "-y+x-z" "+y-x-z" "-z+y-x" "-z-y+x"
    "+z-y-x" "+z+y+x" "-x+z-y" "+x+z+y" "-x-z+y" "+x-z-y" "+y+x+z" "-y-x+z";
static_assert(sizeof(kOhOps) == 48 * 6 + 1, "48 operations of six characters");

void oh_apply(int op, const double* in, double* out, int natoms) {
  const char* c = kOhOps + 6 * op;
  for (int a = 0; a < natoms; ++a)
    for (int r = 0; r < 3; ++r)
      out[3 * a + r] = (c[2 * r] == '-' ? -1.0 : 1.0) * in[3 * a + (c[2 * r + 1] - 'x')];
}

// Batches of the candidate search, kept like the single-pair batch: the
// per-structure coefficients, the overlap grids and the refinement of every
// (operation, displacement) candidate.
CachedBatch g_coef_cache, g_grid_cache, g_ref_cache;

int cached_per_batch(CachedBatch& cb, fo_context* ctx, const fo_per_batch_desc& d,
                     const std::vector<double>& key, fo_per_batch** out) {
  if (!(cb.per && cb.key == key)) {
    if (cb.per) fo_per_batch_destroy(cb.per);
    cb.per = nullptr;
    int rc = fo_per_batch_create(ctx, &d, &cb.per);
    if (rc) return rc;
    cb.key = key;
  }
  *out = cb.per;
  return FO_OK;
}

// Descriptor of P pairs of one system (the single-pair template d); the
// vectors hold the per-pair arrays it points into.
struct PerPairs {
  std::vector<int> nat, ngr, gs, ga;
  std::vector<double> sc, box;
  fo_per_batch_desc d;
  PerPairs(const fo_per_batch_desc& d1, const std::vector<int>& sizes,
           const std::vector<int>& atoms, int P, int flags)
      : nat(P, d1.natoms[0]), ngr(P, (int)sizes.size()), sc(P, d1.scale[0]), d(d1) {
    for (int p = 0; p < P; ++p) {
      gs.insert(gs.end(), sizes.begin(), sizes.end());
      ga.insert(ga.end(), atoms.begin(), atoms.end());
      box.insert(box.end(), d1.box, d1.box + 3);
    }
    d.npairs = P;
    d.natoms = nat.data();
    d.scale = sc.data();
    d.box = box.data();
    d.ngroups = ngr.data();
    d.group_sizes = gs.data();
    d.group_atoms = ga.data();
    d.flags = flags;
  }
};

// P pairs (B[p] onto A[p]): candidate displacements of each pair's overlap
// (the interpolated maximum for ndisps <= 1, else up to ndisps
// Gaussian-fitted peaks, falling back to the maximum), then one refinement
// per (operation, displacement), operations outer, and per pair the first
// strictly smallest distance (ALIGNCOEFFS, f90/fastbulk.f90:501-565).  With
// ohcell the operations are the 48 images of A[p]; the loop of
// f90/fastbulk.f90:458-481 hands ALIGNCOEFFS the coefficients of the
// untransformed COORDSA for every operation (the transformed copy,
// FCOEFFDUMMYA, is never read), so every operation starts from the
// displacements of the untransformed pair.  With struct_coords the overlap
// of pair p is formed from per-structure coefficients (structures idx1[p],
// idx2[p] of struct_coords), each computed once.
int bulk_align_pairs(fo_context* ctx, const fo_per_batch_desc& d1, const std::vector<int>& sizes,
                     const std::vector<int>& atoms, const std::vector<const double*>& B,
                     const std::vector<const double*>& A, const std::vector<const double*>* structs,
                     const int* idx1, const int* idx2, int ndisps, bool ohcell, double* dist_out,
                     double* X_out, int* perm_out, double* disp_out) {
  const int natoms = d1.natoms[0], P = (int)B.size();
  const size_t n3 = 3 * (size_t)natoms;
  const std::vector<double> skey = shape_key(
      {(double)natoms, d1.box[0], d1.box[1], d1.box[2], d1.scale[0], (double)d1.nk, (double)d1.F},
      sizes, atoms);
  auto key_with = [&](std::initializer_list<double> extra) {
    std::vector<double> k = skey;
    k.insert(k.end(), extra);
    return k;
  };
  std::vector<double> X1((size_t)P * n3), X2((size_t)P * n3);
  for (int p = 0; p < P; ++p) {
    std::copy(B[p], B[p] + n3, X1.begin() + p * n3);
    std::copy(A[p], A[p] + n3, X2.begin() + p * n3);
  }
  // overlap grids
  int rc;
  fo_per_batch* gb = nullptr;
  if (structs) {
    const int ns = (int)structs->size();
    PerPairs pc(d1, sizes, atoms, ns, FO_PER_KEEP_COEFFS);
    fo_per_batch* cb = nullptr;
    rc = cached_per_batch(g_coef_cache, ctx, pc.d, key_with({1.0, (double)ns}), &cb);
    if (rc) return rc;
    std::vector<double> S((size_t)ns * n3);
    for (int i = 0; i < ns; ++i) std::copy((*structs)[i], (*structs)[i] + n3, S.begin() + i * n3);
    rc = fo_per_batch_set_positions(cb, S.data(), S.data());
    if (!rc) rc = fo_per_batch_run(cb, FO_PSTAGE_SF, 0);
    if (rc) return rc;
    const int cube = (2 * d1.nk + 1) * (2 * d1.nk + 1) * (2 * d1.nk + 1);
    std::vector<double> Cs((size_t)ns * sizes.size() * cube * 2);
    rc = fo_per_batch_coeffs(cb, Cs.data(), nullptr);
    if (rc) return rc;
    PerPairs pg(d1, sizes, atoms, P, FO_PER_KEEP_GRID | FO_PER_INPUT_COEFFS);
    rc = cached_per_batch(g_grid_cache, ctx, pg.d, key_with({2.0, (double)P}), &gb);
    if (!rc) rc = fo_per_batch_set_positions(gb, X1.data(), X2.data());
    if (!rc) rc = fo_per_batch_set_coeffs_indexed(gb, Cs.data(), ns, idx1, idx2);
  } else {
    PerPairs pg(d1, sizes, atoms, P, FO_PER_KEEP_GRID);
    rc = cached_per_batch(g_grid_cache, ctx, pg.d, key_with({3.0, (double)P}), &gb);
    if (!rc) rc = fo_per_batch_set_positions(gb, X1.data(), X2.data());
  }
  if (!rc) rc = fo_per_batch_run(gb, FO_PSTAGE_SF | FO_PSTAGE_GRID, 0);
  if (rc) return rc;
  // candidate displacements per pair
  std::vector<std::vector<double>> disps(P);
  if (ndisps > 1) {
    std::vector<double> fit((size_t)P * ndisps * 11);
    std::vector<int> cnt(P);
    rc = fo_per_batch_find_peaks(gb, ndisps, 2, fit.data(), cnt.data(), nullptr);
    if (rc) return rc;
    for (int p = 0; p < P; ++p)
      for (int k = 0; k < cnt[p]; ++k)
        for (int a = 0; a < 3; ++a)
          disps[p].push_back(fit[((size_t)p * ndisps + k) * 11 + 8 + a] * d1.box[a] / d1.F);
  }
  std::vector<double> d0((size_t)P * 3);
  rc = fo_per_batch_peaks(gb, nullptr, d0.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  for (int p = 0; p < P; ++p)
    if (disps[p].empty()) disps[p].assign(d0.begin() + 3 * p, d0.begin() + 3 * p + 3);
  // refinement of every (pair, operation, displacement)
  const int nops = ohcell ? 48 : 1;
  std::vector<int> first(P + 1, 0);
  for (int p = 0; p < P; ++p) first[p + 1] = first[p] + nops * (int)(disps[p].size() / 3);
  const int R = first[P];
  std::vector<double> RB((size_t)R * n3), RA((size_t)R * n3), RD((size_t)R * 3), img(n3);
  for (int p = 0; p < P; ++p) {
    const int nd = (int)disps[p].size() / 3;
    for (int o = 0; o < nops; ++o) {
      oh_apply(o, A[p], img.data(), natoms);
      for (int k = 0; k < nd; ++k) {
        const size_t r = (size_t)first[p] + (size_t)o * nd + k;
        std::copy(B[p], B[p] + n3, RB.begin() + r * n3);
        std::copy(img.begin(), img.end(), RA.begin() + r * n3);
        std::copy(disps[p].begin() + 3 * k, disps[p].begin() + 3 * k + 3, RD.begin() + 3 * r);
      }
    }
  }
  PerPairs pr(d1, sizes, atoms, R, 0);
  fo_per_batch* rb = nullptr;
  rc = cached_per_batch(g_ref_cache, ctx, pr.d, key_with({4.0, (double)R}), &rb);
  if (rc) return rc;
  std::vector<double> dist(R), Xo((size_t)R * n3), dout((size_t)R * 3);
  std::vector<int> perm((size_t)R * natoms);
  rc = fo_per_batch_set_positions(rb, RB.data(), RA.data());
  if (!rc) rc = fo_per_batch_set_disp(rb, RD.data());
  if (!rc) rc = fo_per_batch_run(rb, FO_PSTAGE_REFINE, 0);
  if (!rc) rc = fo_per_batch_results(rb, dist.data(), nullptr, Xo.data(), perm.data(), dout.data());
  if (rc) return rc;
  for (int p = 0; p < P; ++p) {
    int best = first[p];
    for (int r = first[p] + 1; r < first[p + 1]; ++r)
      if (dist[r] < dist[best]) best = r;
    if (dist_out) dist_out[p] = dist[best];
    if (X_out) std::copy(Xo.begin() + best * n3, Xo.begin() + (best + 1) * n3, X_out + p * n3);
    if (perm_out)
      for (int i = 0; i < natoms; ++i) perm_out[(size_t)p * natoms + i] = perm[(size_t)best * natoms + i];
    if (disp_out)
      for (int k = 0; k < 3; ++k) disp_out[3 * p + k] = dout[3 * best + k];
  }
  return FO_OK;
}

// The cluster ALIGN of both coefficient methods: Gaussian overlap
// (harm_nmax < 0) or the Fortran module's harmonic basis (ALIGNHARM).
int cluster_align(double* coordsb, double* coordsa, int natoms, int L, double* kwidth,
                  int harm_nmax, double hwidth, int invert, int ngroups, const int* npermsize,
                  const int* permgroup, double* distance, double* dist2, double* rmatbest,
                  int* bestperm, int* inverted) {
  if (!coordsa || !coordsb || !kwidth || natoms < 1)
    return fo::set_error(FO_ERR_INVALID, "bad argument");
  fo_context* ctx = nullptr;
  int rc = default_ctx(&ctx);
  if (rc) return rc;
  std::vector<int> sizes, atoms;
  rc = groups_from_f90(natoms, ngroups, npermsize, permgroup, sizes, atoms);
  if (rc) return rc;
  int nat = natoms, ng = (int)sizes.size();
  double sc = *kwidth;
  fo_sph_batch_desc d{};
  d.npairs = 1;
  d.grid_n = 2 * (L + 1);
  d.natoms = &nat;
  d.L = &L;
  d.scale = &sc;
  d.ngroups = &ng;
  d.group_sizes = sizes.data();
  d.group_atoms = atoms.data();
  d.flags = invert ? 0 : FO_SPH_NO_INVERT;
  std::lock_guard<std::mutex> lk(g_call_mu);
  const std::vector<double> key = shape_key(
      {(double)natoms, (double)L, sc, (double)d.flags, (double)harm_nmax, hwidth}, sizes, atoms);
  const bool reused = g_sph_cache.sph && g_sph_cache.key == key;
  if (!reused) {
    if (g_sph_cache.sph) fo_sph_batch_destroy(g_sph_cache.sph);
    g_sph_cache.sph = nullptr;
    rc = harm_nmax >= 0 ? fo_sph_batch_create_harmonic(ctx, &d, harm_nmax, hwidth, 1,
                                                       &g_sph_cache.sph)
                        : fo_sph_batch_create(ctx, &d, &g_sph_cache.sph);
    if (rc) return rc;
    g_sph_cache.key = key;
  }
  fo_sph_batch* b = g_sph_cache.sph;
  std::vector<double> X1(3 * (size_t)natoms), X2(3 * (size_t)natoms);
  std::vector<int> perm(natoms);
  double dist = 0, R[3], rmat[9];
  int inv = 0;
  rc = fo_sph_batch_set_positions(b, coordsb, coordsa);
  if (!rc) rc = fo_sph_batch_run(b, FO_STAGE_ALL, reused ? 1 : 0);
  if (!rc) rc = fo_sph_batch_results(b, &dist, X1.data(), X2.data(), perm.data(), R, rmat, &inv);
  if (!rc) rc = fo_sph_batch_scales(b, &sc);
  if (rc) return rc;
  for (size_t i = 0; i < X1.size(); ++i) {
    coordsb[i] = X1[i];
    coordsa[i] = X2[i];
  }
  *kwidth = sc;
  if (distance) *distance = dist;
  if (dist2) *dist2 = dist * dist;
  const double sg = inv ? -1.0 : 1.0;
  // row-vector map X2c -> X2out is rmat (row-major); as a column-vector map
  // of the (inverted) input it is (rmat/sg)^T, whose column-major storage is
  // the row-major rmat / sg
  if (rmatbest)
    for (int k = 0; k < 9; ++k) rmatbest[k] = rmat[k] / sg;
  if (bestperm)
    for (int i = 0; i < natoms; ++i) bestperm[i] = perm[i] + 1;
  if (inverted) *inverted = inv;
  return FO_OK;
}

}  // namespace

extern "C" {

// COORDSB (target) and COORDSA are returned centred; COORDSA is replaced by
// the aligned, permuted (and possibly inverted) structure.  KWIDTH <= 0
// selects the automatic width and returns it.  RMATBEST (column-major 3x3)
// is the rotation applied to the (inverted, if *inverted) COORDSA.
int fo_f90_cluster_align(double* coordsb, double* coordsa, int natoms, int L, double* kwidth,
                         int invert, int ngroups, const int* npermsize, const int* permgroup,
                         double* distance, double* dist2, double* rmatbest, int* bestperm,
                         int* inverted) {
  return cluster_align(coordsb, coordsa, natoms, L, kwidth, -1, 1.0, invert, ngroups, npermsize,
                       permgroup, distance, dist2, rmatbest, bestperm, inverted);
}

// ALIGNHARM (f90/fastclusters.f90:271-391): as the cluster ALIGN, with the
// overlap coefficients of the harmonic-oscillator basis of radial degree
// <= n and length scale hwidth (HARMONICCOEFFS, :654-687).
int fo_f90_cluster_align_harm(double* coordsb, double* coordsa, int natoms, int n, int L,
                              double hwidth, double* kwidth, int invert, int ngroups,
                              const int* npermsize, const int* permgroup, double* distance,
                              double* dist2, double* rmatbest, int* bestperm, int* inverted) {
  if (n < 0 || !(hwidth > 0)) return fo::set_error(FO_ERR_INVALID, "bad harmonic basis");
  return cluster_align(coordsb, coordsa, natoms, L, kwidth, n, hwidth, invert, ngroups,
                       npermsize, permgroup, distance, dist2, rmatbest, bestperm, inverted);
}

// CALCOVERLAPMATRICES (f90/fastclusters.f90:853-866): for every pair i <= j
// of the nlist structures (coordinates used as given, column-major
// [3 natoms][nlist]) the sum of |C|^2 (NORMS) and the maximum of the real
// overlap grid at bandwidth L + 1 (MAXOVERS), with the harmonic basis; both
// matrices symmetric, column-major [nlist][nlist].  All pairs run as one
// batch.
int fo_f90_harm_overlap_matrices(const double* coords, int natoms, int nlist, int n, int L,
                                 double hwidth, double kwidth, int ngroups, const int* npermsize,
                                 const int* permgroup, double* norms, double* maxovers) {
  if (!coords || natoms < 1 || nlist < 1 || n < 0 || !(hwidth > 0) || !(kwidth > 0))
    return fo::set_error(FO_ERR_INVALID, "bad argument");
  fo_context* ctx = nullptr;
  int rc = default_ctx(&ctx);
  if (rc) return rc;
  std::vector<int> sizes, atoms;
  rc = groups_from_f90(natoms, ngroups, npermsize, permgroup, sizes, atoms);
  if (rc) return rc;
  std::vector<int> pi, pj;
  for (int i = 0; i < nlist; ++i)
    for (int j = i; j < nlist; ++j) {
      pi.push_back(i);
      pj.push_back(j);
    }
  const int P = (int)pi.size();
  const size_t n3 = 3 * (size_t)natoms;
  std::vector<int> nat(P, natoms), Ls(P, L), ngr(P, (int)sizes.size()), gs, ga;
  std::vector<double> sc(P, kwidth), X1((size_t)P * n3), X2((size_t)P * n3);
  for (int p = 0; p < P; ++p) {
    gs.insert(gs.end(), sizes.begin(), sizes.end());
    ga.insert(ga.end(), atoms.begin(), atoms.end());
    std::copy(coords + pi[p] * n3, coords + (pi[p] + 1) * n3, X1.begin() + p * n3);
    std::copy(coords + pj[p] * n3, coords + (pj[p] + 1) * n3, X2.begin() + p * n3);
  }
  fo_sph_batch_desc d{};
  d.npairs = P;
  d.grid_n = 2 * (L + 1);
  d.natoms = nat.data();
  d.L = Ls.data();
  d.scale = sc.data();
  d.ngroups = ngr.data();
  d.group_sizes = gs.data();
  d.group_atoms = ga.data();
  d.flags = FO_SPH_NO_INVERT | FO_SPH_NO_CENTER;
  fo_sph_batch* b = nullptr;
  rc = fo_sph_batch_create_harmonic(ctx, &d, n, hwidth, 1, &b);
  if (rc) return rc;
  std::vector<double> nrm(P), val(P);
  rc = fo_sph_batch_set_positions(b, X1.data(), X2.data());
  if (!rc) rc = fo_sph_batch_run(b, FO_STAGE_PREP | FO_STAGE_COEFF | FO_STAGE_GRID, 0);
  if (!rc) rc = fo_sph_batch_coef_norm2(b, nrm.data());
  if (!rc) rc = fo_sph_batch_peaks(b, nullptr, nullptr, val.data(), nullptr, nullptr, nullptr,
                                   nullptr);
  fo_sph_batch_destroy(b);
  if (rc) return rc;
  for (int p = 0; p < P; ++p) {
    const size_t a1 = pi[p] + (size_t)nlist * pj[p], a2 = pj[p] + (size_t)nlist * pi[p];
    if (norms) norms[a1] = norms[a2] = nrm[p];
    if (maxovers) maxovers[a1] = maxovers[a2] = val[p];
  }
  return FO_OK;
}

// COORDSA is replaced by the aligned, permuted structure (wrapped into the
// box); COORDSB is unchanged.  KERNELWIDTH <= 0 selects (V/N)^(1/3)/3;
// NWAVE = ceil(1.3 N^(1/3)) and the FFT side is the next 2,3,5-smooth size
// >= 4 NWAVE + 3 (f90/fastbulk.f90:293-319).  NDISPLACEMENTS <= 1 refines
// the interpolated overlap maximum; above 1, up to NDISPLACEMENTS
// Gaussian-fitted peaks of the overlap are each refined and the first
// strictly best kept (ALIGNCOEFFS, f90/fastbulk.f90:501-565).  ohcell 1
// (OHCELLT) repeats the refinement for the 48 octahedral images of COORDSA
// from the displacements of the untransformed pair, as
// f90/fastbulk.f90:458-481 does; ohcell 2 takes each image's displacements
// from its own overlap (the transformed coefficients the reference
// computes but does not use).
int fo_f90_bulk_align(double* coordsb, double* coordsa, int natoms, double boxlx, double boxly,
                      double boxlz, double kernelwidth, int ndisps, int ohcell, int ngroups,
                      const int* npermsize, const int* permgroup, double* distance, double* dist2,
                      int* bestperm, double* disp_out) {
  if (!coordsa || !coordsb || natoms < 1) return fo::set_error(FO_ERR_INVALID, "bad argument");
  if (ndisps > 64) return fo::set_error(FO_ERR_INVALID, "NDISPLACEMENTS must be <= 64");
  fo_context* ctx = nullptr;
  int rc = default_ctx(&ctx);
  if (rc) return rc;
  std::vector<int> sizes, atoms;
  rc = groups_from_f90(natoms, ngroups, npermsize, permgroup, sizes, atoms);
  if (rc) return rc;
  int nat = natoms, ng = (int)sizes.size();
  double box[3] = {boxlx, boxly, boxlz};
  double sc = kernelwidth > 0 ? kernelwidth : std::cbrt(boxlx * boxly * boxlz / natoms) / 3.0;
  const int nk = (int)std::ceil(1.3 * std::cbrt((double)natoms));
  fo_per_batch_desc d{};
  d.npairs = 1;
  d.nk = nk;
  d.F = nfspace_for(nk);
  d.natoms = &nat;
  d.box = box;
  d.scale = &sc;
  d.ngroups = &ng;
  d.group_sizes = sizes.data();
  d.group_atoms = atoms.data();
  d.niter = 10;
  d.flags = 0;
  std::lock_guard<std::mutex> lk(g_call_mu);
  if (ndisps > 1 || ohcell) {
    // ohcell 2: each octahedral image gets the candidates of its own overlap
    const int n3 = 3 * natoms, nimg = ohcell == 2 ? 48 : 1;
    std::vector<double> imgs((size_t)nimg * n3), X((size_t)nimg * n3), disp(3 * nimg),
        dist(nimg);
    std::vector<int> perm((size_t)nimg * natoms);
    std::vector<const double*> B(nimg, coordsb), A(nimg);
    for (int o = 0; o < nimg; ++o) {
      if (nimg > 1)
        oh_apply(o, coordsa, imgs.data() + (size_t)o * n3, natoms);
      else
        std::copy(coordsa, coordsa + n3, imgs.begin());
      A[o] = imgs.data() + (size_t)o * n3;
    }
    rc = bulk_align_pairs(ctx, d, sizes, atoms, B, A, nullptr, nullptr, nullptr, ndisps,
                          ohcell == 1, dist.data(), X.data(), perm.data(), disp.data());
    if (rc) return rc;
    int best = 0;
    for (int o = 1; o < nimg; ++o)
      if (dist[o] < dist[best]) best = o;
    std::copy(X.begin() + (size_t)best * n3, X.begin() + (size_t)(best + 1) * n3, coordsa);
    if (distance) *distance = dist[best];
    if (dist2) *dist2 = dist[best] * dist[best];
    if (bestperm)
      for (int i = 0; i < natoms; ++i) bestperm[i] = perm[(size_t)best * natoms + i] + 1;
    if (disp_out)
      for (int k = 0; k < 3; ++k) disp_out[k] = disp[3 * best + k];
    return FO_OK;
  }
  const std::vector<double> key =
      shape_key({(double)natoms, boxlx, boxly, boxlz, sc, (double)nk, (double)d.F}, sizes, atoms);
  const bool reused = g_per_cache.per && g_per_cache.key == key;
  if (!reused) {
    if (g_per_cache.per) fo_per_batch_destroy(g_per_cache.per);
    g_per_cache.per = nullptr;
    rc = fo_per_batch_create(ctx, &d, &g_per_cache.per);
    if (rc) return rc;
    g_per_cache.key = key;
  }
  fo_per_batch* b = g_per_cache.per;
  std::vector<double> X2(3 * (size_t)natoms);
  std::vector<int> perm(natoms);
  double dist = 0, disp[3];
  rc = fo_per_batch_set_positions(b, coordsb, coordsa);
  if (!rc) rc = fo_per_batch_run(b, FO_PSTAGE_ALL, reused ? 1 : 0);
  if (!rc) rc = fo_per_batch_results(b, &dist, nullptr, X2.data(), perm.data(), disp);
  if (rc) return rc;
  for (size_t i = 0; i < X2.size(); ++i) coordsa[i] = X2[i];
  if (distance) *distance = dist;
  if (dist2) *dist2 = dist * dist;
  if (bestperm)
    for (int i = 0; i < natoms; ++i) bestperm[i] = perm[i] + 1;
  if (disp_out)
    for (int k = 0; k < 3; ++k) disp_out[k] = disp[k];
  return FO_OK;
}

// CALCDEFAULTS (f90/fastbulk.f90:253-277): kernel width (V/N)^(1/3)/3,
// NWAVE = ceil(2 pi / min(box) * 1.5 / width), NFSPACE as ALIGN.
int fo_f90_bulk_defaults(int natoms, double boxlx, double boxly, double boxlz, double* kwidth,
                         int* nwave, int* nfspace) {
  if (natoms < 1 || !(boxlx > 0 && boxly > 0 && boxlz > 0))
    return fo::set_error(FO_ERR_INVALID, "bad argument");
  const double kw = std::cbrt(boxlx * boxly * boxlz / natoms) / 3.0;
  const double maxk = 1.5 / kw;
  // the reference's own value of 2 pi
  const int nw = (int)std::ceil(2 * 3.14159265359 / std::min(boxlx, std::min(boxly, boxlz)) * maxk);
  if (kwidth) *kwidth = kw;
  if (nwave) *nwave = nw;
  if (nfspace) *nfspace = nfspace_for(nw);
  return FO_OK;
}

// ALIGNGROUP (f90/fastbulk.f90:336-412): COORDS2LIST[j] aligned onto
// COORDS1LIST[i] for every (i, j) (i >= j only when sym, where the overlap
// of the second structure is formed from COORDS1LIST[j], as the
// reference's FCOEFF2 = CONJG(FCOEFF1)).  Column-major outputs:
// distmat[i + n1 j], aligned[(i + n1 j) 3 natoms + c]; entries sym leaves
// out are zero.  Coefficients are computed once per structure; all pairs
// run as one batch per stage.
int fo_f90_bulk_align_group(const double* coords1, int n1, const double* coords2, int n2,
                            int natoms, double boxlx, double boxly, double boxlz, double kwidth,
                            int ndisps, int nwave, int nfspace, int sym, int ngroups,
                            const int* npermsize, const int* permgroup, double* distmat,
                            double* aligned) {
  if (!coords1 || n1 < 1 || natoms < 1 || (!sym && (!coords2 || n2 < 1)) || nwave < 1 ||
      !(kwidth > 0))
    return fo::set_error(FO_ERR_INVALID, "bad argument");
  if (ndisps > 64) return fo::set_error(FO_ERR_INVALID, "NDISPLACEMENTS must be <= 64");
  if (sym) {
    n2 = n1;
    coords2 = coords2 ? coords2 : coords1;
  }
  fo_context* ctx = nullptr;
  int rc = default_ctx(&ctx);
  if (rc) return rc;
  std::vector<int> sizes, atoms;
  rc = groups_from_f90(natoms, ngroups, npermsize, permgroup, sizes, atoms);
  if (rc) return rc;
  int nat = natoms, ng = (int)sizes.size();
  double box[3] = {boxlx, boxly, boxlz}, sc = kwidth;
  fo_per_batch_desc d{};
  d.npairs = 1;
  d.nk = nwave;
  d.F = nfspace > 0 ? nfspace : nfspace_for(nwave);
  d.natoms = &nat;
  d.box = box;
  d.scale = &sc;
  d.ngroups = &ng;
  d.group_sizes = sizes.data();
  d.group_atoms = atoms.data();
  d.niter = 10;
  const size_t n3 = 3 * (size_t)natoms;
  std::vector<const double*> structs, B, A;
  for (int i = 0; i < n1; ++i) structs.push_back(coords1 + i * n3);
  if (!sym)
    for (int j = 0; j < n2; ++j) structs.push_back(coords2 + j * n3);
  std::vector<int> i1, i2, cell;
  for (int j = 0; j < n2; ++j)
    for (int i = sym ? j : 0; i < n1; ++i) {
      B.push_back(coords1 + i * n3);
      A.push_back(coords2 + j * n3);
      i1.push_back(i);
      i2.push_back(sym ? j : n1 + j);
      cell.push_back(i + n1 * j);
    }
  const int P = (int)B.size();
  std::vector<double> dist(P), X((size_t)P * n3);
  std::lock_guard<std::mutex> lk(g_call_mu);
  rc = bulk_align_pairs(ctx, d, sizes, atoms, B, A, &structs, i1.data(), i2.data(),
                        ndisps == 0 ? 10 : ndisps, false, dist.data(), X.data(), nullptr, nullptr);
  if (rc) return rc;
  if (distmat) std::fill(distmat, distmat + (size_t)n1 * n2, 0.0);
  if (aligned) std::fill(aligned, aligned + (size_t)n1 * n2 * n3, 0.0);
  for (int p = 0; p < P; ++p) {
    if (distmat) distmat[cell[p]] = dist[p];
    if (aligned) std::copy(X.begin() + p * n3, X.begin() + (p + 1) * n3, aligned + cell[p] * n3);
  }
  return FO_OK;
}

int fo_oh_operation(int op, double* rmat) {
  if (op < 0 || op >= 48 || !rmat) return fo::set_error(FO_ERR_INVALID, "operation must be in [0, 48)");
  const char* c = kOhOps + 6 * op;
  for (int k = 0; k < 9; ++k) rmat[k] = 0.0;
  for (int r = 0; r < 3; ++r) rmat[3 * r + (c[2 * r + 1] - 'x')] = c[2 * r] == '-' ? -1.0 : 1.0;
  return FO_OK;
}

}  // extern "C"
