// Internal host-side interface between the C ABI (fo_api.cpp) and the
// serving engine (fo_server.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include "fastoverlap_hip.h"
#include "fo_internal.h"

namespace fo {

// Plan a batch: validate the descriptor, build host-side descriptors and the
// memory layout, and resolve device tables; no batch memory is allocated.
// build_tables = false: the grid's tables must already be resident (the
// allocation-free entry point plans this way).
// harm_nmax >= 0: the coefficient stage projects on the harmonic-oscillator
// basis of radial degree <= harm_nmax and length scale harm_r0
// (SphericalHarmonicAlign) instead of contracting the Gaussian overlap.
int sph_batch_plan(fo_context* ctx, const fo_sph_batch_desc* d, fo_sph_batch** out,
                   bool build_tables = true, int harm_nmax = -1, double harm_r0 = 1.0,
                   int harm_basis = 0);
int per_batch_plan(fo_context* ctx, const fo_per_batch_desc* d, fo_per_batch** out);
// Device bytes the planned batch needs.
size_t sph_batch_bytes(const fo_sph_batch* b);
size_t per_batch_bytes(const fo_per_batch* b);
// Bind memory and a stream.  mem == nullptr allocates (counted against the
// context budget); stream == nullptr creates a private stream.  Uploads the
// descriptors asynchronously on the stream.
int sph_batch_bind(fo_sph_batch* b, void* mem, hipStream_t stream);
int per_batch_bind(fo_per_batch* b, void* mem, hipStream_t stream);

// Device-resident outputs of a bound spherical batch.
struct SphOutputs {
  const SphResult* res;  // [npairs]
  const double* X2;      // [3 * natoms] aligned second structures
  const int* perm;       // [natoms]
  long long natoms;      // atoms over all pairs
  int npairs;
};
void sph_batch_outputs(const fo_sph_batch* b, SphOutputs* o);

// Context allocation with budget accounting.
int ctx_alloc_bytes(fo_context* c, size_t bytes, void** p);
void ctx_free_bytes(fo_context* c, void* p, size_t bytes);
int ctx_device(const fo_context* c);
hipStream_t ctx_stream(const fo_context* c);

// Resolve (once per device, with that device current) the device address of
// the allocation-free entry point's scratch array (fo_noalloc.hip).
void noalloc_attach(int device);

int set_error(int code, const char* msg);

}  // namespace fo
