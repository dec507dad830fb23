// semx_hip.h -- HIP-side helpers shared by the kernel sources of libsemx:
// status-checked HIP calls, device uploads, and the launch timing registry
// behind semx_profile_*.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>

#include "semx_internal.h"

#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                              \
      fprintf(stderr, "semx: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return SEMX_E_HIP;                                                 \
    }                                                                    \
  } while (0)

namespace semx {

// HIP events around a launch on its stream, accumulated per kernel name
// while profiling is enabled (no-ops otherwise).
void prof_begin(hipStream_t s, hipEvent_t* a);
void prof_end(hipStream_t s, const char* name, hipEvent_t a);
void prof_flush();
bool prof_enabled();

template <class T>
inline int dev_upload(T** dptr, const T* src, size_t count) {
  *dptr = nullptr;
  const size_t alloc = count ? count : 1;
  HIPCHK(hipMalloc((void**)dptr, alloc * sizeof(T)));
  if (src && count) HIPCHK(hipMemcpy(*dptr, src, count * sizeof(T), hipMemcpyHostToDevice));
  else HIPCHK(hipMemset(*dptr, 0, alloc * sizeof(T)));
  return SEMX_OK;
}

template <class T>
inline void dev_free(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// Field::solve over all planes of one system (the body of semx_helm_solve).
int helm_solve(Helm& H, double* f, const double* bc, double* u, double tol_rel, int step_max,
               int* iters, hipStream_t s);
// The same for nf <= SEMX_MAX_FIELDS fields sharing the system: the element
// kernels run per field, the band factor is streamed once for all of them.
int helm_solve_fields(Helm& H, int nf, double* const* fs, const double* const* bcs,
                      double* const* us, double tol_rel, int step_max, int* iters,
                      hipStream_t s);
// Device address (current device) of the static scratch array of the
// allocation-free entry points; looked up once per device.
int noalloc_scratch(double** out);
// Waits for the stream, then reports (SEMX_E_TIMEOUT) and clears a
// cooperative band sweep's expired hand-off wait since the last check.
int helm_check(Helm& H, hipStream_t s);

}  // namespace semx


