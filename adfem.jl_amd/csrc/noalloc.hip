// The allocation-free variant of the adjoint-gradient entry point.
//
// admfem_system_kappa_gradient stages u and lam, given in the caller's node
// order, as (u, lam) pairs in the internal order (one 16-byte write per node)
// in a buffer the system allocated when it was created, then runs the
// gradient kernel over it.  The variant here stages them in `scratch`, a
// file-scope device array of fixed size, so a call makes no allocation of
// any kind: for callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid it).
//
// Each GPU holds its own copy of the array, at its own address, until the
// process ends.  Its device address is looked up once per device, with that
// device current, and kept; host code uses only that address (the array's
// name is a host-side stand-in).  Every call on one GPU shares its array:
// calls that overlap on one GPU (other streams or threads) race, and each
// call overwrites what the last one left (it is fully rewritten before it is
// read, so nothing needs clearing).
#include <mutex>

#include "internal.h"

namespace admfem {

namespace {

constexpr size_t SCRATCH_BYTES = size_t(32) << 20;  // 32 MiB: 2,097,152 (u, lam) pairs
constexpr int MAX_DEVICES = 64;

__device__ __align__(256) char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};

// The scratch array's device address on the current device.
void *scratch_ptr() {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  REQUIRE(dev >= 0 && dev < MAX_DEVICES, ADMFEM_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    REQUIRE(p != nullptr, ADMFEM_ERR_HIP, "no device address for the scratch array");
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

}  // namespace

size_t kappa_gradient_noalloc_max_nodes() { return SCRATCH_BYTES / sizeof(double2); }

void kappa_gradient_noalloc(const System &s, const double *u, const double *lam, double scale,
                            double *gk, hipStream_t st) {
  REQUIRE((size_t)s.n <= kappa_gradient_noalloc_max_nodes(), ADMFEM_ERR_UNSUPPORTED,
          "system too large for the allocation-free gradient: more than 2,097,152 nodes "
          "(32 MiB of (u, lam) pairs)");
  launch_kappa_gradient_user(s, u, lam, static_cast<double2 *>(scratch_ptr()), scale, gk, st);
}

}  // namespace admfem
