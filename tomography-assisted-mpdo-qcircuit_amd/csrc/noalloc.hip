// mpdo_pauli_expect_noalloc: the Pauli-string expectations of mpdo_pauli_expect through a fixed
// device scratch array of the library instead of its grow-only workspace, for callers that cannot
// allocate device memory. No hipMalloc/hipFree on this path: the strings' environments and
// intermediates, their operator codes and the per-string results live in `scratch`.
#include "mpdo.h"
#include "mpdo_internal.h"

#include <algorithm>
#include <mutex>

namespace mpdo {
mpdo_status layer_fail(mpdo_status s, const char* msg);
}  // namespace mpdo

namespace {

// One array per GPU, held until the process ends. Host code never uses the array's name as an
// address: the device address comes from hipGetSymbolAddress, once per device with that device
// current, and is kept per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
__device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
char* g_scratch_ptr[64] = {};

mpdo_status scratch_address(char** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return mpdo::layer_fail(MPDO_EHIP, "hipGetDevice failed");
  if (dev < 0 || dev >= 64) return mpdo::layer_fail(MPDO_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return mpdo::layer_fail(MPDO_EHIP, "hipGetSymbolAddress of the fixed scratch failed");
    g_scratch_ptr[dev] = (char*)p;
  }
  *out = g_scratch_ptr[dev];
  return MPDO_OK;
}

}  // namespace

extern "C" {

mpdo_status mpdo_pauli_expect_noalloc(int n, const mpdo_site* sites, int nstrings, const int* codes, double* out,
                                      void* stream) {
  size_t per = 0;
  mpdo_status st = mpdo::expect_check(n, sites, nstrings, codes, out, &per);
  if (st != MPDO_OK) return st;
  if (n <= 0 || nstrings <= 0) return MPDO_OK;
  if (per > NOALLOC_SCRATCH_BYTES)
    return mpdo::layer_fail(MPDO_ENOMEM,
                            "mpdo_pauli_expect_noalloc: one string's environments at these site shapes exceed the "
                            "fixed scratch of 32 MiB (use mpdo_pauli_expect)");
  int chunk = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nstrings, (size_t)mpdo::EXPECT_MAX_CHUNK,
                                                           NOALLOC_SCRATCH_BYTES / per}));
  while (chunk > 1 && mpdo::expect_chunk_bytes(n, sites, chunk) > NOALLOC_SCRATCH_BYTES) --chunk;
  char* base = nullptr;
  st = scratch_address(&base);
  if (st != MPDO_OK) return st;
  return mpdo::expect_run(n, sites, nstrings, codes, out, base, chunk, (hipStream_t)stream);
}

}  // extern "C"
