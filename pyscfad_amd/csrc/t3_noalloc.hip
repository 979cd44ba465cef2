// Device scratch of the allocation-free entry point
// (pyscfad_amd_ccsd_t_energy_vjp_noalloc): one file-scope device array that
// the code object holds on every GPU the library runs on, for the life of
// the process.  Host code never uses the array's name as a device address:
// the address is resolved once per device with hipGetSymbolAddress (with
// that device current) and kept per device.
#include <hip/hip_runtime.h>

#include <mutex>

namespace {

constexpr size_t SCRATCH_BYTES = size_t(1) << 30;   // 1 GiB
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};   // resolved device address per device id

}  // namespace

// Address and size of the current device's scratch (resolved on the first
// call on that device).
extern "C" hipError_t t3_noalloc_scratch(void** ptr, size_t* bytes) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p || reinterpret_cast<uintptr_t>(p) % 256 != 0) return hipErrorInvalidSymbol;
    g_scratch[dev] = p;
  }
  *ptr = g_scratch[dev];
  *bytes = SCRATCH_BYTES;
  return hipSuccess;
}

extern "C" size_t t3_noalloc_scratch_bytes(void) { return SCRATCH_BYTES; }
