// bf_scratch.hip -- the library's static device scratch array, for the
// entry point that allocates nothing (dsaxbf_execute_noalloc): the fused
// kernel's per-32-sample beam-power partial sums go here instead of to a
// buffer the plan allocated.
//
// The array is a file-scope __device__ object, so every GPU the library runs
// on holds its own copy (SCRATCH_BYTES) from module load until the process
// ends.  Host code never passes the array's name to a kernel (on the host it
// is only a stand-in symbol): its device address is looked up once per
// device, with that device current, by hipGetSymbolAddress at set-up
// (dsaxbf_create), and that address is what the kernels receive.
#include <mutex>

#include "bf_common.h"

namespace {
constexpr size_t SCRATCH_BYTES = (size_t)8 << 20;   // 8 MiB per device
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES];
}  // namespace

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

extern "C" size_t bfk_scratch_bytes(void) { return SCRATCH_BYTES; }

// Device address of `scratch` on `device`, which must be the current device.
extern "C" int bfk_scratch_init(int device, void **ptr) {
  if (device < 0 || device >= MAX_DEVICES) {
    bfk_set_error("bfk_scratch_init: device %d outside 0..%d", device, MAX_DEVICES - 1);
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      bfk_set_error("hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
      return -3;   // DSAXBF_EHIP
    }
    g_ptr[device] = p;
  }
  *ptr = g_ptr[device];
  return 0;
}
