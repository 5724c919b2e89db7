// mg_static.hip -- the fixed device scratch of the evaluation that allocates nothing (mg_eval_static).
//
// One file-scope device array per GPU holds every intermediate buffer of such an evaluation (model tables,
// structure tables, neighbour rows, descriptors, dE/dG, per-edge forces, reductions).  The array's host-side name
// is only a stand-in for the symbol: its device address is looked up once per device, with that device current,
// and only that address is ever handed to kernels or copies.  Every call on a device shares the array and leaves
// its contents behind; calls that overlap on one device race.
#include <hip/hip_runtime.h>
#include <stddef.h>

#define MG_STATIC_BYTES ((size_t)128 << 20)  // 128 MiB per device, held until the process ends
#define MG_STATIC_MAXDEV 64

static __device__ __align__(256) unsigned char scratch[MG_STATIC_BYTES];

static void* g_scratch_ptr[MG_STATIC_MAXDEV];  // device address of scratch, per device (NULL until looked up)

// Address and size of this device's scratch.  The first call on a device (which must be current) looks the symbol
// up; the address is kept for the later calls.  Returns 0, or -1 when the lookup fails.
extern "C" int mgs_scratch(int device, void** ptr, size_t* bytes) {
  if (device < 0 || device >= MG_STATIC_MAXDEV) return -1;
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return -1;
    }
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  *bytes = MG_STATIC_BYTES;
  return 0;
}
