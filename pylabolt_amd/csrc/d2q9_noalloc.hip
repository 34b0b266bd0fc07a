// d2q9_mass_fp64_static: the FP64 mass reduction of d2q9_mass_fp64 with no
// caller workspace and no allocation, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// The per-workgroup partial sums live in `scratch`, a file-scope device
// array of SCRATCH_DOUBLES doubles (16 KiB) that each GPU the library runs
// on holds until the process ends.  Host code never passes the array's name
// to a kernel: on the first call on a device, with that device current,
// hipGetSymbolAddress gives the device's own address of the array, which is
// kept per device and passed to the kernels like the workspace of
// d2q9_mass_fp64.  Each partial-sum workgroup writes its own slot and the
// final workgroup reads only the slots of the same call, so nothing has to
// start at zero; calls that overlap on one device share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "d2q9_device.h"
#include "d2q9_reduce.h"

using namespace d2q9;

namespace {

constexpr int SCRATCH_DOUBLES = RED_MAX_BLOCKS;

}  // namespace

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {

constexpr int MAX_DEVICES = 64;
double* g_scratch[MAX_DEVICES];
std::mutex g_scratch_mutex;

int fail(int code, const char* what) {
  set_error(what);
  return code;
}

// Device address of `scratch` on the current device, looked up once per
// device.
int scratch_for_current_device(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail((int)e, hipGetErrorString(e));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(D2Q9_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail((int)e, hipGetErrorString(e));
    if (!p) return fail(D2Q9_E_ARG, "hipGetSymbolAddress gave a null address");
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" int d2q9_mass_fp64_static(const d2q9_lattice* L, const void* f, double* out,
                                     void* stream) {
  if (int r = validate_lattice(L)) return r;
  if (!f || !out) return fail(D2Q9_E_ARG, "null buffer");
  if (red_blocks(L) > SCRATCH_DOUBLES)
    return fail(D2Q9_E_TOO_LARGE, "more partial sums than the static scratch holds (2048)");
  double* work = nullptr;
  if (int r = scratch_for_current_device(&work)) return r;
  return launch_mass(L, f, work, out, static_cast<hipStream_t>(stream));
}
