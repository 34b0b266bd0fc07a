// C-ABI entry of the batched persistent sweep (kernel in d2q9_batched.h,
// one translation unit per storage type).
#include <hip/hip_runtime.h>

#include "d2q9_batched.h"

using namespace d2q9;

extern "C" int d2q9_batched_step(const d2q9_batch_entry* entries,
                                 const d2q9_batch_item* items,
                                 const int32_t* queue_bounds, int32_t* counters,
                                 int parity, int storage, int tile_i, int grid,
                                 int policy, int chunk, int waves, int has_general,
                                 int write_moments, void* stream) {
  if (!entries || !items || !queue_bounds || !counters) {
    set_error("batched step: null buffer");
    return D2Q9_E_ARG;
  }
  if (grid < 8 || grid % 8 != 0) {
    set_error("batched step: grid must be a positive multiple of 8");
    return D2Q9_E_ARG;
  }
  if (tile_i < 1 || tile_i > 4 || (parity != 0 && parity != 1) || policy < 0 || policy > 1 ||
      chunk < 1 || (waves != 0 && waves != 6 && waves != 8)) {
    set_error("batched step: bad tile_i, parity, policy, chunk or waves");
    return D2Q9_E_ARG;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e;
  switch (storage) {
    case D2Q9_STORE_F64:
      e = batched::launch<D2Q9_STORE_F64>(entries, items, queue_bounds, counters, parity,
                                          tile_i, grid, policy, chunk, waves, has_general, write_moments, s);
      break;
    case D2Q9_STORE_F32:
      e = batched::launch<D2Q9_STORE_F32>(entries, items, queue_bounds, counters, parity,
                                          tile_i, grid, policy, chunk, waves, has_general, write_moments, s);
      break;
    default:
      set_error("batched step: FP64 or FP32 storage only");
      return D2Q9_E_UNSUPPORTED;
  }
  if (e != hipSuccess) {
    set_error(hipGetErrorString(e));
    return (int)e;
  }
  return 0;
}
