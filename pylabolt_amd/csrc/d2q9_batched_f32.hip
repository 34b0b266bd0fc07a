// Batched persistent sweep, f32 population storage.
#include "d2q9_batched.h"

namespace d2q9 {
namespace batched {
template hipError_t launch<D2Q9_STORE_F32>(const d2q9_batch_entry*, const d2q9_batch_item*,
                                 const int32_t*, int32_t*, int, int, int, int, int, int, int, int,
                                 hipStream_t);
}  // namespace batched
}  // namespace d2q9
