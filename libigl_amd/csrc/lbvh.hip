// lbvh.hip -- device LBVH build (placeholder until the Morton/Karras build lands).
#include <hip/hip_runtime.h>

#include "gwn_internal.h"

namespace gwn {
int64_t lbvh_build(const float*, int64_t, const float*, DevBuf<uint32_t>&, hipStream_t) {
  throw Error{GWN_ERR_INVALID, "GWN_TREE_LBVH is not available in this build"};
}
}  // namespace gwn
