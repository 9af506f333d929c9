// cloud.hip -- point-cloud path (placeholder until the octree kernels land).
#include <hip/hip_runtime.h>

#include "gwn_internal.h"

namespace gwn {
void cloud_build_device(gwn_cloud*, hipStream_t) { throw Error{GWN_ERR_INVALID, "point cloud path not available in this build"}; }
void cloud_query_device(const gwn_cloud*, const float*, int64_t, double, float*, hipStream_t) {
  throw Error{GWN_ERR_INVALID, "point cloud path not available in this build"};
}
}  // namespace gwn
