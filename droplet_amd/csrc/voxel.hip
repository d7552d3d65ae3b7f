// Cooperative voxelizer (placeholder until the device path lands).
#include <hip/hip_runtime.h>

#include "dl_internal.h"

struct dl_voxel_result {
  int unused;
};

using namespace dl;

extern "C" {
int dl_voxelize(dl_context *, const dl_splats *, const dl_voxel_opts *, dl_voxel_result **out) {
  return guard([&] {
    if (out) *out = nullptr;
    fail(DL_ERR_INTERNAL, "voxelize: device path not built");
  });
}
uint64_t dl_voxel_result_count(const dl_voxel_result *) { return 0; }
int dl_voxel_result_blocks(const dl_voxel_result *, uint32_t *, uint32_t *, uint32_t *) { return DL_ERR_INTERNAL; }
int dl_voxel_result_bounds(const dl_voxel_result *, double *, double *, double *) { return DL_ERR_INTERNAL; }
float dl_voxel_result_device_ms(const dl_voxel_result *) { return 0.0f; }
void dl_voxel_result_free(dl_voxel_result *r) { delete r; }
}
