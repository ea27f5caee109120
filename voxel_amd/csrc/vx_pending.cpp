// vx_pending.cpp -- entry points whose GPU kernels are not built yet; they
// report an error instead of computing anything on the host.
#include "../../include/voxel_c.h"
#include "vx_runtime.hpp"

extern "C" {
void* voxel_agg_create(void) { vx::set_error("voxel_agg_*: GPU group-by not built"); return nullptr; }
void voxel_agg_destroy(void*) {}
void voxel_agg_init(void*, size_t) { vx::set_error("voxel_agg_*: GPU group-by not built"); }
void voxel_agg_accumulate(void*, const uint32_t*, const double*, size_t) { vx::set_error("voxel_agg_*: GPU group-by not built"); }
size_t voxel_agg_group_count(const void*) { vx::set_error("voxel_agg_*: GPU group-by not built"); return 0; }
}
