// List-path KSG instances (build and walk kernels) for d = 3, one
// translation unit per dimension so that they compile in parallel.
#include "ksg_walk.h"

namespace nsx {
namespace ksgw {

BuildFn build_d3(int width) { return build_lookup<3>(width); }
WalkFn walk_d3(int k) { return walk_lookup<3>(k); }
TieFn tie_d3(int k, int kind) { return tie_lookup<3>(k, kind); }
WalkFn bigk_d3() { return bigk_kernel<3>; }

}  // namespace ksgw
}  // namespace nsx
