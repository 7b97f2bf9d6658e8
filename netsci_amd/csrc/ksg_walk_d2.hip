// List-path KSG instances (build and walk kernels) for d = 2, one
// translation unit per dimension so that they compile in parallel.
#include "ksg_walk.h"

namespace nsx {
namespace ksgw {

BuildFn build_d2(int width) { return build_lookup<2>(width); }
WalkFn walk_d2(int k) { return walk_lookup<2>(k); }
TieFn tie_d2(int k, int kind) { return tie_lookup<2>(k, kind); }
WalkFn bigk_d2() { return bigk_kernel<2>; }

}  // namespace ksgw
}  // namespace nsx
