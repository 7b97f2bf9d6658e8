// List-path KSG instances (build and walk kernels) for d = 1, one
// translation unit per dimension so that they compile in parallel.
#include "ksg_walk.h"

namespace nsx {
namespace ksgw {

BuildFn build_d1(int width) { return build_lookup<1>(width); }
WalkFn walk_d1(int k) { return walk_lookup<1>(k); }
TieFn tie_d1(int k, int kind) { return tie_lookup<1>(k, kind); }
WalkFn bigk_d1() { return bigk_kernel<1>; }

}  // namespace ksgw
}  // namespace nsx
