// KSG kernel instances for d = 2 (one translation unit per dimension so
// that the instances compile in parallel).
#include "ksg_kernel.h"

namespace nsx {
namespace ksg {

KernelFn kernel_d2(int k, int qpl, bool sym) { return kernel_lookup<2>(k, qpl, sym); }

}  // namespace ksg
}  // namespace nsx
