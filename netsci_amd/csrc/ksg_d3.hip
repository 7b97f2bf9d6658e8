// KSG kernel instances for d = 3 (one translation unit per dimension so
// that the instances compile in parallel).
#include "ksg_kernel.h"

namespace nsx {
namespace ksg {

KernelFn kernel_d3(int k, int qpl, bool sym) { return kernel_lookup<3>(k, qpl, sym); }

}  // namespace ksg
}  // namespace nsx
