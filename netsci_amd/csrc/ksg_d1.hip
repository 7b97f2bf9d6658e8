// KSG kernel instances for d = 1 (one translation unit per dimension so
// that the instances compile in parallel).
#include "ksg_kernel.h"

namespace nsx {
namespace ksg {

KernelFn kernel_d1(int k, int qpl, bool sym) { return kernel_lookup<1>(k, qpl, sym); }

}  // namespace ksg
}  // namespace nsx
