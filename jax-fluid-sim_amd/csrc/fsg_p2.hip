// Kernel instantiations for polynomial degree 2.
#include "fsg_dispatch.h"

namespace fsg {
FSG_DEFINE_LAUNCHERS(2, 5, double)
FSG_DEFINE_LAUNCHERS(2, 5, float)
FSG_DEFINE_LAUNCHERS(2, 16, double)
FSG_DEFINE_LAUNCHERS(2, 16, float)
}  // namespace fsg
