// Kernel instantiations for polynomial degree 1.
#include "fsg_dispatch.h"

namespace fsg {
FSG_DEFINE_LAUNCHERS(1, 5, double)
FSG_DEFINE_LAUNCHERS(1, 5, float)

}  // namespace fsg
