// Kernel instantiations for polynomial degree 3.
#include "fsg_dispatch.h"

namespace fsg {
FSG_DEFINE_LAUNCHERS(3, 5, double)
FSG_DEFINE_LAUNCHERS(3, 5, float)

}  // namespace fsg
