// Kernels of the catalyst_cylinder plugin family (one translation unit per family keeps
// the template instantiations compiling in parallel).
#include "mf_kernels.inc.h"

namespace mfd {
const KernelSet ks_catalyst_cylinder = Launch<PCatalystCylinder>::set();
}
