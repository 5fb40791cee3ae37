// Kernels of the porous5 plugin family (one translation unit per family keeps
// the template instantiations compiling in parallel), with its discrete adjoint.
#include "mf_adjoint_kernels.inc.h"

namespace mfd {
static const AdjSet adj_porous5 = AdjLaunch<PPorous<5>>::set();
const KernelSet ks_porous5 = Launch<PPorous<5>>::set(&adj_porous5);
}
