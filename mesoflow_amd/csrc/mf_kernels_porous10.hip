// Kernels of the porous10 plugin family (one translation unit per family keeps
// the template instantiations compiling in parallel), with its discrete adjoint.
#include "mf_adjoint_kernels.inc.h"

namespace mfd {
static const AdjSet adj_porous10 = AdjLaunch<PPorous<10>>::set();
const KernelSet ks_porous10 = Launch<PPorous<10>>::set(&adj_porous10);
}
