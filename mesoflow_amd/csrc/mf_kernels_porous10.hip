// Kernels of the porous10 plugin family (one translation unit per family keeps
// the template instantiations compiling in parallel); its discrete adjoint is
// in mf_adjoint_porous10.hip (compiled with the default instruction scheduler).
#include "mf_kernels.inc.h"

namespace mfd {
extern const AdjSet adj_porous10;
const KernelSet ks_porous10 = Launch<PPorous<10>>::set(&adj_porous10);
}
