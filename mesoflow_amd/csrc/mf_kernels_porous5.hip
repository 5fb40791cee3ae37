// Kernels of the porous5 plugin family (one translation unit per family keeps
// the template instantiations compiling in parallel); its discrete adjoint is
// in mf_adjoint_porous5.hip (compiled with the default instruction scheduler).
#include "mf_kernels.inc.h"

namespace mfd {
extern const AdjSet adj_porous5;
const KernelSet ks_porous5 = Launch<PPorous<5>>::set(&adj_porous5);
}
