// Kernels of the porous5 plugin family (one translation unit per family keeps
// the template instantiations compiling in parallel).
#include "mf_kernels.inc.h"

namespace mfd {
const KernelSet ks_porous5 = Launch<PPorous<5>>::set();
}
