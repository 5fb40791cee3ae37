// Kernels of the porous10 plugin family (one translation unit per family keeps
// the template instantiations compiling in parallel).
#include "mf_kernels.inc.h"

namespace mfd {
const KernelSet ks_porous10 = Launch<PPorous<10>>::set();
}
