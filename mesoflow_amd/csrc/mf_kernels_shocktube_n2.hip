// Kernels of the shocktube_n2 plugin family (one translation unit per family keeps
// the template instantiations compiling in parallel).
#include "mf_kernels.inc.h"

namespace mfd {
const KernelSet ks_shocktube_n2 = Launch<PShockTubeN2>::set();
}
