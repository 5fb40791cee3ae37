// Discrete-adjoint kernels of the porous10 plugin family.
#include "mf_adjoint_kernels.inc.h"

namespace mfd {
extern const AdjSet adj_porous10;
const AdjSet adj_porous10 = AdjLaunch<PPorous<10>>::set();
}
