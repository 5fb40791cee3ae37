// Discrete-adjoint kernels of the porous5 plugin family.
#include "mf_adjoint_kernels.inc.h"

namespace mfd {
extern const AdjSet adj_porous5;
const AdjSet adj_porous5 = AdjLaunch<PPorous<5>>::set();
}
