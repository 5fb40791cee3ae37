// Kernels of one case written against mf_plugin_api.h (a plugin family of its
// own). The Makefile compiles this file once per case directory with
// -DMF_CASE_NS=<namespace> -DMF_CASE_HEADER="<dir>/mflo_case.H"; the family
// registers itself when the library loads.
#include "mf_kernels.inc.h"
#include MF_CASE_HEADER

namespace mfd {
namespace {

using CaseT = MF_CASE_NS::Case;
static_assert(CaseT::PLUGIN_ID >= 100, "case plugin ids start at 100");
static_assert(CaseT::NUM_SPECIES >= 1 && CaseT::NUM_SPECIES <= MF_MAX_SPECIES, "species count out of range");
static_assert(CaseT::NUM_GAS_SPECIES >= 0 && CaseT::NUM_GAS_SPECIES <= CaseT::NUM_SPECIES, "gas species count");
static_assert(CaseT::GAMMA_IGNORES_PRESSURE, "the WENO5 sweeps stage one gamma per cell: it must not depend on p");

const KernelSet ks_case = Launch<PRef<CaseT>>::set();

void init_species(mf_plugin_params *p) { CaseT::init_species(*p); }

#define MF_CASE_STR2(x) #x
#define MF_CASE_STR(x) MF_CASE_STR2(x)
const int registered = register_case_family(CaseFamily{CaseT::PLUGIN_ID, CaseT::NUM_SPECIES, CaseT::HAS_BG_GAS,
                                                       &ks_case, &init_species, MF_CASE_STR(MF_CASE_NS)});

}  // namespace
}  // namespace mfd
