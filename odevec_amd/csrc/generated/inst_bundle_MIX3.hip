// Generated by odevec_amd.build -- do not edit.
// bundle MIX3: ROBER/NONE/DENSE FP64, ROBER/NONE/DENSE mixed, OREGO/NONE/DENSE FP64, OREGO/NONE/DENSE mixed, PRIMORDIAL/CMK/CSC FP64, PRIMORDIAL/CMK/CSC mixed
#include "odevec_registry.h"
#include "generated/net_ROBER_NONE_DENSE.h"
#include "generated/net_OREGO_NONE_DENSE.h"
#include "generated/net_PRIMORDIAL_CMK_CSC.h"
namespace {
using Bundle_MIX3 = odv::Bundle<
    odv::Entry<odv_net_ROBER_NONE_DENSE, double, false>,
    odv::Entry<odv_net_ROBER_NONE_DENSE, float, true>,
    odv::Entry<odv_net_OREGO_NONE_DENSE, double, false>,
    odv::Entry<odv_net_OREGO_NONE_DENSE, float, true>,
    odv::Entry<odv_net_PRIMORDIAL_CMK_CSC, double, false>,
    odv::Entry<odv_net_PRIMORDIAL_CMK_CSC, float, true>>;
const char* const nets_MIX3[] = {"ROBER/NONE/DENSE", "ROBER/NONE/DENSE", "OREGO/NONE/DENSE", "OREGO/NONE/DENSE", "PRIMORDIAL/CMK/CSC", "PRIMORDIAL/CMK/CSC"};
const int precs_MIX3[] = {0, 1, 0, 1, 0, 1};
const int neq_MIX3[] = {odv_net_ROBER_NONE_DENSE::NEQ, odv_net_ROBER_NONE_DENSE::NEQ, odv_net_OREGO_NONE_DENSE::NEQ, odv_net_OREGO_NONE_DENSE::NEQ, odv_net_PRIMORDIAL_CMK_CSC::NEQ, odv_net_PRIMORDIAL_CMK_CSC::NEQ};
const int npar_MIX3[] = {odv_net_ROBER_NONE_DENSE::NPARAM, odv_net_ROBER_NONE_DENSE::NPARAM, odv_net_OREGO_NONE_DENSE::NPARAM, odv_net_OREGO_NONE_DENSE::NPARAM, odv_net_PRIMORDIAL_CMK_CSC::NPARAM, odv_net_PRIMORDIAL_CMK_CSC::NPARAM};
void (*const defs_MIX3[])(double*) = {&odv_net_ROBER_NONE_DENSE::param_defaults, &odv_net_ROBER_NONE_DENSE::param_defaults, &odv_net_OREGO_NONE_DENSE::param_defaults, &odv_net_OREGO_NONE_DENSE::param_defaults, &odv_net_PRIMORDIAL_CMK_CSC::param_defaults, &odv_net_PRIMORDIAL_CMK_CSC::param_defaults};
}  // namespace
ODV_REGISTER_BUNDLE(MIX3, Bundle_MIX3, nets_MIX3, precs_MIX3, neq_MIX3, npar_MIX3, defs_MIX3)
