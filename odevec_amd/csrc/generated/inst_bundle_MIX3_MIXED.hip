// Generated by odevec_amd.build -- do not edit.
// bundle MIX3_MIXED: ROBER/NONE/DENSE FP64, OREGO/NONE/DENSE FP64, PRIMORDIAL/CMK/CSC mixed
#include "odevec_registry.h"
#include "generated/net_ROBER_NONE_DENSE.h"
#include "generated/net_OREGO_NONE_DENSE.h"
#include "generated/net_PRIMORDIAL_CMK_CSC.h"
namespace {
using Bundle_MIX3_MIXED = odv::Bundle<
    odv::Entry<odv_net_ROBER_NONE_DENSE, double, false>,
    odv::Entry<odv_net_OREGO_NONE_DENSE, double, false>,
    odv::Entry<odv_net_PRIMORDIAL_CMK_CSC, float, true>>;
const char* const nets_MIX3_MIXED[] = {"ROBER/NONE/DENSE", "OREGO/NONE/DENSE", "PRIMORDIAL/CMK/CSC"};
const int precs_MIX3_MIXED[] = {0, 0, 1};
const int neq_MIX3_MIXED[] = {odv_net_ROBER_NONE_DENSE::NEQ, odv_net_OREGO_NONE_DENSE::NEQ, odv_net_PRIMORDIAL_CMK_CSC::NEQ};
const int npar_MIX3_MIXED[] = {odv_net_ROBER_NONE_DENSE::NPARAM, odv_net_OREGO_NONE_DENSE::NPARAM, odv_net_PRIMORDIAL_CMK_CSC::NPARAM};
void (*const defs_MIX3_MIXED[])(double*) = {&odv_net_ROBER_NONE_DENSE::param_defaults, &odv_net_OREGO_NONE_DENSE::param_defaults, &odv_net_PRIMORDIAL_CMK_CSC::param_defaults};
}  // namespace
ODV_REGISTER_BUNDLE(MIX3_MIXED, Bundle_MIX3_MIXED, nets_MIX3_MIXED, precs_MIX3_MIXED, neq_MIX3_MIXED, npar_MIX3_MIXED, defs_MIX3_MIXED)
