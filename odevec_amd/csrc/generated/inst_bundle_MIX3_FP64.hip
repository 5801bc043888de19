// Generated by odevec_amd.build -- do not edit.
// bundle MIX3_FP64: ROBER/NONE/DENSE FP64, OREGO/NONE/DENSE FP64, PRIMORDIAL/CMK/CSC FP64
#include "odevec_registry.h"
#include "generated/net_ROBER_NONE_DENSE.h"
#include "generated/net_OREGO_NONE_DENSE.h"
#include "generated/net_PRIMORDIAL_CMK_CSC.h"
namespace {
using Bundle_MIX3_FP64 = odv::Bundle<
    odv::Entry<odv_net_ROBER_NONE_DENSE, double, false>,
    odv::Entry<odv_net_OREGO_NONE_DENSE, double, false>,
    odv::Entry<odv_net_PRIMORDIAL_CMK_CSC, double, false>>;
const char* const nets_MIX3_FP64[] = {"ROBER/NONE/DENSE", "OREGO/NONE/DENSE", "PRIMORDIAL/CMK/CSC"};
const int precs_MIX3_FP64[] = {0, 0, 0};
const int neq_MIX3_FP64[] = {odv_net_ROBER_NONE_DENSE::NEQ, odv_net_OREGO_NONE_DENSE::NEQ, odv_net_PRIMORDIAL_CMK_CSC::NEQ};
const int npar_MIX3_FP64[] = {odv_net_ROBER_NONE_DENSE::NPARAM, odv_net_OREGO_NONE_DENSE::NPARAM, odv_net_PRIMORDIAL_CMK_CSC::NPARAM};
void (*const defs_MIX3_FP64[])(double*) = {&odv_net_ROBER_NONE_DENSE::param_defaults, &odv_net_OREGO_NONE_DENSE::param_defaults, &odv_net_PRIMORDIAL_CMK_CSC::param_defaults};
}  // namespace
ODV_REGISTER_BUNDLE(MIX3_FP64, Bundle_MIX3_FP64, nets_MIX3_FP64, precs_MIX3_FP64, neq_MIX3_FP64, npar_MIX3_FP64, defs_MIX3_FP64)
