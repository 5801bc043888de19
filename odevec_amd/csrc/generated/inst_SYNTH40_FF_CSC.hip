// Generated by odevec_amd.pre_odevec -- do not edit.
#include "../odevec_registry.h"
#include "net_SYNTH40_FF_CSC.h"
ODV_REGISTER(odv_net_SYNTH40_FF_CSC)
