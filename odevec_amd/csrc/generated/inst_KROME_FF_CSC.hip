// Generated by odevec_amd.pre_odevec -- do not edit.
#include "../odevec_registry.h"
#include "net_KROME_FF_CSC.h"
ODV_REGISTER(odv_net_KROME_FF_CSC)
