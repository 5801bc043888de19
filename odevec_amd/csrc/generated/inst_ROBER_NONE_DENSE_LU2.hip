// Generated by odevec_amd.pre_odevec -- do not edit.
#include "../odevec_registry.h"
#include "net_ROBER_NONE_DENSE_LU2.h"
ODV_REGISTER(odv_net_ROBER_NONE_DENSE_LU2)
