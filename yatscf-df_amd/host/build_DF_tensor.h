#ifndef _YATSCF_BUILD_DF_TENSOR_H_
#define _YATSCF_BUILD_DF_TENSOR_H_

#include "TinySCF.h"

void TinySCF_build_DF_tensor(TinySCF_t TinySCF);

#endif
