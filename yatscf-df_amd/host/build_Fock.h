#ifndef _YATSCF_BUILD_FOCK_H_
#define _YATSCF_BUILD_FOCK_H_

#include "TinySCF.h"

void TinySCF_build_FockMat(TinySCF_t TinySCF);

#endif
