#ifndef _YATSCF_DIIS_H_
#define _YATSCF_DIIS_H_

#include "TinySCF.h"

void TinySCF_DIIS(TinySCF_t TinySCF);

#endif
