// step2d_fused.h -- tile shape of the fused barotropic substep.
#pragma once

#ifndef SF_FUSED_TI
#define SF_FUSED_TI 64
#endif
#ifndef SF_FUSED_TJ
#define SF_FUSED_TJ 8
#endif
