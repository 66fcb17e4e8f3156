// Fast build of the CDKLM16 substep kernel (reciprocal division, FMA contraction).
#define CDKLM_FAST 1
#include "cdklm_step.inc"
