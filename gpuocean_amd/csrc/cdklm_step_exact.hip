// Exact build of the CDKLM16 substep kernel (IEEE division/sqrt, no contraction).
#define CDKLM_FAST 0
#include "cdklm_step.inc"
