// Engine kernels with active-list capacity 512, hashed storage.
#define ADM_CAP 512
#define ADM_HASHED 1
#include "engine_inst.inc"
