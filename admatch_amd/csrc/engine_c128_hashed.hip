// Engine kernels with active-list capacity 128, hashed storage.
#define ADM_CAP 128
#define ADM_HASHED 1
#include "engine_inst.inc"
