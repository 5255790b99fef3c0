// Engine kernels with active-list capacity 256, hashed storage.
#define ADM_CAP 256
#define ADM_HASHED 1
#include "engine_inst.inc"
