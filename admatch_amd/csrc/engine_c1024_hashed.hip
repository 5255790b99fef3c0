// Engine kernels with active-list capacity 1024, hashed storage.
#define ADM_CAP 1024
#define ADM_HASHED 1
#include "engine_inst.inc"
