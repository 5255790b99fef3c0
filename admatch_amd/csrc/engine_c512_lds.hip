// Engine kernels with active-list capacity 512, LDS-bitmap storage.
#define ADM_CAP 512
#define ADM_HASHED 0
#include "engine_inst.inc"
