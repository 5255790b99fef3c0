// Engine kernels with active-list capacity 128, LDS-bitmap storage.
#define ADM_CAP 128
#define ADM_HASHED 0
#include "engine_inst.inc"
