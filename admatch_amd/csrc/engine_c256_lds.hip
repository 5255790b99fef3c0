// Engine kernels with active-list capacity 256, LDS-bitmap storage.
#define ADM_CAP 256
#define ADM_HASHED 0
#include "engine_inst.inc"
