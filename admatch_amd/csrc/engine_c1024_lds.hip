// Engine kernels with active-list capacity 1024, LDS-bitmap storage.
#define ADM_CAP 1024
#define ADM_HASHED 0
#include "engine_inst.inc"
