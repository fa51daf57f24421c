// vr_vbx.cpp -- .vbx container I/O (gvdb-voxel-src/GVDB_FILESPEC.txt).
#include "../../include/vrestir.h"

extern "C" int32_t vrestir_write_vbx(const char*, const float*, int32_t, int32_t, int32_t, const float*) {
    return VRESTIR_EUNSUPPORTED;
}
extern "C" int32_t vrestir_read_vbx(const char*, float*, int32_t*, int32_t*, int32_t*, float*) {
    return VRESTIR_EUNSUPPORTED;
}
