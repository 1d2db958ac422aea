// Library-wide entry points: ABI version, error text, stream synchronisation.
#include "common.h"

namespace cvfes {
std::string &last_error() {
    static thread_local std::string msg;
    return msg;
}
}  // namespace cvfes

extern "C" int cvfes_abi_version(void) { return CVFES_ABI_VERSION; }

extern "C" int64_t cvfes_struct_size(int which) {
    switch (which) {
        case 0: return sizeof(cvfes_system_t);
        case 1: return sizeof(cvfes_load_t);
        case 2: return sizeof(cvfes_batch_mesh_t);
        case 3: return sizeof(cvfes_geometry_t);
        case 4: return sizeof(cvfes_update_t);
        default: return -1;
    }
}

extern "C" const char *cvfes_last_error(void) { return cvfes::last_error().c_str(); }

extern "C" int cvfes_stream_sync(void *stream) {
    hipError_t err = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (err != hipSuccess)
        return cvfes::fail(CVFES_ERR_HIP, "hipStreamSynchronize: %s", hipGetErrorString(err));
    return CVFES_OK;
}
