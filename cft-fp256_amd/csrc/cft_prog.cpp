// placeholder replaced by the program module
#include "cft_internal.hpp"
extern "C" {
CFT_API cft_status cft_program_load(cft_device *, const void *, size_t, cft_program **o) { if (o) *o = nullptr; return CFT_ERR_UNSUPPORTED; }
CFT_API void cft_program_free(cft_program *) {}
CFT_API cft_status cft_program_get_info(cft_program *, cft_program_info *) { return CFT_ERR_INVALID_ARGUMENT; }
CFT_API cft_status cft_program_run(cft_program *, const void *, const void *, const void *, void *, uint32_t *, size_t, uint32_t *, uint32_t *) { return CFT_ERR_INVALID_ARGUMENT; }
CFT_API cft_status cft_program_run_bank(cft_program *, const void *, size_t, const void *, const void *, const void *, void *, uint32_t *, size_t, uint32_t *, uint32_t *) { return CFT_ERR_INVALID_ARGUMENT; }
CFT_API cft_status cft_program_run_ex(cft_program *, const cft_run_args *) { return CFT_ERR_INVALID_ARGUMENT; }
}
