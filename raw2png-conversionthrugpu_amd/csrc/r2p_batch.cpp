// r2p_batch.cpp -- batched form (placeholder until the batched launch lands).
#include "r2p_common.h"

using namespace r2p;

extern "C" {
int r2p_batch_create(int, size_t, size_t, const r2p_options*, r2p_batch**) { set_error("batch not built"); return R2P_ERR_UNSUPPORTED; }
int r2p_batch_destroy(r2p_batch*) { return R2P_OK; }
int r2p_batch_set(r2p_batch*, int, const r2p_image*, const void* const*) { return R2P_ERR_UNSUPPORTED; }
int r2p_batch_run(r2p_batch*, int) { return R2P_ERR_UNSUPPORTED; }
int r2p_batch_sync(r2p_batch*) { return R2P_ERR_UNSUPPORTED; }
int r2p_batch_result(r2p_batch*, int, const void**, size_t*) { return R2P_ERR_UNSUPPORTED; }
int r2p_batch_copy_result(r2p_batch*, int, void*, size_t, size_t*) { return R2P_ERR_UNSUPPORTED; }
int r2p_batch_timings(r2p_batch*, r2p_timings*) { return R2P_ERR_UNSUPPORTED; }
int r2p_batch_autotune(r2p_batch*, r2p_options*, double*) { return R2P_ERR_UNSUPPORTED; }
int r2p_batch_options(r2p_batch*, const r2p_options*) { return R2P_ERR_UNSUPPORTED; }
int r2p_server_create(const r2p_server_config*, r2p_server**) { return R2P_ERR_UNSUPPORTED; }
int r2p_server_destroy(r2p_server*) { return R2P_OK; }
int r2p_server_submit(r2p_server*, const r2p_image*, const void*, int64_t*) { return R2P_ERR_UNSUPPORTED; }
int r2p_server_wait(r2p_server*, int64_t, int, void*, size_t, size_t*) { return R2P_ERR_UNSUPPORTED; }
int r2p_server_stats_get(r2p_server*, r2p_server_stats*) { return R2P_ERR_UNSUPPORTED; }
int r2p_server_loadtest(r2p_server*, int, const double*, const int*, const r2p_image*, const void* const*, double*, uint64_t*, double*) { return R2P_ERR_UNSUPPORTED; }
}
