// Continuous-batching request server (placeholder until the server lands).
#include <hip/hip_runtime.h>
#include "../../../include/weave.h"
extern "C" {
WS_API int ws_server_create(ws_ctx*, const ws_server_opts*, ws_server**) { return WS_ERR_STATE; }
WS_API int ws_server_destroy(ws_server*) { return WS_ERR_STATE; }
WS_API int ws_submit(ws_server*, const ws_model*, const ws_state*, int32_t, int64_t*) { return WS_ERR_STATE; }
WS_API int ws_cancel(ws_server*, int64_t) { return WS_ERR_STATE; }
WS_API int ws_pump(ws_server*, int32_t*, int32_t*) { return WS_ERR_STATE; }
WS_API int ws_poll(ws_server*, ws_result*, int32_t, int32_t*) { return WS_ERR_STATE; }
WS_API int ws_fetch(ws_server*, int64_t, ws_state*) { return WS_ERR_STATE; }
WS_API double ws_server_clock_ms(ws_server*) { return 0.0; }
}
