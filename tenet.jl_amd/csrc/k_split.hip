// Split-precision contraction (filled in below).
#include "../../include/tenet_hip.h"
#include "internal.h"

namespace tenet {

int launch_split_gemm(const GemmLaunch& g, int dtype, int prec, hipStream_t s, void* ws, size_t ws_bytes) {
  (void)g; (void)dtype; (void)prec; (void)s; (void)ws; (void)ws_bytes;
  set_error("split precision not built yet");
  return TENET_ERR_UNSUPPORTED;
}
size_t gemm_workspace_bytes(const GemmLaunch& g, int dtype, int prec) {
  (void)g; (void)dtype; (void)prec;
  return 0;
}

}  // namespace tenet
