// fused.hip - fused collide + interpolated-stream path (placeholder until implemented).
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace ipl {

bool fused_supported(const Graph&, const ipl_config&, std::string& why) {
    why = "fused path not built";
    return false;
}

std::unique_ptr<Engine> make_fused_engine(const Graph&, const ipl_config&, const HostState&, hipStream_t,
                                          Status& status) {
    status = Status::err(IPL_ERR_UNSUPPORTED, "fused path not built");
    return nullptr;
}

}  // namespace ipl
