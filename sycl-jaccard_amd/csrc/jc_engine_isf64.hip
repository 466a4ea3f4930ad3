// Owner-engine kernels and launch sequence for one result mode (see jc_engine.hpp).
#include "jc_engine.hpp"

namespace jc {
template hipError_t run_engine_t<kModeIsF64, true>(const EngineArgs&, void*, bool, hipStream_t);
}  // namespace jc
