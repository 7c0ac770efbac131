// Wavefront-cooperative traversal engine (FCLX_ENGINE_FAST).
#include "engine.h"

namespace fclx {

struct EngineWorkspace {
  int unused;
};

EngineWorkspace* engine_workspace_create() { return new EngineWorkspace{}; }
void engine_workspace_destroy(EngineWorkspace* ws) { delete ws; }

hipError_t engine_collide(EngineWorkspace*, const CollideLaunch&, int, hipStream_t) {
  return hipErrorNotSupported;
}

}  // namespace fclx
