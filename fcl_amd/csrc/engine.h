// Wavefront-cooperative traversal engine (FCLX_ENGINE_FAST).
#pragma once
#include "kernels.h"

namespace fclx {

struct EngineWorkspace;

EngineWorkspace* engine_workspace_create();
void engine_workspace_destroy(EngineWorkspace* ws);
hipError_t engine_collide(EngineWorkspace* ws, const CollideLaunch& L, int stack_need,
                          hipStream_t s);

}  // namespace fclx
