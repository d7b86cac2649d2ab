#pragma once
#include <cstdint>
#include <vector>
namespace megamol::core::view {
// core/include/mmcore/view/Framebuffer.h + CPUFramebuffer.h
struct CPUFramebuffer { bool depthBufferActive = false; std::vector<uint32_t> colorBuffer; std::vector<float> depthBuffer; unsigned int width = 0, height = 0; int x = 0, y = 0; };
}
