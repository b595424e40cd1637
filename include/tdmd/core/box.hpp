// box.hpp -- orthorhombic box (src/core/box.hpp:17-26).
#pragma once

#include "types.hpp"

namespace tdmd {

struct Box {
  Vec3D lo{0, 0, 0};
  Vec3D hi{0, 0, 0};
  std::array<bool, 3> periodic{true, true, true};

  Vec3D size() const noexcept { return Vec3D{hi.x - lo.x, hi.y - lo.y, hi.z - lo.z}; }
};

}  // namespace tdmd
