// system_state.hpp -- SystemState (src/core/system_state.hpp): the host
// state the reference's readers fill and its drivers hand to the schedulers.
#pragma once

#include <string>
#include <vector>

#include "box.hpp"
#include "types.hpp"

namespace tdmd {

struct SystemState {
  i64 natoms{0};
  Box box;
  std::vector<PositionVec> positions;
  std::vector<VelocityVec> velocities;
  std::vector<ForceVec> forces;
  std::vector<i32> types;
  std::vector<i32> ids;
  std::vector<real> masses;  // by type; index 0 unused
  std::vector<std::string> type_names;
  i64 step{0};
  real time{0};

  void resize(i64 n) {
    natoms = n;
    const auto k = static_cast<std::size_t>(n);
    positions.resize(k);
    velocities.resize(k);
    forces.resize(k);
    types.resize(k);
    ids.resize(k);
  }
};

}  // namespace tdmd
