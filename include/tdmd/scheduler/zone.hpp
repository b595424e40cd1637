// zone.hpp -- the zone and its state machine (src/scheduler/zone.hpp:17-64):
// the unit of work of the zone pipeline, an index range of atoms with its
// own time step.
#pragma once

#include <array>
#include <cstdint>

#include "../core/box.hpp"
#include "../core/error.hpp"

namespace tdmd::scheduler {

enum class ZoneState : std::uint8_t {
  Free = 0,
  Receiving = 1,
  Received = 2,
  Ready = 3,
  Computing = 4,
  Done = 5,
  Sending = 6,
};

// The cycle Free -> Receiving -> Received -> Ready -> Computing -> Done ->
// Sending -> Free; nothing else is allowed.
constexpr bool is_legal_transition(ZoneState from, ZoneState to) noexcept {
  switch (from) {
    case ZoneState::Free: return to == ZoneState::Receiving;
    case ZoneState::Receiving: return to == ZoneState::Received;
    case ZoneState::Received: return to == ZoneState::Ready;
    case ZoneState::Ready: return to == ZoneState::Computing;
    case ZoneState::Computing: return to == ZoneState::Done;
    case ZoneState::Done: return to == ZoneState::Sending;
    case ZoneState::Sending: return to == ZoneState::Free;
  }
  return false;
}

struct Zone {
  i32 id{-1};
  std::array<i32, 3> lattice_index{0, 0, 0};
  Aabb bbox{};
  i32 atom_offset{0};
  i32 natoms_in_zone{0};
  i32 time_step{0};
  ZoneState state{ZoneState::Free};
  i32 owner_rank{-1};

  // Computing -> Done finishes a step. An illegal transition throws
  // tdmd::Error (TDMD_E_STATE) where the reference asserts.
  void transition_to(ZoneState next) {
    if (!is_legal_transition(state, next)) throw Error(TDMD_E_STATE, "illegal zone transition");
    state = next;
    if (next == ZoneState::Done) ++time_step;
  }
};

}  // namespace tdmd::scheduler
