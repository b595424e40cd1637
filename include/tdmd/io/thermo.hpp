// thermo.hpp -- the driver's per-step thermo line (print_thermo,
// src/drivers/tdmd_main.cpp:110-122) and the KE it reports
// (tdmd_main.cpp:97-108): T from 3N degrees of freedom, T = 2 KE / (3 N kB).
//
// The values come from FastPipelineScheduler::observables() (device FP64
// reductions); these functions only format. Header-only host code;
// tdmd_claude_amd/output.py formats the same line.
#pragma once

#include <cstdio>
#include <string>

#include "../core/types.hpp"

namespace tdmd::io {

inline constexpr double kBoltzmannEv = 8.617333262e-5;  // eV/K (src/core/constants.hpp:15)

[[nodiscard]] inline std::string thermo_line(i64 step, double pe, double ke, i64 natoms) {
  const double te = pe + ke;
  const double temp =
      natoms > 0 ? 2.0 * ke / (3.0 * static_cast<double>(natoms) * kBoltzmannEv) : 0.0;
  char buf[160];
  std::snprintf(buf, sizeof(buf), "Step %8lld  PE %20.14f  KE %20.14f  TE %20.14f  T %10.4f",
                static_cast<long long>(step), pe, ke, te, temp);
  return buf;
}

inline void print_thermo(i64 step, double pe, double ke, i64 natoms) {
  std::printf("%s\n", thermo_line(step, pe, ke, natoms).c_str());
}

}  // namespace tdmd::io
