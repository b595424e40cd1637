// dump.hpp -- LAMMPS custom force dumps, `id type x y z fx fy fz`.
//
//   * write_lammps_force_dump: the driver's writer
//     (src/drivers/tdmd_main.cpp:65-95): reals with 17 significant digits,
//     box bounds tagged `pp pp pp`, atoms in the given order;
//   * read_lammps_dump: the reader (src/io/dump_reader.cpp:15-54, header
//     src/io/dump_reader.hpp:13-23): atoms sorted by id; a missing count, a
//     short file or a malformed line throws tdmd::Error.
//
// Header-only host code; tdmd_claude_amd/output.py writes the same bytes.
#pragma once

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../core/box.hpp"
#include "../core/error.hpp"
#include "../core/types.hpp"

namespace tdmd::io {

struct DumpAtom {
  i32 id;
  i32 type;
  PositionVec pos;
  ForceVec force;
};

[[nodiscard]] inline std::vector<DumpAtom> read_lammps_dump(const std::string& filename) {
  std::ifstream file(filename);
  if (!file.is_open()) throw Error(TDMD_E_INVALID, "cannot open dump file: " + filename);
  std::string line;
  i64 natoms = 0;
  while (std::getline(file, line)) {
    if (line.find("ITEM: NUMBER OF ATOMS") != std::string::npos) {
      if (!std::getline(file, line)) break;
      natoms = std::stoll(line);
    }
    if (line.find("ITEM: ATOMS") != std::string::npos) break;
  }
  if (natoms <= 0) throw Error(TDMD_E_INVALID, "no atoms in dump file");
  std::vector<DumpAtom> atoms;
  atoms.reserve(static_cast<std::size_t>(natoms));
  for (i64 k = 0; k < natoms; ++k) {
    if (!std::getline(file, line)) throw Error(TDMD_E_INVALID, "premature end of dump");
    std::istringstream in(line);
    DumpAtom a{};
    double f[3];
    in >> a.id >> a.type >> a.pos.x >> a.pos.y >> a.pos.z >> f[0] >> f[1] >> f[2];
    if (in.fail()) throw Error(TDMD_E_INVALID, "bad dump line: " + line);
    a.force = ForceVec{static_cast<force_t>(f[0]), static_cast<force_t>(f[1]),
                       static_cast<force_t>(f[2])};
    atoms.push_back(a);
  }
  std::sort(atoms.begin(), atoms.end(),
            [](const DumpAtom& x, const DumpAtom& y) { return x.id < y.id; });
  return atoms;
}

// Writes `natoms` atoms from caller arrays (ids, 1-based types, positions,
// forces), e.g. as FastPipelineScheduler::download returns them.
inline void write_lammps_force_dump(const std::string& path, i64 step, const Box& box,
                                    i64 natoms, const i32* ids, const i32* types,
                                    const PositionVec* pos, const ForceVec* forces) {
  std::ofstream out(path);
  if (!out) throw Error(TDMD_E_INVALID, "cannot open " + path + " for writing");
  out.precision(17);
  out << "ITEM: TIMESTEP\n" << step << "\n";
  out << "ITEM: NUMBER OF ATOMS\n" << natoms << "\n";
  out << "ITEM: BOX BOUNDS pp pp pp\n";
  out << box.lo.x << " " << box.hi.x << "\n";
  out << box.lo.y << " " << box.hi.y << "\n";
  out << box.lo.z << " " << box.hi.z << "\n";
  out << "ITEM: ATOMS id type x y z fx fy fz\n";
  for (i64 k = 0; k < natoms; ++k) {
    const PositionVec& p = pos[k];
    const ForceVec& f = forces[k];
    out << ids[k] << " " << types[k] << " " << p.x << " " << p.y << " " << p.z << " "
        << static_cast<double>(f.x) << " " << static_cast<double>(f.y) << " "
        << static_cast<double>(f.z) << "\n";
  }
  if (!out) throw Error(TDMD_E_INVALID, "write failed: " + path);
}

}  // namespace tdmd::io
