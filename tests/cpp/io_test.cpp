// io_test.cpp -- host-only checks of include/tdmd/io/ (no device):
//   io_test thermo <step> <pe> <ke> <natoms>     prints the thermo line
//   io_test copy <in.dump> <out.dump> <step> <lox> <hix> <loy> <hiy> <loz> <hiz>
//                                                 reads a dump, writes it back
//   io_test read <in.dump>                        prints the atom count
// A tdmd::Error prints "error: <what>" and exits with status 2.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tdmd/io/dump.hpp"
#include "tdmd/io/thermo.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 1;
  const std::string mode = argv[1];
  try {
    if (mode == "thermo" && argc == 6) {
      std::printf("%s\n", tdmd::io::thermo_line(std::atoll(argv[2]), std::strtod(argv[3], nullptr),
                                                std::strtod(argv[4], nullptr),
                                                std::atoll(argv[5]))
                              .c_str());
      return 0;
    }
    if (mode == "read" && argc == 3) {
      const auto atoms = tdmd::io::read_lammps_dump(argv[2]);
      std::printf("%zu\n", atoms.size());
      return 0;
    }
    if (mode == "copy" && argc == 11) {
      const auto atoms = tdmd::io::read_lammps_dump(argv[2]);
      tdmd::Box box;
      box.lo = {std::strtod(argv[5], nullptr), std::strtod(argv[7], nullptr),
                std::strtod(argv[9], nullptr)};
      box.hi = {std::strtod(argv[6], nullptr), std::strtod(argv[8], nullptr),
                std::strtod(argv[10], nullptr)};
      std::vector<tdmd::i32> ids, types;
      std::vector<tdmd::PositionVec> pos;
      std::vector<tdmd::ForceVec> f;
      for (const auto& a : atoms) {
        ids.push_back(a.id);
        types.push_back(a.type);
        pos.push_back(a.pos);
        f.push_back(a.force);
      }
      tdmd::io::write_lammps_force_dump(argv[3], std::atoll(argv[4]), box,
                                        static_cast<tdmd::i64>(atoms.size()), ids.data(),
                                        types.data(), pos.data(), f.data());
      return 0;
    }
  } catch (const tdmd::Error& e) {
    std::printf("error: %s\n", e.what());
    return 2;
  }
  return 1;
}
