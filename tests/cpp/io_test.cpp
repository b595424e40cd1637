// io_test.cpp -- host-only checks of include/tdmd/io/ (no device):
//   io_test thermo <step> <pe> <ke> <natoms>     prints the thermo line
//   io_test copy <in.dump> <out.dump> <step> <lox> <hix> <loy> <hiy> <loz> <hiz>
//                                                 reads a dump, writes it back
//   io_test read <in.dump>                        prints the atom count
//   io_test data_copy <in.data> <out.data>        reads a LAMMPS data file,
//                                                 writes it back, prints N
//   io_test setfl_write <tables.txt> <out.setfl>  writes setfl tables given
//                                                 as plain text (below)
// A tdmd::Error prints "error: <what>" and exits with status 2.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <fstream>

#include "tdmd/io/dump.hpp"
#include "tdmd/io/lammps_data.hpp"
#include "tdmd/io/setfl.hpp"
#include "tdmd/io/thermo.hpp"

namespace {

// tables.txt: "ne nrho drho nr dr cutoff", the element names, the masses,
// then F, rho per element and r*phi per pair, whitespace separated.
tdmd::io::SetflData read_tables(const char* path) {
  std::ifstream in(path);
  tdmd::io::SetflData t;
  std::size_t ne = 0;
  in >> ne >> t.nrho >> t.drho >> t.nr >> t.dr >> t.cutoff;
  t.elements.resize(ne);
  t.masses.resize(ne);
  for (auto& e : t.elements) in >> e;
  for (auto& m : t.masses) in >> m;
  auto vec = [&](std::size_t n) {
    std::vector<double> v(n);
    for (auto& x : v) in >> x;
    return v;
  };
  for (std::size_t e = 0; e < ne; ++e) {
    t.F.push_back(vec(static_cast<std::size_t>(t.nrho)));
    t.rho.push_back(vec(static_cast<std::size_t>(t.nr)));
  }
  for (std::size_t p = 0; p < ne * (ne + 1) / 2; ++p) t.rphi.push_back(vec(static_cast<std::size_t>(t.nr)));
  if (!in) throw tdmd::Error(TDMD_E_IO, "bad tables file");
  return t;
}

}  // namespace

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
    if (mode == "data_copy" && argc == 4) {
      const auto st = tdmd::io::read_lammps_data(argv[2]);
      tdmd::io::write_lammps_data(argv[3], st);
      std::printf("%lld\n", static_cast<long long>(st.natoms));
      return 0;
    }
    if (mode == "setfl_write" && argc == 4) {
      tdmd::io::write_setfl(argv[3], read_tables(argv[2]));
      return 0;
    }
  } catch (const tdmd::Error& e) {
    std::printf("error: %s\n", e.what());
    return 2;
  }
  return 1;
}
