// setfl_potential.hpp -- an EamAlloy-shaped holder of setfl tables read by
// tdmd_setfl_read (EamAlloy::read_setfl, src/potentials/eam_alloy.cpp:
// 90-192): knot values only, phi = (r phi)/r with phi(0) = 0, pair tables in
// phi_index order. For callers without the reference's EamAlloy.
#pragma once

#include <string>
#include <vector>

#include "../core/error.hpp"
#include "../core/types.hpp"

namespace tdmd::potentials {

struct SetflTable {
  i32 n{0};
  real dr{0};
  real rmin{0};
  std::vector<real> values;
};

class SetflPotential {
 public:
  void read_setfl(const std::string& filename) {
    i32 nt = 0, nrho = 0, nr = 0;
    double drho = 0, dr = 0, rc = 0;
    TDMD_HIP_CALL(tdmd_setfl_read(filename.c_str(), &nt, &nrho, &drho, &nr, &dr, &rc, nullptr,
                                  nullptr, nullptr));
    const i32 np = nt * (nt + 1) / 2;
    std::vector<double> F(static_cast<std::size_t>(nt) * nrho);
    std::vector<double> rho(static_cast<std::size_t>(nt) * nr);
    std::vector<double> rphi(static_cast<std::size_t>(np) * nr);
    TDMD_HIP_CALL(tdmd_setfl_read(filename.c_str(), &nt, &nrho, &drho, &nr, &dr, &rc, F.data(),
                                  rho.data(), rphi.data()));
    ntypes_ = nt;
    cutoff_ = static_cast<real>(rc);
    embedding_.assign(static_cast<std::size_t>(nt), {});
    density_.assign(static_cast<std::size_t>(nt), {});
    phi_.assign(static_cast<std::size_t>(np), {});
    for (i32 t = 0; t < nt; ++t) {
      fill(embedding_[static_cast<std::size_t>(t)], &F[static_cast<std::size_t>(t) * nrho], nrho,
           drho);
      fill(density_[static_cast<std::size_t>(t)], &rho[static_cast<std::size_t>(t) * nr], nr, dr);
    }
    for (i32 p = 0; p < np; ++p) {
      SetflTable& tab = phi_[static_cast<std::size_t>(p)];
      const double* rp = &rphi[static_cast<std::size_t>(p) * nr];
      tab.n = nr;
      tab.dr = static_cast<real>(dr);
      tab.values.assign(static_cast<std::size_t>(nr), real{0});
      for (i32 k = 1; k < nr; ++k)
        tab.values[static_cast<std::size_t>(k)] =
            static_cast<real>(rp[k] / (static_cast<double>(k) * dr));
    }
  }

  real cutoff() const noexcept { return cutoff_; }
  i32 ntypes() const noexcept { return ntypes_; }
  const std::vector<SetflTable>& embedding() const noexcept { return embedding_; }
  const std::vector<SetflTable>& density() const noexcept { return density_; }
  const std::vector<SetflTable>& phi() const noexcept { return phi_; }

 private:
  static void fill(SetflTable& t, const double* v, i32 n, double dr) {
    t.n = n;
    t.dr = static_cast<real>(dr);
    t.values.assign(v, v + n);
  }
  i32 ntypes_{0};
  real cutoff_{0};
  std::vector<SetflTable> embedding_, density_, phi_;
};

}  // namespace tdmd::potentials
