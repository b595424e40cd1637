// dropin_test.cpp -- the C++ drop-in interface (include/tdmd/tdmd.hpp)
// exercised the way the reference's own code calls it: DeviceNeighborList,
// compute_morse_gpu, DeviceEam, the integrator functions, the NHC
// reductions, FastPipelineScheduler and the zone pipeline. Checks against host brute force and
// against the reference's step sequence assembled from the free functions
// (fast_pipeline_scheduler.cu:101-166,195-232). Built twice, once per
// precision macro. Prints one line per check; exit status 0 = all passed.
//
// usage: dropin_test <path to a setfl file> [dump file to write]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <string>
#include <vector>

#include "tdmd/tdmd.hpp"

using namespace tdmd;

namespace {

int g_failures = 0;
std::string g_dump_path = "dropin_test.dump";

void report(const char* name, bool ok, const std::string& detail) {
  std::printf("%s %s: %s\n", ok ? "PASS" : "FAIL", name, detail.c_str());
  std::fflush(stdout);
  if (!ok) ++g_failures;
}

constexpr bool kMixed = sizeof(real) == 4;
constexpr double kMvv2e = 1.0364269e-4;

struct Lcg {  // deterministic jitter, no library RNG
  unsigned long long s;
  double next() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return static_cast<double>(s >> 11) * (1.0 / 9007199254740992.0);
  }
};

struct System {
  Box box;
  std::vector<PositionVec> pos;
  std::vector<VelocityVec> vel;
  std::vector<i32> types, ids;
};

System fcc(int n, double a, double jitter, double vscale, unsigned long long seed) {
  System s;
  s.box.lo = Vec3D{0, 0, 0};
  s.box.hi = Vec3D{n * a, n * a, n * a};
  const double basis[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
  Lcg g{seed};
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x)
        for (auto& b : basis) {
          s.pos.push_back(PositionVec{(x + b[0]) * a + jitter * (2 * g.next() - 1),
                                      (y + b[1]) * a + jitter * (2 * g.next() - 1),
                                      (z + b[2]) * a + jitter * (2 * g.next() - 1)});
          s.vel.push_back(VelocityVec{vscale * (2 * g.next() - 1), vscale * (2 * g.next() - 1),
                                      vscale * (2 * g.next() - 1)});
          s.types.push_back(1);
          s.ids.push_back(static_cast<i32>(s.ids.size()) + 1);
        }
  return s;
}

double mimg(double d, double L) {
  if (d > 0.5 * L) return d - L;
  if (d < -0.5 * L) return d + L;
  return d;
}

double r2_of(const System& s, int i, int j) {
  const Vec3D L = s.box.size();
  const double dx = mimg(s.pos[j].x - s.pos[i].x, L.x);
  const double dy = mimg(s.pos[j].y - s.pos[i].y, L.y);
  const double dz = mimg(s.pos[j].z - s.pos[i].z, L.z);
  const double xx = dx * dx, yy = dy * dy, zz = dz * dz;
  return (xx + yy) + zz;
}

template <class T>
std::vector<T> to_host(const DeviceBuffer<T>& b, std::size_t n) {
  std::vector<T> h(n);
  b.copy_to_host(h.data(), n);
  return h;
}

template <class T>
std::vector<T> to_host_ptr(const T* d, std::size_t n) {
  std::vector<T> h(n);
  TDMD_HIP_CALL(tdmd_memcpy_dtoh(h.data(), d, n * sizeof(T), nullptr));
  return h;
}

// ---------------------------------------------------------------------------

void check_neighbor_list(const System& s) {
  const int N = static_cast<int>(s.pos.size());
  DeviceBuffer<PositionVec> d_pos(N);
  d_pos.copy_from_host(s.pos.data(), N);
  neighbors::DeviceNeighborList nl;
  nl.build(d_pos.data(), N, s.box, real(6.0), real(1.0));
  TDMD_HIP_CALL(tdmd_device_synchronize());
  auto cnt = to_host_ptr(nl.d_counts(), N);
  auto off = to_host_ptr(nl.d_offsets(), N);
  const int total = off[N - 1] + cnt[N - 1];
  auto nbr = to_host_ptr(nl.d_neighbors(), total);
  const double rl2 = 7.0 * 7.0;
  int mismatched = 0, maxc = 0;
  for (int i = 0; i < N; ++i) {
    std::set<int> want, got(nbr.begin() + off[i], nbr.begin() + off[i] + cnt[i]);
    for (int j = 0; j < N; ++j)
      if (j != i && r2_of(s, i, j) < rl2) want.insert(j);
    if (want != got || static_cast<int>(got.size()) != cnt[i]) ++mismatched;
    maxc = std::max(maxc, cnt[i]);
  }
  report("DeviceNeighborList::build vs brute force", mismatched == 0 && maxc == nl.max_neighbors(),
         std::to_string(N) + " atoms, " + std::to_string(total) + " pairs, " +
             std::to_string(mismatched) + " rows differ, max " + std::to_string(maxc));
}

void host_morse(const System& s, const potentials::MorseParams& p, std::vector<Vec3D>* f,
                double* pe) {
  const int N = static_cast<int>(s.pos.size());
  const Vec3D L = s.box.size();
  f->assign(N, Vec3D{0, 0, 0});
  *pe = 0;
  const double D = p.d, al = p.alpha, r0 = p.r0, rc2 = static_cast<double>(p.rc) * p.rc;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      if (j == i) continue;
      const double dx = mimg(s.pos[i].x - s.pos[j].x, L.x);
      const double dy = mimg(s.pos[i].y - s.pos[j].y, L.y);
      const double dz = mimg(s.pos[i].z - s.pos[j].z, L.z);
      const double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 >= rc2) continue;
      const double r = std::sqrt(r2);
      const double e = std::exp(-al * (r - r0));
      const double fpair = -2.0 * D * al * (1.0 - e) * e / r;
      (*f)[i].x += fpair * dx;
      (*f)[i].y += fpair * dy;
      (*f)[i].z += fpair * dz;
      *pe += 0.5 * D * ((1.0 - e) * (1.0 - e) - 1.0);
    }
}

double max_rel(const std::vector<Vec3D>& a, const std::vector<ForceVec>& b) {
  double rms = 0, mx = 0;
  for (std::size_t i = 0; i < a.size(); ++i) {
    rms += a[i].x * a[i].x + a[i].y * a[i].y + a[i].z * a[i].z;
    const double dx = a[i].x - b[i].x, dy = a[i].y - b[i].y, dz = a[i].z - b[i].z;
    mx = std::max(mx, std::sqrt(dx * dx + dy * dy + dz * dz));
  }
  rms = std::sqrt(rms / static_cast<double>(a.size()));
  return mx / rms;
}

potentials::MorseParams morse_c1() {
  // Cu Morse of the reference tests (tests/unit/test_pipeline_scheduler.cu:52):
  // D 0.3429 eV, alpha 1.3588 1/A, r0 2.866 A, rc 6.0 A.
  return potentials::MorseParams{real(0.3429), real(1.3588), real(2.866), real(6.0), real(36.0)};
}

void check_morse(const System& s) {
  const int N = static_cast<int>(s.pos.size());
  DeviceBuffer<PositionVec> d_pos(N);
  d_pos.copy_from_host(s.pos.data(), N);
  neighbors::DeviceNeighborList nl;
  nl.build(d_pos.data(), N, s.box, real(6.0), real(1.0));
  DeviceBuffer<ForceVec> d_f(N);
  DeviceBuffer<accum_t> d_e(1);
  d_e.zero();
  const auto p = morse_c1();
  potentials::compute_morse_gpu(d_pos.data(), d_f.data(), nl.d_neighbors(), nl.d_offsets(),
                                nl.d_counts(), N, s.box, p, d_e.data());
  TDMD_HIP_CALL(tdmd_device_synchronize());
  auto f = to_host(d_f, N);
  auto e = to_host(d_e, 1);
  std::vector<Vec3D> fh;
  double peh = 0;
  host_morse(s, p, &fh, &peh);
  const double err = max_rel(fh, f), tol = kMixed ? 1e-5 : 1e-12;
  const double perr = std::fabs(e[0] - peh) / std::fabs(peh);
  report("compute_morse_gpu vs host FP64", err <= tol && perr <= (kMixed ? 1e-6 : 1e-12),
         "max|dF|/rms " + std::to_string(err) + ", PE rel " + std::to_string(perr));
}

// The reference step sequence (fast_pipeline_scheduler.cu:101-166,195-232)
// assembled from the drop-in free functions: fixed-interval rebuild,
// bootstrap half kick + drift, cross-step full kick + drift, finalize half
// kick at the end of the window.
void reference_sequence(const System& s0, const potentials::MorseParams& p, int nsteps,
                        int rebuild_every, real dt, std::vector<PositionVec>* pos_out,
                        std::vector<VelocityVec>* vel_out) {
  const int N = static_cast<int>(s0.pos.size());
  DeviceBuffer<PositionVec> x(N);
  DeviceBuffer<VelocityVec> v(N);
  DeviceBuffer<ForceVec> f(N);
  DeviceBuffer<i32> types(N);
  DeviceBuffer<real> masses(2);
  const real m[2] = {real(1.0), real(63.546)};
  x.copy_from_host(s0.pos.data(), N);
  v.copy_from_host(s0.vel.data(), N);
  types.copy_from_host(s0.types.data(), N);
  masses.copy_from_host(m, 2);
  neighbors::DeviceNeighborList nl;
  nl.build(x.data(), N, s0.box, p.rc, real(1.0));
  potentials::compute_morse_gpu(x.data(), f.data(), nl.d_neighbors(), nl.d_offsets(),
                                nl.d_counts(), N, s0.box, p, nullptr);
  int since = 0;
  for (int k = 0; k < nsteps; ++k) {
    if (k == 0)
      integrator::device_fused_half_kick_drift(v.data(), x.data(), f.data(), types.data(),
                                               masses.data(), N, dt, s0.box);
    else
      integrator::device_fused_full_kick_drift(v.data(), x.data(), f.data(), types.data(),
                                               masses.data(), N, dt, s0.box);
    if (since >= rebuild_every) {
      nl.build(x.data(), N, s0.box, p.rc, real(1.0));
      since = 0;
    }
    potentials::compute_morse_gpu(x.data(), f.data(), nl.d_neighbors(), nl.d_offsets(),
                                  nl.d_counts(), N, s0.box, p, nullptr);
    ++since;
  }
  integrator::device_half_kick(v.data(), f.data(), types.data(), masses.data(), N, dt);
  TDMD_HIP_CALL(tdmd_device_synchronize());
  *pos_out = to_host(x, N);
  *vel_out = to_host(v, N);
}

void check_scheduler_morse(const System& s) {
  const int N = static_cast<int>(s.pos.size());
  const auto p = morse_c1();
  scheduler::FastPipelineConfig cfg;  // reference defaults: dt 1 fs, skin 1, every 10
  scheduler::FastPipelineScheduler sch(s.box, N, p, cfg);
  std::vector<ForceVec> f0(N, ForceVec{0, 0, 0});
  const real m[2] = {real(1.0), real(63.546)};
  sch.upload(s.pos.data(), s.vel.data(), f0.data(), s.types.data(), s.ids.data(), m, 2);
  const int nsteps = 50;
  sch.run_until(nsteps);
  std::vector<PositionVec> x(N);
  std::vector<VelocityVec> v(N);
  std::vector<ForceVec> f(N);
  std::vector<i32> ty(N), id(N);
  sch.download(x.data(), v.data(), f.data(), ty.data(), id.data(), N);
  std::vector<PositionVec> xr;
  std::vector<VelocityVec> vr;
  reference_sequence(s, p, nsteps, cfg.rebuild_every, cfg.dt, &xr, &vr);
  double dx = 0, dv = 0;
  const Vec3D L = s.box.size();
  bool order = true;
  for (int i = 0; i < N; ++i) {
    dx = std::max(dx, std::fabs(mimg(x[i].x - xr[i].x, L.x)));
    dx = std::max(dx, std::fabs(mimg(x[i].y - xr[i].y, L.y)));
    dx = std::max(dx, std::fabs(mimg(x[i].z - xr[i].z, L.z)));
    dv = std::max({dv, std::fabs(v[i].x - vr[i].x), std::fabs(v[i].y - vr[i].y),
                   std::fabs(v[i].z - vr[i].z)});
    order = order && id[i] == s.ids[i] && ty[i] == s.types[i];
  }
  const double tol = kMixed ? 1e-6 : 1e-10;
  const auto& st = sch.stats();
  report("FastPipelineScheduler(Morse) vs reference step sequence",
         dx <= tol && order && sch.current_step() == nsteps && st.ticks == nsteps &&
             sch.potential_kind() == scheduler::PotentialKind::Morse,
         "max |dx| " + std::to_string(dx) + " A, max |dv| " + std::to_string(dv) +
             ", rebuilds " + std::to_string(st.rebuilds) + ", launches " +
             std::to_string(st.kernel_launches));
}

// The zone pipeline: ZonePartition against the reference's partition tests
// (tests/unit/test_zone_partition.cpp), the deterministic PipelineScheduler
// against the zone sweep assembled from the drop-in free functions (the
// zone step of pipeline_scheduler.cu:128-168 in index order), and the
// pipelined mode's zones all reaching the target
// (tests/unit/test_pipeline_scheduler.cu:184-216).
void check_pipeline(const System& s) {
  const int N = static_cast<int>(s.pos.size());
  const auto p = morse_c1();
  {
    const System lat = fcc(4, 3.615, 0.0, 0.0, 1);  // the reference's cu_fcc_256 lattice
    domain::ZonePartition zp;
    zp.build(lat.box, real(6.0));
    auto x = lat.pos;
    auto v = lat.vel;
    std::vector<ForceVec> f(N, ForceVec{0, 0, 0});
    auto ty = lat.types;
    auto id = lat.ids;
    zp.assign_atoms(x.data(), v.data(), f.data(), ty.data(), id.data(), N);
    std::set<i32> seen(id.begin(), id.end());
    bool inside = true;
    int total = 0;
    for (const auto& z : zp.zones()) {
      total += z.natoms_in_zone;
      for (int k = z.atom_offset; k < z.atom_offset + z.natoms_in_zone; ++k)
        inside = inside && x[k].x >= z.bbox.lo.x && x[k].x < z.bbox.hi.x + 1e-10;
    }
    Box b20;
    b20.hi = Vec3D{20, 20, 20};
    domain::ZonePartition z4;
    z4.build(b20, real(5.0), 4);
    z4.build_zone_neighbors(real(6.5));
    const bool all4 = z4.zone_neighbors(0).size() == 4 && z4.zone_neighbors(3).size() == 4;
    z4.build_zone_neighbors(real(4.0));
    const bool three = z4.zone_neighbors(1).size() == 3 && z4.zone_neighbors(1)[0] == 0;
    bool threw = false;
    scheduler::Zone zone;
    try {
      zone.transition_to(scheduler::ZoneState::Computing);  // Free -> Computing
    } catch (const Error& e) {
      threw = e.status() == TDMD_E_STATE;
    }
    report("ZonePartition / zone state machine",
           zp.n_zones() == 3 && total == N && static_cast<int>(seen.size()) == N && inside &&
               all4 && three && threw,
           std::to_string(zp.n_zones()) + " zones of " + std::to_string(zp.zone_width()) + " A");
  }
  const real m[2] = {real(1.0), real(63.546)};
  const std::vector<ForceVec> f0(N, ForceVec{0, 0, 0});
  scheduler::PipelineConfig cfg;
  cfg.deterministic = true;
  scheduler::PipelineScheduler pipe(s.box, N, p, cfg);
  pipe.upload(s.pos.data(), s.vel.data(), f0.data(), s.types.data(), s.ids.data(), m, 2);
  const int nsteps = 40;
  pipe.run_until(nsteps);
  std::vector<PositionVec> x(N);
  std::vector<VelocityVec> v(N);
  std::vector<ForceVec> f(N);
  std::vector<i32> ty(N), id(N);
  pipe.download(x.data(), v.data(), f.data(), ty.data(), id.data(), N);
  // the sweep from the free functions, on the same zone order
  domain::ZonePartition zp;
  zp.build(s.box, p.rc);
  auto xr = s.pos;
  auto vr = s.vel;
  auto fr = f0;
  auto tr = s.types;
  auto ir = s.ids;
  zp.assign_atoms(xr.data(), vr.data(), fr.data(), tr.data(), ir.data(), N);
  DeviceBuffer<PositionVec> dx(N);
  DeviceBuffer<VelocityVec> dv(N);
  DeviceBuffer<ForceVec> df(N);
  DeviceBuffer<i32> dt(N);
  DeviceBuffer<real> dm(2);
  dx.copy_from_host(xr.data(), N);
  dv.copy_from_host(vr.data(), N);
  dt.copy_from_host(tr.data(), N);
  dm.copy_from_host(m, 2);
  neighbors::DeviceNeighborList nl;
  nl.build(dx.data(), N, s.box, p.rc, cfg.r_skin);
  df.zero();
  potentials::compute_morse_gpu(dx.data(), df.data(), nl.d_neighbors(), nl.d_offsets(),
                                nl.d_counts(), N, s.box, p, nullptr);
  for (int step = 0; step < nsteps; ++step)
    for (const auto& z : zp.zones()) {
      const i32 a = z.atom_offset, c = z.natoms_in_zone;
      integrator::device_half_kick_zone(dv.data(), df.data(), dt.data(), dm.data(), a, c, cfg.dt);
      integrator::device_drift_zone(dx.data(), dv.data(), a, c, cfg.dt, s.box);
      integrator::device_zero_forces_zone(df.data(), a, c);
      potentials::compute_morse_gpu_zone(dx.data(), df.data(), nl.d_neighbors(), nl.d_offsets(),
                                         nl.d_counts(), a, c, s.box, p, nullptr);
      integrator::device_half_kick_zone(dv.data(), df.data(), dt.data(), dm.data(), a, c, cfg.dt);
    }
  dx.copy_to_host(xr.data(), N);
  dv.copy_to_host(vr.data(), N);
  bool same = true;
  for (int i = 0; i < N; ++i)
    same = same && x[i].x == xr[i].x && x[i].y == xr[i].y && x[i].z == xr[i].z &&
           v[i].x == vr[i].x && v[i].y == vr[i].y && v[i].z == vr[i].z && id[i] == ir[i];
  const auto& st = pipe.stats();
  report("PipelineScheduler(deterministic) vs zone sweep of the free functions",
         same && pipe.min_time_step() == nsteps &&
             st.kernel_launches == static_cast<i64>(nsteps) * pipe.partition().n_zones(),
         std::string(same ? "bitwise equal" : "differ") + ", " +
             std::to_string(pipe.partition().n_zones()) + " zones, " +
             std::to_string(st.ticks) + " ticks");
  scheduler::PipelineConfig c2;
  c2.n_streams = 2;
  scheduler::PipelineScheduler pipe2(s.box, N, p, c2);
  pipe2.upload(s.pos.data(), s.vel.data(), f0.data(), s.types.data(), s.ids.data(), m, 2);
  pipe2.run_until(10);
  bool all = pipe2.min_time_step() >= 10;
  for (const auto& z : pipe2.partition().zones()) all = all && z.time_step >= 10;
  report("PipelineScheduler(2 streams): zones reach the target", all,
         "min time step " + std::to_string(pipe2.min_time_step()));
}

void check_eam(const System& s, const std::string& setfl) {
  const int N = static_cast<int>(s.pos.size());
  potentials::SetflPotential pot;
  pot.read_setfl(setfl);
  DeviceBuffer<PositionVec> d_pos(N);
  d_pos.copy_from_host(s.pos.data(), N);
  DeviceBuffer<i32> d_types(N);
  d_types.copy_from_host(s.types.data(), N);
  neighbors::DeviceNeighborList nl;
  potentials::DeviceEam eam;
  eam.upload_tables(pot);
  nl.build(d_pos.data(), N, s.box, eam.cutoff(), real(1.0));
  DeviceBuffer<ForceVec> d_f(N);
  DeviceBuffer<accum_t> d_e(1);
  d_e.zero();
  eam.compute(d_pos.data(), d_f.data(), d_types.data(), nl.d_neighbors(), nl.d_offsets(),
              nl.d_counts(), N, s.box, d_e.data());
  TDMD_HIP_CALL(tdmd_device_synchronize());
  auto f = to_host(d_f, N);
  auto e = to_host(d_e, 1);
  if (kMixed) {
    // Mixed forces against the library's FP64 path on the same tables (the
    // energy is too coarse in FP32 for a finite difference).
    std::vector<double> fv, rv, pv;
    for (const auto& t : pot.embedding()) fv.insert(fv.end(), t.values.begin(), t.values.end());
    for (const auto& t : pot.density()) rv.insert(rv.end(), t.values.begin(), t.values.end());
    for (const auto& t : pot.phi()) pv.insert(pv.end(), t.values.begin(), t.values.end());
    tdmd_eam_tables* t64 = nullptr;
    TDMD_HIP_CALL(tdmd_eam_tables_create_values(
        &t64, TDMD_FP64, pot.ntypes(), pot.embedding()[0].n, pot.embedding()[0].dr,
        pot.density()[0].n, pot.density()[0].dr, pot.cutoff(), fv.data(), rv.data(), pv.data()));
    DeviceBuffer<Vec3D> f64(N);
    DeviceBuffer<double> rho(N), fp(N), e64(1);
    e64.zero();
    const tdmd_box b = hip::to_c(s.box);
    TDMD_HIP_CALL(tdmd_eam_compute(t64, reinterpret_cast<const double*>(d_pos.data()), f64.data(),
                                   d_types.data(), nl.d_neighbors(), nl.d_offsets(),
                                   nl.d_counts(), N, &b, rho.data(), fp.data(), e64.data(),
                                   nullptr, nullptr));
    TDMD_HIP_CALL(tdmd_device_synchronize());
    auto fr = to_host(f64, N);
    const double pe64 = to_host(e64, 1)[0];
    (void)tdmd_eam_tables_destroy(t64);
    const double err = max_rel(fr, f);
    report("DeviceEam::compute (mixed) vs FP64 path", err <= 1e-5,
           "max|dF|/rms " + std::to_string(err) + ", PE " + std::to_string(e[0]) + " vs " +
               std::to_string(pe64));
  } else {
    // Finite difference of the device energy along one coordinate of atom 7.
    const int a = 7;
    const double h = 1e-4;
    double ep = 0, em = 0;
    for (int sgn : {+1, -1}) {
      System t = s;
      t.pos[a].x += sgn * h;
      DeviceBuffer<PositionVec> dp(N);
      dp.copy_from_host(t.pos.data(), N);
      neighbors::DeviceNeighborList nl2;
      nl2.build(dp.data(), N, t.box, eam.cutoff(), real(1.0));
      DeviceBuffer<ForceVec> ff(N);
      DeviceBuffer<accum_t> ee(1);
      ee.zero();
      eam.compute(dp.data(), ff.data(), d_types.data(), nl2.d_neighbors(), nl2.d_offsets(),
                  nl2.d_counts(), N, t.box, ee.data());
      TDMD_HIP_CALL(tdmd_device_synchronize());
      (sgn > 0 ? ep : em) = to_host(ee, 1)[0];
    }
    const double fd = -(ep - em) / (2 * h);
    const double fa = static_cast<double>(f[a].x);
    const double err = std::fabs(fd - fa) / std::max(1e-3, std::fabs(fa));
    report("DeviceEam::compute force vs finite difference of its energy", err < 1e-6,
           "F " + std::to_string(fa) + " vs FD " + std::to_string(fd) + ", PE/atom " +
               std::to_string(e[0] / N));
  }

  // FastPipelineScheduler(EAM) from the EamAlloy-shaped tables: NVE energy.
  scheduler::FastPipelineConfig cfg;
  scheduler::FastPipelineScheduler sch(s.box, N, pot, cfg);
  std::vector<ForceVec> f0(N, ForceVec{0, 0, 0});
  const real m[2] = {real(1.0), real(63.546)};
  sch.upload(s.pos.data(), s.vel.data(), f0.data(), s.types.data(), s.ids.data(), m, 2);
  sch.run_until(0);
  auto th0 = sch.thermo();
  sch.run_until(200);
  auto th1 = sch.thermo();
  const double e0 = th0[0] + th0[1], e1 = th1[0] + th1[1];
  const double de = std::fabs(e1 - e0) / N;
  report("FastPipelineScheduler(EAM) NVE energy over 200 steps",
         de < 1e-4 && sch.potential_kind() == scheduler::PotentialKind::Eam &&
             std::fabs(sch.interaction_cutoff() - pot.cutoff()) == 0 &&
             std::fabs(th0[0] - e[0]) <= 1e-6 * std::fabs(e[0]),
         "|dE|/N " + std::to_string(de) + " eV, PE0 " + std::to_string(th0[0]) +
             " vs DeviceEam " + std::to_string(e[0]));

  // driver outputs: observables -> thermo line; download -> force dump,
  // read back with the reference's reader semantics (sorted by id)
  const tdmd_observables ob = sch.observables();
  std::vector<PositionVec> p1(N);
  std::vector<VelocityVec> v1(N);
  std::vector<ForceVec> f1(N);
  std::vector<i32> t1(N), id1(N);
  sch.download(p1.data(), v1.data(), f1.data(), t1.data(), id1.data(), N);
  io::write_lammps_force_dump(g_dump_path, ob.step, s.box, N, id1.data(), t1.data(), p1.data(),
                              f1.data());
  const auto back = io::read_lammps_dump(g_dump_path);
  bool same = static_cast<int>(back.size()) == N;
  for (int k = 0; same && k < N; ++k) {
    const auto& b = back[static_cast<size_t>(k)];
    const int q = static_cast<int>(std::find(id1.begin(), id1.end(), b.id) - id1.begin());
    same = q < N && b.type == t1[q] && b.pos.x == p1[q].x && b.pos.y == p1[q].y &&
           b.pos.z == p1[q].z && b.force.x == f1[q].x && b.force.z == f1[q].z &&
           (k == 0 || back[static_cast<size_t>(k) - 1].id < b.id);
  }
  const std::string line = io::thermo_line(ob.step, ob.pe, ob.ke, ob.natoms);
  const bool th_ok = ob.step == 200 && std::fabs(ob.pe - th1[0]) <= 1e-9 * std::fabs(th1[0]) &&
                     std::fabs(ob.ke - th1[1]) <= 1e-9 * th1[1] &&
                     line.rfind("Step      200  PE ", 0) == 0 && std::isfinite(ob.pressure);
  report("observables, thermo line and force dump round trip", same && th_ok,
         line + "; P " + std::to_string(ob.pressure) + " bar");
}

void check_nhc_reductions(const System& s) {
  const int N = static_cast<int>(s.pos.size());
  DeviceBuffer<VelocityVec> v(N);
  v.copy_from_host(s.vel.data(), N);
  DeviceBuffer<i32> t(N);
  t.copy_from_host(s.types.data(), N);
  DeviceBuffer<real> m(2);
  const real mh[2] = {real(1.0), real(63.546)};
  m.copy_from_host(mh, 2);
  const double ke = integrator::device_compute_ke(v.data(), t.data(), m.data(), N);
  double keh = 0, vmax = 0;
  for (int i = 0; i < N; ++i) {
    const double v2 = s.vel[i].x * s.vel[i].x + s.vel[i].y * s.vel[i].y + s.vel[i].z * s.vel[i].z;
    keh += 0.5 * static_cast<double>(mh[1]) * kMvv2e * v2;
    vmax = std::max(vmax, std::sqrt(v2));
  }
  const double vm = integrator::device_compute_vmax(v.data(), N);
  integrator::NoseHooverChain nhc(integrator::NHCConfig{real(300), real(0.1), 3}, real(0.001),
                                  3 * N - 3);
  const real scale = nhc.half_step(ke);
  report("device_compute_ke / vmax / NoseHooverChain",
         std::fabs(ke - keh) <= 1e-12 * keh && std::fabs(vm - vmax) <= 1e-12 * vmax &&
             scale > real(0.9) && scale < real(1.1),
         "KE " + std::to_string(ke) + " vs " + std::to_string(keh) + ", vmax " +
             std::to_string(vm) + ", scale " + std::to_string(scale));
}

void check_errors() {
  bool threw = false;
  try {
    neighbors::DeviceNeighborList nl;
    Box b;
    b.hi = Vec3D{10, 10, 10};
    nl.build(nullptr, 10, b, real(-1.0), real(1.0));  // negative cutoff: bad input
  } catch (const Error& e) {
    threw = e.status() == TDMD_E_INVALID;
  }
  report("bad input throws tdmd::Error", threw, threw ? "TDMD_E_INVALID" : "no exception");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <setfl file>\n", argv[0]);
    return 2;
  }
  if (argc > 2) g_dump_path = argv[2];
  std::printf("precision: %s\n", kMixed ? "mixed" : "fp64");
  std::fflush(stdout);
  const System s = fcc(4, 3.615, 0.05, 4.0, 42);  // 256 atoms, box 14.46 A
  check_neighbor_list(s);
  check_morse(s);
  check_scheduler_morse(s);
  check_pipeline(s);
  check_eam(s, argv[1]);
  check_nhc_reductions(s);
  check_errors();
  std::printf("%d failure(s)\n", g_failures);
  return g_failures == 0 ? 0 : 1;
}
