// Host sanitizer run of the oracle (test infrastructure only): built by `make sanitize` with
// AddressSanitizer and UndefinedBehaviorSanitizer, it builds a mixed scene through the oracle's C
// API -- spheres, boxes, capsules, unions of spheres and planes -- and steps it with every
// relaxation model, both trig modes, the HashGrids and exhaustive detectors, forced sweep orders
// (plain and ranked), an abort threshold, forces, damping and the member round trip, so that
// every allocation, index and table of the oracle is exercised.
// Exit status 0 when the sanitizers report nothing (they abort the process on a finding).
#include <cmath>
#include <cstdio>
#include <vector>

#include "oracle_api.h"

static orc_world* build(int model, int exact) {
  orc_world* w = orc_create();
  orc_set_gravity(w, 0.0, 0.0, -9.81);
  orc_set_max_iterations(w, 12);
  orc_set_relaxation_model(w, model);
  orc_set_trig_mode(w, exact);
  const int m0 = orc_create_material(w, 2500.0, 0.1, 0.4, 0.3);
  const int m1 = orc_create_material(w, 7800.0, 0.2, 0.6, 0.5);
  orc_set_pair_friction(w, m0, m1, 0.5, 0.45);
  orc_create_plane(w, 0.0, 0.0, 1.0, 0.0, m0);
  orc_create_plane(w, 1.0, 0.0, 0.0, -4.0, m1);
  orc_create_plane(w, -0.2, 0.1, 0.97, -3.0, m0);
  unsigned s = 12345u;
  auto u = [&s](double lo, double hi) {
    s = s * 1664525u + 1013904223u;
    return lo + (hi - lo) * ((s >> 8) / 16777216.0);
  };
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j)
      for (int k = 0; k < 4; ++k) {
        const double x = -2.2 + 0.8 * i + u(-0.02, 0.02), y = -2.2 + 0.8 * j + u(-0.02, 0.02);
        const double z = 0.42 + 0.8 * k;
        switch ((i + j + k) % 4) {
          case 0: orc_create_sphere(w, x, y, z, u(0.3, 0.45), m0); break;
          case 1: orc_create_box(w, x, y, z, u(0.4, 0.8), u(0.4, 0.8), u(0.4, 0.8), m1); break;
          case 2: orc_create_capsule(w, x, y, z, u(0.15, 0.3), u(0.2, 0.5), m0); break;
          default: {
            const double disp[9] = {0.0, 0.0, 0.0, 0.2, 0.05, 0.0, -0.1, 0.18, 0.05};
            const double rad[3] = {0.25, 0.2, 0.18};
            const int mat[3] = {m0, m1, m0};
            orc_create_union(w, x, y, z, 3, disp, rad, mat);
          }
        }
      }
  // small random velocities and spins: pe's orthogonal-projection models divide by the tangential
  // impulse's norm, which is exactly zero for a body falling straight onto a plane (HCSITS.h:2694-2698:
  // the dual-cone test fsq < p_n / mu^2 cannot hold for p_n < 0, so such a contact reaches the
  // division and yields NaN in pe as in its restatement)
  const size_t nb = orc_num_bodies(w);
  std::vector<double> pos(3 * nb), q(4 * nb), v(3 * nb), om(3 * nb), nrm(3 * nb), d(nb);
  orc_get_state(w, pos.data(), q.data(), v.data(), om.data(), nrm.data(), d.data());
  std::vector<int> shape(nb), mat(nb), fixed(nb);
  std::vector<uint64_t> sid(nb);
  std::vector<double> param(3 * nb), im(nb), ib(9 * nb);
  orc_get_body_info(w, shape.data(), sid.data(), mat.data(), fixed.data(), param.data(), im.data(), ib.data());
  for (size_t i = 0; i < nb; ++i)
    if (!fixed[i])
      for (int k = 0; k < 3; ++k) { v[3 * i + k] = u(-0.3, 0.3); om[3 * i + k] = u(-1.0, 1.0); }
  orc_set_state(w, pos.data(), q.data(), v.data(), om.data(), nrm.data(), d.data());
  return w;
}

int main() {
  size_t total = 0;
  for (int model = 0; model <= 5; ++model)
    for (int exact = 0; exact <= 1; ++exact) {
      orc_world* w = build(model, exact);
      if (orc_step(w, 2e-3, 150) != 0) return 2;
      const size_t nc = orc_detect(w, 0);   // the contacts the next step will relax
      total += nc;
      std::vector<uint32_t> b1(nc), b2(nc);
      std::vector<double> pos(3 * nc), nrm(3 * nc), dist(nc), mu(nc), p(3 * nc);
      orc_get_contacts(w, b1.data(), b2.data(), pos.data(), nrm.data(), dist.data(), mu.data());
      orc_get_impulses(w, p.data());
      // the next step in reversed contact order, then to an abort threshold
      std::vector<uint32_t> r1(b1.rbegin(), b1.rend()), r2(b2.rbegin(), b2.rend());
      orc_set_sweep_order(w, nc, r1.data(), r2.data());
      if (orc_step(w, 2e-3, 1) != 0) return 2;
      orc_set_abort_threshold(w, 1e-9);
      orc_set_max_iterations(w, 200);
      if (orc_step(w, 2e-3, 2) != 0) return 2;
      const size_t a = orc_detect(w, 0), b = orc_detect(w, 1);
      if (a != b) {
        std::printf("model %d: HashGrids %zu contacts, exhaustive %zu\n", model, a, b);
        return 3;
      }
      // the impulses of a detection are zero, one per contact
      std::vector<double> pd(3 * b + 3, -1.0);
      orc_get_impulses(w, pd.data());
      for (size_t k = 0; k < 3 * b; ++k)
        if (pd[k] != 0.0) return 5;
      // forces for one step, members read and written back, a damped step, the ranked order
      const size_t nb = orc_num_bodies(w), nm = orc_num_members(w);
      std::vector<double> force(3 * nb, 0.0), torque(3 * nb, 0.0);
      for (size_t i = 0; i < nb; ++i) { force[3 * i] = 0.5; torque[3 * i + 2] = 0.1; }
      orc_set_forces(w, force.data(), torque.data());
      std::vector<uint32_t> owner(nm);
      std::vector<uint64_t> msid(nm);
      std::vector<int> mmat(nm);
      std::vector<double> mrad(nm), mrpos(3 * nm), mgpos(3 * nm);
      orc_get_members(w, owner.data(), msid.data(), mmat.data(), mrad.data(), mrpos.data(), mgpos.data());
      orc_set_members(w, mgpos.data());
      orc_set_damping(w, 0.99);
      orc_set_erp(w, 0.5);
      const size_t nc2 = orc_detect(w, 0);
      std::vector<uint32_t> c1(nc2), c2(nc2), rank(nc2, 0);
      std::vector<double> cp(3 * nc2), cn(3 * nc2), cd(nc2), cm(nc2);
      orc_get_contacts(w, c1.data(), c2.data(), cp.data(), cn.data(), cd.data(), cm.data());
      for (size_t k = 1; k < nc2; ++k)   // the rank of each contact among its pair's, in emission order
        for (size_t j = 0; j < k; ++j)
          if (c1[j] == c1[k] && c2[j] == c2[k]) ++rank[k];
      orc_set_sweep_order_ranked(w, nc2, c1.data(), c2.data(), rank.data());
      if (orc_step(w, 2e-3, 1) != 0) return 2;
      std::vector<uint32_t> pa(nb * nb), pb(nb * nb);
      orc_candidate_pairs(w, pa.data(), pb.data(), nb * nb);
      orc_destroy(w);
    }
  std::printf("sanitized oracle runs: %zu contacts over 12 runs\n", total);
  return total > 0 ? 0 : 4;
}
