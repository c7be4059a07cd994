// pe's pose utilities in the C++ drop-in (include/pe_amd/pe.hpp) for the CPU tests: applies a
// fixed sequence of translate / rotate / rotateAroundOrigin / rotateAroundPoint / setPosition /
// setOrientation calls to a box, a sphere, a capsule, a plane and a union of three spheres, and
// prints each body's pose after every call, one line each (%.17g):
//   <step> pos[3] q[4] normal[3] d member-positions[3 per member of that body]
// followed by the members' body-frame positions, one pointFromBFtoWF result and draws of pe's
// random-number functions.  Nothing is stepped, so no device is needed.
#include <cstdio>
#include <string>
#include <vector>

#include <pe_amd/pe.hpp>

using namespace pe;

static void dump(const std::string& step, BodyID b) {
  const detail::Engine& e = detail::E();
  const size_t i = b->index();
  std::printf("%s", step.c_str());
  for (int k = 0; k < 3; ++k) std::printf(" %.17g", e.pos[3 * i + k]);
  for (int k = 0; k < 4; ++k) std::printf(" %.17g", e.q[4 * i + k]);
  for (int k = 0; k < 3; ++k) std::printf(" %.17g", e.normal[3 * i + k]);
  std::printf(" %.17g", e.d[i]);
  for (size_t m = 0; m < e.m_owner.size(); ++m)
    if (e.m_owner[m] == i)
      for (int k = 0; k < 3; ++k) std::printf(" %.17g", e.m_gpos[3 * m + k]);
  std::printf("\n");
}

int main() {
  MaterialID m = createMaterial("granular", 2500, 0, 0.2, 0.2, 0.5, 1, 1, 0, 0);
  BoxID box = createBox(1, Vec3(1.25, -0.5, 2), Vec3(0.5, 1.25, 0.8), m);
  SphereID sph = createSphere(2, Vec3(-1, 0.75, 1.5), 0.4, m);
  CapsuleID cap = createCapsule(3, Vec3(0.5, 2, 3), 0.2, 1.1, m);
  PlaneID pl = createPlane(4, Vec3(0, 0, 1), -0.5, m);
  UnionID un = createUnion(5, Vec3(2, 2, 3),
                           {UnionSphere{Vec3(0.3, 0, 0), 0.25, m}, UnionSphere{Vec3(-0.1, 0.2, 0.05), 0.3, m},
                            UnionSphere{Vec3(0, -0.15, 0.2), 0.2, m}});
  dump("box0", box); dump("sph0", sph); dump("cap0", cap); dump("pl0", pl); dump("un0", un);

  box->rotate(Vec3(1, 2, 3), 0.4);                                dump("box1", box);
  box->rotate(0.3, -0.2, 0.5);                                    dump("box2", box);
  box->rotate(Quat(Vec3(-1, 0.5, 2), 1.1));                       dump("box3", box);
  box->rotateAroundOrigin(Vec3(0, 1, 1), 0.7);                    dump("box4", box);
  box->rotateAroundOrigin(0.2, 0.1, -0.3);                        dump("box5", box);
  box->rotateAroundPoint(Vec3(1, -1, 0.5), Vec3(0.3, 0.3, 1), 0.9); dump("box6", box);
  box->rotateAroundPoint(Vec3(-0.5, 0.25, 1), Vec3(0.4, -0.6, 0.2)); dump("box7", box);
  box->translate(Vec3(0.1, 0.2, -0.3));                           dump("box8", box);

  sph->rotate(1, 0, 1, 0.6);                                      dump("sph1", sph);
  cap->rotate(Vec3(0.7, -0.1, 0.25));                             dump("cap1", cap);

  pl->rotate(Vec3(1, 0.5, 0), 0.3);                               dump("pl1", pl);
  pl->rotateAroundOrigin(Vec3(0, 1, 0.2), 0.25);                  dump("pl2", pl);
  pl->setPosition(Vec3(0.2, 0.3, 1.5));                           dump("pl3", pl);
  pl->setOrientation(Quat(Vec3(1, 1, 0), 0.3));                   dump("pl4", pl);
  pl->translate(0.05, -0.1, 0.2);                                 dump("pl5", pl);

  un->rotate(0.3, 0.6, -0.4);                                     dump("un1", un);
  un->rotateAroundOrigin(Vec3(0.5, -1, 0.3), 0.45);               dump("un2", un);
  un->translate(Vec3(-0.2, 0.1, 0.35));                           dump("un3", un);
  un->rotateAroundPoint(Vec3(1, 1, 1), Vec3(0, 0, 1), -0.8);      dump("un4", un);

  const detail::Engine& e = detail::E();
  std::printf("m_rpos");
  for (real x : e.m_rpos) std::printf(" %.17g", x);
  std::printf("\n");
  const Vec3 w = box->pointFromBFtoWF(Vec3(0.25, -0.5, 0.125));
  std::printf("box_point %.17g %.17g %.17g\n", w[0], w[1], w[2]);

  // the turned box's rotation and inertias (RigidBody.h:152-158, 842-869) and World's getters
  auto mat = [](const char* name, const Mat3& M) {
    std::printf("%s", name);
    for (int k = 0; k < 9; ++k) std::printf(" %.17g", M[k]);
    std::printf("\n");
  };
  mat("box_R", box->getRotation());
  mat("box_Ib", box->getBodyInertia());
  mat("box_Ibinv", box->getInvBodyInertia());
  mat("box_Iw", box->getInertia());
  mat("box_Iwinv", box->getInvInertia());
  theWorld()->setGravity(0, 0, -9.81);
  theWorld()->setDamping(0.875);
  const Vec3 g = theWorld()->getGravity();
  std::printf("world %.17g %.17g %.17g %.17g %llu\n", g[0], g[1], g[2], theWorld()->getDamping(),
              (unsigned long long)theWorld()->getBody(1)->getID());

  // pe's random numbers: the default seed, a seed that repeats its sequence, the ranges, and
  // granular particles that follow the seed
  std::printf("seed0 %u\n", getSeed());
  for (int pass = 0; pass < 2; ++pass) {
    setSeed(7);
    std::printf("draws%d %u", pass, getSeed());
    for (int k = 0; k < 4; ++k) std::printf(" %.17g", rand<real>());
    for (int k = 0; k < 4; ++k) std::printf(" %.17g", rand<real>(-2, 3));
    for (int k = 0; k < 4; ++k) std::printf(" %d", rand<int>(1, 6));
    UnionID gp = createGranularParticle(100 + pass, Vec3(0, 0, 10 + 5 * pass), 0.5, m);
    std::printf(" %zu", gp->size());
    const size_t gi = gp->index();
    for (size_t k = 0; k < e.m_owner.size(); ++k)
      if (e.m_owner[k] == gi) std::printf(" %.17g", e.m_rad[k]);
    std::printf("\n");
  }
  // pe's wall-clock timer: three intervals, then their summary
  timing::WcTimer t;
  for (int k = 0; k < 3; ++k) {
    t.start();
    volatile double x = 0;
    for (int j = 0; j < 200000 * (k + 1); ++j) x = x + 1e-9 * j;
    t.end();
  }
  std::printf("timer %zu %.17g %.17g %.17g %.17g %.17g\n", t.getCounter(), t.total(), t.average(), t.min(), t.max(),
              t.last());
  return 0;
}
