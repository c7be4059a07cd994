// A pe-style program on the drop-in API: spheres on a lattice fall into a box of planes
// (the setup of examples/lattice/Lattice.cpp:197-217 with a deterministic jitter).
// Usage: granular_dropin <nx> <ny> <nz> <steps> [capsules]; the optional capsules (radius 0.25,
// length 0.8, axis along x) drop onto the spheres.  Prints the final state as hex floats.
#include <cstdio>
#include <cstdlib>
#include <exception>

#include <pe_amd/pe.hpp>

using namespace pe;

static int run(int argc, char** argv) {
  const int nx = argc > 1 ? std::atoi(argv[1]) : 5;
  const int ny = argc > 2 ? std::atoi(argv[2]) : 5;
  const int nz = argc > 3 ? std::atoi(argv[3]) : 5;
  const unsigned steps = argc > 4 ? (unsigned)std::atoi(argv[4]) : 100;
  const int ncaps = argc > 5 ? std::atoi(argv[5]) : 0;
  const int nunions = argc > 6 ? std::atoi(argv[6]) : 0;
  const real r = 0.5, spacing = 1.2, dt = 1e-3;
  const real lx = 2.0 + nx * spacing, ly = 2.0 + ny * spacing, lz = 40.0;

  WorldID world = theWorld();
  world->setGravity(0, 0, -9.81);
  world->setDamping(1);
  CollisionSystemID cs = theCollisionSystem();
  cs->setMaxIterations(50);
  cs->setRelaxationModel(CollisionSystem::InelasticCoulombContactByDecoupling);

  MaterialID granular = createMaterial("granular", 2500, 0, 0.2, 0.2, 0.5, 1, 1, 0, 0);
  createPlane(0, Vec3(0, 0, +1), 0, granular);
  createPlane(1, Vec3(0, 0, -1), -lz, granular);
  createPlane(2, Vec3(0, +1, 0), 0, granular);
  createPlane(3, Vec3(0, -1, 0), -ly, granular);
  createPlane(4, Vec3(+1, 0, 0), 0, granular);
  createPlane(5, Vec3(-1, 0, 0), -lx, granular);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const real jit = 0.001 * (real)((i * 7 + j * 13 + k * 17) % 11 - 5);
        createSphere(6, Vec3(1.5 + spacing * i + jit, 1.5 + spacing * j - jit, 0.6 + spacing * k + jit), r, granular);
      }
  for (int c = 0; c < ncaps; ++c)
    createCapsule(7, Vec3(1.6 + 1.1 * c, 0.5 * ly + 0.05 * c, 0.6 + spacing * (nz - 1) + 0.77), 0.25, 0.8, granular);
  for (int u = 0; u < nunions; ++u)  // dumbbells and triplets of spheres dropped on the pile
    createUnion(8, Vec3(1.7 + 1.0 * u, 0.5 * ly - 0.6 + 0.07 * u, 0.6 + spacing * (nz - 1) + 1.9),
                {UnionSphere{Vec3(-0.25, 0.0, 0.02), 0.22, granular}, UnionSphere{Vec3(0.25, 0.03, 0.0), 0.2, granular},
                 UnionSphere{Vec3(0.0, 0.2, -0.1 * (u % 2)), 0.18, granular}});

  world->run(steps, dt);

  std::printf("contacts %zu maxpen %a schedule %d\n", cs->getNumberOfContacts(), cs->getMaximumPenetration(),
              cs->getLastSweepSchedule());
  for (size_t b = 0; b < world->size(); ++b) {
    BodyID body = (*world)[b];
    const Vec3& p = body->getPosition();
    const Vec3& v = body->getLinearVel();
    std::printf("%zu %a %a %a %a %a %a\n", b, p[0], p[1], p[2], v[0], v[1], v[2]);
  }
  return 0;
}

int main(int argc, char** argv) {
  // pe reports failures as exceptions (std::invalid_argument, std::runtime_error)
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "granular_dropin: %s\n", e.what());
    return 1;
  }
}
