// pe's binary body files through the C++ drop-in (include/pe_amd/pe.hpp) for the CPU tests.
//   dropin_body_file write <file> <fp-size>   builds a mixed scene (planes, spheres, a fixed box,
//                                             a capsule, unions) and writes it with BodyBinaryWriter
//   dropin_body_file read <file>              reads the file with BodyBinaryReader into a world
//                                             holding the scene's two materials and prints the
//                                             body and member arrays, one "name values" line each
//   dropin_body_file tryread <file> <nmat>    reads it into a world of nmat (1 or 2) materials and
//                                             prints whether it was refused and the body count
// Nothing is stepped, so no device is needed.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <pe_amd/pe.hpp>

using namespace pe;

template <class T>
static void put(const char* name, const std::vector<T>& a) {
  std::printf("%s", name);
  for (const T& x : a) std::printf(" %.17g", static_cast<double>(x));
  std::printf("\n");
}

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  MaterialID gran = createMaterial("granular", 2500, 0, 0.2, 0.2, 0.5, 1, 1, 0, 0);
  if (std::strcmp(argv[1], "tryread") == 0) {
    // a world holding only the first material: a file whose bodies use the second, or a file cut
    // short, is refused, and the world keeps no body of it
    if (argc > 3 && std::strcmp(argv[3], "2") == 0) createMaterial("wood", 700, 0.1, 0.35, 0.3, 0.5, 1, 1, 0, 0);
    try {
      BodyBinaryReader r;
      r.readFile(argv[2]);
      std::printf("read\n");
    } catch (const std::exception& ex) {
      std::printf("refused %s\n", ex.what());
    }
    std::printf("bodies %zu members %zu\n", theWorld()->size(), detail::E().m_owner.size());
    return 0;
  }
  MaterialID wood = createMaterial("wood", 700, 0.1, 0.35, 0.3, 0.5, 1, 1, 0, 0);
  if (std::strcmp(argv[1], "write") == 0) {
    createPlane(100, Vec3(0, 0, 1), 0, gran);
    createPlane(101, Vec3(1, -2, 0.5), -3.25, wood);
    SphereID s1 = createSphere(1, Vec3(1.5, -0.25, 2), 0.375, gran);
    BoxID b = createBox(2, Vec3(-1, 2, 0.75), Vec3(0.5, 1.25, 0.8), wood);
    createCapsule(3, Vec3(0.3, 0.1, 4), 0.2, 1.1, gran);
    UnionID u = createUnion(4, Vec3(2, 2, 3),
                            {UnionSphere{Vec3(0.3, 0, 0), 0.25, gran}, UnionSphere{Vec3(-0.1, 0.2, 0.05), 0.3, wood},
                             UnionSphere{Vec3(0, -0.15, 0.2), 0.2, gran}});
    createUnion(5, Vec3(-2, -1, 1.5), {UnionSphere{Vec3(0.1, 0.1, 0.1), 0.4, wood}});
    createSphere(6, Vec3(0, 0, 7), 0.5, wood);
    b->setFixed(true);
    s1->setLinearVel(Vec3(0.5, -0.25, 1.0));
    s1->setAngularVel(Vec3(2, 0, -1));
    s1->setOrientation(Quat(Vec3(1, 2, 3), 0.4));
    u->setLinearVel(Vec3(0, 0, -2));
    BodyBinaryWriter w;
    if (argc > 3) w.setFloatingPointSize(std::atoi(argv[3]));
    w.writeFile(argv[2]);
    return 0;
  }
  BodyBinaryReader r;
  r.readFile(argv[2]);
  const detail::Engine& e = detail::E();
  put("shape", e.shape);
  put("sid", e.sid);
  put("uid", e.uid);
  put("mat", e.mat);
  put("fixed", e.fixed);
  put("param", e.param);
  put("pos", e.pos);
  put("q", e.q);
  put("v", e.v);
  put("w", e.w);
  put("inv_mass", e.inv_mass);
  put("I", e.I);
  put("Iinv", e.Iinv);
  put("normal", e.normal);
  put("d", e.d);
  put("m_owner", e.m_owner);
  put("m_sid", e.m_sid);
  put("m_uid", e.m_uid);
  put("m_mat", e.m_mat);
  put("m_rad", e.m_rad);
  put("m_rpos", e.m_rpos);
  put("m_gpos", e.m_gpos);
  std::printf("next_sid %llu\n", (unsigned long long)e.next_sid);
  return 0;
}
