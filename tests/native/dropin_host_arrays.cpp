// Host side of the C++ drop-in (include/pe_amd/pe.hpp) for the CPU tests: creates a fixed set of
// bodies through pe's factory functions and prints the body and union-member arrays the library
// would receive, one "name value value ..." line each (%.17g), then moves and turns a union and
// prints its members again, then destroys two bodies and prints every array again.  Nothing is
// stepped, so no device is needed.
#include <cstdio>
#include <vector>

#include <pe_amd/pe.hpp>

using namespace pe;

template <class T>
static void put(const char* name, const std::vector<T>& a) {
  std::printf("%s", name);
  for (const T& x : a) std::printf(" %.17g", static_cast<double>(x));
  std::printf("\n");
}

int main() {
  MaterialID gran = createMaterial("granular", 2500, 0, 0.2, 0.2, 0.5, 1, 1, 0, 0);
  MaterialID wood = createMaterial("wood", 700, 0.1, 0.35, 0.3, 0.5, 1, 1, 0, 0);
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
  // World::Iterator walks every body in body order; a cast iterator only the bodies of its type
  std::vector<double> all, boxes, spheres, capsules, planes, unions, union_sizes;
  for (World::Iterator it = theWorld()->begin(); it != theWorld()->end(); ++it) all.push_back((double)it->getID());
  for (World::Bodies::CastIterator<Box> it = theWorld()->begin<Box>(); it != theWorld()->end<Box>(); ++it)
    boxes.push_back((double)(*it)->getID());
  for (auto it = theWorld()->begin<Sphere>(); it != theWorld()->end<Sphere>(); it++) spheres.push_back((double)it->getID());
  for (auto it = theWorld()->begin<Capsule>(); it != theWorld()->end<Capsule>(); ++it) capsules.push_back((double)it->getID());
  for (auto it = theWorld()->begin<Plane>(); it != theWorld()->end<Plane>(); ++it) planes.push_back((double)it->getID());
  for (auto it = theWorld()->begin<Union>(); it != theWorld()->end<Union>(); ++it) {
    unions.push_back((double)it->getID());
    union_sizes.push_back((double)it->size());
  }
  put("iter_all", all);
  put("iter_box", boxes);
  put("iter_sphere", spheres);
  put("iter_capsule", capsules);
  put("iter_plane", planes);
  put("iter_union", unions);
  put("iter_union_size", union_sizes);
  const detail::Engine& e = detail::E();
  put("shape", e.shape);
  put("sid", e.sid);
  put("mat", e.mat);
  put("fixed", e.fixed);
  put("param", e.param);
  put("pos", e.pos);
  put("q", e.q);
  put("inv_mass", e.inv_mass);
  put("I", e.I);
  put("Iinv", e.Iinv);
  put("normal", e.normal);
  put("d", e.d);
  put("m_owner", e.m_owner);
  put("m_sid", e.m_sid);
  put("m_mat", e.m_mat);
  put("m_rad", e.m_rad);
  put("m_rpos", e.m_rpos);
  put("m_gpos", e.m_gpos);
  // a union moved and then turned (Union::setPosition / setOrientation): its members follow
  u->setPosition(Vec3(2.5, 1.75, 3.25));
  put("moved_m_gpos", e.m_gpos);
  u->setOrientation(Quat(Vec3(0.3, -0.5, 0.8), 0.7));
  put("turned_q", e.q);
  put("turned_m_gpos", e.m_gpos);
  put("turned_pos", e.pos);
  // destroy: a sphere from the middle (the last body, a sphere, takes its place), then the turned
  // union (the other union, now last, takes its place with its member)
  destroy(s1);
  destroy(u);
  put("after_shape", e.shape);
  put("after_sid", e.sid);
  put("after_mat", e.mat);
  put("after_fixed", e.fixed);
  put("after_param", e.param);
  put("after_pos", e.pos);
  put("after_q", e.q);
  put("after_inv_mass", e.inv_mass);
  put("after_I", e.I);
  put("after_Iinv", e.Iinv);
  put("after_normal", e.normal);
  put("after_d", e.d);
  put("after_m_owner", e.m_owner);
  put("after_m_sid", e.m_sid);
  put("after_m_mat", e.m_mat);
  put("after_m_rad", e.m_rad);
  put("after_m_rpos", e.m_rpos);
  put("after_m_gpos", e.m_gpos);
  std::vector<double> idx, uids;
  for (size_t k = 0; k < theWorld()->size(); ++k) {
    idx.push_back((double)(*theWorld())[k]->index());
    uids.push_back((double)(*theWorld())[k]->getID());
  }
  put("after_handle_index", idx);
  put("after_handle_uid", uids);
  return 0;
}
