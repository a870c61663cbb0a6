// Oracle: sphere and capsule narrow phase (placeholder until restated).
// TEST INFRASTRUCTURE ONLY (see ref_math.h).
#include <cstdio>
#include <cstdlib>

#include "ref_world.h"

namespace nslref {

static void unsupported(const char* what) {
  std::fprintf(stderr, "oracle: %s is not restated yet\n", what);
  std::abort();
}
void collide_spheres(LocalManifold&, Sphere, Sphere, Xf) { unsupported("collide_spheres"); }
void collide_capsule_and_sphere(LocalManifold&, Capsule, Sphere, Xf) { unsupported("collide_capsule_and_sphere"); }
void collide_capsules(LocalManifold&, Capsule, Capsule, Xf) { unsupported("collide_capsules"); }
void collide_hull_and_sphere(LocalManifold&, const Hull&, Sphere, Xf, SimplexCache&) { unsupported("collide_hull_and_sphere"); }
void collide_hull_and_capsule(LocalManifold&, const Hull&, Capsule, Xf, SimplexCache&) { unsupported("collide_hull_and_capsule"); }

}  // namespace nslref
