// orc_pending.cpp -- oracle (test infrastructure): entry points of subsystems whose restatement
// is not written yet. Each one fails loudly instead of guessing: scenes that need them are
// rejected by orc_create, so no parity claim rests on these bodies.
#include <cstdio>
#include <cstdlib>
#include "orc_world.h"

namespace orc {

[[noreturn]] static void Missing(const char* what) {
    std::fprintf(stderr, "oracle: %s is not restated yet\n", what);
    std::abort();
}

void MeshContactFindPairs(World&, Contact*) { Missing("mesh contacts"); }
void MeshContactSynchronize(World&, Contact*) { Missing("mesh contacts"); }
void MeshContactCollide(World&, Contact*) { Missing("mesh contacts"); }

}  // namespace orc

#include "../../include/bxstep.h"
namespace orc {
bool BuildMesh(Mesh&, const bx_mesh_desc&) { Missing("meshes"); }
}  // namespace orc
