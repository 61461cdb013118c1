// CPU build of the device library (test tool): steps a workload written by check.py --dump through the
// C ABI under MemorySanitizer, which reports any branch, index or address that depends on memory never
// written (a device compiler may resolve such a read to any value). Prints each world's status and a
// hash of all body states after the last step.
//
//   make -C tests/cpu_build msan && tests/cpu_build/_out/bxcpu_msan scene.bin steps
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bxstep.h"

namespace {

struct Reader {
    std::vector<char> buf;
    size_t at = 0;
    template <class T> T Get() {
        T v;
        Need(sizeof(T));
        std::memcpy(&v, buf.data() + at, sizeof(T));
        at += sizeof(T);
        return v;
    }
    template <class T> const T* Array(size_t n) {
        Need(n * sizeof(T));
        const T* p = reinterpret_cast<const T*>(buf.data() + at);
        at += n * sizeof(T);
        return p;
    }
    void Need(size_t n) {
        if (at + n > buf.size()) {
            std::fprintf(stderr, "msan_driver: truncated scene file\n");
            std::exit(2);
        }
    }
};

void Check(int rc, const char* what) {
    if (rc < 0) {
        std::fprintf(stderr, "msan_driver: %s: %d %s\n", what, rc, bx_last_error());
        std::exit(1);
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s scene.bin steps\n", argv[0]);
        return 2;
    }
    Reader r;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::fseek(f, 0, SEEK_END);
    r.buf.resize(size_t(std::ftell(f)));
    std::fseek(f, 0, SEEK_SET);
    if (std::fread(r.buf.data(), 1, r.buf.size(), f) != r.buf.size()) return 2;
    std::fclose(f);
    if (std::memcmp(r.Array<char>(4), "BXSC", 4) != 0 || r.Get<uint32_t>() != 1) return 2;
    std::vector<bx_hull_desc> hulls(r.Get<uint32_t>());
    for (bx_hull_desc& h : hulls) {
        h.vertex_count = r.Get<uint32_t>();
        h.edge_count = r.Get<uint32_t>();
        h.face_count = r.Get<uint32_t>();
        h.vertices = r.Array<float>(3 * h.vertex_count);
        h.edges = r.Array<uint32_t>(5 * h.edge_count);
        h.faces = r.Array<uint32_t>(h.face_count);
        h.planes = r.Array<float>(4 * h.face_count);
        std::memcpy(h.centroid, r.Array<float>(3), sizeof(h.centroid));
    }
    std::vector<bx_mesh_desc> meshes(r.Get<uint32_t>());
    for (bx_mesh_desc& m : meshes) {
        m.vertex_count = r.Get<uint32_t>();
        m.triangle_count = r.Get<uint32_t>();
        m.vertices = r.Array<float>(3 * m.vertex_count);
        m.triangles = r.Array<uint32_t>(3 * m.triangle_count);
    }
    std::vector<bx_world_desc> worlds(r.Get<uint32_t>());
    for (bx_world_desc& w : worlds) {
        std::memcpy(w.gravity, r.Array<float>(3), sizeof(w.gravity));
        w.flags = r.Get<uint32_t>();
        w.body_count = r.Get<uint32_t>();
        w.bodies = r.Array<bx_body_desc>(w.body_count);
        w.fixture_count = r.Get<uint32_t>();
        w.fixtures = r.Array<bx_fixture_desc>(w.fixture_count);
        w.joint_count = r.Get<uint32_t>();
        w.joints = r.Array<bx_joint_desc>(w.joint_count);
    }
    bx_shape_library lib{uint32_t(hulls.size()), hulls.data(), uint32_t(meshes.size()), meshes.data()};
    bx_context* ctx = nullptr;
    Check(bx_context_create(0, &ctx), "bx_context_create");
    bx_batch_opts opts{};
    opts.solver_mode = 1;
    bx_batch* batch = nullptr;
    Check(bx_batch_create(ctx, &lib, worlds.data(), uint32_t(worlds.size()), &opts, &batch), "bx_batch_create");
    bx_step_cfg cfg{1.0f / 60.0f, 8, 2};
    const int steps = std::atoi(argv[2]);
    for (int s = 0; s < steps; ++s) {
        Check(bx_batch_step(batch, &cfg, 1), "bx_batch_step");
        std::printf("[msan] step %d\n", s + 1);
        std::fflush(stdout);
    }
    uint64_t hash = 1469598103934665603ull;
    uint32_t bad = 0;
    for (uint32_t w = 0; w < worlds.size(); ++w) {
        bx_world_counts c;
        Check(bx_batch_get_counts(batch, w, &c), "bx_batch_get_counts");
        if (c.status) ++bad;
        std::vector<bx_body_state> b(c.body_count);
        Check(bx_batch_get_bodies(batch, w, b.data(), c.body_count), "bx_batch_get_bodies");
        const unsigned char* p = reinterpret_cast<const unsigned char*>(b.data());
        for (size_t i = 0; i < b.size() * sizeof(bx_body_state); ++i) hash = (hash ^ p[i]) * 1099511628211ull;
    }
    std::printf("[msan] %zu worlds, %u with nonzero status, body hash %016llx\n", worlds.size(), bad,
                (unsigned long long)hash);
    bx_batch_destroy(batch);
    bx_context_destroy(ctx);
    return bad ? 1 : 0;
}
