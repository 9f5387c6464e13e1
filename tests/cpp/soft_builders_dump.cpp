// Writes the facade's cloth and hair builders' output (include/pxg_soft.hpp) for the cases
// tests/test_soft_builders.py compares with the oracle's builders: raw records, one case after
// another, to the file named by argv[1].
#include "pxg_soft.hpp"

#include <cstdio>
#include <vector>

namespace {

FILE* g_out = nullptr;
template <class T>
void put(const std::vector<T>& v) {
    const uint32_t n = static_cast<uint32_t>(v.size());
    std::fwrite(&n, 4, 1, g_out);
    if (n) std::fwrite(v.data(), sizeof(T), n, g_out);
}
void put_cloth(const pxg::Cloth& c) {
    put(c.particles);
    put(c.stretch);
    put(c.bend);
    put(c.stretch_batches);
    put(c.bend_batches);
    put(c.tethers);
    put(c.anchors);
    put(c.triangles);
    std::fwrite(&c.params, sizeof(c.params), 1, g_out);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    g_out = std::fopen(argv[1], "wb");
    if (!g_out) return 2;
    {  // a 9 x 7 sheet, turned, shear on, pinned at two corners (one static, one to a body), tethered
        pxg::GridClothDesc d;
        d.columns = 9;
        d.rows = 7;
        d.width = 2.0f;
        d.height = 1.5f;
        d.center[0] = 0.3f; d.center[1] = -0.2f; d.center[2] = 2.0f;
        d.orientation[0] = 0.0f; d.orientation[1] = 0.38268343f; d.orientation[2] = 0.0f; d.orientation[3] = 0.92387953f;
        d.total_mass = 1.7f;
        pxg::Cloth c = pxg::make_grid_cloth(d);
        const float p0[3] = {c.particles[0].position[0], c.particles[0].position[1], c.particles[0].position[2]};
        pxg::pin_static(c, p0, 0.3f);
        pxg_body pose;
        std::memset(&pose, 0, sizeof(pose));
        pose.position[0] = 1.0f; pose.position[1] = 0.5f; pose.position[2] = 2.5f;
        pose.orientation[0] = 0.1f; pose.orientation[1] = 0.2f; pose.orientation[2] = -0.3f; pose.orientation[3] = 0.9273618f;
        const float* last = c.particles.back().position;
        const float p1[3] = {last[0], last[1], last[2]};
        pxg::pin_to_body(c, 3u, 7u, pose, p1, 0.4f);
        pxg::build_tethers(c);
        put_cloth(c);
    }
    {  // a 21 x 21 sheet without shear, default pose
        pxg::GridClothDesc d;
        d.shear = false;
        put_cloth(pxg::make_grid_cloth(d));
    }
    {  // two strands of 5 and 9 points, turned roots, one of a single point (skipped)
        std::vector<pxg::StrandDesc> s(3);
        for (int k = 0; k < 5; ++k) s[0].points.insert(s[0].points.end(), {0.1f * k, 0.02f * k * k, 0.5f});
        s[0].root_orientation[2] = 0.38268343f;
        s[0].root_orientation[3] = 0.92387953f;
        s[1].points = {1.0f, 1.0f, 1.0f};
        for (int k = 0; k < 9; ++k) s[2].points.insert(s[2].points.end(), {-0.3f, 0.05f * k, 1.0f - 0.03f * k});
        s[2].particle_mass = 0.002f;
        pxg_strand_params p = pxg::default_strand_params();
        p.thickness = 0.02f;
        pxg::StrandSet set = pxg::make_strands(s, p);
        put(set.particles);
        put(set.stretch);
        put(set.stretch_batches);
        put(set.rest_local);
        put(set.rest_offset);
        put(set.strands);
    }
    std::fclose(g_out);
    return 0;
}
