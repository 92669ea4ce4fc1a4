// Tree invariants of the host SAH builder (ramp_amd/csrc/bvh_sah.h), run by
// tests/test_bvh_sah.py: every triangle in exactly one leaf of 1 to
// SAH_LEAF_TRIS triangles, 2 leaves - 1 nodes, parents and subtree sizes
// consistent, every child box inside its parent's, every leaf box holding
// its triangles' vertices, depth under the 96 levels the preorder pass walks.
// Usage: bvh_sah_check <float32 points file, 9 per triangle> <triangles>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "../../ramp_amd/csrc/bvh_types.h"
#include "../../ramp_amd/csrc/bvh_sah.h"

int main(int argc, char **argv)
{
    if (argc != 3) return 2;
    const int n = std::atoi(argv[2]);
    std::vector<float> P(9 * (size_t)n);
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(P.data(), sizeof(float) * 9, n, f) != (size_t)n) return 2;
    std::fclose(f);
    SahBuilder b;
    b.P = P.data();
    b.n = n;
    b.margin = 1e-6f;
    b.init();
    int nn = 0;
    b.build(0, n, -1, nn, 0);
    std::vector<int> seen(n, 0);
    int bad = 0, leaves = 0, maxd = 0;
    for (int x = 0; x < nn; x++) {
        const BvhNode &d = b.nodes[x];
        if (d.left < 0) {
            leaves++;
            if (d.count != 1 || d.ntri < 1 || d.ntri > SAH_LEAF_TRIS || d.right < 0 || d.right + d.ntri > n) {
                bad++;
                continue;
            }
            for (int k = 0; k < d.ntri; k++) {
                const uint32_t t = b.order[d.right + k];
                seen[t]++;
                for (int v = 0; v < 3; v++)
                    for (int a = 0; a < 3; a++) {
                        const float c = P[9 * (size_t)t + 3 * v + a];
                        if (c < d.lo[a] || c > d.hi[a]) bad++;
                    }
            }
        } else {
            if (d.left >= nn || d.right >= nn || b.parent[d.left] != x || b.parent[d.right] != x) { bad++; continue; }
            if (d.count != 1 + b.nodes[d.left].count + b.nodes[d.right].count) bad++;
            for (int c : {d.left, d.right})
                for (int a = 0; a < 3; a++)
                    if (b.nodes[c].lo[a] < d.lo[a] || b.nodes[c].hi[a] > d.hi[a]) bad++;
        }
        int dep = 0;
        for (int y = x; b.parent[y] >= 0 && dep < 200; y = b.parent[y]) dep++;
        if (dep > maxd) maxd = dep;
    }
    int miss = 0;
    for (int t = 0; t < n; t++) miss += seen[t] != 1;
    const bool ok = bad == 0 && miss == 0 && nn == 2 * leaves - 1 && b.nodes[0].count == nn && b.parent[0] == -1 &&
                    maxd < 96;
    std::printf("nodes %d leaves %d depth %d bad %d miss %d %s\n", nn, leaves, maxd, bad, miss, ok ? "ok" : "FAIL");
    return ok ? 0 : 1;
}
