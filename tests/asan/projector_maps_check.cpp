// CPU check of the host-side projector maps (csrc/projector_maps.h), built
// with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_projector_maps_asan.py.  Random projector sets on small periodic
// grids: spheres that wrap across the cell, overlapping atoms, atoms without
// voxels, up to 32 projectors, and points covered by more atoms than a
// scatter block stages.  Every map is checked against its definition.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <set>
#include <vector>

#include "projector_maps.h"

using namespace qm;

static int fails = 0;
#define CHECK(c)                                                                   \
    do {                                                                           \
        if (!(c)) {                                                                \
            if (fails++ < 20) std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
        }                                                                          \
    } while (0)

static bool is_perm(const std::vector<int> &p, int n) {
    if ((int)p.size() != n) return false;
    std::vector<char> seen(n, 0);
    for (int x : p) {
        if (x < 0 || x >= n || seen[x]) return false;
        seen[x] = 1;
    }
    return true;
}

static void run_case(std::mt19937 &rng, int nx, int ny, int nz, int natoms, int maxm, int pile) {
    std::uniform_int_distribution<int> ux(0, nx - 1), uy(0, ny - 1), uz(0, nz - 1), um(1, maxm), ur(0, 3);
    std::uniform_real_distribution<double> ut(-1.0, 1.0);
    std::vector<int> nproj(natoms), proj_off(natoms + 1, 0), xyz;
    std::vector<int64_t> vox_off(natoms + 1, 0), tab_off(natoms + 1, 0);
    for (int a = 0; a < natoms; ++a) {
        nproj[a] = um(rng);
        const int r = ur(rng);
        // `pile` atoms share one centre; the others are spread over the cell
        // (their spheres wrap across the faces)
        const int cx = a < pile ? 1 : ux(rng), cy = a < pile ? 0 : uy(rng), cz = a < pile ? nz - 1 : uz(rng);
        int64_t V = 0;
        if (a % 7 != 6)  // every seventh atom has no voxels
            for (int z = -r; z <= r; ++z)
                for (int y = -r; y <= r; ++y)
                    for (int x = -r; x <= r; ++x)
                        if (x * x + y * y + z * z <= r * r) {
                            xyz.push_back(cx + x);
                            xyz.push_back(cy + y);
                            xyz.push_back(cz + z);
                            ++V;
                        }
        vox_off[a + 1] = vox_off[a] + V;
        tab_off[a + 1] = tab_off[a] + V * nproj[a];
        proj_off[a + 1] = proj_off[a] + nproj[a];
    }
    std::vector<double> table(tab_off[natoms]);
    for (double &t : table) t = ut(rng);
    int max_m = 0;
    for (int a = 0; a < natoms; ++a) max_m = std::max(max_m, nproj[a]);
    const int mpad = max_m > 16 ? 32 : 16;
    if (xyz.empty()) xyz.push_back(0);  // keep the pointer valid when no atom has voxels
    ProjectorMaps m;
    build_projector_maps(nx, ny, nz, natoms, nproj.data(), vox_off.data(), xyz.data(), table.data(),
                         tab_off.data(), proj_off.data(), mpad, m);

    const int64_t vtot = vox_off[natoms];
    const int ng = nx * ny * nz;
    // voxels: wrapped grid index and owner
    CHECK((int64_t)m.grid.size() == vtot && (int64_t)m.atom.size() == vtot);
    for (int a = 0; a < natoms; ++a)
        for (int64_t v = vox_off[a]; v < vox_off[a + 1]; ++v) {
            const int gx = ((xyz[3 * v] % nx) + nx) % nx, gy = ((xyz[3 * v + 1] % ny) + ny) % ny,
                      gz = ((xyz[3 * v + 2] % nz) + nz) % nz;
            CHECK(m.grid[v] == gx + nx * (gy + ny * gz));
            CHECK(m.grid[v] >= 0 && m.grid[v] < ng);
            CHECK(m.atom[v] == a);
        }
    // covered points
    CHECK(is_perm(m.cov_vox, (int)vtot));
    const int ncov = (int)m.cov_grid.size();
    CHECK((int)m.cov_off.size() == ncov + 1);
    CHECK(m.cov_off.front() == 0 && m.cov_off.back() == (int)vtot);
    std::set<int> distinct(m.grid.begin(), m.grid.end());
    CHECK((int)distinct.size() == ncov);
    for (int i = 0; i < ncov; ++i) {
        CHECK(m.cov_off[i] < m.cov_off[i + 1]);
        if (i > 0) CHECK(m.cov_grid[i - 1] < m.cov_grid[i]);
        for (int e = m.cov_off[i]; e < m.cov_off[i + 1]; ++e) CHECK(m.grid[m.cov_vox[e]] == m.cov_grid[i]);
    }
    // atom order: heaviest first
    CHECK(is_perm(m.atom_order, natoms));
    for (int i = 1; i < natoms; ++i) {
        const int a = m.atom_order[i - 1], b = m.atom_order[i];
        CHECK((vox_off[a + 1] - vox_off[a]) * nproj[a] >= (vox_off[b + 1] - vox_off[b]) * nproj[b]);
    }
    // padded tables and projector owners
    CHECK(m.tabp64.size() == (size_t)vtot * mpad && m.tabp32.size() == (size_t)vtot * mpad);
    for (int a = 0; a < natoms; ++a) {
        const int64_t V = vox_off[a + 1] - vox_off[a];
        for (int64_t v = 0; v < V; ++v)
            for (int w = 0; w < mpad; ++w) {
                const double want = w < nproj[a] ? table[tab_off[a] + w * V + v] : 0.0;
                CHECK(m.tabp64[(size_t)(vox_off[a] + v) * mpad + w] == want);
                CHECK(m.tabp32[(size_t)(vox_off[a] + v) * mpad + w] == (float)want);
            }
        for (int w = 0; w < nproj[a]; ++w) CHECK(m.proj_atom[proj_off[a] + w] == a);
    }
    // projection work items
    const int nchunk = (int)m.chunk_atom.size();
    CHECK((int)m.chunk_v0.size() == nchunk && (int)m.atom_chunk_off.size() == natoms + 1);
    CHECK(is_perm(m.item_order, nchunk));
    for (int a = 0; a < natoms; ++a) {
        const int64_t V = vox_off[a + 1] - vox_off[a];
        CHECK(m.atom_chunk_off[a + 1] - m.atom_chunk_off[a] == (int)((V + kProjChunk - 1) / kProjChunk));
        for (int it = m.atom_chunk_off[a]; it < m.atom_chunk_off[a + 1]; ++it) {
            CHECK(m.chunk_atom[it] == a);
            CHECK(m.chunk_v0[it] == (it - m.atom_chunk_off[a]) * kProjChunk);
        }
    }
    // scatter blocks and slots
    const int nb = (ncov + kScatterBlock - 1) / kScatterBlock;
    CHECK((int)m.block_atom_off.size() == nb + 1 && m.block_atom_off[0] == 0);
    CHECK((int64_t)m.cov_slot.size() == vtot && !m.block_atoms.empty());
    for (int b = 0; b < nb; ++b) {
        const int o0 = m.block_atom_off[b], o1 = m.block_atom_off[b + 1];
        CHECK(o0 <= o1 && o1 - o0 <= kScatterSlots && o1 <= (int)m.block_atoms.size());
        std::set<int> staged(m.block_atoms.begin() + o0, m.block_atoms.begin() + o1);
        CHECK((int)staged.size() == o1 - o0);
        const int i0 = b * kScatterBlock, i1 = std::min(ncov, i0 + kScatterBlock);
        for (int e = m.cov_off[i0]; e < m.cov_off[i1]; ++e) {
            const int s = m.cov_slot[e], a = m.atom[m.cov_vox[e]];
            CHECK(s >= -1 && s < kScatterSlots);
            CHECK(s < o1 - o0);
            if (s >= 0 && s < o1 - o0) CHECK(m.block_atoms[o0 + s] == a);
            if (s >= 0) continue;
            CHECK(o1 - o0 == kScatterSlots && !staged.count(a));
        }
    }
}

int main() {
    std::mt19937 rng(20260917u);
    int cases = 0;
    for (int rep = 0; rep < 40; ++rep) {
        const int nx = 6 + rep % 7, ny = 5 + rep % 5, nz = 4 + rep % 6;
        const int natoms = 1 + rep % 14, maxm = rep % 3 == 0 ? 32 : 9;
        const int pile = rep % 4 == 0 ? std::min(natoms, 12) : 0;  // up to 12 atoms on one point
        run_case(rng, nx, ny, nz, natoms, maxm, pile);
        ++cases;
    }
    // a larger cell with more covered points than one scatter block
    run_case(rng, 24, 20, 18, 40, 16, 10);
    ++cases;
    if (fails) {
        std::fprintf(stderr, "%d checks failed\n", fails);
        return 1;
    }
    std::printf("projector maps: %d cases ok\n", cases);
    return 0;
}
