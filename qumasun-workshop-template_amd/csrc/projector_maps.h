// Host-side index maps of a projector set: the part of qm_set_projectors that
// needs no device, kept free of HIP so that it also builds with a plain C++
// compiler (the CPU tests build it with AddressSanitizer).
//
// Inputs follow qm_set_projectors (include/qumasun_hip.h): atom a owns voxels
// vox_off[a] .. vox_off[a+1]-1 with unwrapped coordinates vox_xyz[3 v ..],
// nproj[a] projectors whose table is table[tab_off[a] + w V_a + v] (V_a voxels,
// projector-major, as the reference's SvF tables, PseudoPotNonlocal_SvF2_kpoint.h
// :911-1178), and projector ids proj_off[a] .. proj_off[a+1]-1.
#pragma once

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

namespace qm {

constexpr int kProjChunk = 512;     // voxels per projection work item (4 waves x 128)
constexpr int kScatterBlock = 256;  // covered points per scatter workgroup
constexpr int kScatterSlots = 8;    // atoms whose coefficients a scatter workgroup stages in LDS

// periodic wrap of a grid coordinate into 0 .. n-1
inline int wrap_index(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

struct ProjectorMaps {
    std::vector<int> grid, atom;  // [vtot] wrapped grid index and owner atom of every voxel
    // owner-computes scatter: covered grid points in increasing order, each with
    // its voxels cov_vox[cov_off[i] .. cov_off[i+1]-1]
    std::vector<int> cov_grid, cov_off, cov_vox;
    std::vector<int> atom_order;  // heaviest atoms (V_a M_a) first
    // voxel-major tables padded to mpad projectors: tabp[v mpad + w]
    std::vector<double> tabp64;
    std::vector<float> tabp32;
    // projection work items (chunks of kProjChunk voxels of one atom), the
    // atoms' ranges of them, each projector's atom, the launch order
    std::vector<int> chunk_atom, chunk_v0, atom_chunk_off, proj_atom, item_order;
    // per scatter block of kScatterBlock covered points: its distinct atoms
    // (at most kScatterSlots staged) and each voxel entry's slot (-1: not staged)
    std::vector<int> block_atom_off, block_atoms, cov_slot;
};

inline void build_projector_maps(int nx, int ny, int nz, int natoms, const int *nproj, const int64_t *vox_off,
                                 const int *vox_xyz, const double *table, const int64_t *tab_off,
                                 const int *proj_off, int mpad, ProjectorMaps &m) {
    const int64_t vtot = vox_off[natoms];
    const int ptot = proj_off[natoms];
    m.grid.assign(vtot, 0);
    m.atom.assign(vtot, 0);
    for (int a = 0; a < natoms; ++a)
        for (int64_t v = vox_off[a]; v < vox_off[a + 1]; ++v) {
            m.grid[v] = wrap_index(vox_xyz[3 * v], nx) +
                        nx * (wrap_index(vox_xyz[3 * v + 1], ny) + ny * wrap_index(vox_xyz[3 * v + 2], nz));
            m.atom[v] = a;
        }
    // covered-point map: voxels sorted by grid point
    m.cov_vox.resize(vtot);
    std::iota(m.cov_vox.begin(), m.cov_vox.end(), 0);
    std::stable_sort(m.cov_vox.begin(), m.cov_vox.end(), [&](int x, int y) { return m.grid[x] < m.grid[y]; });
    m.cov_grid.clear();
    m.cov_off.clear();
    m.cov_off.reserve(vtot + 1);
    for (int64_t i = 0; i < vtot; ++i) {
        if (i == 0 || m.grid[m.cov_vox[i]] != m.grid[m.cov_vox[i - 1]]) {
            m.cov_grid.push_back(m.grid[m.cov_vox[i]]);
            m.cov_off.push_back((int)i);
        }
    }
    m.cov_off.push_back((int)vtot);
    const int ncov = (int)m.cov_grid.size();
    m.atom_order.resize(natoms);
    std::iota(m.atom_order.begin(), m.atom_order.end(), 0);
    std::stable_sort(m.atom_order.begin(), m.atom_order.end(), [&](int x, int y) {
        return (vox_off[x + 1] - vox_off[x]) * nproj[x] > (vox_off[y + 1] - vox_off[y]) * nproj[y];
    });
    // padded voxel-major tables and the projection work list
    m.tabp64.assign((size_t)vtot * mpad, 0.0);
    m.tabp32.assign((size_t)vtot * mpad, 0.0f);
    m.chunk_atom.clear();
    m.chunk_v0.clear();
    m.atom_chunk_off.assign(natoms + 1, 0);
    m.proj_atom.assign(ptot, 0);
    for (int a = 0; a < natoms; ++a) {
        const int64_t V = vox_off[a + 1] - vox_off[a];
        for (int w = 0; w < nproj[a]; ++w) {
            m.proj_atom[proj_off[a] + w] = a;
            for (int64_t v = 0; v < V; ++v) {
                const double val = table[tab_off[a] + w * V + v];
                m.tabp64[(size_t)(vox_off[a] + v) * mpad + w] = val;
                m.tabp32[(size_t)(vox_off[a] + v) * mpad + w] = (float)val;
            }
        }
        for (int64_t v0 = 0; v0 < V; v0 += kProjChunk) {
            m.chunk_atom.push_back(a);
            m.chunk_v0.push_back((int)v0);
        }
        m.atom_chunk_off[a + 1] = (int)m.chunk_atom.size();
    }
    // launch order of the work items: by the (wrapped) z, then y, of each
    // chunk's first voxel, so that the chunks of neighbouring atoms that read
    // the same grid lines run together on one XCD and share its L2
    const int nchunk = (int)m.chunk_atom.size();
    std::vector<int64_t> key(nchunk);
    for (int it = 0; it < nchunk; ++it) {
        const int64_t v = vox_off[m.chunk_atom[it]] + m.chunk_v0[it];
        key[it] = (int64_t)wrap_index(vox_xyz[3 * v + 2], nz) * ny + wrap_index(vox_xyz[3 * v + 1], ny);
    }
    m.item_order.resize(nchunk);
    std::iota(m.item_order.begin(), m.item_order.end(), 0);
    std::stable_sort(m.item_order.begin(), m.item_order.end(), [&](int a, int b) { return key[a] < key[b]; });
    // scatter blocks: distinct atoms of their entries, and every entry's slot
    const int nb = (ncov + kScatterBlock - 1) / kScatterBlock;
    m.block_atom_off.assign(nb + 1, 0);
    m.block_atoms.clear();
    m.cov_slot.assign(vtot, -1);
    for (int b = 0; b < nb; ++b) {
        std::vector<int> list;
        const int i0 = b * kScatterBlock, i1 = std::min(ncov, i0 + kScatterBlock);
        for (int i = i0; i < i1; ++i)
            for (int e = m.cov_off[i]; e < m.cov_off[i + 1]; ++e) {
                const int a = m.atom[m.cov_vox[e]];
                int s = -1;
                for (int q = 0; q < (int)list.size(); ++q)
                    if (list[q] == a) s = q;
                if (s < 0 && (int)list.size() < kScatterSlots) {
                    list.push_back(a);
                    s = (int)list.size() - 1;
                }
                m.cov_slot[e] = s;
            }
        m.block_atoms.insert(m.block_atoms.end(), list.begin(), list.end());
        m.block_atom_off[b + 1] = (int)m.block_atoms.size();
    }
    if (m.block_atoms.empty()) m.block_atoms.push_back(0);
}

}  // namespace qm
