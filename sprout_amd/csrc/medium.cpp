// Host construction of the majorant gridtree of a density grid: sprout's
// Octree_builder for a one-channel texture
// (core/scene/material/volumetric/volumetric_octree_builder.cpp:13-296).
// Top grid of 32^3-voxel cells; each cell splits into octants down to 8^3
// voxels while the density range (one voxel of border included, for the
// linear filter) is 0.1 or more; leaves keep {minorant, majorant} of mu_a and
// mu_s.  The uniform variant splits every cell down to its 8^3-voxel leaves:
// a flat grid of majorants (32^3 of them for 256^3 voxels) in the same node
// format, so the tracking walks it unchanged.  Cells are built one after the other (the reference hands them to its
// thread pool; serialization order does not depend on it).

#include "medium.hpp"

#include <algorithm>
#include <memory>

namespace mi {

namespace {

struct Build_node {
    std::unique_ptr<Build_node[]> children;
    Medium_cm                     data{0.f, 0.f, 0.f, 0.f};
};

bool is_empty(Medium_cm const& c) {
    return 0.f == c.majorant_mu_a && 0.f == c.majorant_mu_s;
}

struct Splitter {
    float const* density;
    int32_t      d[3];
    bool         uniform = false;  // split every cell to its 8^3-voxel leaves
    Medium_cm    cm;  // CM(cc): min / max components of the unit coefficients
    uint32_t     num_nodes = 0;
    uint32_t     num_data  = 0;

    void split(Build_node& node, int32_t const lo[3], int32_t const hi[3], uint32_t depth) {
        int32_t constexpr w = (dev::Cell_dim >> (dev::Log2_cell_dim - 3)) + 1;

        int32_t minb[3], maxb[3];
        for (int k = 0; k < 3; ++k) {
            minb[k] = std::max(lo[k] - 1, 0);
            maxb[k] = std::min(hi[k] + 1, d[k]);
        }
        float min_density = 1.f;
        float max_density = 0.f;
        for (int32_t z = minb[2]; z < maxb[2]; ++z) {
            for (int32_t y = minb[1]; y < maxb[1]; ++y) {
                for (int32_t x = minb[0]; x < maxb[0]; ++x) {
                    float const v = density[(size_t(z) * d[1] + size_t(y)) * d[0] + size_t(x)];
                    min_density   = std::min(v, min_density);
                    max_density   = std::max(v, max_density);
                }
            }
        }
        if (min_density > max_density) {
            min_density = 0.f;
            max_density = 0.f;
        }
        float const diff = max_density - min_density;
        if (dev::Log2_cell_dim - 3 == int32_t(depth) || (!uniform && diff < 0.1f) || maxb[0] - minb[0] < w ||
            maxb[1] - minb[1] < w || maxb[2] - minb[2] < w) {
            node.children.reset();
            node.data = {min_density * cm.minorant_mu_a, min_density * cm.minorant_mu_s,
                         max_density * cm.majorant_mu_a, max_density * cm.majorant_mu_s};
            if (!is_empty(node.data)) {
                ++num_data;
            }
            return;
        }
        ++depth;
        int32_t center[3];
        for (int k = 0; k < 3; ++k) {
            center[k] = lo[k] + ((hi[k] - lo[k]) >> 1);
        }
        node.children.reset(new Build_node[8]);
        for (int i = 0; i < 8; ++i) {
            int32_t slo[3], shi[3];
            for (int k = 0; k < 3; ++k) {
                bool const upper = (i >> k) & 1;
                slo[k]           = upper ? center[k] : lo[k];
                shi[k]           = upper ? hi[k] : center[k];
            }
            split(node.children[i], slo, shi, depth);
        }
        num_nodes += 8;
    }
};

void serialize(Build_node const& node, uint32_t current, uint32_t& next, uint32_t& data,
               HostGridtree& out) {
    if (node.children) {
        out.nodes[current] = (next << 1) | 1u;
        uint32_t const first = next;
        next += 8;
        for (uint32_t i = 0; i < 8; ++i) {
            serialize(node.children[i], first + i, next, data, out);
        }
    } else if (!is_empty(node.data)) {
        out.nodes[current] = data << 1;
        out.data[data++]   = node.data;
    } else {
        out.nodes[current] = dev::Grid_empty << 1;
    }
}

}  // namespace

void build_gridtree(float const* density, int32_t const dims[3], float const cc_a[3],
                    float const cc_s[3], HostGridtree& out, bool uniform) {
    out = HostGridtree();
    int32_t nc[3];
    for (int k = 0; k < 3; ++k) {
        nc[k] = dims[k] >> dev::Log2_cell_dim;
        nc[k] += std::min(dims[k] - (nc[k] << dev::Log2_cell_dim), 1);
        out.num_cells[k] = nc[k];
    }
    uint32_t const cells = uint32_t(nc[0]) * uint32_t(nc[1]) * uint32_t(nc[2]);
    std::unique_ptr<Build_node[]> grid(new Build_node[cells]);

    Splitter sp;
    sp.density = density;
    sp.uniform = uniform;
    for (int k = 0; k < 3; ++k) {
        sp.d[k] = dims[k];
    }
    sp.cm = {std::min(std::min(cc_a[0], cc_a[1]), cc_a[2]), std::min(std::min(cc_s[0], cc_s[1]), cc_s[2]),
             std::max(std::max(cc_a[0], cc_a[1]), cc_a[2]), std::max(std::max(cc_s[0], cc_s[1]), cc_s[2])};

    int32_t const area = nc[0] * nc[1];
    for (uint32_t i = 0; i < cells; ++i) {
        int32_t c[3];
        c[2]           = int32_t(i) / area;
        int32_t const t = c[2] * area;
        c[1]           = (int32_t(i) - t) / nc[0];
        c[0]           = int32_t(i) - (t + c[1] * nc[0]);
        int32_t lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = c[k] << dev::Log2_cell_dim;
            hi[k] = lo[k] + dev::Cell_dim;
        }
        sp.split(grid[i], lo, hi, 0);
    }

    out.nodes.assign(cells + sp.num_nodes, 0u);
    out.data.resize(sp.num_data);
    uint32_t next = cells;
    uint32_t data = 0;
    for (uint32_t i = 0; i < cells; ++i) {
        serialize(grid[i], i, next, data, out);
    }
}

}  // namespace mi
