#include "geometry.hpp"
#include <algorithm>
#include <cstdlib>
#include <stdexcept>

namespace dmhm {

Geometry::Geometry(int dim_, const int* sizes_, int numLevels_, bool strong_)
    : dim(dim_), numLevels(numLevels_), strong(strong_)
{
    if (dim != 2 && dim != 3) throw std::invalid_argument("dim must be 2 or 3");
    if (numLevels < 1) throw std::invalid_argument("numLevels must be >= 1");
    n = 1;
    for (int a = 0; a < 3; ++a) {
        sizes[a] = (a < dim) ? sizes_[a] : 1;
        if (sizes[a] < 1) throw std::invalid_argument("grid sizes must be positive");
        n *= sizes[a];
    }
    // Boxes, level by level in hierarchical order.
    boxes.assign(numLevels, {});
    Box root;
    root.off = 0; root.size = n;
    for (int a = 0; a < 3; ++a) root.ext[a] = sizes[a];
    boxes[0].push_back(root);
    for (int l = 0; l + 1 < numLevels; ++l) {
        for (int b = 0; b < (int)boxes[l].size(); ++b) {
            int off = boxes[l][b].off;
            for (int c = 0; c < nc(); ++c) {
                Box ch;
                ch.parent = b;
                int size = 1;
                for (int a = 0; a < 3; ++a) {
                    if (a >= dim) { ch.ext[a] = 1; ch.coord[a] = 0; continue; }
                    const int e = boxes[l][b].ext[a], half = e / 2;
                    const bool hi = (c >> a) & 1;
                    ch.ext[a] = hi ? e - half : half;
                    ch.coord[a] = 2 * boxes[l][b].coord[a] + (hi ? 1 : 0);
                    size *= ch.ext[a];
                }
                ch.off = off; ch.size = size;
                off += size;
                boxes[l][b].child[c] = (int)boxes[l + 1].size();
                boxes[l + 1].push_back(ch);
            }
        }
    }
    // Atoms from leaf boxes.
    auto& leafBoxes = boxes[numLevels - 1];
    for (auto& b : leafBoxes) {
        b.atomBegin = (int)atomStart.size();
        for (int r = 0; r < b.size; r += kAtom) {
            atomStart.push_back(b.off + r);
            atomSize.push_back(std::min(kAtom, b.size - r));
        }
        b.atomEnd = (int)atomStart.size();
    }
    for (int l = numLevels - 2; l >= 0; --l)
        for (auto& b : boxes[l]) {
            b.atomBegin = boxes[l + 1][b.child[0]].atomBegin;
            b.atomEnd = boxes[l + 1][b.child[nc() - 1]].atomEnd;
        }
    build(0, 0, 0);
    for (auto& blk : blocks) blk.transposeBlock = find(blk.level, blk.s, blk.t);
}

bool Geometry::admissible(int level, int t, int s) const
{
    const Box& bt = boxes[level][t];
    const Box& bs = boxes[level][s];
    if (strong) {
        int d = 0;
        for (int a = 0; a < dim; ++a) d = std::max(d, std::abs(bt.coord[a] - bs.coord[a]));
        return d > 1;
    }
    for (int a = 0; a < dim; ++a)
        if (bt.coord[a] != bs.coord[a]) return true;
    return false;
}

int Geometry::build(int level, int t, int s)
{
    const int id = (int)blocks.size();
    blocks.emplace_back();
    blocks[id].level = level; blocks[id].t = t; blocks[id].s = s;
    index_[key(level, t, s)] = id;
    blocks[id].leafBegin = (int)leaves.size();
    if (admissible(level, t, s)) {
        blocks[id].type = BT_LOWRANK;
        blocks[id].leafIndex = (int)leaves.size();
        blocks[id].lrIndex = (int)lrBlocks.size();
        leaves.push_back(id);
        lrBlocks.push_back(id);
    } else if (level < numLevels - 1) {
        blocks[id].type = BT_NODE;
        const int cb = (int)children.size();
        blocks[id].childBegin = cb;
        children.resize(cb + nc() * nc(), -1);
        for (int tc = 0; tc < nc(); ++tc)
            for (int sc = 0; sc < nc(); ++sc) {
                const int ct = boxes[level][t].child[tc], cs = boxes[level][s].child[sc];
                const int c = build(level + 1, ct, cs);
                children[cb + sc + nc() * tc] = c;
            }
    } else {
        blocks[id].type = BT_DENSE;
        blocks[id].leafIndex = (int)leaves.size();
        blocks[id].denseIndex = (int)denseBlocks.size();
        leaves.push_back(id);
        denseBlocks.push_back(id);
    }
    blocks[id].leafEnd = (int)leaves.size();
    blocks[id].blockEnd = (int)blocks.size();
    return id;
}

int Geometry::find(int level, int t, int s) const
{
    auto it = index_.find(key(level, t, s));
    return it == index_.end() ? -1 : it->second;
}

std::vector<int> Geometry::naturalToHierarchical() const
{
    std::vector<int> map(n);
    const auto& leafBoxes = boxes[numLevels - 1];
    for (const auto& b : leafBoxes) {
        // Corner of the box in grid coordinates, from its per-level coordinates:
        // reconstruct by walking the parents.
        int lo[3] = {0, 0, 0};
        {
            // Walk up: at each level the box's low corner is the parent's low
            // corner plus the parent's low-half extent when the axis bit is 1.
            std::vector<int> chain;
            int l = numLevels - 1, id = (int)(&b - &leafBoxes[0]);
            while (l > 0) { chain.push_back(id); id = boxes[l][id].parent; --l; }
            int cur = 0; l = 0;
            for (int i = (int)chain.size() - 1; i >= 0; --i) {
                const Box& p = boxes[l][cur];
                const Box& c = boxes[l + 1][chain[i]];
                for (int a = 0; a < dim; ++a)
                    if (c.coord[a] & 1) lo[a] += p.ext[a] / 2;
                cur = chain[i]; ++l;
            }
        }
        int idx = b.off;
        for (int z = 0; z < b.ext[2]; ++z)
            for (int y = 0; y < b.ext[1]; ++y)
                for (int x = 0; x < b.ext[0]; ++x) {
                    const int nat = (lo[0] + x) + sizes[0] * ((lo[1] + y) + sizes[1] * (lo[2] + z));
                    map[nat] = idx++;
                }
    }
    return map;
}

void Geometry::counts(std::vector<long>& dense, std::vector<long>& lowrank) const
{
    dense.assign(numLevels, 0); lowrank.assign(numLevels, 0);
    for (const auto& b : blocks) {
        if (b.type == BT_DENSE) dense[b.level]++;
        else if (b.type == BT_LOWRANK) lowrank[b.level]++;
    }
}

Layout Layout::make(const Geometry& g, const std::vector<int>& caps, int id)
{
    auto align = [](int64_t x) { return (x + 31) & ~int64_t(31); };
    Layout L;
    L.id = id;
    L.cap = caps;
    L.denseOff.resize(g.denseBlocks.size());
    int64_t o = 0;
    for (size_t i = 0; i < g.denseBlocks.size(); ++i) {
        const int b = g.denseBlocks[i];
        L.denseOff[i] = o;
        o = align(o + (int64_t)g.tbox(b).size * g.sbox(b).size);
    }
    L.denseElems = o;
    L.uOff.resize(g.lrBlocks.size());
    L.vOff.resize(g.lrBlocks.size());
    int64_t u = o;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        L.uOff[i] = u;
        u = align(u + (int64_t)g.tbox(b).size * caps[i]);
    }
    L.uElems = u - o;
    int64_t v = u;
    for (size_t i = 0; i < g.lrBlocks.size(); ++i) {
        const int b = g.lrBlocks[i];
        L.vOff[i] = v;
        v = align(v + (int64_t)g.sbox(b).size * caps[i]);
    }
    L.vElems = v - u;
    return L;
}

} // namespace dmhm
