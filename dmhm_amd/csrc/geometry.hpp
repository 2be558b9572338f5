// Block-tree geometry of a square H-matrix on a uniform quadtree (2D) or
// octree (3D), flattened into per-level box arrays and one block list.
//
// Follows the reference's recursion (src/HMat2d/HMat2d.cpp:961-1033,
// include/dmhm/hmat2d.hpp:446-464, src/HMat3d/HMat3d.cpp:918-956):
//  * a box of per-axis extent e splits into e/2 (axis bit 0) and e - e/2
//    (axis bit 1); child c has axis bits (c&1, (c>>1)&1, (c>>2)&1) and its
//    unknowns follow those of children c' < c (hierarchical ordering);
//  * block (t,s) at level l is admissible when max_axis |t_a - s_a| > 1
//    (strong, HMat2d.cpp:800-803) or when t != s (weak);
//  * a non-admissible block is a NODE above the leaf level (numLevels-1) and
//    DENSE on it; the children of NODE (t,s) are (t',s') for t' in children(t),
//    s' in children(s), stored at index s' + nc*t' (hmat2d.hpp:489).
// The tree is identical for every H-matrix of the same geometry, so A, X = A^-1
// iterates and all products share one Geometry object.
#pragma once
#include <cstdint>
#include <vector>
#include <unordered_map>

namespace dmhm {

struct Box {
    int off = 0, size = 0;        // hierarchical index range [off, off+size)
    int ext[3] = {1, 1, 1};       // per-axis extent (number of grid points)
    int coord[3] = {0, 0, 0};     // box coordinate at its level
    int parent = -1;
    int child[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    int atomBegin = 0, atomEnd = 0;   // row atoms covered by the box
};

enum BlockType : int { BT_LOWRANK = 0, BT_DENSE = 1, BT_NODE = 2 };

struct Block {
    int level = 0, t = 0, s = 0;  // box ids at `level`
    int type = BT_NODE;
    int childBegin = -1;          // index into Geometry::children (nc*nc entries)
    int leafBegin = 0, leafEnd = 0; // DFS leaf range (leaves() order)
    int blockEnd = 0;             // the subtree is blocks [this block, blockEnd) (DFS pre-order)
    int leafIndex = -1;           // position in leaves() for DENSE / LOW_RANK
    int denseIndex = -1;          // index among dense leaves
    int lrIndex = -1;             // index among low-rank blocks
    int transposeBlock = -1;      // block (level, s, t)
};

class Geometry {
public:
    Geometry(int dim, const int* sizes, int numLevels, bool strong);

    int dim, numLevels, n;        // n = number of unknowns
    int sizes[3];
    bool strong;
    int nc() const { return 1 << dim; }

    std::vector<std::vector<Box>> boxes;    // boxes[level][box id]
    // Row atoms: every leaf box split into chunks of at most kAtom rows. Every
    // block's row and column ranges are unions of atoms.
    static constexpr int kAtom = 64;
    std::vector<int> atomStart, atomSize;

    std::vector<Block> blocks;              // all blocks, DFS pre-order
    std::vector<int> children;              // child block ids of NODE blocks
    std::vector<int> leaves;                // DENSE/LOW_RANK block ids, DFS order
    std::vector<int> denseBlocks, lrBlocks; // block ids
    int root() const { return 0; }

    int find(int level, int t, int s) const;    // -1 if the block does not exist
    int child(int b, int tc, int sc) const { return children[blocks[b].childBegin + sc + nc() * tc]; }
    bool admissible(int level, int t, int s) const;
    bool near(int level, int t, int s) const { return !admissible(level, t, s); }
    const Box& tbox(int b) const { return boxes[blocks[b].level][blocks[b].t]; }
    const Box& sbox(int b) const { return boxes[blocks[b].level][blocks[b].s]; }

    // Natural index x + nx*y (+ nx*ny*z) -> hierarchical index
    // (HMat2d::BuildNaturalToHierarchicalMap, HMat2d.cpp:27-87).
    std::vector<int> naturalToHierarchical() const;

    // Per-level counts for reports.
    void counts(std::vector<long>& dense, std::vector<long>& lowrank) const;

private:
    std::unordered_map<uint64_t, int> index_;
    static uint64_t key(int level, int t, int s) {
        return ((uint64_t)level << 58) | ((uint64_t)(uint32_t)t << 29) | (uint64_t)(uint32_t)s;
    }
    int build(int level, int t, int s);
};

// Device storage layout of one H-matrix over a Geometry: per dense leaf an
// m x n column-major block, per low-rank block U (m x cap) and V (n x cap)
// column-major and one int rank. All pools live in one allocation:
//   [dense values | U values | V values] (elements), ranks separately.
struct Layout {
    int id = 0;
    std::vector<int64_t> denseOff;       // per dense leaf
    std::vector<int64_t> uOff, vOff;     // per low-rank block
    std::vector<int> cap;                // per low-rank block
    int64_t denseElems = 0, uElems = 0, vElems = 0;
    int64_t totalElems() const { return denseElems + uElems + vElems; }
    int nLowRank() const { return (int)cap.size(); }
    static Layout make(const Geometry& g, const std::vector<int>& caps, int id);
};

} // namespace dmhm
