// gfx950 kernels of the CSR import (HMat2d/HMat3d::ImportSparse with
// ConvertSubmatrix, HMat2d.cpp:961-1033, ConvertSubmatrix.cpp:57-119): every
// nonzero walks the block tree from the root to its leaf block; dense leaves
// receive the value, admissible blocks an exact low-rank term (a unit column
// of U at the row, the value in V at the column), one term per nonzero in
// CSR order. Also the natural -> hierarchical index map.
#include "kcommon.hpp"

namespace dmhm {

// Child box of `box` (flattened index) whose index range holds position i:
// children are contiguous and ordered, so it is the last one starting at or
// before i.
__device__ __forceinline__ int child_box_of(const GeoDev& g, int box, int i)
{
    int c = 0;
    for (int k = 1; k < g.nc; ++k)
        if (g.boxOff[g.boxChild[box * g.nc + k]] <= i) c = k;
    return c;
}

__device__ __forceinline__ int locate_block(const GeoDev& g, int i, int j)
{
    int b = 0;
    while (g.blkType[b] == 2) {                       // NODE
        const int l = g.blkLevel[b];
        const int tb = g.lvlBase[l] + g.blkT[b], sb = g.lvlBase[l] + g.blkS[b];
        const int tc = child_box_of(g, tb, i), sc = child_box_of(g, sb, j);
        b = g.children[g.blkChildBegin[b] + sc + g.nc * tc];
    }
    return b;
}

// Pass 1: leaf block of every nonzero; per admissible block, its count.
__global__ void csr_locate_kernel(GeoDev g, const int* __restrict__ rowptr, const int* __restrict__ colind, int n,
                                  int* __restrict__ nzBlock, int* __restrict__ lrCount, int* __restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int j = colind[k];
        if (j < 0 || j >= n) { atomicAdd(bad, 1); nzBlock[k] = -1; continue; }
        const int b = locate_block(g, i, j);
        nzBlock[k] = b;
        if (g.blkType[b] == 0) atomicAdd(&lrCount[g.blkLr[b]], 1);
    }
}

// Pass 2: dense values in place; low-rank entries staged per block (slot
// order arbitrary, fixed by the sort of pass 3).
template <typename T>
__global__ void csr_scatter_kernel(GeoDev g, LayDev lay, const int* __restrict__ rowptr,
                                   const int* __restrict__ colind, const double* __restrict__ vals, int n,
                                   const int* __restrict__ nzBlock, const int* __restrict__ lrBase,
                                   int* __restrict__ lrFill, int* __restrict__ stageK, T* __restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const int b = nzBlock[k];
        if (b < 0) continue;
        const int l = g.blkLevel[b];
        const int tb = g.lvlBase[l] + g.blkT[b];
        if (g.blkType[b] == 1) {                     // DENSE
            const int sb = g.lvlBase[l] + g.blkS[b];
            const int r = i - g.boxOff[tb], c = colind[k] - g.boxOff[sb];
            const int m = g.boxSize[tb];
            dst[lay.denseOff[g.blkDense[b]] + r + (int64_t)c * m] = (T)vals[k];
        } else {
            const int lr = g.blkLr[b];
            const int slot = atomicAdd(&lrFill[lr], 1);
            stageK[lrBase[lr] + slot] = k;
        }
    }
}

// Pass 3: one thread per admissible block sorts its nonzeros into CSR order
// (ascending nonzero index) and writes term p: U(row, p) = 1, V(col, p) = value.
template <typename T>
__global__ void csr_lowrank_kernel(GeoDev g, LayDev lay, const int* __restrict__ rowOf,
                                   const int* __restrict__ colind, const double* __restrict__ vals,
                                   const int* __restrict__ lrBase, const int* __restrict__ lrCount,
                                   int* __restrict__ stageK, int nLr, T* __restrict__ dst)
{
    const int lr = blockIdx.x * blockDim.x + threadIdx.x;
    if (lr >= nLr) return;
    const int cnt = lrCount[lr];
    if (cnt == 0) return;
    int* K = stageK + lrBase[lr];
    for (int a = 1; a < cnt; ++a) {                   // insertion sort (counts are small)
        const int x = K[a];
        int p = a - 1;
        while (p >= 0 && K[p] > x) { K[p + 1] = K[p]; --p; }
        K[p + 1] = x;
    }
    const int b = g.lrBlock[lr];
    const int l = g.blkLevel[b];
    const int tb = g.lvlBase[l] + g.blkT[b], sb = g.lvlBase[l] + g.blkS[b];
    const int m = g.boxSize[tb], nn = g.boxSize[sb];
    for (int p = 0; p < cnt; ++p) {
        const int k = K[p];
        const int r = rowOf[k] - g.boxOff[tb], c = colind[k] - g.boxOff[sb];
        dst[lay.uOff[lr] + r + (int64_t)p * m] = T(1);
        dst[lay.vOff[lr] + c + (int64_t)p * nn] = (T)vals[k];
    }
}

// Row of every nonzero (for pass 3).
__global__ void csr_rows_kernel(const int* __restrict__ rowptr, int n, int* __restrict__ rowOf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) rowOf[k] = i;
}

void launch_csr_locate(const GeoDev& g, const int* rowptr, const int* colind, int n, int* nzBlock, int* lrCount,
                       int* bad, hipStream_t s)
{
    if (n <= 0) return;
    csr_locate_kernel<<<(n + 255) / 256, 256, 0, s>>>(g, rowptr, colind, n, nzBlock, lrCount, bad);
    HIP_LAUNCH_CHECK();
}

template <typename T>
void launch_csr_fill(const GeoDev& g, const LayDev& lay, const int* rowptr, const int* colind, const double* vals,
                     int n, const int* nzBlock, const int* lrBase, int* lrFill, const int* lrCount, int* stageK,
                     int* rowOf, int nLr, T* dst, hipStream_t s)
{
    if (n <= 0) return;
    csr_rows_kernel<<<(n + 255) / 256, 256, 0, s>>>(rowptr, n, rowOf);
    csr_scatter_kernel<T><<<(n + 255) / 256, 256, 0, s>>>(g, lay, rowptr, colind, vals, n, nzBlock, lrBase, lrFill,
                                                         stageK, dst);
    if (nLr > 0)
        csr_lowrank_kernel<T><<<(nLr + 255) / 256, 256, 0, s>>>(g, lay, rowOf, colind, vals, lrBase, lrCount,
                                                               stageK, nLr, dst);
    HIP_LAUNCH_CHECK();
}

// Natural index (x + nx y [+ nx ny z]) -> hierarchical index: descend the
// box tree, choosing at each level the child whose per-axis range holds the
// point (children split every axis at e/2), and add the point's offset
// inside the leaf box (natural order within the leaf).
__global__ void natural_map_kernel(GeoDev g, int nx, int ny, int nz, int* __restrict__ map)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = nx * ny * nz;
    if (p >= n) return;
    int x = p % nx, y = (p / nx) % ny, z = p / (nx * ny);
    int box = 0;                                      // flattened root
    int lo[3] = {0, 0, 0}, ext[3] = {nx, ny, nz};
    for (int l = 0; l + 1 < g.numLevels; ++l) {
        int c = 0;
        for (int a = 0; a < g.dim; ++a) {
            const int half = ext[a] / 2;
            const int q = a == 0 ? x : (a == 1 ? y : z);
            if (q - lo[a] >= half) { c |= 1 << a; lo[a] += half; ext[a] -= half; }
            else ext[a] = half;
        }
        box = g.boxChild[box * g.nc + c];
    }
    const int rx = x - lo[0], ry = y - lo[1], rz = z - lo[2];
    map[p] = g.boxOff[box] + rx + ext[0] * (ry + ext[1] * rz);
}

void launch_natural_map(const GeoDev& g, int nx, int ny, int nz, int* map, hipStream_t s)
{
    const int n = nx * ny * nz;
    if (n <= 0) return;
    natural_map_kernel<<<(n + 255) / 256, 256, 0, s>>>(g, nx, ny, nz, map);
    HIP_LAUNCH_CHECK();
}

template void launch_csr_fill<double>(const GeoDev&, const LayDev&, const int*, const int*, const double*, int,
                                      const int*, const int*, int*, const int*, int*, int*, int, double*,
                                      hipStream_t);
template void launch_csr_fill<float>(const GeoDev&, const LayDev&, const int*, const int*, const double*, int,
                                     const int*, const int*, int*, const int*, int*, int*, int, float*,
                                     hipStream_t);

} // namespace dmhm
