// Generic LBVH over float boxes on gfx950, shared by stage 3 (contact and
// closest-point trees) and the surface metrics: Morton codes of box centres,
// the radix sort of prim.h, Karras's hierarchy and bottom-up node boxes
// (the stage-2 tree kernels of stage2_kernels.h), escape links for
// stackless walks.
#pragma once

#include <algorithm>

#include "pamo_hip.h"
#include "prim.h"
#include "stage2_kernels.h"
#include "stage3_math.h"

namespace pamo {
namespace s3 {
namespace {

constexpr int kB = 256;

// ------------------------------------------------------------ generic LBVH over boxes
struct Tree {
    int n = 0, cap = 0;
    float4* item = nullptr;  // 2 per item (lo, hi)
    uint32_t *key = nullptr, *val = nullptr, *key2 = nullptr, *val2 = nullptr;
    int *left = nullptr, *right = nullptr, *parent = nullptr;
    int2* range = nullptr;
    float4 *box = nullptr, *pre = nullptr, *suf = nullptr, *blk = nullptr, *blkpre = nullptr, *blksuf = nullptr,
           *sblk = nullptr, *blktab = nullptr;
    long long tab_stride = 0;  // entries per level of blktab
    float* part = nullptr;
    int* cnt = nullptr;  // [C_COUNT]; cnt[C_NLIVE] = n
};

template <typename T>
static int dalloc(T** q, size_t n, hipStream_t st) {
    if (*q) PAMO_HIP(hipFree(*q));
    *q = nullptr;
    PAMO_HIP(hipMalloc((void**)q, sizeof(T) * std::max<size_t>(n, 1)));
    // zeroed in order on the owning stream: a null-stream hipMemset is not
    // ordered with the context's non-blocking stream and may complete after
    // the kernels that fill the buffer
    PAMO_HIP(hipMemsetAsync(*q, 0, sizeof(T) * std::max<size_t>(n, 1), st));
    return 0;
}

static int tree_alloc(Tree& t, int n, hipStream_t st) {
    if (n <= t.cap) return 0;
    size_t a = (size_t)n + n / 4 + 16;
    t.cap = (int)a;
    PAMO_TRY(dalloc(&t.item, 2 * a, st));
    PAMO_TRY(dalloc(&t.key, a, st)); PAMO_TRY(dalloc(&t.val, a, st));
    PAMO_TRY(dalloc(&t.key2, a, st)); PAMO_TRY(dalloc(&t.val2, a, st));
    PAMO_TRY(dalloc(&t.left, a, st)); PAMO_TRY(dalloc(&t.right, a, st)); PAMO_TRY(dalloc(&t.parent, 2 * a, st));
    PAMO_TRY(dalloc(&t.range, a, st));
    PAMO_TRY(dalloc(&t.box, 4 * a, st));
    PAMO_TRY(dalloc(&t.pre, 2 * a, st)); PAMO_TRY(dalloc(&t.suf, 2 * a, st));
    size_t nb = a / kBlock + 2, ns = nb / kBlock + 2;
    PAMO_TRY(dalloc(&t.blk, 2 * nb, st)); PAMO_TRY(dalloc(&t.blkpre, 2 * nb, st)); PAMO_TRY(dalloc(&t.blksuf, 2 * nb, st));
    PAMO_TRY(dalloc(&t.sblk, 2 * ns, st));
    PAMO_TRY(dalloc(&t.blktab, 2 * kTabLevels * nb, st));
    t.tab_stride = (long long)nb;
    PAMO_TRY(dalloc(&t.part, 6 * (a / kBlock + 4), st));
    if (!t.cnt) PAMO_TRY(dalloc(&t.cnt, C_COUNT, st));
    return 0;
}

static void tree_free(Tree& t) {
    void* ptrs[] = {t.item, t.key, t.val, t.key2, t.val2, t.left, t.right, t.parent, t.range, t.box,
                    t.pre,  t.suf, t.blk, t.blkpre, t.blksuf, t.sblk, t.blktab, t.part, t.cnt};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    t = Tree{};
}

__global__ void items_box_part_k(const float4* __restrict__ item, int n, float* __restrict__ part) {
    __shared__ float red[6][kBlock];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const float inf = __builtin_huge_valf();
    float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
    if (i < n) {
        float4 a = item[2 * (long long)i], b = item[2 * (long long)i + 1];
        lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
        hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
    }
    block_minmax6(lo, hi, red);
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void items_morton_k(const float4* __restrict__ item, int n, const float* __restrict__ whole,
                               uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 lo = item[2 * (long long)i], hi = item[2 * (long long)i + 1];
    float c[3] = {(hi.x + lo.x) * 0.5f, (hi.y + lo.y) * 0.5f, (hi.z + lo.z) * 0.5f};
    for (int d = 0; d < 3; ++d) {
        float ext = whole[3 + d] - whole[d];
        c[d] = ext > 0.0f ? (c[d] - whole[d]) / ext : 0.0f;
    }
    key[i] = morton3(c[0], c[1], c[2]);
    val[i] = (uint32_t)i;
}

// Leaf and node boxes of a built hierarchy (box_tab_k, node_box_k) and,
// unless refitting a tree whose links are already written, its rope links.
static int tree_boxes(Tree& t, int n, bool refit, hipStream_t st) {
    const unsigned nb = grid_for(n, kBlock), nb1 = grid_for(nb, kBlock);
    hipLaunchKernelGGL(box_tab_k, dim3(nb), dim3(kBlock), 0, st, (const float4*)t.item, (const uint32_t*)t.val,
                       (const int*)t.cnt, 0, t.box, (const int2*)t.range, t.pre, t.suf, t.blk, (float4*)nullptr, 0LL);
    hipLaunchKernelGGL(box_tab_k, dim3(nb1), dim3(kBlock), 0, st, (const float4*)t.blk, (const uint32_t*)nullptr,
                       (const int*)t.cnt, 1, (float4*)nullptr, (const int2*)nullptr, t.blkpre, t.blksuf, t.sblk,
                       t.blktab, t.tab_stride);
    hipLaunchKernelGGL(node_box_k, dim3(nb), dim3(kBlock), 0, st, (const int2*)t.range, (const int*)t.cnt,
                       (const float4*)t.pre, (const float4*)t.suf, (const float4*)t.blktab, t.tab_stride,
                       (const float4*)t.blkpre, (const float4*)t.blksuf, (const float4*)t.sblk, t.box);
    if (!refit)
        hipLaunchKernelGGL(rope_k, dim3(grid_for(2LL * n, kBlock)), dim3(kBlock), 0, st, (const int*)t.left,
                           (const int2*)t.range, (const int*)t.cnt, (const uint32_t*)t.val, t.box);
    PAMO_HIP(hipGetLastError());
    return 0;
}

// Builds the tree over t.item[0..n) (boxes already written).
static int tree_build(Tree& t, int n, int* ws, hipStream_t st) {
    t.n = n;
    PAMO_TRY(fill(t.cnt + C_NLIVE, n, 1, st));
    if (n <= 0) return 0;
    const unsigned nb = grid_for(n, kBlock);
    hipLaunchKernelGGL(items_box_part_k, dim3(nb), dim3(kBlock), 0, st, (const float4*)t.item, n, t.part);
    hipLaunchKernelGGL(whole_box_k, dim3(1), dim3(kBlock), 0, st, t.part, (int)nb);
    hipLaunchKernelGGL(items_morton_k, dim3(nb), dim3(kBlock), 0, st, (const float4*)t.item, n,
                       (const float*)(t.part + 6 * nb), t.key, t.val);
    PAMO_HIP(hipGetLastError());
    PAMO_TRY(radix_sort_pairs(t.key, t.val, t.key2, t.val2, n, 32, ws, st));
    hipLaunchKernelGGL(karras_k, dim3(nb), dim3(kBlock), 0, st, (const uint32_t*)t.key, (const uint32_t*)t.val,
                       (const int*)t.cnt, t.left, t.right, t.parent, t.range);
    return tree_boxes(t, n, false, st);
}

// Recomputes every box of a tree built by tree_build over the same n items,
// after the items' boxes changed: leaf boxes in the built leaf order, node
// boxes over the built hierarchy; the rope links in the records' .w fields
// stay as built.  The tree stays exact for any item boxes; its quality degrades
// only as far as the items moved since the build.
static int tree_refit(Tree& t, int n, hipStream_t st) {
    PAMO_CHECK(n == t.n, "tree_refit: item count differs from the built tree");
    if (n <= 0) return 0;
    return tree_boxes(t, n, true, st);
}

struct TreeView {
    int n;
    const float4* box;
    const int *left, *right, *parent;
    const uint32_t* val;
};
static TreeView view(const Tree& t) { return TreeView{t.n, t.box, t.left, t.right, t.parent, t.val}; }

__device__ __forceinline__ bool box_hit(V3 lo, V3 hi, float4 blo, float4 bhi) {
    return !(hi.x < blo.x || bhi.x < lo.x || hi.y < blo.y || bhi.y < lo.y || hi.z < blo.z || bhi.z < lo.z);
}

}  // namespace
}  // namespace s3
}  // namespace pamo
