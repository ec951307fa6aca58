// Host-side launchers of the device kernels (one translation unit each).
#pragma once

#include "common.h"

namespace thull
{

// ---- integral image (vh_kernels.hip) ----
// float32 integral image of masks [B,H,W] of element type `dtype`, equal to
// torch::cumsum(.., 1, f32) followed by cumsum(.., 2, f32) on CUDA.
void launch_integral_image(const void* masks, int32_t dtype, int64_t B, int64_t H, int64_t W, float* out,
                           hipStream_t st);

// ---- octree carving (vh_kernels.hip) ----
constexpr int kCarveThreads = 256;

// Classifies the 8 children of each parent; writes an 8-bit occupancy mask per
// parent and the number of occupied children per block of kCarveThreads
// parents.  `tests` accumulates the number of camera tests (early exit).
void launch_carve_classify(const uint64_t* parents, uint32_t n, int32_t level_i, const SceneDev* scenes,
                           uint8_t* child_mask, uint32_t* block_counts, unsigned long long* tests, hipStream_t st);

// Writes the occupied children of every parent, parent-major and child-minor,
// at `block_offsets[block]` + the in-block prefix.
void launch_carve_scatter(const uint64_t* parents, uint32_t n, const uint8_t* child_mask,
                          const uint32_t* block_offsets, uint64_t* children, hipStream_t st);

// Marks the 8 corner vertices of every cell of `cells` whose scene has
// target level `level` in the scene's vertex bitmap.
void launch_mark_corners(const uint64_t* cells, uint32_t n, int32_t level, const SceneDev* scenes, uint32_t* bitmap,
                         hipStream_t st);

// Writes the packed key of every set bit of the bitmap at prefix[word] + rank.
// `scene_word_begin` has nscenes + 1 entries (word offsets of each scene).
void launch_extract_vertices(const uint32_t* bitmap, const uint32_t* prefix, uint64_t nwords,
                             const uint64_t* scene_word_begin, const SceneDev* scenes, int32_t nscenes,
                             uint64_t* vertices, hipStream_t st);

// Field value of every vertex: product over cameras of the bilinear mask sample.
void launch_field(const uint64_t* vertices, uint32_t n, const SceneDev* scenes, float* values, hipStream_t st);

// ---- marching cubes (mc_kernels.hip) ----

// Grid description for marching cubes.  For the bitmap policy, vertex
// positions are ranks in the batch bitmap; for the dense policy the vertex of
// (x, y, z) is at x + dims[0] * (y + dims[1] * z).
struct McGrid
{
    uint32_t dims[3];       // grid vertices per axis
    uint32_t row_words;     // bitmap words per (y, z) row
    uint64_t bitmap_offset; // first bitmap word of this grid
    int32_t out_mode;       // 0: grid units, 1: local [-1, 1] (v * scale - 1), 2: world (fmaf(scale, v, corner))
    float scale[3];
    float corner[3];
    int32_t flip;           // swap face columns 1 and 2
};

enum McValueType : int32_t
{
    kMcF32 = 0,
    kMcF64 = 1,
};

struct McPlan
{
    // inputs
    const McGrid* grids;      // device, one per scene
    const uint64_t* vertices; // packed keys (bitmap policy); null for the dense policy
    uint64_t nverts;          // vertex count (dense: dims product)
    const void* values;       // [nverts] f32 or f64
    int32_t value_type;
    float iso;
    const uint32_t* bitmap;   // bitmap policy
    const uint32_t* prefix;   // bitmap policy: exclusive popcount prefix per word
    bool dense;
    uint32_t dense_dims[3];
};

// Pass 1: per vertex info word (case, complete bit, edge sign-change bits) and
// number of triangle vertices of its cell.
void launch_mc_classify(const McPlan& p, uint32_t* info, uint32_t* ntv, hipStream_t st);
// Pass 2: 3-bit mask of the emitted edges starting at each vertex.
void launch_mc_edges(const McPlan& p, const uint32_t* info, uint32_t* emask, hipStream_t st);
// Pass 3: one output vertex per emitted edge, at vbase[i] + rank within emask.
void launch_mc_vertices(const McPlan& p, const uint32_t* emask, const uint32_t* vbase, float* verts, hipStream_t st);
// Pass 4: faces of every active cell; unique: indices through emask/vbase,
// soup: vertices interpolated per triangle corner.  scene_vstart / scene_tstart
// give the first output vertex / triangle-vertex of each scene.
void launch_mc_faces(const McPlan& p, const uint32_t* info, const uint32_t* ntv_base, const uint32_t* emask,
                     const uint32_t* vbase, const uint32_t* scene_vstart, const uint32_t* scene_tstart, bool unique,
                     float* soup_verts, int64_t* faces, hipStream_t st);

// Grid-vertex bitmap of a sparse COO volume: marks (x, y, z) = (idx[3], idx[2], idx[1]).
void launch_mark_coo(const int64_t* coo, uint64_t nnz, const McGrid* grid, uint32_t* bitmap, hipStream_t st);
// Packed keys (scene 0) of the set bits of a single-grid bitmap.
void launch_extract_grid_vertices(const uint32_t* bitmap, const uint32_t* prefix, uint64_t nwords, const McGrid* grid,
                                  uint64_t* vertices, hipStream_t st);

// ---- gaussian blur (blur_kernels.hip) ----
void launch_gaussian_blur(const void* images, int32_t in_dtype, int64_t B, int64_t H, int64_t W, int kernel_size,
                          float sigma, bool sparse, void* out, int32_t out_dtype, uint8_t* tile_flags,
                          hipStream_t st);
int64_t blur_tile_flag_bytes(int64_t B, int64_t H, int64_t W, int kernel_size);

} // namespace thull
