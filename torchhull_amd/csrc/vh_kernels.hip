// Visual-hull kernels for CDNA4 (gfx950): integral image, level-synchronous
// octree carving with stable compaction, corner-vertex bitmap and the field.
//
// Reference behaviour (vc-bonn/torchhull, src/torchhull/_C/src/visual_hull_cuda.cu,
// abbreviated VH): integral image VH:76-90, classification VH:168-417, level
// loop and compaction VH:627-751, corner dedup VH:770-825, field VH:419-520
// and VH:827-888.

#include <cfloat>

#include <hip/hip_fp16.h>

#include "kernels.h"
#include "scan.h"

namespace thull
{

// ===========================================================================
// Mask element access
// ===========================================================================

template <typename T>
__device__ __forceinline__ float
as_float(T v)
{
    return static_cast<float>(v);
}

template <>
__device__ __forceinline__ float
as_float<__half>(__half v)
{
    return __half2float(v);
}

// Mask sample as float (masks of any element type are used as float32,
// image_utils.h:178).  The two native element types are f32 and u8 (bool);
// other types are converted to f32 by the caller.
__device__ __forceinline__ float
mask_at(const void* masks, int32_t dtype, int64_t idx)
{
    if (dtype == kF32)
        return static_cast<const float*>(masks)[idx];
    return float(static_cast<const uint8_t*>(masks)[idx]);
}

// ===========================================================================
// Integral image (VH:76-90)
//
// torch::cumsum on CUDA scans every non-innermost dimension with
// tensor_kernel_scan_outer_dim (ATen/native/cuda/ScanUtils.cuh): one
// sequential float chain per line, acc = acc + x from acc = 0.  Both passes of
// the reference are of that kind on [B,H,W,1], so the integral image is
// reproduced exactly by a sequential column pass followed by a sequential row
// pass.  The parallelism is across lines: B*W column chains, then B*H row
// chains.
// ===========================================================================

template <typename T>
__global__ __launch_bounds__(256) void
integral_cols_kernel(const T* __restrict__ m, float* __restrict__ out, int32_t H, int32_t W)
{
    const int32_t x = blockIdx.x * 256 + threadIdx.x;
    if (x >= W)
        return;
    const int64_t plane = int64_t(H) * W;
    const T* src = m + blockIdx.y * plane + x;
    float* dst = out + blockIdx.y * plane + x;
    float acc = 0.0f;
    int32_t y = 0;
    constexpr int U = 16;
    for (; y + U <= H; y += U)
    {
        float v[U];
#pragma unroll
        for (int k = 0; k < U; ++k)
            v[k] = as_float(src[int64_t(y + k) * W]);
#pragma unroll
        for (int k = 0; k < U; ++k)
        {
            acc = acc + v[k];
            dst[int64_t(y + k) * W] = acc;
        }
    }
    for (; y < H; ++y)
    {
        acc = acc + as_float(src[int64_t(y) * W]);
        dst[int64_t(y) * W] = acc;
    }
}

// Row pass in place: one wavefront owns 64 rows and walks 64x64 tiles through
// LDS, so loads and stores stay coalesced while every lane runs the
// sequential chain of its own row.
__global__ __launch_bounds__(64) void
integral_rows_kernel(float* __restrict__ io, int32_t H, int32_t W)
{
    __shared__ float tile[64][65];
    const int lane = threadIdx.x;
    const int32_t r0 = blockIdx.x * 64;
    float* img = io + int64_t(blockIdx.y) * H * W;
    const int32_t rows = min(64, H - r0);
    float acc = 0.0f;
    for (int32_t c0 = 0; c0 < W; c0 += 64)
    {
        const int32_t cols = min(64, W - c0);
        if (lane < cols)
            for (int r = 0; r < rows; ++r)
                tile[r][lane] = img[int64_t(r0 + r) * W + c0 + lane];
        __syncthreads();
        if (lane < rows)
            for (int c = 0; c < cols; ++c)
            {
                acc = acc + tile[lane][c];
                tile[lane][c] = acc;
            }
        __syncthreads();
        if (lane < cols)
            for (int r = 0; r < rows; ++r)
                img[int64_t(r0 + r) * W + c0 + lane] = tile[r][lane];
        __syncthreads();
    }
}

void
launch_integral_image(const void* masks, int32_t dtype, int64_t B, int64_t H, int64_t W, float* out, hipStream_t st)
{
    if (B == 0 || H == 0 || W == 0)
        return;
    dim3 gc(blocks_for(W, 256), unsigned(B));
    switch (dtype)
    {
    case kF32:
        integral_cols_kernel<float><<<gc, 256, 0, st>>>(static_cast<const float*>(masks), out, int32_t(H), int32_t(W));
        break;
    case kU8:
    case kBool:
        integral_cols_kernel<uint8_t>
                <<<gc, 256, 0, st>>>(static_cast<const uint8_t*>(masks), out, int32_t(H), int32_t(W));
        break;
    default:
        throw Error("integral image: unsupported mask element type");
    }
    THULL_LAUNCH_CHECK();
    dim3 gr(blocks_for(H, 64), unsigned(B));
    integral_rows_kernel<<<gr, 64, 0, st>>>(out, int32_t(H), int32_t(W));
    THULL_LAUNCH_CHECK();
}

// ===========================================================================
// Octree carving (VH:168-417)
//
// One thread per parent cell.  The 8 children of a parent share a 3x3x3
// lattice of corners; each lattice point is projected once per camera (the
// reference projects 8 corners per child, 64 per parent) with bit-identical
// results, because a shared corner has the same integer grid coordinate and
// hence the same world point.  The fused row products are shared along the
// lattice axes: fmaf(A2, z, A3) per z, fmaf(A1, y, .) per (y, z), and
// fmaf(A0, x, .) per point, exactly the nested fmaf chain of VH:150-153.
//
// Cameras are visited in index order and a child stops at the first camera
// that classifies it empty, which fixes its result (VH:283, VH:415); the
// remaining children of the parent continue.
// ===========================================================================

// isolevel - epsilon = 0.5f - 0.1f (VH:260-267)
constexpr float kMargin = 0.5f - 1e-1f;

struct CarveCamera
{
    float x0, x1, x2, x3; // row 0
    float y0, y1, y2, y3; // row 1
    float d0, d1, d2, d3; // row 3 (OpenGL, w) or row 2 (OpenCV, z)
};

__device__ __forceinline__ CarveCamera
load_camera(const float* __restrict__ T, bool gl)
{
    CarveCamera c;
    const float4 r0 = *reinterpret_cast<const float4*>(T);
    const float4 r1 = *reinterpret_cast<const float4*>(T + 4);
    const float4 rd = *reinterpret_cast<const float4*>(T + (gl ? 12 : 8));
    c.x0 = r0.x;
    c.x1 = r0.y;
    c.x2 = r0.z;
    c.x3 = r0.w;
    c.y0 = r1.x;
    c.y1 = r1.y;
    c.y2 = r1.z;
    c.y3 = r1.w;
    c.d0 = rd.x;
    c.d1 = rd.y;
    c.d2 = rd.z;
    c.d3 = rd.w;
    return c;
}

struct Box
{
    float minx, miny, maxx, maxy;
};

__device__ __forceinline__ Box
box_merge(const Box& a, const Box& b)
{
    return Box{ fminf(a.minx, b.minx), fminf(a.miny, b.miny), fmaxf(a.maxx, b.maxx), fmaxf(a.maxy, b.maxy) };
}

// Classification of one child against one camera.  `bit` is the child's bit.
// Returns nothing; updates alive / refine / inside masks.
__device__ __forceinline__ void
classify_child_camera(Box bb,
                      bool partial,
                      bool fully_inside,
                      const float* __restrict__ I,
                      int32_t H,
                      int32_t W,
                      float Hm1,
                      float Wm1,
                      uint32_t bit,
                      uint32_t& alive,
                      uint32_t& refine,
                      uint32_t& inside_any)
{
    // The reference starts the box at (FLT_MAX, -FLT_MAX) and folds the 8
    // corners in with fminf/fmaxf (VH:201-234); NaN corners are ignored.
    bb.minx = fminf(FLT_MAX, bb.minx);
    bb.miny = fminf(FLT_MAX, bb.miny);
    bb.maxx = fmaxf(-FLT_MAX, bb.maxx);
    bb.maxy = fmaxf(-FLT_MAX, bb.maxy);
    // VH:236-245
    const float rminx = roundf(bb.minx + -0.5f);
    const float rminy = roundf(bb.miny + -0.5f);
    const float rmaxx = roundf(bb.maxx + 0.5f);
    const float rmaxy = roundf(bb.maxy + 0.5f);
    const int32_t x0 = clamp_rounded(rminx, Wm1);
    const int32_t y0 = clamp_rounded(rminy, Hm1);
    const int32_t x1 = clamp_rounded(rmaxx, Wm1);
    const int32_t y1 = clamp_rounded(rmaxy, Hm1);
    const int32_t area = (y1 - y0 + 1) * (x1 - x0 + 1); // VH:247
    // VH:249-257, zero padding outside the image
    const float i11 = I[y1 * W + x1];
    const float i10 = (x0 > 0) ? I[y1 * W + (x0 - 1)] : 0.0f;
    const float i01 = (y0 > 0) ? I[(y0 - 1) * W + x1] : 0.0f;
    const float i00 = (x0 > 0 && y0 > 0) ? I[(y0 - 1) * W + (x0 - 1)] : 0.0f;
    const float sum = ((i11 + i00) - i10) - i01;
    if (!partial)
    {
        // VH:269-280
        if (sum <= kMargin)
            alive &= ~bit;
        else if (!(sum >= float(area) - kMargin))
            refine |= bit;
    }
    else
    {
        // VH:376-412; the unclamped box area uses int64 wrap-around like the GPU reference
        const uint64_t fy = uint64_t(f2i64_sat(rmaxy)) - uint64_t(f2i64_sat(rminy)) + 1u;
        const uint64_t fx = uint64_t(f2i64_sat(rmaxx)) - uint64_t(f2i64_sat(rminx)) + 1u;
        const bool same = int64_t(fy * fx) == int64_t(area);
        if (sum <= kMargin && same)
            alive &= ~bit;
        else if (sum >= float(area) - kMargin && same)
        {
        }
        else if (y1 - y0 > 1 && x1 - x0 > 1)
            refine |= bit;
        if (fully_inside)
            inside_any |= bit;
    }
}

__device__ uint32_t
classify_parent(const SceneDev& s, uint32_t px, uint32_t py, uint32_t pz, int32_t li, bool last, uint32_t& tests)
{
    const float inv = __builtin_ldexpf(1.0f, -(li + 1)); // 1 / resolution_children, exact
    float wx[3], wy[3], wz[3];
#pragma unroll
    for (int u = 0; u < 3; ++u)
    {
        wx[u] = grid_to_world(2 * px + u, inv, s.length, s.corner[0]);
        wy[u] = grid_to_world(2 * py + u, inv, s.length, s.corner[1]);
        wz[u] = grid_to_world(2 * pz + u, inv, s.length, s.corner[2]);
    }
    const bool gl = s.opengl != 0;
    const bool partial = s.partial != 0;
    const int32_t H = s.H, W = s.W;
    const float fH = float(H), fW = float(W);
    const float Hm1 = float(H - 1), Wm1 = float(W - 1);
    const float Hm2 = float(H - 2), Wm2 = float(W - 2);
    const int64_t plane = int64_t(H) * W;

    uint32_t alive = 0xffu, refine = 0u, inside_any = 0u;
    for (int32_t b = 0; b < s.B && alive != 0u; ++b)
    {
        tests += __popc(alive);
        const CarveCamera cam = load_camera(s.transforms + 16 * b, gl);
        const float* I = s.integral + b * plane;

        // window boxes of the three z layers: win[w][a + 2 b] covers lattice
        // points (a..a+1, b..b+1, w)
        Box win[3][4];
        uint32_t inbits[3] = { 0u, 0u, 0u }; // per layer, bit u + 3 v: rounded pixel inside with margin 1
#pragma unroll
        for (int w = 0; w < 3; ++w)
        {
            const float ax = fmaf(cam.x2, wz[w], cam.x3);
            const float ay = fmaf(cam.y2, wz[w], cam.y3);
            const float ad = fmaf(cam.d2, wz[w], cam.d3);
            float ppx[3][3], ppy[3][3]; // [v][u]
#pragma unroll
            for (int v = 0; v < 3; ++v)
            {
                const float bx = fmaf(cam.x1, wy[v], ax);
                const float by = fmaf(cam.y1, wy[v], ay);
                const float bd = fmaf(cam.d1, wy[v], ad);
#pragma unroll
                for (int u = 0; u < 3; ++u)
                {
                    const float cx = fmaf(cam.x0, wx[u], bx);
                    const float cy = fmaf(cam.y0, wx[u], by);
                    const float cd = fmaf(cam.d0, wx[u], bd);
                    const float qx = cx / cd;
                    const float qy = cy / cd;
                    float pxl, pyl;
                    if (gl)
                    {
                        pxl = unnormalize_ndc(qx, fW);
                        pyl = unnormalize_ndc(qy, fH);
                    }
                    else
                    {
                        pxl = qx - 0.5f;
                        pyl = qy - 0.5f;
                    }
                    ppx[v][u] = pxl;
                    ppy[v][u] = pyl;
                    if (partial)
                    {
                        // VH:358-362: roundf(pixel) inside the image with margin 1
                        const float rx = roundf(pxl), ry = roundf(pyl);
                        if (rx >= 1.0f && rx <= Wm2 && ry >= 1.0f && ry <= Hm2)
                            inbits[w] |= 1u << (u + 3 * v);
                    }
                }
            }
            // 2x2 windows in the layer: reduce along u, then along v
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int a = 0; a < 2; ++a)
                {
                    const float mnx0 = fminf(ppx[bb][a], ppx[bb][a + 1]);
                    const float mnx1 = fminf(ppx[bb + 1][a], ppx[bb + 1][a + 1]);
                    const float mny0 = fminf(ppy[bb][a], ppy[bb][a + 1]);
                    const float mny1 = fminf(ppy[bb + 1][a], ppy[bb + 1][a + 1]);
                    const float mxx0 = fmaxf(ppx[bb][a], ppx[bb][a + 1]);
                    const float mxx1 = fmaxf(ppx[bb + 1][a], ppx[bb + 1][a + 1]);
                    const float mxy0 = fmaxf(ppy[bb][a], ppy[bb][a + 1]);
                    const float mxy1 = fmaxf(ppy[bb + 1][a], ppy[bb + 1][a + 1]);
                    win[w][a + 2 * bb] = Box{ fminf(mnx0, mnx1), fminf(mny0, mny1), fmaxf(mxx0, mxx1), fmaxf(mxy0, mxy1) };
                }
        }

#pragma unroll
        for (int c = 0; c < 8; ++c)
        {
            const uint32_t bit = 1u << c;
            if (!(alive & bit))
                continue;
            const int a = c & 1, bb = (c >> 1) & 1, cz = (c >> 2) & 1;
            const Box box = box_merge(win[cz][a + 2 * bb], win[cz + 1][a + 2 * bb]);
            bool inside = false;
            if (partial)
            {
                const uint32_t m = (3u << a) | (3u << (a + 3));
                const uint32_t mm = m << (3 * bb);
                inside = ((inbits[cz] & mm) == mm) && ((inbits[cz + 1] & mm) == mm);
            }
            classify_child_camera(box, partial, inside, I, H, W, Hm1, Wm1, bit, alive, refine, inside_any);
        }
    }
    uint32_t occ = refine & alive; // should_refine && !is_empty
    if (partial && last)
        occ &= inside_any; // (!last_children || fully_inside_one_frame)
    return occ;
}

__global__ __launch_bounds__(kCarveThreads) void
carve_classify_kernel(const uint64_t* __restrict__ parents,
                      uint32_t n,
                      int32_t level_i,
                      const SceneDev* __restrict__ scenes,
                      uint8_t* __restrict__ child_mask,
                      uint32_t* __restrict__ block_counts,
                      unsigned long long* __restrict__ tests_out)
{
    __shared__ uint32_t lds[kCarveThreads / 64];
    const uint32_t t = blockIdx.x * kCarveThreads + threadIdx.x;
    uint32_t mask = 0u, tests = 0u;
    if (t < n)
    {
        const uint64_t key = parents[t];
        const SceneDev s = scenes[key_scene(key)];
        if (level_i < s.level)
            mask = classify_parent(s, key_x(key), key_y(key), key_z(key), level_i, level_i == s.level - 1, tests);
        child_mask[t] = uint8_t(mask);
    }
    uint32_t total;
    block_exclusive_sum(__popc(mask), lds, total);
    if (threadIdx.x == 0)
        block_counts[blockIdx.x] = total;
    uint32_t tsum;
    block_exclusive_sum(tests, lds, tsum);
    if (threadIdx.x == 0 && tests_out != nullptr)
        atomicAdd(tests_out, (unsigned long long)tsum);
}

__global__ __launch_bounds__(kCarveThreads) void
carve_scatter_kernel(const uint64_t* __restrict__ parents,
                     uint32_t n,
                     const uint8_t* __restrict__ child_mask,
                     const uint32_t* __restrict__ block_offsets,
                     uint64_t* __restrict__ children)
{
    __shared__ uint32_t lds[kCarveThreads / 64];
    const uint32_t t = blockIdx.x * kCarveThreads + threadIdx.x;
    const uint32_t mask = (t < n) ? child_mask[t] : 0u;
    uint32_t total;
    uint32_t off = block_offsets[blockIdx.x] + block_exclusive_sum(__popc(mask), lds, total);
    if (mask == 0u)
        return;
    const uint64_t key = parents[t];
    const uint32_t s = key_scene(key), x = 2 * key_x(key), y = 2 * key_y(key), z = 2 * key_z(key);
    // VH:694-711: children of a parent in child order (x bit, y bit, z bit)
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (mask & (1u << c))
            children[off++] = pack_key(s, x + (c & 1), y + ((c >> 1) & 1), z + ((c >> 2) & 1));
}

void
launch_carve_classify(const uint64_t* parents,
                      uint32_t n,
                      int32_t level_i,
                      const SceneDev* scenes,
                      uint8_t* child_mask,
                      uint32_t* block_counts,
                      unsigned long long* tests,
                      hipStream_t st)
{
    if (n == 0)
        return;
    carve_classify_kernel<<<blocks_for(n, kCarveThreads), kCarveThreads, 0, st>>>(parents, n, level_i, scenes,
                                                                                 child_mask, block_counts, tests);
    THULL_LAUNCH_CHECK();
}

void
launch_carve_scatter(const uint64_t* parents,
                     uint32_t n,
                     const uint8_t* child_mask,
                     const uint32_t* block_offsets,
                     uint64_t* children,
                     hipStream_t st)
{
    if (n == 0)
        return;
    carve_scatter_kernel<<<blocks_for(n, kCarveThreads), kCarveThreads, 0, st>>>(parents, n, child_mask,
                                                                                block_offsets, children);
    THULL_LAUNCH_CHECK();
}

// ===========================================================================
// Corner vertices (VH:770-825)
//
// Instead of sorting the 8 corner ids of every leaf and removing duplicates,
// the corners are set in a bitmap of the (R+1)^3 grid whose rows are padded
// to whole 32-bit words.  The bitmap order is the (z, y, x) order of the
// reference's sorted ravel ids, so a popcount prefix over the words yields the
// sorted unique vertex list directly, and also serves as the rank dictionary
// (vertex -> list position) of marching cubes.
// ===========================================================================

__global__ __launch_bounds__(256) void
mark_corners_kernel(const uint64_t* __restrict__ cells,
                    uint32_t n,
                    int32_t level,
                    const SceneDev* __restrict__ scenes,
                    uint32_t* __restrict__ bitmap)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n)
        return;
    const uint64_t key = cells[t];
    const SceneDev& s = scenes[key_scene(key)];
    if (s.level != level)
        return;
    const uint32_t x = key_x(key), y = key_y(key), z = key_z(key);
    const uint64_t R1 = (uint64_t(1) << level) + 1;
    uint32_t* base = bitmap + s.bitmap_offset;
    const uint32_t bit = x & 31u;
#pragma unroll
    for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
        {
            uint32_t* w = base + ((uint64_t(z + dz) * R1 + (y + dy)) * s.row_words + (x >> 5));
            if (bit != 31u)
                atomicOr(w, 3u << bit);
            else
            {
                atomicOr(w, 1u << 31);
                atomicOr(w + 1, 1u);
            }
        }
}

void
launch_mark_corners(const uint64_t* cells, uint32_t n, int32_t level, const SceneDev* scenes, uint32_t* bitmap,
                    hipStream_t st)
{
    if (n == 0)
        return;
    mark_corners_kernel<<<blocks_for(n, 256), 256, 0, st>>>(cells, n, level, scenes, bitmap);
    THULL_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void
extract_vertices_kernel(const uint32_t* __restrict__ bitmap,
                        const uint32_t* __restrict__ prefix,
                        uint64_t nwords,
                        const uint64_t* __restrict__ scene_word_begin,
                        const SceneDev* __restrict__ scenes,
                        int32_t nscenes,
                        uint64_t* __restrict__ vertices)
{
    const uint64_t w = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (w >= nwords)
        return;
    uint32_t word = bitmap[w];
    if (word == 0u)
        return;
    // scene of this word: last s with scene_word_begin[s] <= w
    int32_t lo = 0, hi = nscenes - 1;
    while (lo < hi)
    {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (scene_word_begin[mid] <= w)
            lo = mid;
        else
            hi = mid - 1;
    }
    const SceneDev& s = scenes[lo];
    const uint64_t local = w - scene_word_begin[lo];
    const uint64_t R1 = (uint64_t(1) << s.level) + 1;
    const uint64_t row = local / s.row_words;
    const uint32_t xw = uint32_t(local - row * s.row_words);
    const uint32_t y = uint32_t(row % R1), z = uint32_t(row / R1);
    uint32_t pos = prefix[w];
    while (word != 0u)
    {
        const uint32_t b = __builtin_ctz(word);
        vertices[pos++] = pack_key(uint32_t(lo), xw * 32u + b, y, z);
        word &= word - 1u;
    }
}

void
launch_extract_vertices(const uint32_t* bitmap,
                        const uint32_t* prefix,
                        uint64_t nwords,
                        const uint64_t* scene_word_begin,
                        const SceneDev* scenes,
                        int32_t nscenes,
                        uint64_t* vertices,
                        hipStream_t st)
{
    if (nwords == 0)
        return;
    extract_vertices_kernel<<<blocks_for(int64_t(nwords), 256), 256, 0, st>>>(bitmap, prefix, nwords,
                                                                                scene_word_begin, scenes, nscenes,
                                                                                vertices);
    THULL_LAUNCH_CHECK();
}

// ===========================================================================
// Field (VH:419-520, camera loop VH:847)
//
// The reference launches one kernel per camera, each re-reading and
// re-writing the value array.  Here one thread owns a vertex and multiplies
// the camera samples in camera order 0..B-1 in a register, which gives the
// same float product.
// ===========================================================================

__device__ __forceinline__ float
i64_to_f32(int64_t v)
{
    if (v >= INT32_MIN && v <= INT32_MAX)
        return float(int32_t(v));
    return float(v);
}

__global__ __launch_bounds__(256) void
field_kernel(const uint64_t* __restrict__ vertices, uint32_t n, const SceneDev* __restrict__ scenes,
             float* __restrict__ values)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t key = vertices[i];
    const SceneDev s = scenes[key_scene(key)];
    const float inv = __builtin_ldexpf(1.0f, -s.level);
    const float wx = grid_to_world(key_x(key), inv, s.length, s.corner[0]);
    const float wy = grid_to_world(key_y(key), inv, s.length, s.corner[1]);
    const float wz = grid_to_world(key_z(key), inv, s.length, s.corner[2]);
    const bool gl = s.opengl != 0;
    const bool partial = s.partial != 0;
    const int64_t H = s.H, W = s.W;
    const float fH = float(s.H), fW = float(s.W);
    const float Hm2 = float(s.H - 2), Wm2 = float(s.W - 2);
    const int64_t plane = H * W;
    const float pad = partial ? 1.0f : 0.0f; // ones padding (partial) or zeros padding (full), IU:162-214
    float value = 1.0f;
    for (int32_t b = 0; b < s.B; ++b)
    {
        const float* T = s.transforms + 16 * b;
        const float cx = fmaf(T[0], wx, fmaf(T[1], wy, fmaf(T[2], wz, T[3])));
        const float cy = fmaf(T[4], wx, fmaf(T[5], wy, fmaf(T[6], wz, T[7])));
        float px, py;
        if (gl)
        {
            const float cw = fmaf(T[12], wx, fmaf(T[13], wy, fmaf(T[14], wz, T[15])));
            px = unnormalize_ndc(cx / cw, fW);
            py = unnormalize_ndc(cy / cw, fH);
        }
        else
        {
            const float cz = fmaf(T[8], wx, fmaf(T[9], wy, fmaf(T[10], wz, T[11])));
            px = cx / cz - 0.5f;
            py = cy / cz - 0.5f;
        }
        if (partial)
        {
            // VH:513-518: only cameras whose rounded pixel lies inside with margin 1
            const float rx = roundf(px), ry = roundf(py);
            if (!(rx >= 1.0f && rx <= Wm2 && ry >= 1.0f && ry <= Hm2))
                continue;
        }
        const int64_t y0 = f2i64_sat(roundf(py + -0.5f));
        const int64_t x0 = f2i64_sat(roundf(px + -0.5f));
        const bool y0in = y0 >= 0 && y0 < H, y1in = y0 >= -1 && y0 < H - 1;
        const bool x0in = x0 >= 0 && x0 < W, x1in = x0 >= -1 && x0 < W - 1;
        const int64_t base = b * plane;
        const float v00 = (y0in && x0in) ? mask_at(s.masks, s.mask_dtype, base + y0 * W + x0) : pad;
        const float v10 = (y1in && x0in) ? mask_at(s.masks, s.mask_dtype, base + (y0 + 1) * W + x0) : pad;
        const float v01 = (y0in && x1in) ? mask_at(s.masks, s.mask_dtype, base + y0 * W + x0 + 1) : pad;
        const float v11 = (y1in && x1in) ? mask_at(s.masks, s.mask_dtype, base + (y0 + 1) * W + x0 + 1) : pad;
        const float ty = py - i64_to_f32(y0);
        const float tx = px - i64_to_f32(x0);
        const float l0 = lerp_ref(v00, v10, ty);
        const float l1 = lerp_ref(v01, v11, ty);
        value *= lerp_ref(l0, l1, tx);
    }
    values[i] = value;
}

void
launch_field(const uint64_t* vertices, uint32_t n, const SceneDev* scenes, float* values, hipStream_t st)
{
    if (n == 0)
        return;
    field_kernel<<<blocks_for(n, 256), 256, 0, st>>>(vertices, n, scenes, values);
    THULL_LAUNCH_CHECK();
}

} // namespace thull
