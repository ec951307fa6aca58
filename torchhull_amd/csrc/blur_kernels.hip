// Gaussian mask blur for CDNA4.
//
// Reference (vc-bonn/torchhull, src/torchhull/_C/src/gaussian_blur_cuda.cu, GB):
// sparse band blur GB:288-507 (edge pixels GB:116-142, 3x3 tile window
// GB:144-178, per-pixel convolution GB:180-286), dense blur GB:20-114.
//
// The reference builds the band as a sorted unique list of tile ids (sort +
// unique of 9 ids per edge pixel).  Here the band is a per-tile flag grid: one
// pass marks every tile that holds an edge pixel, and the output pass treats a
// pixel as a band pixel when any tile of the 3x3 window around its tile is
// marked, which is the same set (the reference's clamped window of an edge
// tile never leaves the tile grid, so membership is symmetric).  Band pixels
// get the convolution with the reference's tap order (x offset outer, y
// offset inner), its accumulation in the output type (fused multiply-add for
// float and double, per-operation binary16 rounding for half) and its final
// division by the weight sum; every other pixel is the converted input.

#include <hip/hip_fp16.h>

#include "kernels.h"

namespace thull
{

template <typename T>
struct OutOps;

template <>
struct OutOps<float>
{
    using W = float;
    __device__ static float zero() { return 0.0f; }
    __device__ static float mac(float acc, float w, float x) { return fmaf(w, x, acc); }
    __device__ static float add(float a, float b) { return a + b; }
    __device__ static float div(float a, float b) { return a / b; }
    __device__ static float from_float(float v) { return v; }
    __device__ static float from_double(double v) { return float(v); }
};

template <>
struct OutOps<double>
{
    __device__ static double zero() { return 0.0; }
    __device__ static double mac(double acc, double w, double x) { return fma(w, x, acc); }
    __device__ static double add(double a, double b) { return a + b; }
    __device__ static double div(double a, double b) { return a / b; }
    __device__ static double from_float(float v) { return double(v); }
    __device__ static double from_double(double v) { return v; }
};

template <>
struct OutOps<__half>
{
    // c10::Half arithmetic: compute in float, round to binary16 after every operation
    __device__ static __half zero() { return __float2half(0.0f); }
    __device__ static __half mac(__half acc, __half w, __half x)
    {
        const __half prod = __float2half(__half2float(w) * __half2float(x));
        return __float2half(__half2float(acc) + __half2float(prod));
    }
    __device__ static __half add(__half a, __half b) { return __float2half(__half2float(a) + __half2float(b)); }
    __device__ static __half div(__half a, __half b) { return __float2half(__half2float(a) / __half2float(b)); }
    __device__ static __half from_float(float v) { return __float2half(v); }
    __device__ static __half from_double(double v) { return __float2half(float(v)); }
};

// static_cast<BlurredImageT>(input value)
template <typename To, typename Ti>
__device__ __forceinline__ To
convert_in(Ti v)
{
    if constexpr (std::is_same_v<Ti, double>)
        return OutOps<To>::from_double(v);
    else if constexpr (std::is_same_v<Ti, __half>)
        return OutOps<To>::from_float(__half2float(v));
    else if constexpr (std::is_same_v<To, double>)
        return double(v);
    else
        return OutOps<To>::from_float(float(v));
}

// Pixel comparison in the input type (GB:140); half compares as float.
template <typename Ti>
__device__ __forceinline__ bool
differs(Ti a, Ti b)
{
    if constexpr (std::is_same_v<Ti, __half>)
        return __half2float(a) != __half2float(b);
    else
        return a != b;
}

// image_utils.h:129-160 (border not repeated)
__device__ __forceinline__ int64_t
reflect_index(int64_t i, int64_t n)
{
    if (i >= 0 && i < n)
        return i;
    if (n == 1)
        return 0;
    const int64_t p = n - 1;
    const int64_t flips = i / p + (i < 0 ? 1 : 0);
    int64_t rem = i % p;
    if (rem < 0)
        rem += p;
    return (flips % 2 == 0) ? rem : p - rem;
}

template <typename Ti>
__global__ __launch_bounds__(256) void
blur_tile_flags_kernel(const Ti* __restrict__ img, int64_t H, int64_t W, int32_t ts, int64_t TH, int64_t TW,
                       uint8_t* __restrict__ flags)
{
    const int64_t tx = int64_t(blockIdx.x) * 16 + (threadIdx.x & 15);
    const int64_t ty = int64_t(blockIdx.y) * 16 + (threadIdx.x >> 4);
    const int64_t b = blockIdx.z;
    if (tx >= TW || ty >= TH)
        return;
    const Ti* I = img + b * H * W;
    uint8_t edge = 0;
    for (int32_t ly = 0; ly < ts && !edge; ++ly)
    {
        const int64_t y = ty * ts + ly;
        if (y >= H)
            break;
        for (int32_t lx = 0; lx < ts; ++lx)
        {
            const int64_t x = tx * ts + lx;
            if (x >= W)
                break;
            const Ti c = I[y * W + x];
            const Ti r = (x + 1 < W) ? I[y * W + x + 1] : c;
            const Ti d = (y + 1 < H) ? I[(y + 1) * W + x] : c;
            if (differs(r, c) || differs(d, c))
            {
                edge = 1;
                break;
            }
        }
    }
    flags[(b * TH + ty) * TW + tx] = edge;
}

template <typename Ti, typename To>
__global__ __launch_bounds__(256) void
blur_output_kernel(const Ti* __restrict__ img,
                   int64_t H,
                   int64_t W,
                   int32_t k,
                   int32_t ts,
                   int64_t TH,
                   int64_t TW,
                   const uint8_t* __restrict__ flags,
                   bool dense,
                   bool divide,
                   const To* __restrict__ weights,
                   To* __restrict__ out)
{
    const int64_t x = int64_t(blockIdx.x) * 64 + (threadIdx.x & 63);
    const int64_t y = int64_t(blockIdx.y) * 4 + (threadIdx.x >> 6);
    const int64_t b = blockIdx.z;
    if (x >= W || y >= H)
        return;
    const Ti* I = img + b * H * W;
    bool band = dense;
    if (!band && flags != nullptr)
    {
        const int64_t tx = x / ts, ty = y / ts;
        for (int64_t dy = -1; dy <= 1 && !band; ++dy)
            for (int64_t dx = -1; dx <= 1; ++dx)
            {
                const int64_t ux = tx + dx, uy = ty + dy;
                if (ux < 0 || uy < 0 || ux >= TW || uy >= TH)
                    continue;
                if (flags[(b * TH + uy) * TW + ux])
                {
                    band = true;
                    break;
                }
            }
    }
    To result;
    if (!band)
    {
        result = convert_in<To>(I[y * W + x]); // images.clone().to(dtype) (GB:444)
    }
    else
    {
        const int32_t h = k / 2;
        const bool interior = (x - h >= 0) && (x + h < W) && (y - h >= 0) && (y + h < H);
        To acc = OutOps<To>::zero();
        To norm = OutOps<To>::zero();
        for (int32_t i = -h; i <= h; ++i)
        {
            const int64_t xi = interior ? x + i : reflect_index(x + i, W);
            for (int32_t j = -h; j <= h; ++j)
            {
                const int64_t yj = interior ? y + j : reflect_index(y + j, H);
                const To v = convert_in<To>(I[yj * W + xi]);
                const To wt = weights[(i + h) * k + (j + h)];
                acc = OutOps<To>::mac(acc, wt, v);
                if (divide)
                    norm = OutOps<To>::add(norm, wt);
            }
        }
        result = divide ? OutOps<To>::div(acc, norm) : acc;
    }
    out[(b * H + y) * W + x] = result;
}

int64_t
blur_tile_flag_bytes(int64_t B, int64_t H, int64_t W, int kernel_size)
{
    const int64_t ts = kernel_size / 2;
    if (ts <= 0)
        return 0;
    return B * (H / ts + 1) * (W / ts + 1);
}

template <typename Ti, typename To>
static void
run_blur(const Ti* img, int64_t B, int64_t H, int64_t W, int k, bool sparse, const To* weights, uint8_t* flags,
         To* out, hipStream_t st)
{
    const int32_t ts = k / 2;
    int64_t TH = 0, TW = 0;
    const bool use_band = sparse && k > 1;
    if (use_band)
    {
        TH = H / ts + 1;
        TW = W / ts + 1;
        dim3 gf(blocks_for(TW, 16), blocks_for(TH, 16), unsigned(B));
        blur_tile_flags_kernel<Ti><<<gf, 256, 0, st>>>(img, H, W, ts, TH, TW, flags);
        THULL_LAUNCH_CHECK();
    }
    dim3 go(blocks_for(W, 64), blocks_for(H, 4), unsigned(B));
    const bool dense = !sparse && k > 1;
    blur_output_kernel<Ti, To><<<go, 256, 0, st>>>(img, H, W, k, ts > 0 ? ts : 1, TH, TW, use_band ? flags : nullptr,
                                                   dense, sparse, weights, out);
    THULL_LAUNCH_CHECK();
}

template <typename Ti>
static void
run_blur_out(const Ti* img, int64_t B, int64_t H, int64_t W, int k, bool sparse, const void* weights,
             uint8_t* flags, void* out, int32_t out_dtype, hipStream_t st)
{
    switch (out_dtype)
    {
    case kF16:
        run_blur<Ti, __half>(img, B, H, W, k, sparse, static_cast<const __half*>(weights), flags,
                             static_cast<__half*>(out), st);
        break;
    case kF32:
        run_blur<Ti, float>(img, B, H, W, k, sparse, static_cast<const float*>(weights), flags,
                            static_cast<float*>(out), st);
        break;
    case kF64:
        run_blur<Ti, double>(img, B, H, W, k, sparse, static_cast<const double*>(weights), flags,
                             static_cast<double*>(out), st);
        break;
    default:
        throw Error("gaussian_blur: output dtype must be floating point");
    }
}

// `weights` is a device table of k*k weights in the output type.
void
launch_gaussian_blur_impl(const void* images, int32_t in_dtype, int64_t B, int64_t H, int64_t W, int k, bool sparse,
                          const void* weights, void* out, int32_t out_dtype, uint8_t* flags, hipStream_t st)
{
    if (B == 0 || H == 0 || W == 0)
        return;
    switch (in_dtype)
    {
    case kBool:
    case kU8:
        run_blur_out(static_cast<const uint8_t*>(images), B, H, W, k, sparse, weights, flags, out, out_dtype, st);
        break;
    case kI8:
        run_blur_out(static_cast<const int8_t*>(images), B, H, W, k, sparse, weights, flags, out, out_dtype, st);
        break;
    case kI16:
        run_blur_out(static_cast<const int16_t*>(images), B, H, W, k, sparse, weights, flags, out, out_dtype, st);
        break;
    case kI32:
        run_blur_out(static_cast<const int32_t*>(images), B, H, W, k, sparse, weights, flags, out, out_dtype, st);
        break;
    case kI64:
        run_blur_out(static_cast<const int64_t*>(images), B, H, W, k, sparse, weights, flags, out, out_dtype, st);
        break;
    case kF16:
        run_blur_out(static_cast<const __half*>(images), B, H, W, k, sparse, weights, flags, out, out_dtype, st);
        break;
    case kF32:
        run_blur_out(static_cast<const float*>(images), B, H, W, k, sparse, weights, flags, out, out_dtype, st);
        break;
    case kF64:
        run_blur_out(static_cast<const double*>(images), B, H, W, k, sparse, weights, flags, out, out_dtype, st);
        break;
    default:
        throw Error("gaussian_blur: unsupported input dtype");
    }
}

} // namespace thull
