// fbf_scan_mfma.hip -- approximate distance scans on the matrix cores.
//
// The q.x term of |x|^2 - 2 q.x (the GEMM form of the reference's commented-out
// exact_knn, bruteforce.h:129-198) runs as a 32x32 MFMA tile: 32 candidate rows
// (A operand) against 32 queries of the work item (B operand), FP32
// accumulation, with the row norm in FP32.  Each lane of the output tile holds
// one query (column l % 32) and 16 rows, so the threshold test of the top-k'
// selection is one compare per score against that lane's own threshold.
// Candidates that pass go to the query's list; the merge kernel then reranks
// the k' best per query with the exact FP32 distance.
//
// Variants (one template, four row encodings):
//   BF16: rows as bf16, |x|^2 folded into two spare K slots (hi + lo bf16)
//         with 1.0 in the matching query slots: score = acc.
//   FP8:  rows e4m3 with a power-of-two scale per row, queries e4m3 with a
//         power-of-two scale per query; both scales ride in the MFMA's E8M0
//         scale operands; score = acc + |x|^2 (FP32, epilogue).
//   SQ8:  per-dimension uint8 codes (SQ8Quantizer, sq8_quant.hpp:30-63)
//         dequantised in registers to bf16 (codes 0..255 are exact in bf16);
//         the query side carries dif_i / 255; |xhat|^2 rides in the spare K
//         slots as for BF16: score = acc.
//   SYM:  the symmetric int8 codes (SQ8SymmetricQuantizer, sq8_sym_quant.hpp)
//         on the int8 matrix cores, K = 128 (112 code bytes, zero beyond):
//         score = |x|^2 - 2 x.q, integer-exact (|score| < 2^24, so exact in
//         FP32); + |q|^2 gives L2SqrSQ8_sym (distance.hpp:329-366).
//
// MFMA lane layouts (measured on gfx950):
//   v_mfma_f32_32x32x16_bf16: A lane l holds row l%32, k = 8*(l/32) + e;
//     B lane l holds column l%32, same k; D element j of lane l is
//     row 8*(j/4) + 4*(l/32) + j%4, column l%32.
//   v_mfma_scale_f32_32x32x64_f8f6f4 (e4m3): lane l holds row / column l%32,
//     k = 32*(l/32) + e, e = 0..31; same D layout.
//   v_mfma_i32_32x32x32_i8: lane l holds row / column l%32, k = 16*(l/32) + e,
//     e = 0..15; same D layout (int32).
#include <atomic>
#include <utility>

#include "fbf_select.hpp"

namespace fbf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

constexpr int kMTile = 32;     // rows per MFMA tile
constexpr int kBfSteps = 7;    // K = 112 = 7 x 16 (100 dims + 2 norm slots + pad)
constexpr int kF8Steps = 2;    // K = 128 = 2 x 64
constexpr int kI8Steps = 4;    // K = 128 = 4 x 32
constexpr int kSymBytes = 112; // symmetric int8 code bytes per row (utils: 100 dims + 12 zero)

__device__ inline uint16_t bf16_bits(float f) {  // round to nearest even
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_to_f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// ---------------------------------------------------------------- row encodings (dataset build)

// BF16 tiles: block rb of 32 rows, step s, lane l: 8 bf16 = row rb*32 + l%32,
// k = 16 s + 8 (l/32) + e.  k < 100: x_k; k = 100 / 101: hi / lo of |x|^2.
__global__ void build_bf16_kernel(const float *__restrict__ rows, const uint32_t *__restrict__ perm,
                                  const float *__restrict__ norms, uint64_t n, uint64_t n_blocks,
                                  uint4 *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_blocks * kBfSteps * 64) return;
    const uint32_t l = t % 64, s = (t / 64) % kBfSteps;
    const uint64_t rb = t / (64 * kBfSteps);
    const uint64_t p = rb * kMTile + l % 32;
    uint16_t h[8];
    for (int e = 0; e < 8; ++e) {
        const int k = 16 * s + 8 * (l / 32) + e;
        float v = 0.f;
        if (p < n) {
            const uint32_t id = perm[p];
            if (k < kDim) v = rows[(uint64_t)id * kDim + k];
            else if (k == kDim) v = norms[id];
        }
        uint16_t b = bf16_bits(v);
        if (p < n && k == kDim + 1) {
            const float nn = norms[perm[p]];
            b = bf16_bits(nn - bf16_to_f(bf16_bits(nn)));
        }
        h[e] = b;
    }
    out[t] = make_uint4(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16,
                        h[4] | (uint32_t)h[5] << 16, h[6] | (uint32_t)h[7] << 16);
}

// Power-of-two scale for e4m3: |x| / 2^e lands in (128, 256] for the largest
// element.  Returns the exponent e (0 for an all-zero vector).
__device__ inline int fp8_exponent(float absmax) {
    if (!(absmax > 0.f)) return 0;
    int ex;
    frexpf(absmax, &ex);  // absmax = m * 2^ex, m in [0.5, 1)
    return ex - 8;
}

__device__ inline uint32_t pack_fp8x4(float a, float b, float c, float d) {
    int w = 0;
    w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, w, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
    return (uint32_t)w;
}

// FP8 tiles: block rb, step s (K 64), lane l: 32 bytes = row rb*32 + l%32,
// k = 64 s + 32 (l/32) + e; values x_k / 2^ex.  scale[p] = 127 + ex (E8M0).
__global__ void build_fp8_kernel(const float *__restrict__ rows, const uint32_t *__restrict__ perm,
                                 uint64_t n, uint64_t n_blocks, uint4 *__restrict__ out,
                                 uint8_t *__restrict__ scale) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_blocks * kF8Steps * 64) return;
    const uint32_t l = t % 64, s = (t / 64) % kF8Steps;
    const uint64_t rb = t / (64 * kF8Steps);
    const uint64_t p = rb * kMTile + l % 32;
    uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int ex = 0;
    if (p < n) {
        const float *x = rows + (uint64_t)perm[p] * kDim;
        float am = 0.f;
        for (int k = 0; k < kDim; ++k) am = fmaxf(am, fabsf(x[k]));
        ex = fp8_exponent(am);
        float v[32];
        for (int e = 0; e < 32; ++e) {
            const int k = 64 * s + 32 * (l / 32) + e;
            v[e] = k < kDim ? ldexpf(x[k], -ex) : 0.f;
        }
        for (int q = 0; q < 8; ++q) w[q] = pack_fp8x4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    }
    out[2 * t] = make_uint4(w[0], w[1], w[2], w[3]);
    out[2 * t + 1] = make_uint4(w[4], w[5], w[6], w[7]);
    if (s == 0 && l < 32) scale[p] = (uint8_t)(127 + ex);
}

// SQ8 codes in MFMA order: block rb, step s, lane l: 8 bytes = codes of row
// rb*32 + l%32 at k = 16 s + 8 (l/32) + e (zero beyond d).
// Bytes 4..7 of step 6, lanes 0..31 (k = 100..103, beyond the 100 codes)
// carry the bf16 hi / lo split of |xhat|^2; the scan passes them through
// unconverted, so the norm rides in the MFMA like the BF16 variant's.
__global__ void build_sq8_kernel(const uint8_t *__restrict__ codes_orig, const uint32_t *__restrict__ perm,
                                 const float *__restrict__ norms, uint64_t n, uint64_t n_blocks,
                                 uint2 *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_blocks * kBfSteps * 64) return;
    const uint32_t l = t % 64, s = (t / 64) % kBfSteps;
    const uint64_t rb = t / (64 * kBfSteps);
    const uint64_t p = rb * kMTile + l % 32;
    uint32_t lo = 0, hi = 0;
    if (p < n) {
        const uint8_t *c = codes_orig + (uint64_t)perm[p] * kDim;
        for (int e = 0; e < 8; ++e) {
            const int k = 16 * s + 8 * (l / 32) + e;
            const uint32_t v = k < kDim ? c[k] : 0u;
            if (e < 4) lo |= v << (8 * e); else hi |= v << (8 * (e - 4));
        }
        if (s == kBfSteps - 1 && l < 32) {
            const float nn = norms[perm[p]];
            const uint16_t nh = bf16_bits(nn), nl = bf16_bits(nn - bf16_to_f(nh));
            hi = nh | (uint32_t)nl << 16;
        }
    }
    out[t] = make_uint2(lo, hi);
}

// Symmetric int8 tiles: block rb, step s (K 32), lane l: 16 bytes = codes of
// row rb*32 + l%32 (layout order, 112-byte rows) at k = 32 s + 16 (l/32) + e,
// zero beyond the 112 code bytes; xn[p] = |x|^2 of the codes (exact integer).
__global__ void build_sym_kernel(const int8_t *__restrict__ codes, uint64_t n, uint64_t n_blocks,
                                 uint4 *__restrict__ out, float *__restrict__ xn) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_blocks * kI8Steps * 64) return;
    const uint32_t l = t % 64, s = (t / 64) % kI8Steps;
    const uint64_t rb = t / (64 * kI8Steps);
    const uint64_t p = rb * kMTile + l % 32;
    uint32_t w[4] = {0, 0, 0, 0};
    if (p < n) {
        const int8_t *c = codes + p * kSymBytes;
        for (int e = 0; e < 16; ++e) {
            const int k = 32 * s + 16 * (l / 32) + e;
            const uint32_t v = k < kSymBytes ? (uint32_t)(uint8_t)c[k] : 0u;
            w[e / 4] |= v << (8 * (e % 4));
        }
        if (s == 0 && l < 32) {
            int32_t sq = 0;
            for (int k = 0; k < kSymBytes; ++k) sq += (int32_t)c[k] * (int32_t)c[k];
            xn[p] = (float)sq;
        }
    } else if (s == 0 && l < 32) {
        xn[p] = 0.f;
    }
    out[t] = make_uint4(w[0], w[1], w[2], w[3]);
}

// Per-position FP32 values gathered into a layout (row norms).
__global__ void gather_f32_kernel(const float *__restrict__ src, const uint32_t *__restrict__ perm,
                                  uint64_t n, uint64_t n_pad, float *__restrict__ dst) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pad) return;
    dst[p] = p < n ? src[perm[p]] : 0.f;
}

// |x|^2 of each original row (FP32, any order: these feed approximate scores).
__global__ void row_norms_kernel(const float *__restrict__ rows, uint64_t n, float *__restrict__ out) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float4 *x = reinterpret_cast<const float4 *>(rows + r * kDim);
    float s = 0.f;
    for (int g = 0; g < kDim / 4; ++g) {
        const float4 v = x[g];
        s = __builtin_fmaf(v.x, v.x, s);
        s = __builtin_fmaf(v.y, v.y, s);
        s = __builtin_fmaf(v.z, v.z, s);
        s = __builtin_fmaf(v.w, v.w, s);
    }
    out[r] = s;
}

// Per-dimension min / max for SQ8 (SQ8Quantizer::train, sq8_quant.hpp:30-47):
// one thread per dimension, strided over rows (a build step, not the hot path).
__global__ void sq8_minmax_kernel(const float *__restrict__ rows, uint64_t n, float *__restrict__ mi,
                                  float *__restrict__ mx) {
    __shared__ float smi[256], smx[256];
    const int dim = blockIdx.x;
    float a = __builtin_inff(), b = -__builtin_inff();
    for (uint64_t r = threadIdx.x; r < n; r += blockDim.x) {
        const float v = rows[r * kDim + dim];
        a = fminf(a, v);
        b = fmaxf(b, v);
    }
    smi[threadIdx.x] = a;
    smx[threadIdx.x] = b;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smi[threadIdx.x] = fminf(smi[threadIdx.x], smi[threadIdx.x + s]);
            smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        mi[dim] = smi[0];
        mx[dim] = smx[0];
    }
}

// SQ8Quantizer::encode (sq8_quant.hpp:51-63): trunc(clamp((x-mi)/dif, 0, 1)*255);
// a constant dimension (dif = 0) encodes to 0.  Also |xhat|^2 of the
// dequantised row xhat_i = mi_i + (c_i + 0.5) dif_i / 255 (L2SqrSQ8_ext's
// reconstruction, distance.hpp:284-288).
__global__ void sq8_encode_kernel(const float *__restrict__ rows, uint64_t n, const float *__restrict__ mi,
                                  const float *__restrict__ dif, uint8_t *__restrict__ codes,
                                  float *__restrict__ xhat_norm) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float s = 0.f;
    for (int k = 0; k < kDim; ++k) {
        uint8_t c = 0;
        if (dif[k] > 0.f) {
            float v = (rows[r * kDim + k] - mi[k]) / dif[k];
            v = fminf(fmaxf(v, 0.f), 1.f);
            c = (uint8_t)(v * 255.0f);
        }
        codes[r * kDim + k] = c;
        const float xh = mi[k] + ((float)c + 0.5f) * dif[k] * (1.0f / 255.0f);
        s = __builtin_fmaf(xh, xh, s);
    }
    xhat_norm[r] = s;
}

hipError_t launch_build_mfma(int variant, const float *rows, const uint32_t *perm, const float *norms,
                             const uint8_t *sq8_codes, uint64_t n, void *out, uint8_t *scale,
                             hipStream_t st) {
    const uint64_t n_blocks = (n + kMTile - 1) / kMTile;
    // Every input a variant reads must be present (BF16 and SQ8 fold the
    // norms into their fragments; FP8 writes a scale per row).
    if (perm == nullptr || out == nullptr) return hipErrorInvalidValue;
    if (variant == 0 && (rows == nullptr || norms == nullptr)) return hipErrorInvalidValue;
    if (variant == 1 && (rows == nullptr || scale == nullptr)) return hipErrorInvalidValue;
    if (variant == 2 && (sq8_codes == nullptr || norms == nullptr)) return hipErrorInvalidValue;
    if (variant == 3) {
        // sq8_codes: the layout's symmetric int8 codes; norms: out, [n_blocks * 32]
        if (sq8_codes == nullptr || norms == nullptr) return hipErrorInvalidValue;
        const uint64_t total = n_blocks * kI8Steps * 64;
        hipLaunchKernelGGL(build_sym_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st,
                           reinterpret_cast<const int8_t *>(sq8_codes), n, n_blocks, (uint4 *)out,
                           const_cast<float *>(norms));
        return hipGetLastError();
    }
    if (variant == 0) {
        const uint64_t total = n_blocks * kBfSteps * 64;
        hipLaunchKernelGGL(build_bf16_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, rows,
                           perm, norms, n, n_blocks, (uint4 *)out);
    } else if (variant == 1) {
        const uint64_t total = n_blocks * kF8Steps * 64;
        hipLaunchKernelGGL(build_fp8_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, rows,
                           perm, n, n_blocks, (uint4 *)out, scale);
    } else {
        const uint64_t total = n_blocks * kBfSteps * 64;
        hipLaunchKernelGGL(build_sq8_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st,
                           sq8_codes, perm, norms, n, n_blocks, (uint2 *)out);
    }
    return hipGetLastError();
}

__global__ void bf16_err_max_kernel(const float *__restrict__ rows, uint64_t n, uint32_t *__restrict__ out_bits) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float *x = rows + r * kDim;
    float e2 = 0.f, x2 = 0.f;
    for (int d = 0; d < kDim; ++d) {
        const float v = x[d], dv = bf16_round(v) - v;
        e2 = __builtin_fmaf(dv, dv, e2);
        x2 = __builtin_fmaf(v, v, x2);
    }
    const float rho = x2 > 0.f ? __builtin_sqrtf(e2 / x2) : 0.f;
    atomicMax(out_bits, __builtin_bit_cast(uint32_t, rho));
}

hipError_t launch_bf16_err_max(const float *rows, uint64_t n, uint32_t *out_bits, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(bf16_err_max_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, rows, n, out_bits);
    return hipGetLastError();
}

hipError_t launch_row_norms(const float *rows, uint64_t n, float *out, hipStream_t st) {
    hipLaunchKernelGGL(row_norms_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, rows, n, out);
    return hipGetLastError();
}

hipError_t launch_gather_f32(const float *src, const uint32_t *perm, uint64_t n, uint64_t n_pad, float *dst,
                             hipStream_t st) {
    hipLaunchKernelGGL(gather_f32_kernel, dim3((uint32_t)((n_pad + 255) / 256)), dim3(256), 0, st, src, perm, n,
                       n_pad, dst);
    return hipGetLastError();
}

__global__ void sq8_dif_kernel(const float *__restrict__ mi, float *__restrict__ dif) {
    const int k = threadIdx.x;
    if (k < kDim) dif[k] = dif[k] - mi[k];  // dif held the per-dimension max
}

// SQ8 training and encoding on the device: mi, dif (d floats each), codes
// [n][d] in original row order and |xhat|^2 per original row.
hipError_t launch_sq8_train_encode(const float *rows, uint64_t n, float *mi, float *dif, uint8_t *codes,
                                   float *xhat_norm, hipStream_t st) {
    hipLaunchKernelGGL(sq8_minmax_kernel, dim3(kDim), dim3(256), 0, st, rows, n, mi, dif);
    hipLaunchKernelGGL(sq8_dif_kernel, dim3(1), dim3(128), 0, st, (const float *)mi, dif);
    hipLaunchKernelGGL(sq8_encode_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, rows, n,
                       (const float *)mi, (const float *)dif, codes, xhat_norm);
    return hipGetLastError();
}


// ---------------------------------------------------------------- the scan

template <int VAR> struct RowFrag;
template <> struct RowFrag<0> { bf16x8 v[kBfSteps]; };                 // BF16
template <> struct RowFrag<1> { i32x8 v[kF8Steps]; int scale; };      // FP8
template <> struct RowFrag<2> { bf16x8 v[kBfSteps]; };                 // SQ8 (dequantised)
template <> struct RowFrag<3> { i32x4 v[kI8Steps]; };                  // SYM int8
template <int VAR> struct QFrag;
template <> struct QFrag<0> { bf16x8 v[kBfSteps]; };
template <> struct QFrag<1> { i32x8 v[kF8Steps]; int scale; };
template <> struct QFrag<2> { bf16x8 v[kBfSteps]; };
template <> struct QFrag<3> { i32x4 v[kI8Steps]; };

__device__ inline bf16x8 bf16x8_from_bits(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 u = {a, b, c, d};
    return __builtin_bit_cast(bf16x8, u);
}

// bf16 bits of an exact small integer given as float: the top half.
__device__ inline uint32_t pack_bf16_exact(float lo, float hi) {
    return (__builtin_bit_cast(uint32_t, lo) >> 16) | (__builtin_bit_cast(uint32_t, hi) & 0xFFFF0000u);
}

// Fragment bytes of one 32-row tile (contiguous in the layout's fragment
// array): BF16 7 x 64 x 16 B, FP8 2 x 64 x 32 B, SQ8 7 x 64 x 8 B.
template <int VAR> struct TileBytes;
template <> struct TileBytes<0> { static constexpr int value = kBfSteps * 64 * 16; };
template <> struct TileBytes<1> { static constexpr int value = kF8Steps * 64 * 32; };
template <> struct TileBytes<2> { static constexpr int value = kBfSteps * 64 * 8; };
template <> struct TileBytes<3> { static constexpr int value = kI8Steps * 64 * 16; };

// This lane's A fragments of a tile staged in LDS (layout as in HBM).
template <int VAR>
__device__ inline void frag_from_lds(const uint4 *sA, uint32_t lane, RowFrag<VAR> &r) {
    if constexpr (VAR == 0) {
#pragma unroll
        for (int s = 0; s < kBfSteps; ++s) {
            const uint4 u = sA[s * 64 + lane];
            r.v[s] = bf16x8_from_bits(u.x, u.y, u.z, u.w);
        }
    } else if constexpr (VAR == 3) {
#pragma unroll
        for (int s = 0; s < kI8Steps; ++s) {
            const uint4 u = sA[s * 64 + lane];
            r.v[s] = i32x4{(int)u.x, (int)u.y, (int)u.z, (int)u.w};
        }
    } else if constexpr (VAR == 1) {
#pragma unroll
        for (int s = 0; s < kF8Steps; ++s) {
            const uint4 u0 = sA[(s * 64 + lane) * 2], u1 = sA[(s * 64 + lane) * 2 + 1];
            i32x8 w = {(int)u0.x, (int)u0.y, (int)u0.z, (int)u0.w, (int)u1.x, (int)u1.y, (int)u1.z, (int)u1.w};
            r.v[s] = w;
        }
    } else {
        const uint2 *src = reinterpret_cast<const uint2 *>(sA);
#pragma unroll
        for (int s = 0; s < kBfSteps; ++s) {
            const uint2 u = src[s * 64 + lane];
            // uint8 codes -> bf16 (exact): (float)byte compiles to
            // v_cvt_f32_ubyteN; the top half of the float is the bf16.
            const float c0 = (float)(u.x & 0xFFu), c1 = (float)((u.x >> 8) & 0xFFu);
            const float c2 = (float)((u.x >> 16) & 0xFFu), c3 = (float)(u.x >> 24);
            const float c4 = (float)(u.y & 0xFFu), c5 = (float)((u.y >> 8) & 0xFFu);
            const float c6 = (float)((u.y >> 16) & 0xFFu), c7 = (float)(u.y >> 24);
            uint32_t w2 = pack_bf16_exact(c4, c5), w3 = pack_bf16_exact(c6, c7);
            if (s == kBfSteps - 1 && lane < 32) {  // k = 100..103: bf16 norm split, not codes
                w2 = u.y;
                w3 = 0u;
            }
            r.v[s] = bf16x8_from_bits(pack_bf16_exact(c0, c1), pack_bf16_exact(c2, c3), w2, w3);
        }
    }
}

// B-operand fragments of one query, built once per search (kQFragU4 uint4
// per query): half h (lanes 32 h .. 32 h + 31 of the MFMA) holds
//   BF16 / SQ8: u4[7 h + s], s < 7: bf16 of k = 16 s + 8 h + e (e < 8) --
//               -2 q_k (SQ8: -2 q_k dif_k / 255), 1.0 in the two |x|^2 slots
//               k = 100, 101, zero beyond;
//   FP8:        u4[4 h + 2 s + w], s < 2: e4m3 bytes 16 w .. 16 w + 15 of
//               k = 64 s + 32 h + e, the values -q_k / 2^ex; u4[15].x the
//               E8M0 scale 2^(ex + 1) in every byte (restores -2 q);
//   SYM:        u4[4 h + s], s < 4: code bytes k = 32 s + 16 h + e (zero
//               beyond the 112 code bytes).
// One thread per (query, half).
__global__ void qfrag_kernel(int variant, const float *__restrict__ queries, const float *__restrict__ sq8_w,
                             const int8_t *__restrict__ qcodes, uint64_t nq, uint4 *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * nq) return;
    const uint64_t qi = t >> 1;
    const uint32_t half = (uint32_t)(t & 1);
    const float *q = queries + qi * kDim;
    uint4 *f = out + qi * kQFragU4;
    if (variant == 0 || variant == 2) {
        for (int s = 0; s < kBfSteps; ++s) {
            uint32_t w[4];
            for (int e2 = 0; e2 < 4; ++e2) {
                uint16_t h[2];
                for (int u = 0; u < 2; ++u) {
                    const int k = 16 * s + 8 * half + 2 * e2 + u;
                    float v = 0.f;
                    if (k < kDim) v = variant == 0 ? -2.0f * q[k] : -2.0f * q[k] * sq8_w[k];
                    else if (k == kDim || k == kDim + 1) v = 1.0f;  // |x|^2 slots
                    h[u] = bf16_bits(v);
                }
                w[e2] = h[0] | (uint32_t)h[1] << 16;
            }
            f[7 * half + s] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    } else if (variant == 3) {
        const int8_t *c = qcodes + qi * kSymBytes;
        for (int s = 0; s < kI8Steps; ++s) {
            uint32_t w[4] = {0, 0, 0, 0};
            for (int e = 0; e < 16; ++e) {
                const int k = 32 * s + 16 * half + e;
                const uint32_t v = k < kSymBytes ? (uint32_t)(uint8_t)c[k] : 0u;
                w[e / 4] |= v << (8 * (e % 4));
            }
            f[4 * half + s] = make_uint4(w[0], w[1], w[2], w[3]);
        }
    } else {
        float am = 0.f;
        for (int k = 0; k < kDim; ++k) am = fmaxf(am, fabsf(q[k]));
        const int ex = fp8_exponent(am);
        for (int s = 0; s < kF8Steps; ++s) {
            uint32_t w[8];
            for (int g = 0; g < 8; ++g) {
                float v[4];
                for (int u = 0; u < 4; ++u) {
                    const int k = 64 * s + 32 * half + 4 * g + u;
                    v[u] = k < kDim ? -ldexpf(q[k], -ex) : 0.f;
                }
                w[g] = pack_fp8x4(v[0], v[1], v[2], v[3]);
            }
            f[4 * half + 2 * s] = make_uint4(w[0], w[1], w[2], w[3]);
            f[4 * half + 2 * s + 1] = make_uint4(w[4], w[5], w[6], w[7]);
        }
        if (half == 0) f[15] = make_uint4((uint32_t)(127 + ex + 1) * 0x01010101u, 0u, 0u, 0u);
    }
}

hipError_t launch_qfrag(int variant, const float *queries, const float *sq8_w, const int8_t *qcodes, uint64_t nq,
                        uint4 *out, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    if (out == nullptr || (variant == 3 ? qcodes == nullptr : queries == nullptr) || (variant == 2 && sq8_w == nullptr))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(qfrag_kernel, dim3((uint32_t)((2 * nq + 255) / 256)), dim3(256), 0, st, variant, queries,
                       sq8_w, qcodes, nq, out);
    return hipGetLastError();
}

// This lane's B fragments (its column's query qi, its half), or zeros.
template <int VAR>
__device__ inline void load_query(const ScanMfmaArgs &a, bool valid, uint32_t qi, uint32_t lane, QFrag<VAR> &f) {
    const uint32_t half = lane >> 5;
    const uint4 *src = a.qfrag + (uint64_t)qi * kQFragU4;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    if constexpr (VAR == 0 || VAR == 2) {
#pragma unroll
        for (int s = 0; s < kBfSteps; ++s) {
            const uint4 u = valid ? src[7 * half + s] : z;
            f.v[s] = bf16x8_from_bits(u.x, u.y, u.z, u.w);
        }
    } else if constexpr (VAR == 3) {
#pragma unroll
        for (int s = 0; s < kI8Steps; ++s) {
            const uint4 u = valid ? src[4 * half + s] : z;
            f.v[s] = i32x4{(int)u.x, (int)u.y, (int)u.z, (int)u.w};
        }
    } else {
#pragma unroll
        for (int s = 0; s < kF8Steps; ++s) {
            const uint4 u0 = valid ? src[4 * half + 2 * s] : z, u1 = valid ? src[4 * half + 2 * s + 1] : z;
            f.v[s] = i32x8{(int)u0.x, (int)u0.y, (int)u0.z, (int)u0.w, (int)u1.x, (int)u1.y, (int)u1.z, (int)u1.w};
        }
        // An invalid lane's values are zero; its scale only needs to be a
        // valid E8M0 exponent.
        f.scale = valid ? (int)src[15].x : 127 * 0x01010101;
    }
}

template <int VAR>
__device__ inline f32x16 tile_mma(const RowFrag<VAR> &r, const QFrag<VAR> &q) {
    f32x16 acc = {};
    if constexpr (VAR == 3) {
        i32x16 ia = {};
#pragma unroll
        for (int s = 0; s < kI8Steps; ++s) ia = __builtin_amdgcn_mfma_i32_32x32x32_i8(r.v[s], q.v[s], ia, 0, 0, 0);
        // -2 x.q: exact in FP32 (|x.q| <= 112 * 127^2 < 2^21)
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = -2.0f * (float)ia[j];
    } else if constexpr (VAR == 1) {
#pragma unroll
        for (int s = 0; s < kF8Steps; ++s)
            acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(r.v[s], q.v[s], acc, 0, 0, 0, r.scale, 0, q.scale);
    } else {
#pragma unroll
        for (int s = 0; s < kBfSteps; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(r.v[s], q.v[s], acc, 0, 0, 0);
    }
    return acc;
}

// Row tiles reach LDS by direct loads (global_load_lds_dwordx4): each lane
// fetches 16 bytes and the wave writes them as one 1 KiB chunk at a
// wave-uniform LDS address, with no VGPR staging.  A tile is kChunks chunks;
// an LDS tile buffer is padded to whole chunks (the SQ8 tile is 3.5 chunks:
// the pad takes copies of the tile's last piece and is never read).
template <int VAR> struct TileChunks {
    static constexpr int value = (TileBytes<VAR>::value / 16 + 63) / 64;
};

// Issue this wave's share of one tile's chunks: chunk w, w + NW, ...; every
// wave issues the same number of loads (indices clamped: a clamped chunk is
// written twice with the same bytes), so each step's wait count is a constant.
template <int VAR, int NW>
__device__ __attribute__((always_inline)) inline void dma_tile(const uint4 *src, uint4 *dst, uint32_t wave,
                                                               uint32_t lane) {
    constexpr int kU4 = TileBytes<VAR>::value / 16;
    constexpr int kC = TileChunks<VAR>::value;
    constexpr int kPerW = (kC + NW - 1) / NW;
#pragma unroll
    for (int i = 0; i < kPerW; ++i) {
        const uint32_t c = min(wave + (uint32_t)(i * NW), (uint32_t)kC - 1);
        const uint32_t e = min(c * 64 + lane, (uint32_t)kU4 - 1);
        __builtin_amdgcn_global_load_lds(src + e, (__attribute__((address_space(3))) void *)(dst + c * 64), 16, 0, 0);
    }
}

// s_waitcnt vmcnt(N) alone (gfx9 encoding: vmcnt in bits 3:0 and 15:14,
// expcnt and lgkmcnt at their maxima).
template <int N> __device__ __attribute__((always_inline)) inline void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Workgroup barrier for the tile ring.  The direct loads count in vmcnt and
// each wave has waited for its own share of the next tile before arriving, so
// after the barrier every wave may read the whole tile.  No memory fence: a
// fence would also wait for the tiles still in flight.
__device__ __attribute__((always_inline)) inline void ring_barrier() {
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

__device__ inline float tau_as_float(uint64_t tau) {
    return tau == kKeyMax ? __builtin_inff() : ord2f((uint32_t)(tau >> 32));
}

// One workgroup of NW waves per work item.  The item's rows stream through a
// ring of LDS tiles (loaded once, read by every wave); wave w owns the item's
// queries [32 QB w, 32 QB (w + 1)) as QB (two or four) 32-query B blocks in
// VGPRs, its own candidate lists and thresholds, and synchronises with the
// other waves only at the one barrier per pair of tiles that hands the next
// pair over.  Four blocks per wave read each LDS fragment for four MFMA
// chains instead of two, and put 256 queries on two waves instead of four (two
// waves meet at each barrier); they take 256 VGPRs, so the prefix pass, whose
// open thresholds make it append and compact far more, keeps two.
//
// Candidate lists.  In the MFMA output tile a query's 32 rows sit in two lanes
// (l and l + 32, 16 rows each), so each query's list of B keys is split in two
// halves, one per lane: a lane appends to its own half at its own count (a
// register), with no atomics and no cross-lane traffic.  When a half could
// overflow on the next tile, both halves of that query are sorted together and
// cut to the K smallest keys, split evenly between the halves; the threshold
// becomes the K-th key.  At the end of the item each list is made contiguous
// (its second half moved up, or the whole list cut to K).
constexpr uint32_t kShareTiles = 512;  // tiles between threshold exchanges without a compaction

// Bit j (score j of the lane: row r0 + 8 (j / 4) + (j % 4), r0 = tile row +
// 4 * half) set when that row lies in [qs, qe).
__device__ __attribute__((always_inline)) inline uint32_t range_bits(uint32_t t, uint32_t half, uint32_t qs, uint32_t qe) {
    const uint32_t lo = qs <= t ? 0u : min(qs - t, 32u);
    const uint32_t hi = qe <= t ? 0u : min(qe - t, 32u);
    const uint32_t below_hi = hi >= 32 ? 0xFFFFFFFFu : ((1u << hi) - 1u);
    const uint32_t below_lo = lo >= 32 ? 0xFFFFFFFFu : ((1u << lo) - 1u);
    const uint32_t sh = (below_hi & ~below_lo) >> (4 * half);
    return (sh & 0xFu) | ((sh >> 4) & 0xF0u) | ((sh >> 8) & 0xF00u) | ((sh >> 12) & 0xF000u);
}

// Per-lane select: lanes in m take b, the others a.  Written as the
// instruction itself: a select tree over an array, written in C, is turned
// into an indexed load from a stack copy of the array.
__device__ __attribute__((always_inline)) inline float lane_select(float a, float b, uint64_t m) {
    float r;
    __asm__("v_cndmask_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(m));
    return r;
}

// Store the keys of the set bits of pm (score sc[j], row r0 + 8 (j / 4) +
// j % 4) to list[cnt++], lowest bit first.  A lane picks its score with a
// four-level select on the bit index, so the loop runs only as many times as
// the lane with the most set bits needs.
__device__ __attribute__((always_inline)) inline void append_bits(uint32_t pm, const float (&sc)[16], uint32_t r0,
                                                                  uint32_t lay_bit, uint64_t *list, uint32_t &cnt) {
    while (pm) {
        const uint32_t j = (uint32_t)__builtin_ctz(pm);
        pm &= pm - 1;
        const uint64_t m8 = __ballot(j & 8u), m4 = __ballot(j & 4u), m2 = __ballot(j & 2u), m1 = __ballot(j & 1u);
        float v8[8], v4[4], v2[2];
#pragma unroll
        for (int i = 0; i < 8; ++i) v8[i] = lane_select(sc[i], sc[8 + i], m8);
#pragma unroll
        for (int i = 0; i < 4; ++i) v4[i] = lane_select(v8[i], v8[4 + i], m4);
#pragma unroll
        for (int i = 0; i < 2; ++i) v2[i] = lane_select(v4[i], v4[2 + i], m2);
        const float v = lane_select(v2[0], v2[1], m1);
        const uint32_t row = r0 + 8 * (j >> 2) + (j & 3u);
        list[cnt++] = ((uint64_t)f2ord(v) << 32) | lay_bit | row;
    }
}

// Sort the two halves of one split list (counts c0, c1) together and keep the
// K smallest keys, ceil(keep / 2) in the first half and the rest in the second.
template <int B>
__device__ inline void compact_split(uint64_t *list, uint32_t &c0, uint32_t &c1, uint64_t &tau, uint32_t K,
                                     uint64_t *sbuf) {
    constexpr uint32_t H = B / 2;
    const uint32_t lane = threadIdx.x & 63;
    WaveSync::sync();
#pragma unroll 2
    for (int i0 = 0; i0 < B; i0 += 64) {
        const uint32_t i = i0 + lane;
        const bool in = i < H ? i < c0 : i - H < c1;
        sbuf[i] = in ? list[i] : kKeyMax;
    }
    WaveSync::sync();
    wave_sort_lds<B, WaveSync>(sbuf);
    const uint32_t total = c0 + c1;
    const uint32_t keep = total < K ? total : K;
    const uint32_t h0 = (keep + 1) / 2;
#pragma unroll 2
    for (int i0 = 0; i0 < B; i0 += 64) {
        const uint32_t i = i0 + lane;
        if (i < keep) list[i < h0 ? i : H + (i - h0)] = sbuf[i];
    }
    tau = keep == K ? uniform_u64(sbuf[K - 1]) : kKeyMax;
    c0 = h0;
    c1 = keep - h0;
    WaveSync::sync();
}

template <int VAR, int NW, int B, int QB>
__device__ __attribute__((always_inline)) void scan_mfma_item(const ScanMfmaArgs &a, const WorkItem &W, uint4 *sT0,
                                                               uint4 *sT1, uint4 *sT2, uint4 *sT3, uint64_t *sbuf,
                                                               uint64_t *s_tau) {
    static_assert(QB == 2 || QB == 4, "two or four query blocks per wave");
    constexpr int kQB = QB;                       // 32-query blocks per wave
    constexpr uint32_t kWaveQ = 32u * QB;         // queries per wave
    constexpr int kU4 = TileBytes<VAR>::value / 16;
    constexpr uint32_t H = B / 2;                          // keys per list half
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t col = lane & 31, half = lane >> 5;
    const uint32_t nq_all = W.q_count;
    const uint32_t q0 = kWaveQ * wave;                        // this wave's first query of the item
    const uint32_t nq = nq_all > q0 ? (nq_all - q0 < kWaveQ ? nq_all - q0 : kWaveQ) : 0;
    uint64_t *tau_w = s_tau + q0;
    const LayoutMfma &L = a.lay[W.layout];
    const uint4 *gA = reinterpret_cast<const uint4 *>(L.tiles);

    // Per lane and block: the column's query (j = 32 b + col of the wave's
    // queries), its range, its list slot and this lane's half of the list.
    // Lane l holds query l + 64 r in block 2 r + l / 32 (col = l % 32,
    // half = l / 32).
    QFrag<VAR> qf[kQB];
    float tauf[kQB];
    uint32_t qs[kQB], qe[kQB], qib[kQB];
    bool qv[kQB];
    uint32_t cnt[kQB];          // keys in this lane's half of its query's list
    // List slot of the lane's query (fewer than 2^32 slots: each takes at
    // least 1 KiB of device memory); the lane's half starts at key
    // myslot * list_stride + half * H.
    uint32_t myslot[kQB];
#pragma unroll
    for (int b = 0; b < kQB; ++b) {
        const uint32_t j = 32 * b + col;
        qv[b] = j < nq;
        qib[b] = 0;
        qs[b] = qe[b] = 0;
        myslot[b] = 0;
        if (qv[b]) {
            const ItemQuery e = a.iq[W.q_begin + q0 + j];
            qib[b] = e.qi;
            qs[b] = e.s;
            qe[b] = e.e;
            myslot[b] = (uint32_t)list_slot(e, W);
            if (a.qmask && a.qmask[e.qi] == 0) qv[b] = false;
        }
        cnt[b] = 0;
    }
    auto half_list = [&](int b) { return a.lists + (uint64_t)myslot[b] * a.list_stride + half * H; };
    // The queries j = lane (and j = lane + 64 with four blocks): every
    // per-query loop below runs over exactly those j.
    const uint32_t my_qi0 = half ? qib[1] : qib[0];
    const uint32_t my_qi1 = kQB == 4 ? (half ? qib[kQB - 1] : qib[kQB - 2]) : my_qi0;
    auto qi_of = [&](uint32_t j) { return j >= 64 ? my_qi1 : my_qi0; };
    // Base of the list of query j = 32 b + c (held by lane c, half 0).
    auto list_base = [&](int b, uint32_t c) {
        return a.lists + (uint64_t)__builtin_amdgcn_readlane(myslot[b], c) * a.list_stride;
    };

    // Start from the thresholds earlier items (or the prefix pass) published.
    for (uint32_t j = lane; j < nq; j += 64)
        tau_w[j] = a.gtau ? __hip_atomic_load(&a.gtau[qi_of(j)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kKeyMax;
#pragma unroll
    for (int b = 0; b < kQB; ++b) load_query<VAR>(a, qv[b], qib[b], lane, qf[b]);
    WaveSync::sync();
#pragma unroll
    for (int b = 0; b < kQB; ++b) tauf[b] = qv[b] ? tau_as_float(tau_w[32 * b + col]) : -__builtin_inff();
    // Per block: does it hold a query of this item (blk_on), and the tiles
    // [fl, fh) of the item that every one of its queries covers whole -- in
    // those the block needs no per-lane range test (wave-uniform bounds).
    bool blk_on[kQB];
    uint32_t fl[kQB], fh[kQB];
#pragma unroll
    for (int b = 0; b < kQB; ++b) {
        blk_on[b] = __ballot(qv[b]) != 0;
        uint32_t smax = qv[b] ? qs[b] : 0u, emin = qv[b] ? qe[b] : 0xFFFFFFFFu;
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) {
            smax = max(smax, (uint32_t)__shfl_xor((int)smax, off));
            emin = min(emin, (uint32_t)__shfl_xor((int)emin, off));
        }
        smax = __builtin_amdgcn_readfirstlane(smax);
        emin = __builtin_amdgcn_readfirstlane(emin);
        const uint32_t lo = smax <= W.row_begin ? 0u : (smax - W.row_begin + kMTile - 1) / kMTile;
        const uint32_t hi = emin <= W.row_begin ? 0u : (emin - W.row_begin) / kMTile;
        fl[b] = lo;
        fh[b] = hi > lo ? hi : lo;
    }

    // Rows stream through four LDS tile buffers in pairs of tiles: while the
    // matrix cores work on a pair, the direct loads of the next pair are in
    // flight, and the waves synchronise once per pair (a slow step of one
    // wave is absorbed over two tiles).  The buffers are separate LDS
    // variables and the loop body is unrolled four times, so each step names
    // its buffers at compile time and the compiler can tell the buffers being
    // read from the ones being filled (it adds no wait for the loads in flight
    // before the LDS reads).
    const uint32_t ntile = (W.row_end - W.row_begin + kMTile - 1) / kMTile;
    const uint4 *gbase = gA + (uint64_t)(W.row_begin / kMTile) * kU4;
#define FBF_BUF(i) ((i) == 0 ? sT0 : (i) == 1 ? sT1 : (i) == 2 ? sT2 : sT3)
    // Every item has at least one tile (the planner emits no empty items).
    dma_tile<VAR, NW>(gbase, sT0, wave, lane);
    dma_tile<VAR, NW>(gbase + (uint64_t)min(1u, ntile - 1) * kU4, sT1, wave, lane);
    wait_vm<0>();  // tiles 0 and 1
    ring_barrier();

    const uint32_t lay_bit = W.layout << 31;
    uint32_t since_share = 0;

    // Threshold test of one block's scores of one tile (first row t), and
    // the rare append.  One compare per lane: the smallest of its 16 scores
    // against the lane's threshold, or against -inf when the tile lies wholly
    // outside the lane's query range (the block runs for the other queries).
    // Returns whether any lane appended.
    auto test_block = [&](const f32x16 &acc, const float (&xnv)[16], uint32_t t, int b,
                          bool inner_b) __attribute__((always_inline)) -> bool {
        float sc[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) sc[j] = (VAR == 1 || VAR == 3) ? acc[j] + xnv[j] : acc[j];
        float mn = sc[0];
#pragma unroll
        for (int j = 1; j < 16; ++j) mn = fminf(mn, sc[j]);
        const float tb = tauf[b];
        const uint32_t qsb = qs[b], qeb = qe[b];
        const float th = inner_b ? tb : (t < qeb && t + kMTile > qsb) ? tb : -__builtin_inff();
        if (__ballot(mn <= th) == 0) return false;
        // Rare path: a per-lane bit mask of the scores at or below the lane's
        // threshold (inside its query's range on a boundary tile); each lane
        // then stores its set bits to its half of the list, lowest first.  A
        // score equal to the threshold may be stored needlessly; compaction
        // drops it.  Keys carry the layout position (bit 31: layout), not the
        // id; the merge maps positions to ids.
        uint32_t pm = 0;
#pragma unroll
        for (int j = 0; j < 16; ++j) pm |= (sc[j] <= th ? 1u : 0u) << j;
        if (!inner_b) {
            const bool full = t >= qsb && t + kMTile <= qeb;
            if (__ballot(pm != 0 && !full)) pm &= range_bits(t, half, qsb, qeb);
        }
        append_bits(pm, sc, t + 4 * half, lay_bit, half_list(b), cnt[b]);
        return true;
    };
    // The loop body, once per tile buffer with a literal buffer number.
#define FBF_TILE_STEP(U)                                                                              \
    {                                                                                                 \
        const uint32_t m = m0 + (U);                                                                  \
/* Steps past the last tile (the last pass of the loop runs all four    */                            \
/* steps) compute nothing but keep loading, so that every pass issues    */                           \
/* the same loads in the same order and the loop has no exit in between. */                           \
        const bool live = m < ntile;                                                                  \
        const uint32_t t = W.row_begin + m * kMTile;                                                  \
        bool hit_any = false;                                                                         \
        bool inner[kQB];                                                                              \
        uint64_t hit[kQB];                                                                            \
_Pragma("unroll")                                                                                     \
        for (int b = 0; b < kQB; ++b) {                                                               \
/* Interior tiles (every query of the block covers the tile): a scalar  */                            \
/* test; elsewhere the lanes test their own ranges.                      */                           \
            inner[b] = live && blk_on[b] && m >= fl[b] && m < fh[b];                                  \
            hit[b] = inner[b] ? ~0ull : __ballot(live && qv[b] && t < qe[b] && t + kMTile > qs[b]);   \
            hit_any |= hit[b] != 0;                                                                   \
        }                                                                                             \
/* FP8 row scales, FP8 / SYM row norms: loaded before the tile's      */                            \
/* direct loads are issued, so waiting for them does not wait for the   */                            \
/* tile.                                                                 */                           \
        float xn[16];                                                                                 \
        int rscale = 0;                                                                               \
        if constexpr (VAR == 1 || VAR == 3) {                                                         \
            if (hit_any) {                                                                            \
                if constexpr (VAR == 1) rscale = (int)L.scale[t + col] * 0x01010101;                  \
_Pragma("unroll")                                                                                     \
                for (int g = 0; g < 4; ++g) {                                                         \
                    const float4 v = *reinterpret_cast<const float4 *>(L.xnorm + t + 8 * g + 4 * half); \
                    xn[4 * g + 0] = v.x; xn[4 * g + 1] = v.y; xn[4 * g + 2] = v.z; xn[4 * g + 3] = v.w; \
                }                                                                                     \
            }                                                                                         \
        }                                                                                             \
/* First step of a pair: tiles m + 2 and m + 3 into the buffers the   */                            \
/* previous pair left (every wave is past the barrier that ended it).   */                            \
/* Steps past the last tile re-load the last tile, so every pair issues */                            \
/* the same loads.                                                      */                            \
        if constexpr ((U) % 2 == 0) {                                                                 \
            dma_tile<VAR, NW>(gbase + (uint64_t)min(m + 2, ntile - 1) * kU4, FBF_BUF(((U) + 2) % 4), wave, lane); \
            dma_tile<VAR, NW>(gbase + (uint64_t)min(m + 3, ntile - 1) * kU4, FBF_BUF(((U) + 3) % 4), wave, lane); \
        }                                                                                             \
        bool appended = false;                                                                        \
        if (hit_any) {                                                                                \
            RowFrag<VAR> rf;                                                                          \
            frag_from_lds<VAR>(FBF_BUF(U), lane, rf);                                                 \
            if constexpr (VAR == 1) rf.scale = rscale;                                                \
_Pragma("unroll")                                                                                     \
            for (int b = 0; b < kQB; ++b) {                                                           \
                if (hit[b] == 0) continue;                                                            \
                const f32x16 acc = tile_mma<VAR>(rf, qf[b]);                                          \
                appended |= test_block(acc, xn, t, b, inner[b]);                                      \
            }                                                                                         \
        }                                                                                             \
/* Wave-local list maintenance (no cross-wave synchronisation): halves */                             \
/* that the next tile could overflow (16 keys per lane and block) are   */                            \
/* compacted with their other halves; every 64 tiles the thresholds    */                            \
/* are shared with the other items of the same queries.                */                             \
        since_share += live ? 1u : 0u;                                                                \
        if (appended || since_share >= kShareTiles) {                                                 \
            bool changed = false;                                                                     \
_Pragma("unroll")                                                                                     \
            for (int b = 0; b < kQB; ++b) {                                                           \
                const uint32_t cb = cnt[b];                                                           \
                const uint64_t nb = __ballot(cb > H - 16);                                            \
                uint32_t need = (uint32_t)nb | (uint32_t)(nb >> 32);                                  \
                while (need) {                                                                        \
                    const uint32_t c = (uint32_t)__builtin_ctz(need);                                 \
                    need &= need - 1;                                                                 \
                    const uint32_t jj = 32 * b + c;                                                   \
                    uint32_t c0 = __builtin_amdgcn_readlane(cb, c);                                   \
                    uint32_t c1 = __builtin_amdgcn_readlane(cb, 32 + c);                              \
                    uint64_t tau = tau_w[jj];                                                         \
                    compact_split<B>(list_base(b, c), c0, c1, tau, a.K, sbuf);                        \
                    if (col == c) cnt[b] = half ? c1 : c0;                                            \
                    if (lane == 0) tau_w[jj] = tau;                                                   \
                    changed = true;                                                                   \
                }                                                                                     \
            }                                                                                         \
/* Share after a compaction (a threshold moved; the compaction has      */                           \
/* already waited for this wave's memory traffic) and every kShareTiles  */                           \
/* tiles (the fetched half: other items of the same queries may have     */                           \
/* tightened theirs).  Each share waits for its atomics' results, and    */                           \
/* with them for every row load in flight.                               */                           \
            if (changed || since_share >= kShareTiles) {                                              \
                since_share = 0;                                                                      \
                share_thresholds<WaveSync>(a.gtau, nq, tau_w, qi_of);                                 \
_Pragma("unroll")                                                                                     \
                for (int b = 0; b < kQB; ++b)                                                         \
                    if (qv[b]) tauf[b] = tau_as_float(tau_w[32 * b + col]);                           \
            }                                                                                         \
        }                                                                                             \
/* Second step of a pair: this wave's loads of the next pair (issued at */                           \
/* the start of this pair) have landed; the barrier makes those tiles    */                           \
/* visible to every wave and frees this pair's buffers.                  */                           \
        if constexpr ((U) % 2 == 1) {                                                                 \
            wait_vm<0>();                                                                             \
            ring_barrier();                                                                           \
        }                                                                                             \
    }
#pragma unroll 1
    for (uint32_t m0 = 0; m0 < ntile; m0 += 4) {
        FBF_TILE_STEP(0)
        FBF_TILE_STEP(1)
        FBF_TILE_STEP(2)
        FBF_TILE_STEP(3)
    }
#undef FBF_TILE_STEP
#undef FBF_BUF
    // Final cut: lists longer than K are sorted and cut to K (halves of
    // ceil(K / 2) and the rest); every list's count word names its two
    // halves, [0, c0) and [H, H + c1), which the merge reads as one list.
#pragma unroll
    for (int b = 0; b < kQB; ++b) {
        const uint32_t cb = cnt[b];
        const uint32_t other = (uint32_t)__shfl_xor((int)cb, 32);
        const uint64_t over = __ballot(half == 0 && qv[b] && cb + other > a.K);
        uint32_t need = (uint32_t)over;
        while (need) {
            const uint32_t c = (uint32_t)__builtin_ctz(need);
            need &= need - 1;
            const uint32_t jj = 32 * b + c;
            uint32_t c0 = __builtin_amdgcn_readlane(cb, c), c1 = __builtin_amdgcn_readlane(cb, 32 + c);
            uint64_t tau = tau_w[jj];
            compact_split<B>(list_base(b, c), c0, c1, tau, a.K, sbuf);
            if (col == c) cnt[b] = half ? c1 : c0;
            if (lane == 0) tau_w[jj] = tau;
        }
        const uint32_t cf = cnt[b];
        const uint32_t of = (uint32_t)__shfl_xor((int)cf, 32);
        if (half == 0 && qv[b]) a.counts[myslot[b]] = list_count_word(cf, of, H);
    }
    share_thresholds<WaveSync>(a.gtau, nq, tau_w, qi_of);
}

// Masked launches: does any query of the item take part?  (Uniform over the
// workgroup, so a skipped item returns from every wave together.)
template <int NW>
__device__ inline bool item_active(const ScanMfmaArgs &a, const WorkItem &W) {
    if (a.qmask == nullptr) return true;
    int any = 0;
    for (uint32_t j = threadIdx.x; j < W.q_count; j += 64 * NW) any |= a.qmask[a.iq[W.q_begin + j].qi] != 0;
    return __syncthreads_or(any) != 0;
}

// XCD (XCC) the calling wave runs on: hardware register XCC_ID (id 20),
// bits [3:0].
__device__ inline uint32_t xcc_id() {
    return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) & 0xFu;
}

template <int VAR, int NW, int B, int QB>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(2, 8)))
void scan_mfma_kernel(ScanMfmaArgs a) {
    __shared__ uint4 sT0[64 * TileChunks<VAR>::value], sT1[64 * TileChunks<VAR>::value],
        sT2[64 * TileChunks<VAR>::value], sT3[64 * TileChunks<VAR>::value];
    __shared__ uint64_t sbuf[NW][B];
    __shared__ uint64_t s_tau[32 * QB * NW];
    __shared__ uint32_t s_idx;
    uint64_t *my_sbuf = sbuf[threadIdx.x >> 6];
    // One workgroup per item (queue == null: item blockIdx.x), or the
    // persistent form.  Workgroups are dealt to the 8 XCDs round-robin, and
    // each XCD has its own L2: a persistent workgroup drains its XCD's run of
    // items first (the runs are chunk-major, so the workgroups of one XCD
    // stream the same rows for different query groups at about the same time
    // and share them in L2), then helps the other runs.  One call site keeps
    // a single inlined copy of the item code.
    // A masked launch with no query set (the rescan after a verify that
    // failed no seed) has nothing to do.
    if (a.nmask && __builtin_amdgcn_readfirstlane(*a.nmask) == 0) return;
    const uint32_t me = xcc_id();
    for (uint32_t round = 0;; ++round) {
        if (threadIdx.x == 0) {
            uint32_t idx = kPadId;
            if (a.queue == nullptr) {
                idx = round == 0 ? blockIdx.x : kPadId;
            } else {
                const uint32_t *xo = a.xoff_dev ? a.xoff_dev : a.xoff;
                for (uint32_t t = 0; t < 8; ++t) {
                    const uint32_t x = (me + t) & 7u;
                    const uint32_t n = xo[x + 1] - xo[x];
                    if (n == 0) continue;
                    if (__hip_atomic_load(&a.queue[16 * x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= n) continue;
                    const uint32_t i = atomicAdd(&a.queue[16 * x], 1u);
                    if (i < n) {
                        idx = xo[x] + i;
                        break;
                    }
                }
            }
            s_idx = idx;
        }
        __syncthreads();
        const uint32_t idx = __builtin_amdgcn_readfirstlane(s_idx);
        __syncthreads();
        if (idx == kPadId) break;
        const WorkItem W = as_const(a.items)[idx];
        if (item_active<NW>(a, W)) scan_mfma_item<VAR, NW, B, QB>(a, W, sT0, sT1, sT2, sT3, my_sbuf, s_tau);
        __syncthreads();
    }
}

template <int VAR, int NW, int QB>
static hipError_t launch_mfma_nw(const ScanMfmaArgs &a, int B, uint64_t grid, hipStream_t st) {
    switch (B) {
    case 128: hipLaunchKernelGGL((scan_mfma_kernel<VAR, NW, 128, QB>), dim3((uint32_t)grid), dim3(64 * NW), 0, st, a); break;
    case 256: hipLaunchKernelGGL((scan_mfma_kernel<VAR, NW, 256, QB>), dim3((uint32_t)grid), dim3(64 * NW), 0, st, a); break;
    case 512: hipLaunchKernelGGL((scan_mfma_kernel<VAR, NW, 512, QB>), dim3((uint32_t)grid), dim3(64 * NW), 0, st, a); break;
    case 1024: hipLaunchKernelGGL((scan_mfma_kernel<VAR, NW, 1024, QB>), dim3((uint32_t)grid), dim3(64 * NW), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Item widths: nw = 1 and 2 run as nw waves of two query blocks; nw = 4 (256
// queries; prefix items carry at most 128) as two waves of four blocks, except
// BF16 and SQ8 rows with 1024-key lists: there two four-block waves take 46 KB
// of LDS (16 KB of sort buffers), so three workgroups (six waves) fit a CU,
// while four waves of two blocks take 62 KB and two workgroups (eight waves)
// fit.  Measured on C1: certified exact main pass 92.3 ms as four waves, 101.8
// as two; FP8 (k' = 896, also 1024-key lists) 80.7 ms as two, 92.3 as four.
template <int VAR>
constexpr bool four_waves_of_two(int B) {
    return (VAR == 0 || VAR == 2) && B == 1024;
}

template <int VAR>
static hipError_t launch_mfma_var(const ScanMfmaArgs &a, int B, int nw, uint64_t grid, hipStream_t st) {
    if (nw == 1) return launch_mfma_nw<VAR, 1, 2>(a, B, grid, st);
    if (nw == 2) return launch_mfma_nw<VAR, 2, 2>(a, B, grid, st);
    if (four_waves_of_two<VAR>(B)) {
        hipLaunchKernelGGL((scan_mfma_kernel<VAR, 4, 1024, 2>), dim3((uint32_t)grid), dim3(256), 0, st, a);
        return hipGetLastError();
    }
    return launch_mfma_nw<VAR, 2, 4>(a, B, grid, st);
}

template <int VAR, int NW, int QB>
static int occupancy_nw(int B) {
    int n = 0;
    hipError_t e = hipErrorInvalidValue;
    if (B == 128) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, scan_mfma_kernel<VAR, NW, 128, QB>, 64 * NW, 0);
    if (B == 256) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, scan_mfma_kernel<VAR, NW, 256, QB>, 64 * NW, 0);
    if (B == 512) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, scan_mfma_kernel<VAR, NW, 512, QB>, 64 * NW, 0);
    if (B == 1024) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, scan_mfma_kernel<VAR, NW, 1024, QB>, 64 * NW, 0);
    return (e == hipSuccess && n > 0) ? n : 8 / NW;
}

template <int VAR>
static int occupancy_var(int B, int nw) {
    if (nw == 4 && four_waves_of_two<VAR>(B)) {
        int n = 0;
        const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, scan_mfma_kernel<VAR, 4, 1024, 2>, 256, 0);
        return (e == hipSuccess && n > 0) ? n : 2;
    }
    return nw == 1 ? occupancy_nw<VAR, 1, 2>(B) : nw == 2 ? occupancy_nw<VAR, 2, 2>(B) : occupancy_nw<VAR, 2, 4>(B);
}

// Resident workgroups per CU of a scan instantiation (cached per variant, B, nw).
static int persistent_per_cu(int variant, int B, int nw) {
    static std::atomic<int> cache[4][4][3] = {};
    const int bi = B == 128 ? 0 : B == 256 ? 1 : B == 512 ? 2 : 3, wi = nw == 1 ? 0 : nw == 2 ? 1 : 2;
    std::atomic<int> &c = cache[variant][bi][wi];
    int v = c.load(std::memory_order_relaxed);
    if (v == 0) {
        v = variant == 0   ? occupancy_var<0>(B, nw)
            : variant == 1 ? occupancy_var<1>(B, nw)
            : variant == 2 ? occupancy_var<2>(B, nw)
                           : occupancy_var<3>(B, nw);
        c.store(v, std::memory_order_relaxed);
    }
    return v;
}

void warm_scan_mfma(int variant, int B, int nw) {
    if ((variant >= 0 && variant <= 3) && (B == 128 || B == 256 || B == 512 || B == 1024) && (nw == 1 || nw == 2 || nw == 4))
        (void)persistent_per_cu(variant, B, nw);
}

// nw: item width (1, 2 or 4); an item carries at most 64 * nw queries.
hipError_t launch_scan_mfma(int variant, const ScanMfmaArgs &a, int B, int nw, hipStream_t st) {
    if (a.n_items == 0) return hipSuccess;
    if (nw != 1 && nw != 2 && nw != 4) return hipErrorInvalidValue;
    if ((uint32_t)B > a.list_stride) return hipErrorInvalidValue;
    // After a compaction a list half holds ceil(K / 2) keys and must take one
    // more tile of appends (16 per lane) before the next check.
    if ((a.K + 1) / 2 + 16 > (uint32_t)B / 2) return hipErrorInvalidValue;
    uint64_t grid = a.n_items;
    if (a.queue) {
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        // As many workgroups as fit on the device at once (a multiple of the 8
        // XCDs, so each XCD holds the same number).
        const int per_cu = persistent_per_cu(variant, B, nw);
        grid = std::min<uint64_t>((uint64_t)cus * per_cu, a.n_items);
        if (grid > 8) grid = grid / 8 * 8;
    }
    if (variant == 0) return launch_mfma_var<0>(a, B, nw, grid, st);
    if (variant == 1) return launch_mfma_var<1>(a, B, nw, grid, st);
    if (variant == 2) return launch_mfma_var<2>(a, B, nw, grid, st);
    return launch_mfma_var<3>(a, B, nw, grid, st);
}

}  // namespace fbf
