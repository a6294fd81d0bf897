// radiance.hip -- the radiance payload of a FRAME / PREVIEW message on the
// device: the byte shuffle a compressed encoding starts from, its inverse,
// and the receiver's guarded add of a payload into the running XYZ sums.
//
// Reference: indicatrix-net/src/radiance/{mod.rs, shuffle.rs, payload.rs}
// and client/accumulate/mod.rs (apply_frame).  A payload is width*height
// pixels of three little-endian f32 (12 B); shuffled, its 4n bytes become 4
// planes of n bytes (byte k of every value in plane k).  All three kernels
// are pure byte movement plus one f32 add per valid component: HBM-bound,
// 16-byte vector loads and stores on the aligned body, a byte-wise tail.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <stdio.h>
#include <stdarg.h>
#include "../../include/ixg.h"

extern "C" hipStream_t ixg_internal_stream(ixg_ctx* ctx);
extern "C" int ixg_internal_device(ixg_ctx* ctx);
extern "C" int ixg_internal_lost(ixg_ctx* ctx);
extern "C" ixg_status ixg_internal_fail(ixg_ctx* ctx, ixg_status st, const char* what);
extern "C" ixg_status ixg_internal_hip_fail(ixg_ctx* ctx, hipError_t e, const char* where);

namespace ixg {

constexpr int kRadBlock = 256;

// Plane k of 4 consecutive values: byte k of each, low byte first.
__device__ __forceinline__ uint32_t plane_word(uint4 v, int k) {
    const int s = 8 * k;
    return ((v.x >> s) & 0xffu) | (((v.y >> s) & 0xffu) << 8) | (((v.z >> s) & 0xffu) << 16) |
           (((v.w >> s) & 0xffu) << 24);
}

// Value j (0..3) of 4 consecutive values from the 4 plane words.
__device__ __forceinline__ uint32_t value_of(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3, int j) {
    const int s = 8 * j;
    return ((p0 >> s) & 0xffu) | (((p1 >> s) & 0xffu) << 8) | (((p2 >> s) & 0xffu) << 16) |
           (((p3 >> s) & 0xffu) << 24);
}

// shuffle_into (shuffle.rs:35): n values; the vector body covers groups of 4
// values when n % 4 == 0 (every plane then starts 4-byte aligned), the byte
// loop the rest.
__global__ void __launch_bounds__(kRadBlock) shuffle_kernel(const uint8_t* __restrict__ src, uint64_t n,
                                                            uint8_t* __restrict__ dst, uint64_t vec_groups) {
    const uint64_t stride = (uint64_t)gridDim.x * kRadBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kRadBlock + threadIdx.x; g < vec_groups; g += stride) {
        const uint4 v = reinterpret_cast<const uint4*>(src)[g];
        for (int k = 0; k < 4; ++k) reinterpret_cast<uint32_t*>(dst + k * n)[g] = plane_word(v, k);
    }
    for (uint64_t i = vec_groups * 4 + (uint64_t)blockIdx.x * kRadBlock + threadIdx.x; i < n; i += stride)
        for (int k = 0; k < 4; ++k) dst[k * n + i] = src[4 * i + k];
}

// unshuffle_into (shuffle.rs:60).
__global__ void __launch_bounds__(kRadBlock) unshuffle_kernel(const uint8_t* __restrict__ src, uint64_t n,
                                                              uint8_t* __restrict__ dst, uint64_t vec_groups) {
    const uint64_t stride = (uint64_t)gridDim.x * kRadBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kRadBlock + threadIdx.x; g < vec_groups; g += stride) {
        const uint32_t p0 = reinterpret_cast<const uint32_t*>(src)[g];
        const uint32_t p1 = reinterpret_cast<const uint32_t*>(src + n)[g];
        const uint32_t p2 = reinterpret_cast<const uint32_t*>(src + 2 * n)[g];
        const uint32_t p3 = reinterpret_cast<const uint32_t*>(src + 3 * n)[g];
        reinterpret_cast<uint4*>(dst)[g] = make_uint4(value_of(p0, p1, p2, p3, 0), value_of(p0, p1, p2, p3, 1),
                                                      value_of(p0, p1, p2, p3, 2), value_of(p0, p1, p2, p3, 3));
    }
    for (uint64_t i = vec_groups * 4 + (uint64_t)blockIdx.x * kRadBlock + threadIdx.x; i < n; i += stride)
        for (int k = 0; k < 4; ++k) dst[4 * i + k] = src[k * n + i];
}

// is_valid_sample (shuffle.rs:25): every component finite and not negative
// (-0.0 passes).
__device__ __forceinline__ bool valid_sample(float x, float y, float z) {
    return __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z) && x >= 0.0f && y >= 0.0f &&
           z >= 0.0f;
}

__device__ __forceinline__ uint32_t load_value(const uint8_t* p, uint64_t n, int shuffled, uint64_t i) {
    if (shuffled)
        return (uint32_t)p[i] | ((uint32_t)p[n + i] << 8) | ((uint32_t)p[2 * n + i] << 16) |
               ((uint32_t)p[3 * n + i] << 24);
    return (uint32_t)p[4 * i] | ((uint32_t)p[4 * i + 1] << 8) | ((uint32_t)p[4 * i + 2] << 16) |
           ((uint32_t)p[4 * i + 3] << 24);
}

// The summing arms of PayloadDecoder::decode_and_add (payload.rs:225-250):
// Raw adds the values as they are, Planes (shuffled) through
// add_unshuffled_valid (shuffle.rs:122).  A pixel failing is_valid_sample
// adds nothing and is counted; a valid one is three IEEE f32 adds.  The body
// takes 4 pixels (12 values, 48 B of the running sum) per thread when the
// pixel count is a multiple of 4 and the pointers 16-byte aligned.
template <bool SHUFFLED>
__global__ void __launch_bounds__(kRadBlock) add_kernel(const uint8_t* __restrict__ payload, uint64_t pixels,
                                                        float* __restrict__ acc, uint32_t* __restrict__ dropped,
                                                        uint64_t vec_quads) {
    const uint64_t n = pixels * 3;
    const uint64_t stride = (uint64_t)gridDim.x * kRadBlock;
    uint32_t bad = 0;
    for (uint64_t q = (uint64_t)blockIdx.x * kRadBlock + threadIdx.x; q < vec_quads; q += stride) {
        uint32_t d[12];
        if (SHUFFLED) {
            uint32_t p[4][3];
            for (int k = 0; k < 4; ++k)
                for (int w = 0; w < 3; ++w) p[k][w] = reinterpret_cast<const uint32_t*>(payload + k * n)[3 * q + w];
            for (int w = 0; w < 3; ++w)
                for (int j = 0; j < 4; ++j) d[4 * w + j] = value_of(p[0][w], p[1][w], p[2][w], p[3][w], j);
        } else {
            for (int w = 0; w < 3; ++w) {
                const uint4 v = reinterpret_cast<const uint4*>(payload)[3 * q + w];
                d[4 * w] = v.x; d[4 * w + 1] = v.y; d[4 * w + 2] = v.z; d[4 * w + 3] = v.w;
            }
        }
        float4* a4 = reinterpret_cast<float4*>(acc) + 3 * q;
        float a[12];
        for (int w = 0; w < 3; ++w) {
            const float4 v = a4[w];
            a[4 * w] = v.x; a[4 * w + 1] = v.y; a[4 * w + 2] = v.z; a[4 * w + 3] = v.w;
        }
        for (int px = 0; px < 4; ++px) {
            const float x = __uint_as_float(d[3 * px]), y = __uint_as_float(d[3 * px + 1]),
                        z = __uint_as_float(d[3 * px + 2]);
            if (valid_sample(x, y, z)) {
                a[3 * px] = __fadd_rn(a[3 * px], x);
                a[3 * px + 1] = __fadd_rn(a[3 * px + 1], y);
                a[3 * px + 2] = __fadd_rn(a[3 * px + 2], z);
            } else {
                ++bad;
            }
        }
        for (int w = 0; w < 3; ++w) a4[w] = make_float4(a[4 * w], a[4 * w + 1], a[4 * w + 2], a[4 * w + 3]);
    }
    for (uint64_t px = vec_quads * 4 + (uint64_t)blockIdx.x * kRadBlock + threadIdx.x; px < pixels; px += stride) {
        const float x = __uint_as_float(load_value(payload, n, SHUFFLED, 3 * px));
        const float y = __uint_as_float(load_value(payload, n, SHUFFLED, 3 * px + 1));
        const float z = __uint_as_float(load_value(payload, n, SHUFFLED, 3 * px + 2));
        if (valid_sample(x, y, z)) {
            acc[3 * px] = __fadd_rn(acc[3 * px], x);
            acc[3 * px + 1] = __fadd_rn(acc[3 * px + 1], y);
            acc[3 * px + 2] = __fadd_rn(acc[3 * px + 2], z);
        } else {
            ++bad;
        }
    }
    // One atomic per wave: the 64 lanes' counts summed across the wave.
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor(bad, off);
    if ((threadIdx.x & 63) == 0 && bad) atomicAdd(dropped, bad);
}

__global__ void zero_u32_kernel(uint32_t* p) { *p = 0u; }

static uint32_t grid_for(uint64_t items) {
    // 8 resident blocks of 256 per CU over 256 CUs; grid-stride beyond that.
    const uint64_t want = (items + kRadBlock - 1) / kRadBlock;
    const uint64_t cap = 2048;
    return (uint32_t)(want == 0 ? 1 : (want < cap ? want : cap));
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace ixg

using namespace ixg;

// A HIP error marks the context lost and records where, as every other
// entry point of the library does; the caller recovers with ixg_try_recover.
static ixg_status finish(ixg_ctx* ctx, const char* where, hipStream_t s) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? IXG_OK : ixg_internal_hip_fail(ctx, e, where);
}

static ixg_status rad_stream(ixg_ctx* ctx, hipStream_t* s) {
    if (!ctx) return IXG_INVALID;
    if (ixg_internal_lost(ctx))
        return ixg_internal_fail(ctx, IXG_DEVICE_LOST, "context is lost; call ixg_try_recover");
    const hipError_t e = hipSetDevice(ixg_internal_device(ctx));
    if (e != hipSuccess) return ixg_internal_hip_fail(ctx, e, "hipSetDevice");
    *s = ixg_internal_stream(ctx);
    return IXG_OK;
}

static ixg_status shuffle_common(ixg_ctx* ctx, const void* d_src, uint64_t n_bytes, void* d_dst, bool forward) {
    if (n_bytes % 4 != 0 || n_bytes > IXG_MAX_FRAME_LEN) return IXG_INVALID;
    if (n_bytes == 0) return ctx ? IXG_OK : IXG_INVALID;
    if (!d_src || !d_dst || d_src == d_dst) return IXG_INVALID;
    hipStream_t s;
    ixg_status st = rad_stream(ctx, &s);
    if (st != IXG_OK) return st;
    const uint64_t n = n_bytes / 4;
    const uint64_t groups = (n % 4 == 0 && aligned16(d_src) && aligned16(d_dst)) ? n / 4 : 0;
    const uint32_t grid = grid_for(groups ? groups : n);
    if (forward)
        shuffle_kernel<<<grid, kRadBlock, 0, s>>>((const uint8_t*)d_src, n, (uint8_t*)d_dst, groups);
    else
        unshuffle_kernel<<<grid, kRadBlock, 0, s>>>((const uint8_t*)d_src, n, (uint8_t*)d_dst, groups);
    return finish(ctx, forward ? "radiance shuffle" : "radiance unshuffle", s);
}

static void set_message(ixg_radiance_error* e, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static void set_message(ixg_radiance_error* e, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(e->message, sizeof e->message, fmt, ap);
    va_end(ap);
}

// PayloadEncoding's Debug text (messages/encoding.rs:43-56), which the
// reference's Display of UnsupportedEncoding prints: the zstd arm carries
// its level.
static void encoding_debug(int32_t enc, char* buf, size_t n) {
    switch (enc & 0xFF) {
    case IXG_PAYLOAD_RAW: snprintf(buf, n, "Raw"); break;
    case IXG_PAYLOAD_SHUFFLE_ZSTD: snprintf(buf, n, "ShuffleZstd { level: %d }", (int)IXG_PAYLOAD_ZSTD_LEVEL_OF(enc)); break;
    case IXG_PAYLOAD_SHUFFLE_LZ4: snprintf(buf, n, "ShuffleLz4"); break;
    default: snprintf(buf, n, "Unknown"); break;
    }
}

extern "C" {

ixg_status ixg_radiance_shuffle_device(ixg_ctx* ctx, const void* d_src, uint64_t n_bytes, void* d_dst) {
    return shuffle_common(ctx, d_src, n_bytes, d_dst, true);
}

ixg_status ixg_radiance_unshuffle_device(ixg_ctx* ctx, const void* d_src, uint64_t n_bytes, void* d_dst) {
    return shuffle_common(ctx, d_src, n_bytes, d_dst, false);
}

ixg_status ixg_radiance_add_device(ixg_ctx* ctx, const void* d_payload, int32_t shuffled, uint64_t pixels,
                                   float* d_acc, uint32_t* d_dropped) {
    if (pixels > IXG_MAX_FRAME_LEN / 12 || !d_dropped || (shuffled != 0 && shuffled != 1)) return IXG_INVALID;
    if (pixels && (!d_payload || !d_acc)) return IXG_INVALID;
    if (!shuffled && ((uintptr_t)d_payload & 3u)) return IXG_INVALID;  // RadianceError::Misaligned
    if (((uintptr_t)d_acc & 3u)) return IXG_INVALID;
    hipStream_t s;
    ixg_status st = rad_stream(ctx, &s);
    if (st != IXG_OK) return st;
    zero_u32_kernel<<<1, 1, 0, s>>>(d_dropped);
    if (pixels) {
        const uint64_t quads = (pixels % 4 == 0 && aligned16(d_payload) && aligned16(d_acc)) ? pixels / 4 : 0;
        const uint32_t grid = grid_for(quads ? quads : pixels);
        if (shuffled)
            add_kernel<true><<<grid, kRadBlock, 0, s>>>((const uint8_t*)d_payload, pixels, d_acc, d_dropped, quads);
        else
            add_kernel<false><<<grid, kRadBlock, 0, s>>>((const uint8_t*)d_payload, pixels, d_acc, d_dropped, quads);
    }
    return finish(ctx, "radiance add", s);
}

ixg_status ixg_radiance_check_payload(int32_t encoding, uint32_t raw_len, uint64_t wire_len, uint32_t width,
                                      uint32_t height, int32_t compression, ixg_radiance_error* err) {
    if (!err) return IXG_INVALID;
    memset(err, 0, sizeof *err);
    err->width = width;
    err->height = height;
    err->encoding = encoding;
    const int32_t family = encoding & 0xFF;
    // the level rides in bits 8-15 of the zstd arm only (a u8, encoding.rs:51)
    if (encoding < 0 || family > IXG_PAYLOAD_SHUFFLE_LZ4 || (encoding >> 16) != 0 ||
        (family != IXG_PAYLOAD_SHUFFLE_ZSTD && (encoding >> 8) != 0))
        return IXG_INVALID;
    const uint64_t expected = (uint64_t)width * (uint64_t)height * 12u;
    if (expected > IXG_MAX_FRAME_LEN) {
        err->kind = IXG_RAD_TOO_LARGE;
        set_message(err, "a %ux%u radiance payload would exceed the %u byte frame cap", width, height,
                    IXG_MAX_FRAME_LEN);
        return IXG_INVALID;
    }
    err->expected_bytes = expected;
    err->raw_len = raw_len;
    if ((uint64_t)raw_len != expected) {
        err->kind = IXG_RAD_RAW_LEN_MISMATCH;
        set_message(err, "header raw_len %u disagrees with %ux%u = %llu bytes", raw_len, width, height,
                    (unsigned long long)expected);
        return IXG_INVALID;
    }
    if (family == IXG_PAYLOAD_RAW) {
        if (wire_len != expected) {
            err->kind = IXG_RAD_LENGTH_MISMATCH;
            err->got_bytes = wire_len;
            set_message(err, "radiance payload for %ux%u must be %llu bytes, got %llu", width, height,
                        (unsigned long long)expected, (unsigned long long)wire_len);
            return IXG_INVALID;
        }
        return IXG_OK;
    }
    if (!compression) {
        err->kind = IXG_RAD_UNSUPPORTED_ENCODING;
        char name[48];
        encoding_debug(encoding, name, sizeof name);
        set_message(err, "payload encoding %s is not supported by this build", name);
        return IXG_INVALID;
    }
    return IXG_OK;
}

ixg_status ixg_radiance_check_frame(uint32_t first_sample, uint32_t samples, int32_t has_range, uint32_t start,
                                    uint32_t end, uint32_t samples_done, ixg_radiance_error* err) {
    if (!err) return IXG_INVALID;
    memset(err, 0, sizeof *err);
    err->first_sample = first_sample;
    err->samples = samples;
    err->start = start;
    err->end = end;
    if (samples == 0) {
        err->kind = IXG_RAD_ZERO_SAMPLE_FRAME;
        set_message(err, "FRAME claims zero samples but carries a payload");
        return IXG_INVALID;
    }
    if (!has_range) return IXG_OK;
    const uint64_t frame_end = (uint64_t)first_sample + samples;
    if (frame_end > 0xffffffffull || first_sample < start || frame_end > end) {
        err->kind = IXG_RAD_FRAME_OUTSIDE_RANGE;
        set_message(err, "FRAME [%u, +%u) lies outside the requested sample range [%u, %u)", first_sample, samples,
                    start, end);
        return IXG_INVALID;
    }
    const uint32_t range_len = end - start;
    const uint64_t sum = (uint64_t)samples_done + samples;
    const uint32_t cumulative = sum > 0xffffffffull ? 0xffffffffu : (uint32_t)sum;
    if (cumulative > range_len) {
        err->kind = IXG_RAD_SAMPLE_COUNT_OVERRUN;
        err->cumulative = cumulative;
        err->range_len = range_len;
        set_message(err,
                    "FRAME would bring the cumulative sample count to %u, over the requested range's width %u",
                    cumulative, range_len);
        return IXG_INVALID;
    }
    return IXG_OK;
}

}  // extern "C"
