// Internal helpers of libcuradon_mi: status/error plumbing, storage-type conversions and
// raw buffer-resource loads for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "curadon_mi.h"

namespace cmi {

// ------------------------------------------------------------------ errors -------------

// Records a message for cmi_last_error() on this host thread and returns `s`.
cmi_status fail(cmi_status s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define CMI_HIP_TRY(expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (void)hipGetLastError();                                                        \
            return ::cmi::fail(e_ == hipErrorOutOfMemory ? CMI_ERR_OUT_OF_MEMORY            \
                                                         : CMI_ERR_HIP,                     \
                               "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),       \
                               __FILE__, __LINE__);                                         \
        }                                                                                   \
    } while (0)

#define CMI_TRY(expr)                                                                       \
    do {                                                                                    \
        cmi_status s_ = (expr);                                                             \
        if (s_ != CMI_OK) return s_;                                                        \
    } while (0)

inline bool dtype_ok(int dt) { return dt == CMI_F32 || dt == CMI_F16 || dt == CMI_BF16; }

inline const char *dtype_name(int dt) {
    switch (dt) {
    case CMI_F32: return "f32";
    case CMI_F16: return "f16";
    case CMI_BF16: return "bf16";
    default: return "unknown";
    }
}

inline size_t dtype_size(int dt) { return dt == CMI_F32 ? 4 : 2; }

// Checks that `p` is device memory usable on `device` (or managed / host-mapped memory that
// the device can address). Returns CMI_ERR_DEVICE with a message naming `what` otherwise.
cmi_status check_device_ptr(const void *p, int device, const char *what);

// Makes `device` current for the duration of a call and restores the previous device.
struct DeviceGuard {
    int prev = -1;
    bool changed = false;
    cmi_status enter(int device);
    ~DeviceGuard();
};

// The 2D forward in parts (fan2d.hip), shared by cmi_forward_2d (stream-ordered scratch) and
// cmi_forward_2d_noalloc (the library's static scratch): host validation; the scratch bytes of a
// call; staging and march into caller-provided scratch, with the plan's device current.
cmi_status fan2d_forward_check(const cmi_plan2d *plan, const void *volume, int vol_dtype,
                               const void *sinogram, int sino_dtype, uint64_t batch);
size_t fan2d_forward_scratch_bytes(const cmi_plan2d *plan, int vol_dtype, uint64_t batch);
int fan2d_plan_device(const cmi_plan2d *plan);
cmi_status fan2d_forward_enqueue(const cmi_plan2d *plan, const void *volume, int vol_dtype,
                                 void *sinogram, int sino_dtype, uint64_t batch, void *scratch,
                                 hipStream_t s);

// ------------------------------------------------------------------ storage types ------

// Storage tags. Arithmetic is always FP32; these only say how an element is stored.
struct F32 {
    using raw = float;
};
struct F16 {
    using raw = _Float16;
};
struct BF16 {
    using raw = __bf16;
};

template <class S>
__device__ __forceinline__ float to_f32(typename S::raw v) {
    return static_cast<float>(v);
}

// Round-to-nearest-even conversion for stores (v_cvt_f16_f32 / v_cvt_pk_bf16_f32 on gfx950).
template <class S>
__device__ __forceinline__ typename S::raw from_f32(float v) {
    return static_cast<typename S::raw>(v);
}

// floor(x) as int32 in one instruction (v_cvt_flr_i32_f32) for finite |x| < 2^31; the compiler
// emits a floor and a conversion for (int)floorf(x).
__device__ __forceinline__ int floor_i32(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// floor(x + 0.5) as int32 in one instruction (v_cvt_rpi_i32_f32) for finite |x| < 2^31.
__device__ __forceinline__ int round_i32(float x) {
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// ------------------------------------------------------------------ buffer loads -------
// Raw buffer resources give 32-bit offsets from a scalar base and hardware range checking:
// a load that lies wholly beyond num_records bytes returns 0 instead of touching memory.
// The kernels use that only as a memory-safety net and keep every load they rely on fully
// inside the resource: a multi-dword load that straddles the limit (the compiler merges
// adjacent dword loads into one) comes back as zeros in all of its dwords.

using rsrc_t = __amdgpu_buffer_rsrc_t;

// Dword 3 of a raw (untyped) buffer descriptor on CDNA: DATA_FORMAT = 32, no swizzle.
constexpr int kBufferFlags = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes,
                                             kBufferFlags);
}

template <class S>
__device__ __forceinline__ float buf_load(rsrc_t r, int elem_offset);

template <>
__device__ __forceinline__ float buf_load<F32>(rsrc_t r, int elem_offset) {
    return __builtin_bit_cast(float,
                              __builtin_amdgcn_raw_buffer_load_b32(r, elem_offset * 4, 0, 0));
}
template <>
__device__ __forceinline__ float buf_load<F16>(rsrc_t r, int elem_offset) {
    const unsigned short u = __builtin_amdgcn_raw_buffer_load_b16(r, elem_offset * 2, 0, 0);
    return static_cast<float>(__builtin_bit_cast(_Float16, u));
}
template <>
__device__ __forceinline__ float buf_load<BF16>(rsrc_t r, int elem_offset) {
    const unsigned short u = __builtin_amdgcn_raw_buffer_load_b16(r, elem_offset * 2, 0, 0);
    return __builtin_bit_cast(float, (uint32_t)u << 16);
}

// Two adjacent elements (offsets o and o + 1) in FP32, with one load: 8 bytes for F32 pairs,
// 4 bytes for 16-bit pairs. The load need not be aligned to its size (an F32 pair sits at any
// dword, a 16-bit pair at any 2-byte boundary): the ROCm driver runs gfx9-family shaders in
// unaligned memory-access mode. `soff` is a wave-uniform byte offset.
struct Pair {
    float lo, hi;
};

template <class S>
__device__ __forceinline__ Pair buf_pair(rsrc_t r, int elem_offset, int soff_bytes);

template <>
__device__ __forceinline__ Pair buf_pair<F32>(rsrc_t r, int elem_offset, int soff_bytes) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, elem_offset * 4, soff_bytes, 0);
    const unsigned a = v[0], b = v[1];
    return {__uint_as_float(a), __uint_as_float(b)};
}
template <>
__device__ __forceinline__ Pair buf_pair<F16>(rsrc_t r, int elem_offset, int soff_bytes) {
    const unsigned v = __builtin_amdgcn_raw_buffer_load_b32(r, elem_offset * 2, soff_bytes, 0);
    return {static_cast<float>(__builtin_bit_cast(_Float16, (unsigned short)(v & 0xffffu))),
            static_cast<float>(__builtin_bit_cast(_Float16, (unsigned short)(v >> 16)))};
}
template <>
__device__ __forceinline__ Pair buf_pair<BF16>(rsrc_t r, int elem_offset, int soff_bytes) {
    const unsigned v = __builtin_amdgcn_raw_buffer_load_b32(r, elem_offset * 2, soff_bytes, 0);
    return {__uint_as_float(v << 16), __uint_as_float(v & 0xffff0000u)};
}

// Four consecutive elements (offsets o .. o + 3) in FP32 with one load (16 bytes for F32,
// 8 for 16-bit storage); any element alignment.
struct Quad {
    float v[4];
};

// buf_quad_at takes the byte offset of the first element.
template <class S>
__device__ __forceinline__ Quad buf_quad_at(rsrc_t r, int byte_offset);

template <>
__device__ __forceinline__ Quad buf_quad_at<F32>(rsrc_t r, int byte_offset) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_offset, 0, 0);
    const unsigned a = v[0], b = v[1], c = v[2], d = v[3];
    return {{__uint_as_float(a), __uint_as_float(b), __uint_as_float(c), __uint_as_float(d)}};
}
template <>
__device__ __forceinline__ Quad buf_quad_at<F16>(rsrc_t r, int byte_offset) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_offset, 0, 0);
    const unsigned a = v[0], b = v[1];
    auto h = [](unsigned short x) { return static_cast<float>(__builtin_bit_cast(_Float16, x)); };
    return {{h((unsigned short)(a & 0xffffu)), h((unsigned short)(a >> 16)),
             h((unsigned short)(b & 0xffffu)), h((unsigned short)(b >> 16))}};
}
template <>
__device__ __forceinline__ Quad buf_quad_at<BF16>(rsrc_t r, int byte_offset) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_offset, 0, 0);
    const unsigned a = v[0], b = v[1];
    return {{__uint_as_float(a << 16), __uint_as_float(a & 0xffff0000u), __uint_as_float(b << 16),
             __uint_as_float(b & 0xffff0000u)}};
}

template <class S>
__device__ __forceinline__ Quad buf_quad(rsrc_t r, int elem_offset) {
    return buf_quad_at<S>(r, elem_offset * (int)sizeof(typename S::raw));
}

// Byte size of n elements, for the range check of a resource.
template <class S>
constexpr uint32_t bytes_of(uint64_t n) {
    return (uint32_t)(n * sizeof(typename S::raw));
}

} // namespace cmi
