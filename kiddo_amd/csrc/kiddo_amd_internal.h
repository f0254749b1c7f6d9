// kiddo_amd_internal.h — shared host-side definitions of the engine (not part of the ABI).
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "../../include/kiddo_amd.h"

namespace kd {

constexpr int kMaxLeafLen = 1024;   // kiddo src/leaf_view/mod.rs:18
constexpr int kMaxStemLevels = 31;  // u32 Eytzinger indices on the device

int set_error(int code, const char* fmt, ...);

template <class A> struct AxisHost;
template <> struct AxisHost<float> {
  static float max_value() { return std::numeric_limits<float>::infinity(); }
  static bool is_nan(float v) { return std::isnan(v); }
};
template <> struct AxisHost<double> {
  static double max_value() { return std::numeric_limits<double>::infinity(); }
  static bool is_nan(double v) { return std::isnan(v); }
};
template <> struct AxisHost<uint16_t> {
  static uint16_t max_value() { return 0xFFFFu; }
  static bool is_nan(uint16_t) { return false; }
};
template <> struct AxisHost<uint32_t> {
  static uint32_t max_value() { return 0xFFFFFFFFu; }
  static bool is_nan(uint32_t) { return false; }
};

inline size_t dtype_size(int dtype) {
  switch (dtype) {
    case KD_F32: return 4;
    case KD_F64: return 8;
    case KD_U16: return 2;
    case KD_U32: return 4;
    default: return 0;
  }
}

struct HostLayout {
  int dtype = 0, dims = 0, bucket = 0;
  uint64_t size = 0;
  int max_stem_level = -1;
  std::vector<uint8_t> stems;
  uint64_t stems_len = 0;
  std::vector<uint64_t> extents;
  std::vector<uint8_t> arena;
  uint64_t max_leaf_len = 0;
};

int build_layout(HostLayout& L, int dtype, int dims, int bucket, uint64_t n, const void* points,
                 const uint32_t* items, int threads);

#define KD_DISPATCH_AK(dtype, dims, ...)                                            \
  switch ((dtype) * 8 + (dims)) {                                                   \
    case KD_F32 * 8 + 2: { using A = float; constexpr int KK = 2; __VA_ARGS__; break; }    \
    case KD_F32 * 8 + 3: { using A = float; constexpr int KK = 3; __VA_ARGS__; break; }    \
    case KD_F32 * 8 + 4: { using A = float; constexpr int KK = 4; __VA_ARGS__; break; }    \
    case KD_F64 * 8 + 2: { using A = double; constexpr int KK = 2; __VA_ARGS__; break; }   \
    case KD_F64 * 8 + 3: { using A = double; constexpr int KK = 3; __VA_ARGS__; break; }   \
    case KD_F64 * 8 + 4: { using A = double; constexpr int KK = 4; __VA_ARGS__; break; }   \
    case KD_U16 * 8 + 2: { using A = uint16_t; constexpr int KK = 2; __VA_ARGS__; break; } \
    case KD_U16 * 8 + 3: { using A = uint16_t; constexpr int KK = 3; __VA_ARGS__; break; } \
    case KD_U16 * 8 + 4: { using A = uint16_t; constexpr int KK = 4; __VA_ARGS__; break; } \
    case KD_U32 * 8 + 2: { using A = uint32_t; constexpr int KK = 2; __VA_ARGS__; break; } \
    case KD_U32 * 8 + 3: { using A = uint32_t; constexpr int KK = 3; __VA_ARGS__; break; } \
    case KD_U32 * 8 + 4: { using A = uint32_t; constexpr int KK = 4; __VA_ARGS__; break; } \
    default:                                                                        \
      return ::kd::set_error(KD_ERR_UNSUPPORTED, "unsupported dtype %d / dims %d", (int)(dtype), \
                       (int)(dims));                                                \
  }

// Float trees only (families whose reference requires a float axis).
#define KD_DISPATCH_FK(dtype, dims, ...)                                            \
  switch ((dtype) * 8 + (dims)) {                                                   \
    case KD_F32 * 8 + 2: { using A = float; constexpr int KK = 2; __VA_ARGS__; break; }    \
    case KD_F32 * 8 + 3: { using A = float; constexpr int KK = 3; __VA_ARGS__; break; }    \
    case KD_F32 * 8 + 4: { using A = float; constexpr int KK = 4; __VA_ARGS__; break; }    \
    case KD_F64 * 8 + 2: { using A = double; constexpr int KK = 2; __VA_ARGS__; break; }   \
    case KD_F64 * 8 + 3: { using A = double; constexpr int KK = 3; __VA_ARGS__; break; }   \
    case KD_F64 * 8 + 4: { using A = double; constexpr int KK = 4; __VA_ARGS__; break; }   \
    default:                                                                        \
      return ::kd::set_error(KD_ERR_UNSUPPORTED, "this family needs a float tree (dtype %d / dims %d)", \
                             (int)(dtype), (int)(dims));                            \
  }

}  // namespace kd
