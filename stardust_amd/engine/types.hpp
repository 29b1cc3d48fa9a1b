// Host value types of the STARDUST engine API (src/engine/types.hpp,
// helper_math.hpp of the reference): Scalar, a 4-float vector with the
// component-wise operators the entity code uses, and the 3x3 float9 row-major
// tensor. Host-only: the engine forwards every step to libstardust_hip.so
// through its C ABI, so no GPU header is needed by users of this API.
#ifndef STARDUST_AMD_ENGINE_TYPES_HPP
#define STARDUST_AMD_ENGINE_TYPES_HPP

struct float4 {
    float x, y, z, w;
};

inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline float4 operator+(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
inline float4 operator-(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
inline float4 operator*(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
inline float4 operator/(float4 a, float s) { return make_float4(a.x / s, a.y / s, a.z / s, a.w / s); }
inline void operator+=(float4& a, float4 b) { a = a + b; }

namespace STARDUST {

using Scalar = float;

// Row-major 3x3 tensor (types.hpp:30-40): {xx, xy, xz, yx, yy, yz, zx, zy, zz}.
struct float9 {
    Scalar data[9];
    Scalar& operator[](int i) { return data[i]; }
    const Scalar& operator[](int i) const { return data[i]; }
};

}  // namespace STARDUST

#endif
