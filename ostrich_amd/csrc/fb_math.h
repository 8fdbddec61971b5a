// fb_math.h -- device-side rigid-body algebra for the gfx950 solver kernels.
//
// Transforms are (p, q) with quaternion xyzw, spatial vectors are (linear, angular),
// matching Newton/Warp storage (src/ostrich/mechanics/*.py).  Everything is FP32 and
// inlined; no packed-FP32 instructions are written by hand.
#pragma once

#include <hip/hip_runtime.h>

namespace ofb {

struct v3 { float x, y, z; };
struct qt { float x, y, z, w; };
struct tf { v3 p; qt q; };
struct m3 { float m[9]; };
struct sv { v3 a, b; };  // a = linear (top), b = angular (bottom)

__device__ __forceinline__ v3 mk(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
  return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
__device__ __forceinline__ v3 unit(int i) {
  return i == 0 ? mk(1.f, 0.f, 0.f) : (i == 1 ? mk(0.f, 1.f, 0.f) : mk(0.f, 0.f, 1.f));
}
__device__ __forceinline__ v3 normalize(v3 a) {
  float l = sqrtf(dot(a, a));
  return l > 0.f ? a * (1.f / l) : mk(0.f, 0.f, 0.f);
}

// rotation of v by a unit quaternion: v (2w^2-1) + 2 (u.v) u + 2 w (u x v)
__device__ __forceinline__ v3 qrot(qt q, v3 v) {
  v3 u = mk(q.x, q.y, q.z);
  float c = 2.f * q.w * q.w - 1.f;
  float d = 2.f * dot(u, v);
  return v * c + u * d + cross(u, v) * (2.f * q.w);
}
__device__ __forceinline__ qt qmul(qt a, qt b) {
  return qt{a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y, a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z,
            a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
__device__ __forceinline__ qt qconj(qt a) { return qt{-a.x, -a.y, -a.z, a.w}; }

__device__ __forceinline__ m3 qmat(qt q) {
  m3 R;
  float x = q.x, y = q.y, z = q.z, w = q.w;
  R.m[0] = 1.f - 2.f * (y * y + z * z); R.m[1] = 2.f * (x * y - z * w); R.m[2] = 2.f * (x * z + y * w);
  R.m[3] = 2.f * (x * y + z * w); R.m[4] = 1.f - 2.f * (x * x + z * z); R.m[5] = 2.f * (y * z - x * w);
  R.m[6] = 2.f * (x * z - y * w); R.m[7] = 2.f * (y * z + x * w); R.m[8] = 1.f - 2.f * (x * x + y * y);
  return R;
}
__device__ __forceinline__ v3 mv(const m3& A, v3 v) {
  return v3{A.m[0] * v.x + A.m[1] * v.y + A.m[2] * v.z, A.m[3] * v.x + A.m[4] * v.y + A.m[5] * v.z,
            A.m[6] * v.x + A.m[7] * v.y + A.m[8] * v.z};
}
// R I R^T
__device__ __forceinline__ m3 rot_inertia(qt q, const m3& I) {
  m3 R = qmat(q), RI, out;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      RI.m[i * 3 + j] = R.m[i * 3 + 0] * I.m[0 * 3 + j] + R.m[i * 3 + 1] * I.m[1 * 3 + j] + R.m[i * 3 + 2] * I.m[2 * 3 + j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      out.m[i * 3 + j] = RI.m[i * 3 + 0] * R.m[j * 3 + 0] + RI.m[i * 3 + 1] * R.m[j * 3 + 1] + RI.m[i * 3 + 2] * R.m[j * 3 + 2];
  return out;
}
__device__ __forceinline__ v3 tpoint(const tf& X, v3 p) { return X.p + qrot(X.q, p); }
__device__ __forceinline__ tf tmul(const tf& A, const tf& B) { return tf{A.p + qrot(A.q, B.p), qmul(A.q, B.q)}; }
__device__ __forceinline__ tf tident() { return tf{mk(0.f, 0.f, 0.f), qt{0.f, 0.f, 0.f, 1.f}}; }

__device__ __forceinline__ sv svzero() { return sv{mk(0.f, 0.f, 0.f), mk(0.f, 0.f, 0.f)}; }
__device__ __forceinline__ sv operator*(const sv& s, float k) { return sv{s.a * k, s.b * k}; }
__device__ __forceinline__ float sdot(const sv& x, const sv& y) { return dot(x.a, y.a) + dot(x.b, y.b); }

// orthogonal_basis (mechanics/geometry.py:4-20)
__device__ __forceinline__ void orthogonal_basis(v3 a, v3& b1, v3& b2) {
  v3 v;
  if (fabsf(a.x) <= fabsf(a.y) && fabsf(a.x) <= fabsf(a.z)) v = mk(1.f, 0.f, 0.f);
  else if (fabsf(a.y) <= fabsf(a.z)) v = mk(0.f, 1.f, 0.f);
  else v = mk(0.f, 0.f, 1.f);
  b1 = normalize(cross(a, v));
  b2 = cross(a, b1);
}

}  // namespace ofb
