// nrf_math.h - SE(3) / dual-quaternion algebra for the gfx950 kernels.
//
// The operation sequences follow include/se3_math.cuh of the reference
// (cited per function) so that results agree with it to rounding; the code
// is written for this library: quaternions are float4 (x, y, z, w), a dual
// quaternion is {real, dual}.  All device code is compiled with
// -ffp-contract=off so that the FP32 results do not depend on FMA fusion.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nrf.h"

namespace nrf {

struct DQ {
  float4 r;  // rotation quaternion
  float4 d;  // dual part
};

static_assert(sizeof(DQ) == sizeof(nrf_dualquat), "DQ layout");
static_assert(sizeof(nrf_node) == 96, "DeformNode is 96 bytes");
static_assert(sizeof(nrf_corr) == 76, "Correspondence is 76 bytes");
static_assert(sizeof(nrf_voxel) == 12, "TSDFVoxel is 12 bytes");
static_assert(sizeof(nrf_camera) == 24, "CameraIntrinsics is 24 bytes");

struct M34 {  // rows 0..2 of a rigid 4x4 (row 3 is 0 0 0 1)
  float m[3][4];
};

__host__ __device__ inline float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__host__ __device__ inline float3 f3(const float *p) { return make_float3(p[0], p[1], p[2]); }
__host__ __device__ inline float3 sub(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline float3 cross(float3 a, float3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

__host__ __device__ inline DQ load_dq(const nrf_dualquat &q) {
  DQ o;
  o.r = make_float4(q.real[0], q.real[1], q.real[2], q.real[3]);
  o.d = make_float4(q.dual[0], q.dual[1], q.dual[2], q.dual[3]);
  return o;
}
__device__ inline DQ load_dq_vec(const nrf_dualquat *p) {
  const float4 *v = reinterpret_cast<const float4 *>(p);
  DQ o;
  o.r = v[0];
  o.d = v[1];
  return o;
}
__host__ __device__ inline void store_dq(nrf_dualquat &q, const DQ &v) {
  q.real[0] = v.r.x; q.real[1] = v.r.y; q.real[2] = v.r.z; q.real[3] = v.r.w;
  q.dual[0] = v.d.x; q.dual[1] = v.d.y; q.dual[2] = v.d.z; q.dual[3] = v.d.w;
}

// Hamilton product, se3_math.cuh:205-214.
__host__ __device__ inline float4 qmul(float4 a, float4 b) {
  return make_float4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                     a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                     a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
                     a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
__host__ __device__ inline float4 qconj(float4 q) { return make_float4(-q.x, -q.y, -q.z, q.w); }
__host__ __device__ inline float qdot(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// se3_math.cuh:223-231: unit quaternion, identity below 1e-12.
__host__ __device__ inline float4 qnormalize(float4 q) {
  float n = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  if (n < 1e-12f) return make_float4(0.f, 0.f, 0.f, 1.f);
  return make_float4(q.x / n, q.y / n, q.z / n, q.w / n);
}

__host__ __device__ inline DQ dq_identity() {
  DQ o;
  o.r = make_float4(0.f, 0.f, 0.f, 1.f);
  o.d = make_float4(0.f, 0.f, 0.f, 0.f);
  return o;
}

// se3_math.cuh:302-328: both parts scaled by 1/|real|.
__host__ __device__ inline DQ dq_normalize(DQ q) {
  float n = sqrtf(q.r.x * q.r.x + q.r.y * q.r.y + q.r.z * q.r.z + q.r.w * q.r.w);
  if (n < 1e-12f) return dq_identity();
  float inv = 1.f / n;
  q.r.x *= inv; q.r.y *= inv; q.r.z *= inv; q.r.w *= inv;
  q.d.x *= inv; q.d.y *= inv; q.d.z *= inv; q.d.w *= inv;
  return q;
}

// se3_math.cuh:330-340.
__host__ __device__ inline DQ dq_mul(DQ a, DQ b) {
  DQ o;
  o.r = qmul(a.r, b.r);
  float4 ad = qmul(a.r, b.d);
  float4 bd = qmul(a.d, b.r);
  o.d = make_float4(ad.x + bd.x, ad.y + bd.y, ad.z + bd.z, ad.w + bd.w);
  return dq_normalize(o);
}

// se3_math.cuh:349-360: q p q*.
__host__ __device__ inline float3 dq_rotate(const DQ &q, float3 p) {
  float4 r = qmul(qmul(q.r, make_float4(p.x, p.y, p.z, 0.f)), qconj(q.r));
  return f3(r.x, r.y, r.z);
}

// se3_math.cuh:362-368: t = 2 d r*.
__host__ __device__ inline float3 dq_translation(const DQ &q) {
  float4 t = qmul(q.d, qconj(q.r));
  return f3(2.f * t.x, 2.f * t.y, 2.f * t.z);
}

// se3_math.cuh:370-377.
__host__ __device__ inline float3 dq_transform(const DQ &q, float3 p) {
  float3 r = dq_rotate(q, p);
  float3 t = dq_translation(q);
  return f3(r.x + t.x, r.y + t.y, r.z + t.z);
}

// se3_math.cuh:379-388: rotation about `c`.
__host__ __device__ inline float3 dq_transform_centered(const DQ &q, float3 p, float3 c) {
  float3 rp = dq_transform(q, p);
  float3 rc = dq_rotate(q, c);
  return f3(rp.x + c.x - rc.x, rp.y + c.y - rc.y, rp.z + c.z - rc.z);
}

// se3_math.cuh:390-403.
__host__ __device__ inline DQ dq_from_rt(float4 qr, float3 t) {
  DQ o;
  o.r = qnormalize(qr);
  o.d = qmul(make_float4(t.x, t.y, t.z, 0.f), o.r);
  o.d.x *= 0.5f; o.d.y *= 0.5f; o.d.z *= 0.5f; o.d.w *= 0.5f;
  return dq_normalize(o);
}

// se3_math.cuh:405-415.
__host__ __device__ inline DQ dq_centered(const DQ &q, float3 c) {
  float3 t = dq_translation(q);
  float3 rc = dq_rotate(q, c);
  return dq_from_rt(q.r, f3(t.x + c.x - rc.x, t.y + c.y - rc.y, t.z + c.z - rc.z));
}

// exp_se3, se3_math.cuh:40-98: twist = (omega, v) -> [R | V v].
__host__ __device__ inline M34 exp_se3(const float tw[6]) {
  float ox = tw[0], oy = tw[1], oz = tw[2];
  float vx = tw[3], vy = tw[4], vz = tw[5];
  float th2 = ox * ox + oy * oy + oz * oz;
  float th = sqrtf(th2);
  float A, B, C;
  if (th < 1e-6f) {
    A = 1.f; B = 0.5f; C = 1.f / 6.f;
  } else {
    float it = 1.f / th;
    float it2 = it * it;
    A = sinf(th) * it;
    B = (1.f - cosf(th)) * it2;
    C = (th - sinf(th)) * it2 * it;
  }
  M34 T;
  T.m[0][0] = 1.f - B * (oy * oy + oz * oz);
  T.m[0][1] = B * ox * oy - A * oz;
  T.m[0][2] = B * ox * oz + A * oy;
  T.m[1][0] = B * ox * oy + A * oz;
  T.m[1][1] = 1.f - B * (ox * ox + oz * oz);
  T.m[1][2] = B * oy * oz - A * ox;
  T.m[2][0] = B * ox * oz - A * oy;
  T.m[2][1] = B * oy * oz + A * ox;
  T.m[2][2] = 1.f - B * (ox * ox + oy * oy);
  float Vxx = 1.f - C * (oy * oy + oz * oz);
  float Vxy = C * ox * oy - B * oz;
  float Vxz = C * ox * oz + B * oy;
  float Vyx = C * ox * oy + B * oz;
  float Vyy = 1.f - C * (ox * ox + oz * oz);
  float Vyz = C * oy * oz - B * ox;
  float Vzx = C * ox * oz - B * oy;
  float Vzy = C * oy * oz + B * ox;
  float Vzz = 1.f - C * (ox * ox + oy * oy);
  T.m[0][3] = Vxx * vx + Vxy * vy + Vxz * vz;
  T.m[1][3] = Vyx * vx + Vyy * vy + Vyz * vz;
  T.m[2][3] = Vzx * vx + Vzy * vy + Vzz * vz;
  return T;
}

// mat4_to_dq, se3_math.cuh:244-300 (Shepperd's four branches).
__host__ __device__ inline DQ mat_to_dq(const M34 &T) {
  float tr = T.m[0][0] + T.m[1][1] + T.m[2][2];
  float4 q;
  if (tr > 0) {
    float s = sqrtf(tr + 1.0f) * 2.f;
    q.w = 0.25f * s;
    q.x = (T.m[2][1] - T.m[1][2]) / s;
    q.y = (T.m[0][2] - T.m[2][0]) / s;
    q.z = (T.m[1][0] - T.m[0][1]) / s;
  } else if (T.m[0][0] > T.m[1][1] && T.m[0][0] > T.m[2][2]) {
    float s = sqrtf(1.0f + T.m[0][0] - T.m[1][1] - T.m[2][2]) * 2.f;
    q.w = (T.m[2][1] - T.m[1][2]) / s;
    q.x = 0.25f * s;
    q.y = (T.m[0][1] + T.m[1][0]) / s;
    q.z = (T.m[0][2] + T.m[2][0]) / s;
  } else if (T.m[1][1] > T.m[2][2]) {
    float s = sqrtf(1.0f + T.m[1][1] - T.m[0][0] - T.m[2][2]) * 2.f;
    q.w = (T.m[0][2] - T.m[2][0]) / s;
    q.x = (T.m[0][1] + T.m[1][0]) / s;
    q.y = 0.25f * s;
    q.z = (T.m[1][2] + T.m[2][1]) / s;
  } else {
    float s = sqrtf(1.0f + T.m[2][2] - T.m[0][0] - T.m[1][1]) * 2.f;
    q.w = (T.m[1][0] - T.m[0][1]) / s;
    q.x = (T.m[0][2] + T.m[2][0]) / s;
    q.y = (T.m[1][2] + T.m[2][1]) / s;
    q.z = 0.25f * s;
  }
  q = qnormalize(q);
  DQ o;
  o.r = q;
  float4 qd = qmul(make_float4(T.m[0][3], T.m[1][3], T.m[2][3], 0.f), q);
  o.d = make_float4(qd.x * 0.5f, qd.y * 0.5f, qd.z * 0.5f, qd.w * 0.5f);
  return o;
}

// Blend of pre-centred node transforms (warp_point, se3_math.cuh:450-510):
// running-sign antipodality fix, |w|-weighted sum, normalise.  Returns false
// when the weight sum is below 1e-8 (the caller keeps the point unwarped).
__device__ inline bool dq_blend4(const DQ *table, const int ids[4], const float ws[4], DQ &out) {
  DQ b;
  b.r = make_float4(0.f, 0.f, 0.f, 0.f);
  b.d = make_float4(0.f, 0.f, 0.f, 0.f);
  float wsum = 0.f;
#pragma unroll
  for (int k = 0; k < NRF_K_NEIGHBORS; ++k) {
    int id = ids[k];
    float w = ws[k];
    if (id < 0 || w < 1e-8f) continue;
    DQ q = table[id];
    if (qdot(q.r, b.r) < 0) {
      q.r.x = -q.r.x; q.r.y = -q.r.y; q.r.z = -q.r.z; q.r.w = -q.r.w;
      q.d.x = -q.d.x; q.d.y = -q.d.y; q.d.z = -q.d.z; q.d.w = -q.d.w;
    }
    b.r.x += w * q.r.x; b.r.y += w * q.r.y; b.r.z += w * q.r.z; b.r.w += w * q.r.w;
    b.d.x += w * q.d.x; b.d.y += w * q.d.y; b.d.z += w * q.d.z; b.d.w += w * q.d.w;
    wsum += w;
  }
  if (wsum < 1e-8f) return false;
  out = dq_normalize(b);
  return true;
}

// warp_normal (include/se3_math.cuh:512-558): blends the real parts of the
// *uncentred* node transforms (running-sign fix on the real part only),
// normalises, rotates n and renormalises it.  There is no weight-sum test:
// with no valid neighbour the blend normalises to the identity.
__host__ __device__ inline float3 dq_warp_normal(const DQ *T, int N, const int ids[4], const float ws[4], float3 n) {
  DQ b;
  b.r = make_float4(0.f, 0.f, 0.f, 0.f);
  b.d = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int k = 0; k < 4; ++k) {
    const int id = ids[k];
    const float w = ws[k];
    if (id < 0 || id >= N || w < 1e-8f) continue;
    float4 q = T[id].r;
    if (qdot(q, b.r) < 0) q = make_float4(-q.x, -q.y, -q.z, -q.w);
    b.r.x += w * q.x; b.r.y += w * q.y; b.r.z += w * q.z; b.r.w += w * q.w;
  }
  b = dq_normalize(b);
  const float4 r = qmul(qmul(b.r, make_float4(n.x, n.y, n.z, 0.f)), qconj(b.r));
  const float m = sqrtf(r.x * r.x + r.y * r.y + r.z * r.z);
  return m > 1e-9f ? f3(r.x / m, r.y / m, r.z / m) : f3(r.x, r.y, r.z);
}

// One node's step of the caller: the per-node clamp, then the scale
// (upload_scaled_delta_x, include/fusion_app.h:1353-1375), then
// apply_twist_kernel with update_scale 1 (src/warp_field.cu:250-291), which
// clamps the scaled twist again and composes dq(exp(twist)) on the left.
__host__ __device__ inline DQ scaled_twist_update(const float d[6], float scale, float max_rot, float max_trans,
                                                  const DQ &T) {
  float tw[6];
  for (int k = 0; k < 6; ++k) tw[k] = d[k];
  const float rn = sqrtf(tw[0] * tw[0] + tw[1] * tw[1] + tw[2] * tw[2]);
  const float tn = sqrtf(tw[3] * tw[3] + tw[4] * tw[4] + tw[5] * tw[5]);
  float rsc = 1.0f, tsc = 1.0f;
  if (max_rot > 0.0f && rn > max_rot) rsc = max_rot / rn;
  if (max_trans > 0.0f && tn > max_trans) tsc = max_trans / tn;
  for (int k = 0; k < 3; ++k) tw[k] *= rsc * scale;
  for (int k = 3; k < 6; ++k) tw[k] *= tsc * scale;
  for (int k = 0; k < 6; ++k) tw[k] = 1.0f * tw[k];
  const float rn2 = sqrtf(tw[0] * tw[0] + tw[1] * tw[1] + tw[2] * tw[2]);
  if (max_rot > 0.0f && rn2 > max_rot) {
    const float f = max_rot / rn2;
    tw[0] *= f; tw[1] *= f; tw[2] *= f;
  }
  const float tn2 = sqrtf(tw[3] * tw[3] + tw[4] * tw[4] + tw[5] * tw[5]);
  if (max_trans > 0.0f && tn2 > max_trans) {
    const float f = max_trans / tn2;
    tw[3] *= f; tw[4] *= f; tw[5] *= f;
  }
  return dq_mul(mat_to_dq(exp_se3(tw)), T);
}

// Huber weight, kernels/solver_kernels.cu:13-24.
__host__ __device__ inline float huber_w(float r, float delta) {
  float a = fabsf(r);
  return a <= delta ? 1.0f : delta / a;
}

}  // namespace nrf
