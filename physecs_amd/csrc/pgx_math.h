// Device-side vector math with the association order of GLM 0.9.9.8's scalar path, the
// library the reference computes with. Every kernel is compiled with -ffp-contract=off and
// correctly rounded division/sqrt, so each operation here rounds exactly as the reference's
// x86 build does: the GPU step reproduces the reference's float results bit for bit.
//
//   dot (x*x'+y*y')+z*z'                  glm/detail/func_geometric.inl:45-53
//   normalize v * (1/sqrt(dot))          glm/detail/func_geometric.inl:79-88
//   min/max/abs as ternaries             glm/detail/func_common.inl:15-29, detail/compute_common.hpp:14-24
//   q * v = v + ((uv*w) + uuv) * 2       glm/detail/type_quat.inl:343-350
//   q * p                                glm/detail/type_quat.inl:281-292
//   mat3_cast                            glm/gtc/quaternion.inl:41-66
//   inverse(mat3)                        glm/detail/func_matrix.inl:268-291
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PGX_HD __host__ __device__ __forceinline__

namespace pgx {

struct f3 { float x, y, z; };
struct f2 { float x, y; };
struct qf { float x, y, z, w; };
struct m3 { f3 c0, c1, c2; };        // columns

PGX_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
PGX_HD f3 s3(float s) { return mk3(s, s, s); }
PGX_HD f3 add(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
PGX_HD f3 sub(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
PGX_HD f3 neg(f3 a) { return mk3(-a.x, -a.y, -a.z); }
PGX_HD f3 mul(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
PGX_HD f3 mul(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
PGX_HD f3 divs(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
PGX_HD f3 subs(f3 a, float s) { return mk3(a.x - s, a.y - s, a.z - s); }
PGX_HD f3 adds(f3 a, float s) { return mk3(a.x + s, a.y + s, a.z + s); }
PGX_HD f3 mulv(f3 a, f3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
PGX_HD f3 divv(f3 a, f3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
PGX_HD float at(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
PGX_HD void set(f3& v, int i, float s) { if (i == 0) v.x = s; else if (i == 1) v.y = s; else v.z = s; }

PGX_HD float gmin(float x, float y) { return (y < x) ? y : x; }
PGX_HD float gmax(float x, float y) { return (x < y) ? y : x; }
PGX_HD float gabs(float x) { return x >= 0.f ? x : -x; }
PGX_HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
PGX_HD float gmix(float x, float y, float a) { return x * (1.f - a) + y * a; }
PGX_HD float gsign(float x) { return (float)(0.f < x) - (float)(x < 0.f); }
PGX_HD f3 vmin(f3 a, f3 b) { return mk3(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
PGX_HD f3 vmax(f3 a, f3 b) { return mk3(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }
PGX_HD f3 vabs(f3 a) { return mk3(gabs(a.x), gabs(a.y), gabs(a.z)); }
PGX_HD f3 vclamp(f3 x, f3 lo, f3 hi) { return vmin(vmax(x, lo), hi); }

PGX_HD float dot(f3 a, f3 b) { float x = a.x * b.x, y = a.y * b.y, z = a.z * b.z; return x + y + z; }
PGX_HD f3 cross(f3 a, f3 b) {
    return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
PGX_HD float length(f3 v) { return sqrtf(dot(v, v)); }
PGX_HD f3 normalize(f3 v) { return mul(v, 1.f / sqrtf(dot(v, v))); }

PGX_HD f2 mk2(float x, float y) { f2 r; r.x = x; r.y = y; return r; }
PGX_HD f2 sub2(f2 a, f2 b) { return mk2(a.x - b.x, a.y - b.y); }
PGX_HD f2 add2(f2 a, f2 b) { return mk2(a.x + b.x, a.y + b.y); }
PGX_HD float det2(f2 c0, f2 c1) { return c0.x * c1.y - c1.x * c0.y; }

PGX_HD qf mkq(float w, float x, float y, float z) { qf q; q.x = x; q.y = y; q.z = z; q.w = w; return q; }
PGX_HD qf qmul(qf p, qf q) {
    return mkq(p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z,
               p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y,
               p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z,
               p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x);
}
PGX_HD f3 qrot(qf q, f3 v) {
    f3 qv = mk3(q.x, q.y, q.z);
    f3 uv = cross(qv, v);
    f3 uuv = cross(qv, uv);
    return add(v, mul(add(mul(uv, q.w), uuv), 2.f));
}
PGX_HD float qdot(qf a, qf b) {
    float tx = a.w * b.w, ty = a.x * b.x, tz = a.y * b.y, tw = a.z * b.z;
    return (tx + ty) + (tz + tw);
}
PGX_HD qf qinverse(qf q) {
    float d = qdot(q, q);
    return mkq(q.w / d, -q.x / d, -q.y / d, -q.z / d);
}
PGX_HD qf qnormalize(qf q) {
    float len = sqrtf(qdot(q, q));
    if (len <= 0.f) return mkq(1.f, 0.f, 0.f, 0.f);
    float o = 1.f / len;
    return mkq(q.w * o, q.x * o, q.y * o, q.z * o);
}

PGX_HD f3 col(const m3& m, int i) { return i == 0 ? m.c0 : (i == 1 ? m.c1 : m.c2); }
PGX_HD m3 mk_m3(f3 a, f3 b, f3 c) { m3 m; m.c0 = a; m.c1 = b; m.c2 = c; return m; }
PGX_HD m3 mat3_cast(qf q) {
    float qxx = q.x * q.x, qyy = q.y * q.y, qzz = q.z * q.z;
    float qxz = q.x * q.z, qxy = q.x * q.y, qyz = q.y * q.z;
    float qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
    return mk_m3(mk3(1.f - 2.f * (qyy + qzz), 2.f * (qxy + qwz), 2.f * (qxz - qwy)),
                 mk3(2.f * (qxy - qwz), 1.f - 2.f * (qxx + qzz), 2.f * (qyz + qwx)),
                 mk3(2.f * (qxz + qwy), 2.f * (qyz - qwx), 1.f - 2.f * (qxx + qyy)));
}
PGX_HD f3 mulmv(const m3& m, f3 v) {
    return mk3(m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z,
               m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
               m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z);
}
// column j of a*b is a * (column j of b)
PGX_HD m3 mulmm(const m3& a, const m3& b) { return mk_m3(mulmv(a, b.c0), mulmv(a, b.c1), mulmv(a, b.c2)); }
PGX_HD m3 transpose(const m3& m) {
    return mk_m3(mk3(m.c0.x, m.c1.x, m.c2.x), mk3(m.c0.y, m.c1.y, m.c2.y), mk3(m.c0.z, m.c1.z, m.c2.z));
}
PGX_HD m3 madd(const m3& a, const m3& b) { return mk_m3(add(a.c0, b.c0), add(a.c1, b.c1), add(a.c2, b.c2)); }
PGX_HD m3 msub(const m3& a, const m3& b) { return mk_m3(sub(a.c0, b.c0), sub(a.c1, b.c1), sub(a.c2, b.c2)); }
PGX_HD m3 mscale(float s, const m3& a) { return mk_m3(mul(a.c0, s), mul(a.c1, s), mul(a.c2, s)); }
PGX_HD m3 mzero() { return mk_m3(s3(0.f), s3(0.f), s3(0.f)); }
PGX_HD m3 inverse(const m3& m) {
    float m00 = m.c0.x, m01 = m.c0.y, m02 = m.c0.z;
    float m10 = m.c1.x, m11 = m.c1.y, m12 = m.c1.z;
    float m20 = m.c2.x, m21 = m.c2.y, m22 = m.c2.z;
    float ood = 1.f / (+m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02) + m20 * (m01 * m12 - m11 * m02));
    m3 r;
    r.c0.x = +(m11 * m22 - m21 * m12) * ood;
    r.c1.x = -(m10 * m22 - m20 * m12) * ood;
    r.c2.x = +(m10 * m21 - m20 * m11) * ood;
    r.c0.y = -(m01 * m22 - m21 * m02) * ood;
    r.c1.y = +(m00 * m22 - m20 * m02) * ood;
    r.c2.y = -(m00 * m21 - m20 * m01) * ood;
    r.c0.z = +(m01 * m12 - m11 * m02) * ood;
    r.c1.z = -(m00 * m12 - m10 * m02) * ood;
    r.c2.z = +(m00 * m11 - m10 * m01) * ood;
    return r;
}
// glm/gtx/matrix_cross_product.inl:6-19
PGX_HD m3 matrixCross3(f3 x) {
    return mk_m3(mk3(0.f, x.z, -x.y), mk3(-x.z, 0.f, x.x), mk3(x.y, -x.x, 0.f));
}

// float4 helpers for the SoA buffers
// double vectors (glm::dvec3) for GJK/EPA, same association order as the float ones
struct d3 { double x, y, z; };
PGX_HD d3 mkd(double x, double y, double z) { d3 r; r.x = x; r.y = y; r.z = z; return r; }
PGX_HD d3 wide(f3 v) { return mkd(v.x, v.y, v.z); }
PGX_HD f3 narrow(d3 v) { return mk3((float)v.x, (float)v.y, (float)v.z); }
PGX_HD d3 dadd(d3 a, d3 b) { return mkd(a.x + b.x, a.y + b.y, a.z + b.z); }
PGX_HD d3 dsub(d3 a, d3 b) { return mkd(a.x - b.x, a.y - b.y, a.z - b.z); }
PGX_HD d3 dneg(d3 a) { return mkd(-a.x, -a.y, -a.z); }
PGX_HD d3 dmul(double s, d3 a) { return mkd(s * a.x, s * a.y, s * a.z); }
PGX_HD d3 dmul(d3 a, double s) { return mkd(a.x * s, a.y * s, a.z * s); }
PGX_HD d3 dadds(d3 a, double s) { return mkd(a.x + s, a.y + s, a.z + s); }
PGX_HD d3 ddivs(d3 a, double s) { return mkd(a.x / s, a.y / s, a.z / s); }
PGX_HD double ddot(d3 a, d3 b) { double x = a.x * b.x, y = a.y * b.y, z = a.z * b.z; return x + y + z; }
PGX_HD d3 dcross(d3 a, d3 b) {
    return mkd(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
PGX_HD double dlength(d3 v) { return sqrt(ddot(v, v)); }
PGX_HD d3 dnormalize(d3 v) { return dmul(v, 1.0 / sqrt(ddot(v, v))); }
PGX_HD double dsign(double x) { return (double)(0.0 < x) - (double)(x < 0.0); }

PGX_HD f3 xyz(float4 a) { return mk3(a.x, a.y, a.z); }
PGX_HD float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
PGX_HD qf q4(float4 a) { return mkq(a.w, a.x, a.y, a.z); }
PGX_HD float4 f4q(qf q) { return make_float4(q.x, q.y, q.z, q.w); }

}  // namespace pgx
