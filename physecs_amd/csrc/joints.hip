// Joint rows on the device: row creation per step, world-space rows per substep, preSolve
// (NGS pseudo-velocities) and solve, in the reference's joint colours.
//
// Reference semantics:
//   colouring at creation            src/Physecs.cpp:690-710 (host, pgx_world_add_joints)
//   rows created every step          src/Physecs.cpp:321-355, include/Physecs/Constraint1DContainer.h:133-198
//   world-space data, row values     src/Joint.cpp:5-14, src/JointsUtil.h:8-28, src/Joints/*.cpp
//   prismatic limit row per step     src/Joints/PrismaticJoint.cpp:123-150 (getSolverDesc)
//   gear angles, state per substep   src/Joints/GearJoint.cpp:5-47
//   colours 0-7 (SIMD rows)          src/Constraint1DW.cpp:5-233, per lane
//   colour 8 (scalar overflow)       src/Constraint1D.cpp:5-127
// Joints of one colour 0-7 share no entity, so each is one thread that walks its rows in
// flag-type order; that is exactly what the reference's 4-wide rows compute lane by lane.
// Colour 8 may share bodies and is walked by one thread, flag type by flag type over all
// its joints, as the reference's overflow container does.
#include <cstring>
#include "pgx_internal.h"

namespace pgx {

enum { JF_NONE = 0, JF_SOFT = 1, JF_ANGULAR = 2, JF_LIMITED = 4 };
enum { JT_SPHERICAL = 0, JT_UNIVERSAL = 1, JT_REVOLUTE = 2, JT_FIXED = 3, JT_PRISMATIC = 4, JT_GEAR = 5, JT_SERVO = 6,
       JT_CUSTOM = 7 };
// jRowFlags: 4 flag bits per row (7 rows), then the prismatic limit row's kind
constexpr int JL_SHIFT = 28, JL_UPPER = 1, JL_LOWER = 2;
__constant__ int kFlagOrder[6] = {JF_NONE, JF_ANGULAR, JF_SOFT, JF_LIMITED, JF_ANGULAR | JF_SOFT,
                                  JF_ANGULAR | JF_LIMITED};

// row fields (float4) of row r of joint j: jRow[(r * 8 + f) * Jcap + j]
enum { JR_LIN = 0, JR_ANG0, JR_ANG1, JR_LIN0T, JR_LIN1T, JR_ANG0T, JR_ANG1T, JR_EXTRA, JR_NF };
// JR_LIN.w = targetVelocity, JR_ANG0.w = c, JR_ANG1.w = invEffMass, JR_LIN0T.w = lambda,
// JR_LIN1T.w = min, JR_ANG0T.w = max, JR_ANG1T.w = frequency, JR_EXTRA.x = dampingRatio

struct JointArgs {
    int J, Jcap;
    const int32_t *jType, *jA0, *jA1;
    int32_t *jRowFlags, *jNRows;                              // per step (a prismatic joint's limit row)
    const float4 *jP0, *jQ0, *jP1, *jQ1, *jDrive;            // anchors (body space), drive (enabled, v, maxTorque)
    const float4 *jPar0, *jPar1;                              // (lower, upper, target, stiffness), (damping, ratio)
    float4* jGear;                                            // persistent angle 0, 1, virtual angle 0, 1
    int32_t* jGearInit;
    int32_t *jB0, *jB1;
    float4 *jR0, *jR1;                                        // anchor - com (or anchor), body space
    float4* jRow;
    const float4 *bpos, *bquat, *dcom, *dinvI, *dvel;
    float4 *bvel, *pvel, *pang;
    const int32_t* order;                                     // joints sorted by colour
    float h;
};

#define JROW(r, f) a.jRow[((size_t)((r) * JR_NF + (f))) * a.Jcap + j]

__device__ __forceinline__ int row_flags(const JointArgs& a, int j, int r) { return (a.jRowFlags[j] >> (4 * r)) & 15; }

// The rows PrismaticJoint::getSolverDesc declares for this step (PrismaticJoint.cpp:123-150):
// 2 linear, 3 angular, a limit row when the slide is beyond a limit at the step's start
// pose, a soft drive row when the drive is on.
__device__ void prismatic_layout(const JointArgs& a, int j) {
    int A0 = a.jA0[j], A1 = a.jA1[j];
    qf t0 = q4(a.bquat[A0]), t1 = q4(a.bquat[A1]);
    f3 p0 = add(xyz(a.bpos[A0]), qrot(t0, xyz(a.jP0[j])));
    f3 p1 = add(xyz(a.bpos[A1]), qrot(t1, xyz(a.jP1[j])));
    f3 d = sub(p1, p0);
    f3 u00 = qrot(qmul(t0, q4(a.jQ0[j])), mk3(1.f, 0.f, 0.f));
    float dx = dot(d, u00);
    float4 par = a.jPar0[j];
    int flags = (JF_ANGULAR << 8) | (JF_ANGULAR << 12) | (JF_ANGULAR << 16), n = 5, kind = 0;
    if (dx > par.y) kind = JL_UPPER;
    else if (dx < par.x) kind = JL_LOWER;
    if (kind) flags |= JF_LIMITED << (4 * n++);
    if (a.jDrive[j].x != 0.f) flags |= JF_SOFT << (4 * n++);
    a.jRowFlags[j] = flags | (kind << JL_SHIFT);
    a.jNRows[j] = n;
}

// src/Physecs.cpp:332-352: body indices and the anchors relative to the COM; fresh rows
// (Constraint1D defaults: zeros, min = lowest, max = max, lambda 0)
__global__ void k_joint_setup(JointArgs a) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.J) return;
    // dynamic bodies are the first ones; a kinematic body carries a nonzero dcom.w
    // (the descriptor's actor order is mapped to these internal indices on the host)
    int b0 = a.jB0[j], b1 = a.jB1[j];
    f3 p0 = xyz(a.jP0[j]), p1 = xyz(a.jP1[j]);
    a.jR0[j] = f4(b0 >= 0 ? sub(p0, xyz(a.dcom[b0])) : p0, 0.f);
    a.jR1[j] = f4(b1 >= 0 ? sub(p1, xyz(a.dcom[b1])) : p1, 0.f);
    if (a.jType[j] == JT_PRISMATIC) prismatic_layout(a, j);
    int n = a.jNRows[j];
    for (int r = 0; r < n; ++r) {
        for (int f = 0; f < JR_NF; ++f) JROW(r, f) = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 e = JROW(r, JR_LIN1T); e.w = -3.402823466e+38f; JROW(r, JR_LIN1T) = e;
        float4 m = JROW(r, JR_ANG0T); m.w = 3.402823466e+38f; JROW(r, JR_ANG0T) = m;
    }
}

__device__ __forceinline__ void set_xyz(float4& d, f3 v) { d.x = v.x; d.y = v.y; d.z = v.z; }

// glm::atan(y, x) and glm::acos call the platform's atan2f / acosf, whose last bit differs
// between C libraries; both are computed here as the correctly rounded function, evaluated
// in double and rounded once to float.
__device__ __forceinline__ float atan2_cr(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ __forceinline__ float acos_cr(float x) { return (float)acos((double)x); }

// GearJoint.cpp:5-8: the signed difference of two angles, wrapped to [-pi, pi)
__device__ __forceinline__ float angle_diff(float angle0, float angle1) {
    const float pi = 3.14159265358979323846f, twoPi = 6.28318530717958647692f;
    const float diff = fmodf(angle1 - angle0 + pi, twoPi) - pi;
    return diff < -pi ? diff + twoPi : diff;
}

// row r's linear, angular and c, and optionally its limits or soft parameters
__device__ __forceinline__ void write_row(const JointArgs& a, int j, int r, f3 lin, f3 ang0, f3 ang1, float c) {
    float4 l = JROW(r, JR_LIN), g0 = JROW(r, JR_ANG0), g1 = JROW(r, JR_ANG1);
    set_xyz(l, lin); set_xyz(g0, ang0); g0.w = c; set_xyz(g1, ang1);
    JROW(r, JR_LIN) = l; JROW(r, JR_ANG0) = g0; JROW(r, JR_ANG1) = g1;
}
__device__ __forceinline__ void write_soft(const JointArgs& a, int j, int r, float freq, float damp) {
    float4 t1 = JROW(r, JR_ANG1T), ex = JROW(r, JR_EXTRA);
    t1.w = freq; ex.x = damp;
    JROW(r, JR_ANG1T) = t1; JROW(r, JR_EXTRA) = ex;
}

// JointSolverData::calculateWorldSpaceData (src/Joint.cpp:5-14) at the current poses
struct JWorld { m3 u0, u1; f3 r0, r1, p0, p1; };
__device__ __forceinline__ JWorld joint_world(const JointArgs& a, int j) {
    JWorld s;
    int A0 = a.jA0[j], A1 = a.jA1[j];
    qf t0 = q4(a.bquat[A0]), t1 = q4(a.bquat[A1]);
    s.u0 = mat3_cast(qmul(t0, q4(a.jQ0[j])));
    s.u1 = mat3_cast(qmul(t1, q4(a.jQ1[j])));
    s.r0 = qrot(t0, xyz(a.jR0[j]));
    s.r1 = qrot(t1, xyz(a.jR1[j]));
    s.p0 = add(xyz(a.bpos[A0]), qrot(t0, xyz(a.jP0[j])));
    s.p1 = add(xyz(a.bpos[A1]), qrot(t1, xyz(a.jP1[j])));
    return s;
}

// world-space data and the rows' values for one joint, every substep
__device__ void joint_update(const JointArgs& a, int j) {
    int type = a.jType[j];
    if (type == JT_CUSTOM) return;   // rows from the host (k_custom_rows)
    const JWorld ws = joint_world(a, j);
    const m3 &u0 = ws.u0, &u1 = ws.u1;
    const f3 r0 = ws.r0, r1 = ws.r1, p0 = ws.p0, p1 = ws.p1;
    f3 d = sub(p1, p0);
    auto angular_row = [&](int r, f3 x, f3 y) {   // c = dot(x, y), axis = cross(y, x)
        f3 ax = cross(y, x);
        float4 g0 = JROW(r, JR_ANG0), g1 = JROW(r, JR_ANG1);
        set_xyz(g0, ax); g0.w = dot(x, y); set_xyz(g1, ax);
        JROW(r, JR_ANG0) = g0; JROW(r, JR_ANG1) = g1;
    };
    if (type == JT_PRISMATIC) {                    // PrismaticJoint.cpp:6-92
        write_row(a, j, 0, u0.c1, cross(r0, u0.c1), cross(r1, u0.c1), dot(d, u0.c1));
        write_row(a, j, 1, u0.c2, cross(r0, u0.c2), cross(r1, u0.c2), dot(d, u0.c2));
        angular_row(2, u0.c0, u1.c1);
        angular_row(3, u0.c0, u1.c2);
        angular_row(4, u0.c1, u1.c2);
        float dx = dot(d, u0.c0);
        f3 r0xx = cross(r0, u0.c0), r1xx = cross(r1, u0.c0);
        float4 par = a.jPar0[j];
        int kind = (a.jRowFlags[j] >> JL_SHIFT) & 3;
        int r = 5;
        if (kind) {
            write_row(a, j, r, u0.c0, r0xx, r1xx, kind == JL_UPPER ? dx - par.y : dx - par.x);
            if (kind == JL_UPPER) { float4 e = JROW(r, JR_LIN1T); e.w = 0.f; JROW(r, JR_LIN1T) = e; }
            else { float4 m = JROW(r, JR_ANG0T); m.w = 0.f; JROW(r, JR_ANG0T) = m; }
            ++r;
        }
        if (a.jDrive[j].x != 0.f) {
            write_row(a, j, r, u0.c0, r0xx, r1xx, dx - par.z);
            write_soft(a, j, r, par.w, a.jPar1[j].x);
        }
        return;
    }
    if (type == JT_GEAR) {                         // GearJoint.cpp:10-47
        float angle0, angle1;
        {
            const f3 p1Proj = add(p1, mul(dot(sub(p0, p1), u0.c0), u0.c0));
            const f3 dir = normalize(sub(p0, p1Proj));
            const f3 n = cross(u0.c0, dir);
            const m3 u0t = mulmm(transpose(mk_m3(u0.c0, neg(n), dir)), u0);
            angle0 = atan2_cr(u0t.c1.z, u0t.c2.z);
        }
        {
            const f3 p0Proj = add(p0, mul(dot(sub(p1, p0), u1.c0), u1.c0));
            const f3 dir = normalize(sub(p0Proj, p1));
            const f3 n = cross(u1.c0, dir);
            const m3 u1t = mulmm(transpose(u1), mk_m3(u1.c0, neg(n), dir));
            angle1 = atan2_cr(u1t.c1.z, u1t.c2.z);
        }
        float4 g = a.jGear[j];
        if (!a.jGearInit[j]) {
            g.x = angle0;
            g.y = angle1;
            a.jGearInit[j] = 1;
        }
        g.z = g.z + angle_diff(angle0, g.x);
        g.w = g.w + angle_diff(angle1, g.y);
        g.x = angle0;
        g.y = angle1;
        a.jGear[j] = g;
        float ratio = a.jPar1[j].y;
        write_row(a, j, 0, s3(0.f), mul(u0.c0, ratio), neg(u1.c0), g.z * ratio - g.w);
        return;
    }
    f3 lin[3] = {mk3(1.f, 0.f, 0.f), mk3(0.f, 1.f, 0.f), mk3(0.f, 0.f, 1.f)};
    f3 a0[3] = {mk3(0.f, r0.z, -r0.y), mk3(-r0.z, 0.f, r0.x), mk3(r0.y, -r0.x, 0.f)};
    f3 a1[3] = {mk3(0.f, r1.z, -r1.y), mk3(-r1.z, 0.f, r1.x), mk3(r1.y, -r1.x, 0.f)};
    float dc[3] = {d.x, d.y, d.z};
    for (int k = 0; k < 3; ++k) write_row(a, j, k, lin[k], a0[k], a1[k], dc[k]);
    if (type == JT_UNIVERSAL) angular_row(3, u0.c2, u1.c2);
    if (type == JT_REVOLUTE || type == JT_FIXED || type == JT_SERVO) {
        angular_row(3, u0.c0, u1.c1);
        angular_row(4, u0.c0, u1.c2);
        if (type == JT_FIXED) angular_row(5, u0.c1, u1.c2);
        float4 dr = a.jDrive[j];
        if (type == JT_REVOLUTE && dr.x != 0.f) {
            float4 l = JROW(5, JR_LIN), g0 = JROW(5, JR_ANG0), g1 = JROW(5, JR_ANG1);
            float4 mn = JROW(5, JR_LIN1T), mx = JROW(5, JR_ANG0T);
            set_xyz(g0, u0.c0); set_xyz(g1, u0.c0);
            l.w = dr.y; mx.w = dr.z; mn.w = -dr.z;
            JROW(5, JR_LIN) = l; JROW(5, JR_ANG0) = g0; JROW(5, JR_ANG1) = g1;
            JROW(5, JR_LIN1T) = mn; JROW(5, JR_ANG0T) = mx;
        }
        if (type == JT_SERVO) {                    // ServoJoint.cpp:9-37, glm::orientedAngle
            const f3 x = u0.c2, y = u1.c2, ref = u0.c0;
            const float angle = acos_cr(gclamp(dot(x, y), -1.f, 1.f));
            const float oriented = dot(ref, cross(x, y)) < 0.f ? -angle : angle;
            float4 par = a.jPar0[j];
            float4 g0 = JROW(5, JR_ANG0), g1 = JROW(5, JR_ANG1);
            set_xyz(g0, u0.c0); g0.w = oriented - par.z; set_xyz(g1, u0.c0);
            JROW(5, JR_ANG0) = g0; JROW(5, JR_ANG1) = g1;
            write_soft(a, j, 5, par.w, a.jPar1[j].x);
        }
    }
}

__global__ void k_joint_update(JointArgs a) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < a.J) joint_update(a, j);
}

// ---- user joint types: rows declared and written by host functions
// Staging of n custom joints (device and pinned host, in floats): world-space data
// (CUST_WS per joint), then the layouts (2 ints per joint: rows, packed flags), then the
// rows (JOINT_MAX_ROWS x CUST_ROW per joint, pgx_joint_row's fields).
constexpr int CUST_WS = 30, CUST_LAYOUT = 2, CUST_ROW = 15;
static_assert(sizeof(pgx_joint_world_space) == CUST_WS * sizeof(float), "JointWorldSpaceData layout");
static_assert(sizeof(pgx_joint_row) == CUST_ROW * sizeof(float), "row layout");
constexpr size_t cust_floats(size_t n) { return n * (CUST_WS + CUST_LAYOUT + (size_t)JOINT_MAX_ROWS * CUST_ROW); }

// the listed joints' JointWorldSpaceData, in pgx_joint_world_space's layout
__global__ void k_custom_ws(JointArgs a, const int32_t* __restrict__ list, int n, float* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const JWorld s = joint_world(a, list[k]);
    float* o = out + (size_t)CUST_WS * k;
    const f3 v[10] = {s.p0, s.p1, s.r0, s.r1, s.u0.c0, s.u0.c1, s.u0.c2, s.u1.c0, s.u1.c1, s.u1.c2};
    for (int i = 0; i < 10; ++i) { o[3 * i] = v[i].x; o[3 * i + 1] = v[i].y; o[3 * i + 2] = v[i].z; }
}

// the step's layouts of the listed joints: row count and flags, and fresh rows
// (Constraint1D's defaults, include/Physecs/Constraint1D.h:19-35), as k_joint_setup
__global__ void k_custom_layout(JointArgs a, const int32_t* __restrict__ list, int n, const int32_t* __restrict__ lay) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = list[k], nr = lay[CUST_LAYOUT * k];
    a.jNRows[j] = nr;
    a.jRowFlags[j] = lay[CUST_LAYOUT * k + 1];
    for (int r = 0; r < nr; ++r) {
        for (int f = 0; f < JR_NF; ++f) JROW(r, f) = make_float4(0.f, 0.f, 0.f, 0.f);
        float4 e = JROW(r, JR_LIN1T); e.w = -3.402823466e+38f; JROW(r, JR_LIN1T) = e;
        float4 m = JROW(r, JR_ANG0T); m.w = 3.402823466e+38f; JROW(r, JR_ANG0T) = m;
    }
}

// the rows the host functions wrote this substep, into the fields the solve reads (lambda
// and the preSolve products untouched)
__global__ void k_custom_rows(JointArgs a, const int32_t* __restrict__ list, int n, const float* __restrict__ rows) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int j = list[k], nr = a.jNRows[j];
    for (int r = 0; r < nr; ++r) {
        const float* q = rows + ((size_t)k * JOINT_MAX_ROWS + r) * CUST_ROW;
        float4 l = JROW(r, JR_LIN), g0 = JROW(r, JR_ANG0), g1 = JROW(r, JR_ANG1);
        float4 mn = JROW(r, JR_LIN1T), mx = JROW(r, JR_ANG0T), fr = JROW(r, JR_ANG1T), ex = JROW(r, JR_EXTRA);
        set_xyz(l, mk3(q[0], q[1], q[2])); l.w = q[9];
        set_xyz(g0, mk3(q[3], q[4], q[5])); g0.w = q[10];
        set_xyz(g1, mk3(q[6], q[7], q[8]));
        mn.w = q[11]; mx.w = q[12]; fr.w = q[13]; ex.x = q[14];
        JROW(r, JR_LIN) = l; JROW(r, JR_ANG0) = g0; JROW(r, JR_ANG1) = g1;
        JROW(r, JR_LIN1T) = mn; JROW(r, JR_ANG0T) = mx; JROW(r, JR_ANG1T) = fr; JROW(r, JR_EXTRA) = ex;
    }
}

__device__ __forceinline__ m3 world_inv_inertia_j(const JointArgs& a, int b) {
    m3 rot = mat3_cast(q4(a.bquat[b]));
    m3 invI = mk_m3(xyz(a.dinvI[3 * b]), xyz(a.dinvI[3 * b + 1]), xyz(a.dinvI[3 * b + 2]));
    return mulmm(mulmm(rot, invI), transpose(rot));
}

// The data of one row (JR_* fields), loaded at once so that a joint's next row can be in
// flight while the current one is solved.
struct JRow { float4 L, G0, G1, L0, L1, T0, T1, EX; };
__device__ __forceinline__ JRow load_jrow(const JointArgs& a, int j, int r) {
    JRow d;
    d.L = JROW(r, JR_LIN); d.G0 = JROW(r, JR_ANG0); d.G1 = JROW(r, JR_ANG1);
    d.L0 = JROW(r, JR_LIN0T); d.L1 = JROW(r, JR_LIN1T); d.T0 = JROW(r, JR_ANG0T); d.T1 = JROW(r, JR_ANG1T);
    d.EX = JROW(r, JR_EXTRA);
    return d;
}

// A joint's bodies as its rows' preSolve reads them: inverse mass and the world inverse
// inertia of the substep (massTemp), zero for a static or kinematic side.
struct JBodies { int b0, b1; float im0, im1; m3 I0, I1; };

// From the body records the velocity integration wrote this substep (the same
// R * invI * R^T operations as world_inv_inertia_j).
__device__ __forceinline__ JBodies joint_bodies(const JointArgs& a, int j) {
    JBodies B;
    B.b0 = a.jB0[j]; B.b1 = a.jB1[j];
    B.im0 = 0.f; B.im1 = 0.f; B.I0 = mzero(); B.I1 = mzero();
    if (B.b0 >= 0) {
        B.im0 = a.bvel[bv(B.b0, BV_CV)].w;
        B.I0 = mk_m3(xyz(a.bvel[bv(B.b0, BV_I0)]), xyz(a.bvel[bv(B.b0, BV_I1)]), xyz(a.bvel[bv(B.b0, BV_I2)]));
    }
    if (B.b1 >= 0) {
        B.im1 = a.bvel[bv(B.b1, BV_CV)].w;
        B.I1 = mk_m3(xyz(a.bvel[bv(B.b1, BV_I0)]), xyz(a.bvel[bv(B.b1, BV_I1)]), xyz(a.bvel[bv(B.b1, BV_I2)]));
    }
    return B;
}

__device__ void row_presolve_d(const JointArgs& a, int j, int r, int flags, bool simd, const JRow& d,
                               const JBodies& B, float4& pv0, float4& pw0, float4& pv1, float4& pw1);

// Constraint1D<flags>::preSolve / Constraint1DW<flags>::preSolve for one row, on the
// bodies' pseudo velocities held by the caller (pv, pw; pw.w = constraint count)
__device__ void row_presolve(const JointArgs& a, int j, int r, int flags, bool simd, float4& pv0, float4& pw0,
                             float4& pv1, float4& pw1) {
    JBodies B;
    B.b0 = a.jB0[j]; B.b1 = a.jB1[j];
    B.im0 = 0.f; B.im1 = 0.f; B.I0 = mzero(); B.I1 = mzero();
    if (B.b0 >= 0) { B.im0 = a.dvel[B.b0].w; B.I0 = world_inv_inertia_j(a, B.b0); }
    if (B.b1 >= 0) { B.im1 = a.dvel[B.b1].w; B.I1 = world_inv_inertia_j(a, B.b1); }
    row_presolve_d(a, j, r, flags, simd, load_jrow(a, j, r), B, pv0, pw0, pv1, pw1);
}

__device__ void row_presolve_d(const JointArgs& a, int j, int r, int flags, bool simd, const JRow& d,
                               const JBodies& B, float4& pv0, float4& pw0, float4& pv1, float4& pw1) {
    const int b0 = B.b0, b1 = B.b1;
    const float im0 = B.im0, im1 = B.im1;
    const m3 &I0 = B.I0, &I1 = B.I1;
    bool ang = flags & JF_ANGULAR;
    float4 L = d.L, G0 = d.G0, G1 = d.G1;
    f3 lin = xyz(L), ang0 = xyz(G0), ang1 = xyz(G1);
    float c = G0.w;
    f3 ang0t = mulmv(I0, ang0), ang1t = mulmv(I1, ang1);
    float invEff = dot(ang0, ang0t) + dot(ang1, ang1t);
    f3 lin0t = s3(0.f), lin1t = s3(0.f);
    float4 L0 = d.L0, L1 = d.L1;
    if (!ang) {
        invEff = invEff + dot(lin, lin) * (im0 + im1);
        lin0t = mul(im0, lin);
        lin1t = mul(im1, lin);
        set_xyz(L0, lin0t);
        set_xyz(L1, lin1t);
    }
    float4 T0 = d.T0, T1 = d.T1;
    set_xyz(T0, ang0t);
    set_xyz(T1, ang1t);
    G1.w = invEff;
    JROW(r, JR_ANG1) = G1; JROW(r, JR_LIN0T) = L0; JROW(r, JR_LIN1T) = L1; JROW(r, JR_ANG0T) = T0;
    JROW(r, JR_ANG1T) = T1;
    if (flags & JF_SOFT) return;
    if (c == 0.f || invEff == 0.f) return;
    float lambda = c / invEff;
    if (flags & JF_LIMITED) {
        float mn = L1.w, mx = T0.w;
        if (simd) { lambda = lambda > mn ? lambda : mn; lambda = lambda < mx ? lambda : mx; }
        else lambda = gclamp(lambda, mn, mx);
    }
    if (b0 >= 0) {
        if (!ang) set_xyz(pv0, add(xyz(pv0), mul(lambda, lin0t)));
        set_xyz(pw0, add(xyz(pw0), mul(lambda, ang0t)));
        pw0.w = pw0.w + 1.f;
    }
    if (b1 >= 0) {
        if (!ang) set_xyz(pv1, sub(xyz(pv1), mul(lambda, lin1t)));
        set_xyz(pw1, sub(xyz(pw1), mul(lambda, ang1t)));
        pw1.w = pw1.w + 1.f;
    }
}

// Constraint1DW<flags>::solve for one lane (src/Constraint1DW.cpp:117-233), on the bodies'
// solver velocities held by the caller (bv, bw; an angular row reads and writes only bw)
__device__ void row_solve_simd(const JointArgs& a, int j, int r, int flags, bool warm, const JRow& d, int b0, int b1,
                               f3& bv0, f3& bw0, f3& bv1, f3& bw1) {
    bool ang = flags & JF_ANGULAR;
    float h = a.h;
    f3 v0 = s3(0.f), w0 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
    if (b0 >= 0) { if (!ang) v0 = bv0; w0 = bw0; }
    if (b1 >= 0) { if (!ang) v1 = bv1; w1 = bw1; }
    float4 L = d.L, G0 = d.G0, G1 = d.G1;
    float4 L0 = d.L0, L1 = d.L1, T0 = d.T0, T1 = d.T1;
    float4 EX = d.EX;
    f3 lin0t = xyz(L0), lin1t = xyz(L1), ang0t = xyz(T0), ang1t = xyz(T1);
    float c = G0.w, invEff = G1.w, lam = L0.w;
    if (!(flags & JF_SOFT) && warm) {
        if (fabsf(c) < 1e-4f && fabsf(lam) < 10000.f) {
            float l2 = lam * 0.5f;
            if (!ang) { v0 = add(v0, mul(l2, lin0t)); v1 = sub(v1, mul(l2, lin1t)); }
            w0 = add(w0, mul(l2, ang0t));
            w1 = sub(w1, mul(l2, ang1t));
            lam = l2;
        }
    }
    if (invEff == 0.f) {
        L0.w = lam;
        JROW(r, JR_LIN0T) = L0;
        return;
    }
    float rv = dot(xyz(G1), w1) - dot(xyz(G0), w0);
    if (!ang) rv = rv + (dot(xyz(L), v1) - dot(xyz(L), v0));
    float effMass = 1.f / invEff;
    float lambda;
    if (flags & JF_SOFT) {
        float twoPi = 2.f * 3.14159265358979323846f;
        float af = twoPi * T1.w;
        float stiffness = af * af * effMass;
        float damping = 2.f * af * EX.x * effMass;
        float gamma = 1.f / (damping + h * stiffness);
        float beta = h * stiffness * gamma;
        lambda = (rv + beta * c / h) / (invEff + gamma / h);
    } else {
        float bias = (float)(0.2 / (double)h);
        lambda = (rv - L.w + bias * c) * effMass;
    }
    float prev = lam;
    if (flags & JF_LIMITED) {
        float t = lam + lambda;
        float lo = L1.w * h, hi = T0.w * h;
        t = t > lo ? t : lo;
        t = t < hi ? t : hi;
        lam = t;
        lambda = lam - prev;
    } else {
        lam = lam + lambda;
    }
    if (!ang) { v0 = add(v0, mul(lambda, lin0t)); v1 = sub(v1, mul(lambda, lin1t)); }
    w0 = add(w0, mul(lambda, ang0t));
    w1 = sub(w1, mul(lambda, ang1t));
    L0.w = lam;
    JROW(r, JR_LIN0T) = L0;
    if (b0 >= 0) {
        if (!ang) bv0 = v0;
        bw0 = w0;
    }
    if (b1 >= 0) {
        if (!ang) bv1 = v1;
        bw1 = w1;
    }
}

// Constraint1D<flags>::solve (src/Constraint1D.cpp:54-127)
__device__ void row_solve_scalar(const JointArgs& a, int j, int r, int flags, bool warm) {
    float4 G1 = JROW(r, JR_ANG1);
    float invEff = G1.w;
    if (invEff == 0.f) return;
    int b0 = a.jB0[j], b1 = a.jB1[j];
    bool ang = flags & JF_ANGULAR;
    float h = a.h;
    f3 v0 = s3(0.f), w0 = s3(0.f), v1 = s3(0.f), w1 = s3(0.f);
    if (b0 >= 0) { if (!ang) v0 = xyz(a.bvel[bv(b0, BV_V)]); w0 = xyz(a.bvel[bv(b0, BV_W)]); }
    if (b1 >= 0) { if (!ang) v1 = xyz(a.bvel[bv(b1, BV_V)]); w1 = xyz(a.bvel[bv(b1, BV_W)]); }
    float4 L = JROW(r, JR_LIN), G0 = JROW(r, JR_ANG0);
    float4 L0 = JROW(r, JR_LIN0T), L1 = JROW(r, JR_LIN1T), T0 = JROW(r, JR_ANG0T), T1 = JROW(r, JR_ANG1T);
    float4 EX = JROW(r, JR_EXTRA);
    f3 lin0t = xyz(L0), lin1t = xyz(L1), ang0t = xyz(T0), ang1t = xyz(T1);
    float c = G0.w, lam = L0.w;
    if (!(flags & JF_SOFT)) {
        if (warm && !((double)gabs(c) > 1e-4 || gabs(lam) > 10000.f)) {
            lam = lam * 0.5f;
            if (b0 >= 0) { if (!ang) v0 = add(v0, mul(lam, lin0t)); w0 = add(w0, mul(lam, ang0t)); }
            if (b1 >= 0) { if (!ang) v1 = sub(v1, mul(lam, lin1t)); w1 = sub(w1, mul(lam, ang1t)); }
        }
    }
    float rv = dot(xyz(G1), w1) - dot(xyz(G0), w0);
    if (!ang) rv = rv + (dot(xyz(L), v1) - dot(xyz(L), v0));
    float lambda;
    if (flags & JF_SOFT) {
        float af = 2.f * 3.14159265358979323846f * T1.w;
        float stiffness = af * af / invEff;
        float damping = 2.f * af * EX.x / invEff;
        float gamma = 1.f / (damping + h * stiffness);
        float beta = h * stiffness / (damping + h * stiffness);
        lambda = (rv + beta * c / h) / (invEff + gamma / h);
    } else {
        lambda = (rv - L.w + 0.2f * c / h) / invEff;
    }
    if (flags & JF_LIMITED) {
        float prev = lam;
        lam = lam + lambda;
        lam = gclamp(lam, L1.w, T0.w);
        lambda = lam - prev;
    } else {
        lam = lam + lambda;
    }
    L0.w = lam;
    JROW(r, JR_LIN0T) = L0;
    if (b0 >= 0) {
        if (!ang) { float4 t = a.bvel[bv(b0, BV_V)]; set_xyz(t, add(xyz(t), mul(lambda, lin0t))); a.bvel[bv(b0, BV_V)] = t; }
        float4 t = a.bvel[bv(b0, BV_W)]; set_xyz(t, add(xyz(t), mul(lambda, ang0t))); a.bvel[bv(b0, BV_W)] = t;
    }
    if (b1 >= 0) {
        if (!ang) { float4 t = a.bvel[bv(b1, BV_V)]; set_xyz(t, sub(xyz(t), mul(lambda, lin1t))); a.bvel[bv(b1, BV_V)] = t; }
        float4 t = a.bvel[bv(b1, BV_W)]; set_xyz(t, sub(xyz(t), mul(lambda, ang1t))); a.bvel[bv(b1, BV_W)] = t;
    }
}

// one joint of a SIMD colour, its rows in flag-type order. The joint's bodies belong to no
// other joint of the colour, so their velocities (or pseudo velocities) stay in registers
// across the joint's rows: one load and one store per body. UPFRONT loads every row of the
// joint at once (the loads overlap; about 200 registers, for a kernel bounded to few
// threads); otherwise each row is loaded while the one before it is solved.
template <bool UPFRONT>
__device__ void joint_color_one(const JointArgs& a, int j, int phase, int warm) {
    int n = a.jNRows[j];
    int b0 = a.jB0[j], b1 = a.jB1[j];
    // the joint's rows in solve order (flag type by flag type, creation order inside a
    // type), 4 bits per row
    const int rowFlags = a.jRowFlags[j];
    uint32_t seq = 0;
    int cnt = 0;
    for (int f = 0; f < 6; ++f)
        for (int r = 0; r < n; ++r)
            if (((rowFlags >> (4 * r)) & 15) == kFlagOrder[f]) seq |= (uint32_t)r << (4 * cnt++);
    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    // UPFRONT: rows[i] holds row i of the solve order; otherwise rows[0] the row being solved
    // and rows[1] the next one, in flight meanwhile
    constexpr int NR = UPFRONT ? JOINT_MAX_ROWS : 2;
    JRow rows[NR];
    if constexpr (UPFRONT) {
#pragma unroll
        for (int i = 0; i < NR; ++i)
            if (i < cnt) rows[i] = load_jrow(a, j, (seq >> (4 * i)) & 15);
    } else {
        if (cnt) rows[0] = load_jrow(a, j, seq & 15);
    }
    if (phase == 0) {
        const JBodies B = joint_bodies(a, j);
        float4 pv0 = z, pw0 = z, pv1 = z, pw1 = z;
        if (b0 >= 0) { pv0 = a.pvel[b0]; pw0 = a.pang[b0]; }
        if (b1 >= 0) { pv1 = a.pvel[b1]; pw1 = a.pang[b1]; }
        if constexpr (UPFRONT) {
#pragma unroll
            for (int i = 0; i < NR; ++i) {
                if (i >= cnt) break;
                const int r = (seq >> (4 * i)) & 15;
                row_presolve_d(a, j, r, (rowFlags >> (4 * r)) & 15, true, rows[i], B, pv0, pw0, pv1, pw1);
            }
        } else {
            for (int i = 0; i < cnt; ++i) {
                const int r = (seq >> (4 * i)) & 15;
                if (i + 1 < cnt) rows[1] = load_jrow(a, j, (seq >> (4 * (i + 1))) & 15);
                row_presolve_d(a, j, r, (rowFlags >> (4 * r)) & 15, true, rows[0], B, pv0, pw0, pv1, pw1);
                rows[0] = rows[1];
            }
        }
        if (b0 >= 0) { a.pvel[b0] = pv0; a.pang[b0] = pw0; }
        if (b1 >= 0) { a.pvel[b1] = pv1; a.pang[b1] = pw1; }
        return;
    }
    float4 tv0 = z, tw0 = z, tv1 = z, tw1 = z;
    if (b0 >= 0) { tv0 = a.bvel[bv(b0, BV_V)]; tw0 = a.bvel[bv(b0, BV_W)]; }
    if (b1 >= 0) { tv1 = a.bvel[bv(b1, BV_V)]; tw1 = a.bvel[bv(b1, BV_W)]; }
    f3 v0 = xyz(tv0), w0 = xyz(tw0), v1 = xyz(tv1), w1 = xyz(tw1);
    if constexpr (UPFRONT) {
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            if (i >= cnt) break;
            const int r = (seq >> (4 * i)) & 15;
            row_solve_simd(a, j, r, (rowFlags >> (4 * r)) & 15, warm != 0, rows[i], b0, b1, v0, w0, v1, w1);
        }
    } else {
        for (int i = 0; i < cnt; ++i) {
            const int r = (seq >> (4 * i)) & 15;
            if (i + 1 < cnt) rows[1] = load_jrow(a, j, (seq >> (4 * (i + 1))) & 15);
            row_solve_simd(a, j, r, (rowFlags >> (4 * r)) & 15, warm != 0, rows[0], b0, b1, v0, w0, v1, w1);
            rows[0] = rows[1];
        }
    }
    if (b0 >= 0) { set_xyz(tv0, v0); set_xyz(tw0, w0); a.bvel[bv(b0, BV_V)] = tv0; a.bvel[bv(b0, BV_W)] = tw0; }
    if (b1 >= 0) { set_xyz(tv1, v1); set_xyz(tw1, w1); a.bvel[bv(b1, BV_V)] = tv1; a.bvel[bv(b1, BV_W)] = tw1; }
}

// one SIMD colour: a thread per joint
__global__ void k_joint_color(JointArgs a, int begin, int end, int phase, int warm) {
    int k = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (k < end) joint_color_one<false>(a, a.order[k], phase, warm);
}

// Every SIMD colour (0-7) of a joint pass in one launch: joints that share a dynamic body,
// directly or through other joints, form an island, and one workgroup takes each island,
// its joints colour by colour with a workgroup barrier between colours. Islands share no
// dynamic body, so each body sees its joints in colour order, as with one launch per
// colour, and the floats are the same. islandCS: per island, 9 offsets into islandJoint
// (the start of each colour 0-7, then the island's end).
constexpr int ISLAND_THREADS = 64;
__global__ __launch_bounds__(ISLAND_THREADS) void k_joint_islands(JointArgs a, const int32_t* __restrict__ islandCS,
                                                                  const int32_t* __restrict__ islandJoint, int phase,
                                                                  int warm) {
    const int32_t* cs = islandCS + 9 * blockIdx.x;
    for (int c = 0; c < 8; ++c) {
        const int b = cs[c], e = cs[c + 1];
        if (b == e) continue;   // the same for the whole workgroup
        for (int k = b + (int)threadIdx.x; k < e; k += ISLAND_THREADS)
            joint_color_one<true>(a, islandJoint[k], phase, warm);
        __syncthreads();        // this colour's velocity stores are seen by the next colour's loads
    }
}

// the overflow colour: one thread, flag type by flag type over all its joints (which may
// share bodies), each row on the bodies' velocities in memory
__global__ void k_joint_overflow(JointArgs a, int begin, int end, int phase, int warm) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int f = 0; f < 6; ++f)
        for (int k = begin; k < end; ++k) {
            int j = a.order[k];
            int n = a.jNRows[j];
            int b0 = a.jB0[j], b1 = a.jB1[j];
            for (int r = 0; r < n; ++r) {
                int fl = row_flags(a, j, r);
                if (fl != kFlagOrder[f]) continue;
                if (phase == 0) {
                    float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
                    float4 pv0 = b0 >= 0 ? a.pvel[b0] : z, pw0 = b0 >= 0 ? a.pang[b0] : z;
                    float4 pv1 = b1 >= 0 ? a.pvel[b1] : z, pw1 = b1 >= 0 ? a.pang[b1] : z;
                    row_presolve(a, j, r, fl, false, pv0, pw0, pv1, pw1);
                    if (b0 >= 0) { a.pvel[b0] = pv0; a.pang[b0] = pw0; }
                    if (b1 >= 0) { a.pvel[b1] = pv1; a.pang[b1] = pw1; }
                } else {
                    row_solve_scalar(a, j, r, fl, warm != 0);
                }
            }
        }
}

static JointArgs make_jargs(World& w, float h) {
    JointArgs a;
    a.J = w.J; a.Jcap = w.jCap;
    a.jType = w.jType.p; a.jA0 = w.jA0.p; a.jA1 = w.jA1.p; a.jRowFlags = w.jRowFlags.p; a.jNRows = w.jNRows.p;
    a.jP0 = w.jP0.p; a.jQ0 = w.jQ0.p; a.jP1 = w.jP1.p; a.jQ1 = w.jQ1.p; a.jDrive = w.jDrive.p;
    a.jPar0 = w.jPar0.p; a.jPar1 = w.jPar1.p; a.jGear = w.jGear.p; a.jGearInit = w.jGearInit.p;
    a.jB0 = w.jB0.p; a.jB1 = w.jB1.p; a.jR0 = w.jR0.p; a.jR1 = w.jR1.p; a.jRow = w.jRow.p;
    a.bpos = w.bpos.p; a.bquat = w.bquat.p; a.dcom = w.dcom.p; a.dinvI = w.dinvI.p; a.dvel = w.dvel.p;
    a.bvel = w.bvel.p; a.pvel = w.pvel.p; a.pang = w.pang.p;
    a.order = w.jOrder.p;
    a.h = h;
    return a;
}

// The custom joints' world-space data at the current poses, on the host (waits for the device).
static const pgx_joint_world_space* custom_world_space(World& w, const JointArgs& a) {
    const int n = (int)w.custList.size();
    w.custDev.alloc(cust_floats(n));
    w.custHost.alloc(sizeof(float) * cust_floats(n));
    hipLaunchKernelGGL(k_custom_ws, dim3(div_up(n, 64)), dim3(64), 0, w.stream, a, w.custListDev.p, n, w.custDev.p);
    PGX_LAUNCH_CHECK();
    PGX_HIP(hipMemcpyAsync(w.custHost.p, w.custDev.p, sizeof(float) * CUST_WS * n, hipMemcpyDeviceToHost, w.stream));
    PGX_HIP(hipStreamSynchronize(w.stream));
    return (const pgx_joint_world_space*)w.custHost.p;
}

// Joint::getSolverDesc of every custom joint (src/Physecs.cpp:321-355): the step's rows and
// their flags, fresh rows on the device and on the host.
static void custom_layouts(World& w, const JointArgs& a) {
    if (!w.custLayout) throw PgxError{PGX_ERR_INVALID, "the world has custom joints and no joint callbacks"};
    const int n = (int)w.custList.size();
    const pgx_joint_world_space* ws = custom_world_space(w, a);
    int32_t* lay = (int32_t*)((float*)w.custHost.p + (size_t)CUST_WS * n);
    pgx_joint_row fresh{};
    fresh.min = -3.402823466e+38f;
    fresh.max = 3.402823466e+38f;
    for (int k = 0; k < n; ++k) {
        const int j = w.custList[k];
        int32_t fl[JOINT_MAX_ROWS] = {0};
        const int nr = w.custLayout(w.custUser, j, &ws[k], fl);
        if (nr < 0 || nr > JOINT_MAX_ROWS)
            throw PgxError{PGX_ERR_INVALID, "joint " + std::to_string(j) + ": the layout function returned " +
                                                std::to_string(nr) + " rows"};
        int packed = 0;
        for (int r = 0; r < nr; ++r) {
            const int f = fl[r];
            if (f != 0 && f != 1 && f != 2 && f != 4 && f != 3 && f != 6)
                throw PgxError{PGX_ERR_INVALID, "joint " + std::to_string(j) + ": row " + std::to_string(r) +
                                                    " has flags " + std::to_string(f)};
            packed |= f << (4 * r);
        }
        lay[CUST_LAYOUT * k] = nr;
        lay[CUST_LAYOUT * k + 1] = packed;
        w.jRowsTotal += nr - w.hjNRows[j];
        w.hjNRows[j] = nr;
        w.hjRowFlags[j] = packed;
        for (int r = 0; r < JOINT_MAX_ROWS; ++r) w.custRowsHost[(size_t)k * JOINT_MAX_ROWS + r] = fresh;
    }
    int32_t* layDev = (int32_t*)(w.custDev.p + (size_t)CUST_WS * n);
    PGX_HIP(hipMemcpyAsync(layDev, lay, sizeof(int32_t) * CUST_LAYOUT * n, hipMemcpyHostToDevice, w.stream));
    hipLaunchKernelGGL(k_custom_layout, dim3(div_up(n, 64)), dim3(64), 0, w.stream, a, w.custListDev.p, n, layDev);
    PGX_LAUNCH_CHECK();
}

// MakeConstraintsFunc of every custom joint (src/Physecs.cpp:430-439), its rows to the device
static void custom_rows(World& w, const JointArgs& a) {
    const int n = (int)w.custList.size();
    const pgx_joint_world_space* ws = custom_world_space(w, a);
    for (int k = 0; k < n; ++k)
        w.custRows(w.custUser, w.custList[k], &ws[k], &w.custRowsHost[(size_t)k * JOINT_MAX_ROWS],
                   w.hjNRows[w.custList[k]]);
    const size_t off = (size_t)(CUST_WS + CUST_LAYOUT) * n, bytes = sizeof(pgx_joint_row) * JOINT_MAX_ROWS * n;
    std::memcpy((float*)w.custHost.p + off, w.custRowsHost.data(), bytes);
    PGX_HIP(hipMemcpyAsync(w.custDev.p + off, (float*)w.custHost.p + off, bytes, hipMemcpyHostToDevice, w.stream));
    hipLaunchKernelGGL(k_custom_rows, dim3(div_up(n, 64)), dim3(64), 0, w.stream, a, w.custListDev.p, n,
                       w.custDev.p + off);
    PGX_LAUNCH_CHECK();
}

void jt_step_setup(World& w) {
    if (!w.J) return;
    JointArgs a = make_jargs(w, 0.f);
    hipLaunchKernelGGL(k_joint_setup, dim3(div_up(w.J, 256)), dim3(256), 0, w.stream, a);
    PGX_LAUNCH_CHECK();
    if (!w.custList.empty()) custom_layouts(w, a);
}

void jt_update(World& w, float h) {
    if (!w.J) return;
    JointArgs a = make_jargs(w, h);
    hipLaunchKernelGGL(k_joint_update, dim3(div_up(w.J, 256)), dim3(256), 0, w.stream, a);
    PGX_LAUNCH_CHECK();
    if (!w.custList.empty()) custom_rows(w, a);
}

// phase 0 = preSolve, 1 = solve (warm start on the first iteration)
void jt_colors(World& w, float h, int phase, bool warm) {
    if (!w.J) return;
    JointArgs a = make_jargs(w, h);
    if (w.jIslands > 0) {   // colours 0-7 by island in one launch, then the overflow colour
        hipLaunchKernelGGL(k_joint_islands, dim3(w.jIslands), dim3(ISLAND_THREADS), 0, w.stream, a, w.jIslandCS.p,
                           w.jIslandJoint.p, phase, warm ? 1 : 0);
        const int b = w.jColorStart[8], e = w.jColorStart[9];
        if (e > b)
            hipLaunchKernelGGL(k_joint_overflow, dim3(1), dim3(64), 0, w.stream, a, b, e, phase, warm ? 1 : 0);
        PGX_LAUNCH_CHECK();
        return;
    }
    for (int c = 0; c < 9; ++c) {
        int b = w.jColorStart[c], e = w.jColorStart[c + 1];
        if (e <= b) continue;
        if (c < 8)
            hipLaunchKernelGGL(k_joint_color, dim3(div_up(e - b, 64)), dim3(64), 0, w.stream, a, b, e, phase,
                               warm ? 1 : 0);
        else
            hipLaunchKernelGGL(k_joint_overflow, dim3(1), dim3(64), 0, w.stream, a, b, e, phase, warm ? 1 : 0);
    }
    PGX_LAUNCH_CHECK();
}

}  // namespace pgx
