// World-space AABB of one shape (src/BoundsUtil.cpp:23-85), shared by the broadphase's
// collider bounds and the overlap queries' query bounds.
#pragma once
#include "pgx_internal.h"

namespace pgx {

__device__ __forceinline__ void shape_bounds(int type, float4 prm, f3 pos, qf ori, const float4* __restrict__ cvVerts,
                                             const int4* __restrict__ cvMesh, const float4* __restrict__ tmVerts,
                                             const int4* __restrict__ tmMesh, f3& mn, f3& mx) {
    if (type == PGX_CONVEX_MESH || type == PGX_TRIANGLE_MESH) {
        // src/BoundsUtil.cpp:61-69 (convex: the padded vertex buffer) and :77-85 (triangle mesh)
        bool cvx = type == PGX_CONVEX_MESH;
        int4 mi = cvx ? cvMesh[(int)prm.w] : tmMesh[(int)prm.w];
        const float4* v = (cvx ? cvVerts : tmVerts) + mi.x;
        int n = cvx ? ((mi.y + 3) & ~3) : mi.y;
        f3 sc = cvx ? mk3(prm.x, prm.y, prm.z) : s3(1.f);
        mn = s3(3.402823466e+38f);
        mx = s3(-3.402823466e+38f);
        for (int i = 0; i < n; ++i) {
            f3 t = add(pos, qrot(ori, cvx ? mulv(sc, xyz(v[i])) : xyz(v[i])));
            mn = vmin(mn, t);
            mx = vmax(mx, t);
        }
    } else if (type == PGX_SPHERE) {
        float r = prm.x;
        mn = sub(pos, s3(r));
        mx = add(pos, s3(r));
    } else if (type == PGX_CAPSULE) {
        float hh = prm.x, r = prm.y;
        f3 p0 = add(pos, qrot(ori, mk3(0.f, hh, 0.f)));
        f3 p1 = add(pos, qrot(ori, mk3(0.f, -hh, 0.f)));
        mn = subs(vmin(p0, p1), r);
        mx = adds(vmax(p0, p1), r);
    } else {   // box
        m3 u = mat3_cast(ori);
        m3 au = mk_m3(vabs(u.c0), vabs(u.c1), vabs(u.c2));
        f3 wh = mulmv(au, mk3(prm.x, prm.y, prm.z));
        mn = sub(pos, wh);
        mx = add(pos, wh);
    }
}

}  // namespace pgx
