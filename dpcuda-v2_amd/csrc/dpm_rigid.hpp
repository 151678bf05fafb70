// dpm_rigid.hpp -- per-particle and per-vertex rigid-body arithmetic shared by the
// single-system rigid bumpy kernels (dpm_bumpy.hip) and the batched engine
// (dpm_batch.hip), so that a rigid system carries the same bits on either path.
#pragma once
#include "dpm_device.hpp"
#include "dpm_trig.hpp"

namespace dpm {

// kernelMixRigidVelocitiesAndForces, rigid_bumpy/kernels.cu:89-121
__device__ __forceinline__ void rigid_mix_one(double2& v, double2 f, double& w, double t,
                                              double alpha) {
    double fn = sqrt(f.x * f.x + f.y * f.y);
    double vn = sqrt(v.x * v.x + v.y * v.y);
    double mr = 0.0;
    if (fn > 1e-16) {
        mr = vn / fn * alpha;
    } else {
        v.x = 0.0;
        v.y = 0.0;
    }
    double tn = fabs(t), wn = fabs(w), tmr = 0.0;
    if (tn > 1e-16) {
        tmr = wn / tn * alpha;
    } else {
        w = 0.0;
        tmr = 0.0;
    }
    v = make_double2(v.x * (1 - alpha) + f.x * mr, v.y * (1 - alpha) + f.y * mr);
    w = w * (1 - alpha) + t * tmr;
}

// kernelTranslateAndRotateVertices1 (rigid_bumpy/kernels.cu:152-201) for one vertex: the
// back-step rotation by dth0 about the old centre onto the back-stepped one (when
// uphill), then the rotation by dth about that centre onto the new one.  (s0, c0) and
// (s1, c1) are dpm_sincos of dth0 and dth.
__device__ __forceinline__ double2 bfire_vertex_one(double2 x, double2 old, double2 mid,
                                                    double2 now, double s0, double c0,
                                                    double s1, double c1, int uphill) {
    if (uphill) {
        double rx = x.x - old.x, ry = x.y - old.y;
        double rxr = rx * c0 - ry * s0;
        double ryr = rx * s0 + ry * c0;
        x = make_double2(mid.x + rxr, mid.y + ryr);
    }
    double rx = x.x - mid.x, ry = x.y - mid.y;
    double rxr = rx * c1 - ry * s1;
    double ryr = rx * s1 + ry * c1;
    return make_double2(now.x + rxr, now.y + ryr);
}

// Vertex k of a ring particle of centre c, radius r and angle a with nv vertices of
// radius rv (kernelInitializeVerticesOnParticles, rigid_bumpy/kernels.cu:1137-1162):
// on the circle of radius r - rv at angle k * 2 pi / nv + a.
__device__ __forceinline__ double2 ring_vertex(double2 c, double r, double a, int nv, double rv,
                                               int k) {
    double inc = 2 * M_PI / nv;
    double inner = r - rv;
    double va = k * inc + a;
    double sv, cv;
    dpm_sincos(va, &sv, &cv);
    return make_double2(c.x + inner * cv, c.y + inner * sv);
}

// Whether a vertex v at squared distance d2 from the centre of particle q (radius r_q,
// vertex radius rv) can have one of q's vertices within vc: every vertex of q lies at
// r_q - rv from its centre (to rounding of the rotations, far below the margin), so by
// the triangle inequality none can when d2 >= (r_q - rv + vc)^2; the test keeps a
// relative margin of 1e-6 and so only skips vertex scans that would list nothing.
__device__ __forceinline__ bool vertex_ring_near(double d2, double rq, double rv, double vc) {
    const double reach = rq - rv + vc;
    return d2 < reach * reach * (1.0 + 1e-6);
}

// Vertex rows through the centre's image.  For a vertex v and a particle q that passed
// vertex_ring_near, every vertex u of q lies within ring reach + (r_q - r_v) of v; when
// half the box exceeds that bound (vertex_shift_ok), the minimum image of v - u is the
// image of v - x_q: pbc(v, u) = (v - u) - L k with the same k.  image_index gives that k
// as pbc forms it (0 when |d| < L/2 on the fast path, where pbc returns d itself), and
// pbc_known applies it with pbc's own operations, so the bits are pbc's.
__device__ __forceinline__ double image_index(double d, double L, double invL, double halfL,
                                              int fast) {
    if (fast) return fabs(d) < halfL ? 0.0 : round(d * invL);
    return round(d / L);
}

__device__ __forceinline__ double pbc_known(double a, double b, double L, double k, int fast) {
    const double d = a - b;
    if (fast && k == 0.0) return d;
    return d - L * k;
}

// The half box a system needs for the image shortcut: ring reach plus ring radius of the
// largest particle (vertex_ring_near's relative margin included).
__device__ __host__ inline double vertex_shift_limit(double rmax, double rv, double vc) {
    return (2.0 * rmax - 2.0 * rv + vc) * (1.0 + 1e-6);
}

// The Fire::step moves of one rigid particle (src/integrator/fire.cu:20-40 with the
// rigid overrides, rigid_bumpy/kernels.cu:21-150): [uphill: updatePositions(-dt/2) with
// the decreased dt, zero v and omega] + updateVelocities(dt/2) + mix + updatePositions(dt).
// q, v, a, w are updated in place; old / mid are the centre before and after the
// back-step, dth0 / dth the two rotation angles.
__device__ __forceinline__ void rigid_fire_move(double4& q, double2& v, double& a, double& w,
                                                double2 f, double t, double moi, double dt,
                                                double alpha, int uphill, int rotation,
                                                double2& old, double2& mid, double& dth0,
                                                double& dth) {
    old = make_double2(q.x, q.y);
    dth0 = 0.0;
    if (uphill) {
        double bdt = -dt / 2.0;
        double ddx = v.x * bdt, ddy = v.y * bdt;
        dth0 = w * bdt;
        q.x = q.x + ddx;
        q.y = q.y + ddy;
        a = a + dth0;
        v = make_double2(0.0, 0.0);
        w = 0.0;
    }
    mid = make_double2(q.x, q.y);
    // updateVelocities(dt/2) (kernelUpdateRigidVelocities)
    double hdt = dt / 2.0;
    if (rotation) w += t * hdt / moi;
    else w = 0.0;
    v.x += f.x * hdt / q.w;
    v.y += f.y * hdt / q.w;
    rigid_mix_one(v, f, w, t, alpha);
    // updatePositions(dt) (kernelUpdateRigidPositions)
    double ddx = v.x * dt, ddy = v.y * dt, dda = w * dt;
    dth = dda;
    q.x = q.x + ddx;
    q.y = q.y + ddy;
    a = a + dda;
}

}  // namespace dpm
