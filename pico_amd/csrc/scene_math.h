// Scene arithmetic shared by the host builder and the device TLAS rebuild:
// both sides must produce the same bits (instance inverses feed the
// object-space rays; quantised boxes must contain the exact ones under the
// device's own fma).  Compiled without FP contraction on both sides.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

#include "device_types.h"

namespace pico {

#define PICO_HD __host__ __device__ inline

// Inverse of a column-major 4x4 by cofactor expansion in single precision
// (the reference inverts instance transforms in float, Source/Core/UpperLevelBVH.cpp:24);
// rows 0..2 of the inverse are returned.  The expansion order is fixed so that
// every consumer of the scene (device traversal, CPU checks) sees the same
// object-space rays bit for bit.
// Binned-SAH split decision of both BVH builders: a node of n triangles
// becomes a leaf (n <= 4) when n * area(node) <= kSahTraversal * area(node) +
// the best split's cost (sum over the two sides of count * area), i.e. the
// traversal of one more node costs half a triangle test.  On the instanced
// 1080p frame a traversal cost of 0.25 / 0.5 / 1 / 2 / 3 gave 2534 / 2533 /
// 2487 / 2337 / 2289 Mrays/s: with one triangle tested per traversal step and
// four children tested per node, smaller leaves pay.
constexpr float kSahTraversal = 0.5f;

PICO_HD bool affine_inverse_rows(const float* m, float inv[3][4]) {
    auto e = [&](int r, int c) { return m[c * 4 + r]; };
    auto minor3 = [&](int r0, int r1, int r2, int c0, int c1, int c2) {
        return e(r0, c0) * (e(r1, c1) * e(r2, c2) - e(r2, c1) * e(r1, c2)) -
               e(r0, c1) * (e(r1, c0) * e(r2, c2) - e(r2, c0) * e(r1, c2)) +
               e(r0, c2) * (e(r1, c0) * e(r2, c1) - e(r2, c0) * e(r1, c1));
    };
    const int rs[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    float cof[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            const float mn = minor3(rs[r][0], rs[r][1], rs[r][2], rs[c][0], rs[c][1], rs[c][2]);
            cof[r][c] = ((r + c) & 1) ? -mn : mn;
        }
    const float det = e(0, 0) * cof[0][0] + e(0, 1) * cof[0][1] + e(0, 2) * cof[0][2] + e(0, 3) * cof[0][3];
    if (!(fabsf(det) > 1e-30f) || !(det - det == 0.0f)) return false;  // tiny, or inf / NaN
    const float id = 1.0f / det;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) inv[r][c] = cof[c][r] * id;
    return true;
}

// World boxes of instances.  A point's world position through the
// column-major transform is ((M[r] x + M[4+r] y) + M[8+r] z) + M[12+r] per
// row r (no contraction), on the host and on the device alike; the finished
// box is widened by a relative epsilon so that float rounding of the
// transform never culls a hit.
PICO_HD void grow_world_box(const float* M, const float p[3], float lo[3], float hi[3]) {
    for (int r = 0; r < 3; ++r) {
        const float w = M[r] * p[0] + M[4 + r] * p[1] + M[8 + r] * p[2] + M[12 + r];
        lo[r] = fminf(lo[r], w);
        hi[r] = fmaxf(hi[r], w);
    }
}

PICO_HD void widen_world_box(float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) {
        const float e = 1e-5f * (fabsf(lo[k]) + fabsf(hi[k]) + 1.0f);
        lo[k] -= e;
        hi[k] += e;
    }
}

// World box of an instance from the 8 corners of its object-space box (the
// built-in shapes).  An instance of a triangle mesh is bounded by its
// vertices instead (grow_world_box over them, then widen_world_box): under a
// rotation the corners' box is much larger than the mesh's (on the benchmark
// scene the instance boxes' total area is 2.06 times the exact one), and a
// ray that enters an instance box it would have missed pays for the
// instance's transform and BLAS root.
PICO_HD void instance_world_box(const float* M, const float mlo[3], const float mhi[3], float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
    for (int c = 0; c < 8; ++c) {
        const float p[3] = {(c & 1) ? mhi[0] : mlo[0], (c & 2) ? mhi[1] : mlo[1], (c & 4) ? mhi[2] : mlo[2]};
        grow_world_box(M, p, lo, hi);
    }
    widen_world_box(lo, hi);
}

// Quantises `nc` child boxes into a DNode4 (references left empty): per axis
// origin = the children's union lower bound and the smallest power-of-two
// scale 2^e with 255 * 2^e >= extent, raised until every decoded plane
// fma(q, 2^e, origin) lies outside the exact box.
PICO_HD void quantise_children(const float (*lo)[3], const float (*hi)[3], int nc, DNode4& nd) {
    memset(&nd, 0, sizeof(nd));
    for (int i = 0; i < 4; ++i) nd.ref[i] = kEmptyRef;
    uint8_t* q[3] = {nd.qx, nd.qy, nd.qz};
    for (int k = 0; k < 3; ++k) {
        float o = INFINITY, m = -INFINITY;
        for (int i = 0; i < nc; ++i) { o = fminf(o, lo[i][k]); m = fmaxf(m, hi[i][k]); }
        if (nc == 0 || !(o <= m)) { o = 0.0f; m = 0.0f; }
        nd.origin[k] = o;
        const double ext = (double)m - (double)o;
        int e = -126;
        if (ext > 0.0) {
            int ex = 0;
            const double f = frexp(ext / 255.0, &ex);  // ext / 255 = f 2^ex, f in [0.5, 1)
            e = f > 0.5 ? ex : ex - 1;
        }
        e = e < -126 ? -126 : (e > 126 ? 126 : e);
        for (;; ++e) {
            const float sc = ldexpf(1.0f, e);
            bool ok = true;
            for (int i = 0; i < nc && ok; ++i) {
                int ql = (int)floor(((double)lo[i][k] - o) / sc);
                ql = ql < 0 ? 0 : (ql > 255 ? 255 : ql);
                while (ql > 0 && fmaf((float)ql, sc, o) > lo[i][k]) --ql;
                int qh = (int)ceil(((double)hi[i][k] - o) / sc);
                qh = qh < 0 ? 0 : (qh > 255 ? 255 : qh);
                while (qh < 255 && fmaf((float)qh, sc, o) < hi[i][k]) ++qh;
                if (fmaf((float)ql, sc, o) > lo[i][k] || fmaf((float)qh, sc, o) < hi[i][k]) ok = false;
                q[k][i] = (uint8_t)ql;
                q[k][4 + i] = (uint8_t)qh;
            }
            if (ok || e >= 126) break;
        }
        const float sc = ldexpf(1.0f, e);
        if (k == 0) nd.scale_x = sc; else if (k == 1) nd.scale_y = sc; else nd.scale_z = sc;
    }
}

}  // namespace pico
