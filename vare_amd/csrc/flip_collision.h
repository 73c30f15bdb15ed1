// flip_collision.h -- collision source queries on the device (CollisionSource.h:128-144).
//
// The source is a sparse field on its own block table over the simulation frame's geometry: signed
// distance in voxel units (background `bg` in inactive blocks), velocity in voxel units per second
// and the normalised SDF gradient (both 0 in inactive blocks).  Queries are the trilinear
// SparseField::lerp (SparseField.h:150-173) of those fields.  With no source set, or a source whose
// frame has no active block, every query returns the constant background (5 before the first
// update, CollisionSource.h:56) and zero vectors, exactly as the reference's size()==0 branches do.
#pragma once
#include "flip_device.h"

namespace vare {

struct ColK {
    const int* slot;       // per frame block: active slot of the collision field, or -1
    const float* phi;      // signed distance / voxel size at cell centres
    const float* vel[3];   // collision velocity, voxel units
    const float* nrm[3];   // normalised gradient, voxel space
    float bg;              // signed distance of inactive cells, or the constant when !active
    int active;            // 1: sampled field; 0: constant bg, zero velocity and normal
    // true when some query can return a negative distance (otherwise collisions never fire)
    __host__ __device__ bool any() const { return active != 0 || bg < 0.f; }
};

// Trilinear stencil of a voxel-space point: storage index of the 8 corners (dx + 2dy + 4dz) in the
// collision field (-1 = inactive) and the clamped weights (SparseField.h:152-161).
struct ColStencil {
    int c[8];
    float x, y, z;
};

template <int B>
__device__ __forceinline__ ColStencil col_stencil(const ColK& col, const FrameK& f, float px, float py, float pz) {
    ColStencil s;
    const int i = rclamp((int)(px - 0.5f), 0, f.nx - 2);
    const int j = rclamp((int)(py - 0.5f), 0, f.ny - 2);
    const int k = rclamp((int)(pz - 0.5f), 0, f.nz - 2);
    s.x = rclamp(px - ((float)i + 0.5f), 0.f, 1.f);
    s.y = rclamp(py - ((float)j + 0.5f), 0.f, 1.f);
    s.z = rclamp(pz - ((float)k + 0.5f), 0.f, 1.f);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int a = i + (q & 1), b = j + ((q >> 1) & 1), c = k + (q >> 2);
        const int sl = col.slot[((c / B) * f.fy + b / B) * f.fx + a / B];
        s.c[q] = sl < 0 ? -1 : sl * Blk<B>::B3 + ((c % B) * B + b % B) * B + a % B;
    }
    return s;
}

// x-then-y-then-z blend of 8 corner values (SparseField.h:165-172).
__device__ __forceinline__ float lerp8(const float v[8], float x, float y, float z) {
    const float c00 = v[0] * (1.f - x) + v[1] * x;
    const float c01 = v[4] * (1.f - x) + v[5] * x;
    const float c10 = v[2] * (1.f - x) + v[3] * x;
    const float c11 = v[6] * (1.f - x) + v[7] * x;
    const float c0 = c00 * (1.f - y) + c10 * y;
    const float c1 = c01 * (1.f - y) + c11 * y;
    return c0 * (1.f - z) + c1 * z;
}

// CollisionSource::signedDistance
template <int B>
__device__ __forceinline__ float col_phi(const ColK& col, const FrameK& f, float px, float py, float pz) {
    if (!col.active) return col.bg;
    const ColStencil s = col_stencil<B>(col, f, px, py, pz);
    float v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = s.c[q] >= 0 ? col.phi[s.c[q]] : col.bg;
    return lerp8(v, s.x, s.y, s.z);
}

// CollisionSource::velocity (which = 0) / normal (which = 1)
template <int B>
__device__ __forceinline__ void col_vec(const ColK& col, const FrameK& f, int which, float px, float py, float pz,
                                        float out[3]) {
    if (!col.active) {
        out[0] = 0.f; out[1] = 0.f; out[2] = 0.f;
        return;
    }
    const ColStencil s = col_stencil<B>(col, f, px, py, pz);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float* src = which ? col.nrm[a] : col.vel[a];
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = s.c[q] >= 0 ? src[s.c[q]] : 0.f;
        out[a] = lerp8(v, s.x, s.y, s.z);
    }
}

// Vector3::normalized (Vector3.h:354-358): v * (1 / sqrt(|v|^2 + 1e-30)), all in FP32.
__device__ __forceinline__ void normalize3(float v[3]) {
    const float d = 1.f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-30f);
    v[0] = v[0] * d;
    v[1] = v[1] * d;
    v[2] = v[2] * d;
}

// Classification of one cell (FLIPMethod.cu:242-260, curl disabled): fluid if it holds mass, else
// solid when the source's signed distance at the cell centre is negative, else air; where the
// distance is negative the collision velocity times col_vel_scale is added to the cell velocity.
template <int B>
__device__ __forceinline__ uint8_t classify_cell(const ColK& col, const FrameK& f, float col_scale, int i, int j, int k,
                                                 float m, float& vx, float& vy, float& vz) {
    const float cx = (float)i + 0.5f, cy = (float)j + 0.5f, cz = (float)k + 0.5f;   // Grid::cellCenter
    const float phi = col_phi<B>(col, f, cx, cy, cz);
    uint8_t t = m > 0.f ? kFluid : (phi < 0.f ? kSolid : kAir);
    if (phi < 0.f) {
        float v[3];
        col_vec<B>(col, f, 0, cx, cy, cz, v);
        vx += v[0] * col_scale;
        vy += v[1] * col_scale;
        vz += v[2] * col_scale;
    }
    return t;
}

}  // namespace vare
