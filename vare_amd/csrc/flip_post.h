// flip_post.h -- the frame post-process, FLIPMethod::postProcess (base/source/FLIPMethod.cu:300-405),
// on the sparse blocks of the step's frame: the surface/velocity splat, Godunov redistancing,
// curl, normalised gradient and its divergence, and the per-particle vorticity and curvature.
//
// Field storage follows the reference's use of the step's sparse fields: the level set goes to the
// count field (_sSurfField), the splatted velocity to the P2G velocity (_sVeloField), the curl and
// then the normals to the projected velocity (_sCompField), the curvature to the pressure
// (_sPresField); the splat's weight and position sums stay in registers.  Every per-cell operation is the
// reference's FP32 arithmetic in its order.  The reference's splat accumulates with float atomics
// (Rasterization.cu:343-372), its sums in an unspecified order; here they come in a fixed order.
#pragma once
#include "flip_kernels.h"

namespace vare {

constexpr float kSplatInfluence = 1.5f;   // ParticlesToSurfaceAndVelocity(..., 1.5f, 0.5f), FLIPMethod.cu:318
constexpr float kSplatDroplet = 0.5f;
constexpr float kSurfBackground = 10.f;   // _sSurfField background after the splat (Rasterization.cu:374)

// Rasterization::ParticlesToSurfaceAndVelocity (Rasterization.cu:311-393) as a per-cell gather.  A
// particle at p adds w = max(0, (1 - (d/1.5)^2)^3) (d = distance to the cell centre) times (1, p, v) to
// the 3x3x3 cells around (int)p (Grid::neighborCells with offset (int)1.5 = 1, clamped to the domain,
// Grid.h:353-370), so cell c receives from the particles with |(int)p - c| <= 1 on every axis.  The
// bins hold the particles by their interpolation corner max((int)(p - 0.5), 0), which is (int)p or
// (int)p - 1: the gather reads the bins c-2 .. c+1 on each axis and keeps the particles whose (int)p
// is in range.  The reference adds with float atomics in whatever order they land; here each cell adds
// in a fixed order (bins by z, y, x, each bin in its sorted order), so the splat is the same from run
// to run (except around a cell holding more than kSortMax particles, whose bin keeps its arrival
// order).  kernel2 (:375-392) follows in the same thread: level set |sum(w p)/sum(w) - centre| - 0.5
// (5 without weight) and velocity sum(w v)/sum(w) (0 without weight).
template <int B>
struct k_splat_gather_f {
    static constexpr int kLB = WG;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, const float* __restrict__ px, const float* __restrict__ py,
                                                         const float* __restrict__ pz, VelPtr u,
                                                         const int* __restrict__ cstart, FrameK f, float* __restrict__ surf,
                                                         float* __restrict__ ox, float* __restrict__ oy,
                                                         float* __restrict__ oz) {
        for (int c = vbid * WG + threadIdx.x; c < nc; c += vnbid * WG) {
            int i, j, k;
            cell_coords<B>(f, c, i, j, k);
            const float cxf = (float)i + 0.5f, cyf = (float)j + 0.5f, czf = (float)k + 0.5f;
            float sw = 0.f, sx = 0.f, sy = 0.f, sz = 0.f, su = 0.f, sv = 0.f, sr = 0.f;
            for (int bk = rmax(k - 2, 0); bk <= rmin(k + 1, f.nz - 1); ++bk)
                for (int bj = rmax(j - 2, 0); bj <= rmin(j + 1, f.ny - 1); ++bj)
                    for (int bi = rmax(i - 2, 0); bi <= rmin(i + 1, f.nx - 1); ++bi) {
                        const int bin = cell_index<B>(f, bi, bj, bk);
                        if (bin < 0) continue;                        // inactive block: no particles
                        const int q1 = cstart[bin + 1];
                        for (int q = cstart[bin]; q < q1; ++q) {
                            const float x = px[q], y = py[q], z = pz[q];
                            const int di = (int)x - i, dj = (int)y - j, dk = (int)z - k;
                            if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) continue;
                            const float dx = x - cxf, dy = y - cyf, dz = z - czf;
                            const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
                            const float r = dist / kSplatInfluence;
                            const float tt = 1.f - r * r;
                            const float w = rmax(0.f, tt * tt * tt);
                            if (!(w > 0.f)) continue;                 // adding 0 changes no sum
                            sw += w;
                            sx += w * x; sy += w * y; sz += w * z;
                            su += w * u.ld(0, q); sv += w * u.ld(1, q); sr += w * u.ld(2, q);
                        }
                    }
            if (sw > 1e-30f) {
                const float ex = sx / sw - cxf, ey = sy / sw - cyf, ez = sz / sw - czf;
                surf[c] = sqrtf(ex * ex + ey * ey + ez * ez) - kSplatDroplet;
                ox[c] = su / sw; oy[c] = sv / sw; oz[c] = sr / sw;
            } else {
                surf[c] = 5.f;
                ox[c] = 0.f; oy[c] = 0.f; oz[c] = 0.f;
            }
        }
    }
};
template <int B>
__global__ void __launch_bounds__(WG) k_splat_gather(int nc, const float* __restrict__ px, const float* __restrict__ py,
                                                     const float* __restrict__ pz, VelPtr u,
                                                     const int* __restrict__ cstart, FrameK f, float* __restrict__ surf,
                                                     float* __restrict__ ox, float* __restrict__ oy,
                                                     float* __restrict__ oz) {
    k_splat_gather_f<B>::run(blockIdx.x, gridDim.x, nc, px, py, pz, u, cstart, f, surf, ox, oy, oz);
}

// Value of a scalar field at neighbour q of cell c: the cell's value, the background for an inactive
// block, or (domain face) `edge`.
__device__ __forceinline__ float nbr_val(const float* __restrict__ f, int id, float bg, float edge) {
    return id >= 0 ? f[id] : (id == -1 ? bg : edge);
}

// One step of Rasterization::ReinitializeSurface (Rasterization.cu:405-485): dst = src -
// 0.3 max(S,0) (|grad+| - 1) - 0.3 min(S,0) (|grad-| - 1), S = v / sqrt(v^2 + 1), Godunov upwind
// differences; inactive neighbours read 10; domain faces extrapolate (2v - v_opp), or with the water
// tank hold the level's plane (cell.y - level, -1 / +1 below / above).
template <int B, bool TANK>
struct k_redistance_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, FrameK f, const float* __restrict__ src, float* __restrict__ dst, float level) {
        for (int c = vbid * WG + threadIdx.x; c < nc; c += vnbid * WG) {
            const float v = src[c];
            const float sign = v / sqrtf(v * v + 1.f);
            int nid[6];
    #pragma unroll
            for (int q = 0; q < 6; ++q) nid[q] = cell_neighbor<B>(f, c, q);
            float nv[6];
            if (TANK) {
                int i, j, k;
                cell_coords<B>(f, c, i, j, k);
                const float ll = ((float)j + 0.5f) - level;
    #pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const float e = q == 2 ? ll - 1.f : (q == 3 ? ll + 1.f : ll);
                    nv[q] = nbr_val(src, nid[q], kSurfBackground, e);
                }
            } else {
    #pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float lo = nbr_val(src, nid[2 * a], kSurfBackground, 0.f);
                    const float hi = nbr_val(src, nid[2 * a + 1], kSurfBackground, 0.f);
                    nv[2 * a] = nid[2 * a] == -2 ? 2.f * v - hi : lo;
                    nv[2 * a + 1] = nid[2 * a + 1] == -2 ? 2.f * v - lo : hi;
                }
            }
            const float dx0 = v - nv[0], dx1 = nv[1] - v;
            const float dy0 = v - nv[2], dy1 = nv[3] - v;
            const float dz0 = v - nv[4], dz1 = nv[5] - v;
            auto P2 = [](float a) { return a * a; };
            const float gp = sqrtf(P2(rmax(dx0, 0.f)) + P2(rmin(dx1, 0.f)) + P2(rmax(dy0, 0.f)) + P2(rmin(dy1, 0.f)) +
                                   P2(rmax(dz0, 0.f)) + P2(rmin(dz1, 0.f)));
            const float gm = sqrtf(P2(rmin(dx0, 0.f)) + P2(rmax(dx1, 0.f)) + P2(rmin(dy0, 0.f)) + P2(rmax(dy1, 0.f)) +
                                   P2(rmin(dz0, 0.f)) + P2(rmax(dz1, 0.f)));
            constexpr float dt = 0.3f;
            dst[c] = v - dt * rmax(sign, 0.f) * (gp - 1.f) - dt * rmin(sign, 0.f) * (gm - 1.f);
        }
    }
};
template <int B, bool TANK>
__global__ void k_redistance(int nc, FrameK f, const float* __restrict__ src, float* __restrict__ dst, float level) {
    k_redistance_f<B, TANK>::run(blockIdx.x, gridDim.x, nc, f, src, dst, level);
}

struct Vec3Field {
    const float* x; const float* y; const float* z;
};

// Neighbour values of a vector field along axis a with the reference's domain-face extrapolation
// (v*2 - v_opp) and background 0 for inactive blocks: lo, hi of component m.
template <int B>
__device__ __forceinline__ void vec_pair(const Vec3Field& v, const int* nid, int a, int m, float own, float& lo, float& hi) {
    const float* comp = m == 0 ? v.x : (m == 1 ? v.y : v.z);
    const float l = nbr_val(comp, nid[2 * a], 0.f, 0.f), h = nbr_val(comp, nid[2 * a + 1], 0.f, 0.f);
    lo = nid[2 * a] == -2 ? own * 2.f - h : l;
    hi = nid[2 * a + 1] == -2 ? own * 2.f - l : h;
}

// FieldOP::curl (FieldOP.cu:338-369) of v into out.
template <int B>
struct k_curl_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, FrameK f, Vec3Field v, float* __restrict__ ox, float* __restrict__ oy,
                           float* __restrict__ oz) {
        for (int c = vbid * WG + threadIdx.x; c < nc; c += vnbid * WG) {
            int nid[6];
    #pragma unroll
            for (int q = 0; q < 6; ++q) nid[q] = cell_neighbor<B>(f, c, q);
            const float own[3] = {v.x[c], v.y[c], v.z[c]};
            float x0y, x1y, x0z, x1z, y0x, y1x, y0z, y1z, z0x, z1x, z0y, z1y;
            vec_pair<B>(v, nid, 0, 1, own[1], x0y, x1y);
            vec_pair<B>(v, nid, 0, 2, own[2], x0z, x1z);
            vec_pair<B>(v, nid, 1, 0, own[0], y0x, y1x);
            vec_pair<B>(v, nid, 1, 2, own[2], y0z, y1z);
            vec_pair<B>(v, nid, 2, 0, own[0], z0x, z1x);
            vec_pair<B>(v, nid, 2, 1, own[1], z0y, z1y);
            ox[c] = (y1z - y0z) * 0.5f - (z1y - z0y) * 0.5f;
            oy[c] = (z1x - z0x) * 0.5f - (x1z - x0z) * 0.5f;
            oz[c] = (x1y - x0y) * 0.5f - (y1x - y0x) * 0.5f;
        }
    }
};
template <int B>
__global__ void k_curl(int nc, FrameK f, Vec3Field v, float* __restrict__ ox, float* __restrict__ oy,
                       float* __restrict__ oz) {
    k_curl_f<B>::run(blockIdx.x, gridDim.x, nc, f, v, ox, oy, oz);
}

// FieldOP::normalGradient (FieldOP.cu:80-110): normalised central-difference gradient of s (inactive
// neighbours read the background bg), Vector3::normalized (Vector3.h:350-354).
template <int B>
struct k_normal_gradient_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, FrameK f, const float* __restrict__ s, float bg, float* __restrict__ ox,
                                      float* __restrict__ oy, float* __restrict__ oz) {
        for (int c = vbid * WG + threadIdx.x; c < nc; c += vnbid * WG) {
            const float v = s[c];
            float g[3];
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int il = cell_neighbor<B>(f, c, 2 * a), ih = cell_neighbor<B>(f, c, 2 * a + 1);
                const float l = nbr_val(s, il, bg, 0.f), h = nbr_val(s, ih, bg, 0.f);
                const float lo = il == -2 ? 2.f * v - h : l;
                const float hi = ih == -2 ? 2.f * v - l : h;
                g[a] = (hi - lo) * 0.5f;
            }
            const float d = 1.f / sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + 1e-30f);
            ox[c] = g[0] * d; oy[c] = g[1] * d; oz[c] = g[2] * d;
        }
    }
};
template <int B>
__global__ void k_normal_gradient(int nc, FrameK f, const float* __restrict__ s, float bg, float* __restrict__ ox,
                                  float* __restrict__ oy, float* __restrict__ oz) {
    k_normal_gradient_f<B>::run(blockIdx.x, gridDim.x, nc, f, s, bg, ox, oy, oz);
}

// FieldOP::curvature (FieldOP.cu:273-301): central-difference divergence of n into out.
template <int B>
struct k_divergence_vec_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int nc, FrameK f, Vec3Field n, float* __restrict__ out) {
        for (int c = vbid * WG + threadIdx.x; c < nc; c += vnbid * WG) {
            int nid[6];
    #pragma unroll
            for (int q = 0; q < 6; ++q) nid[q] = cell_neighbor<B>(f, c, q);
            float lo[3], hi[3];
            vec_pair<B>(n, nid, 0, 0, n.x[c], lo[0], hi[0]);
            vec_pair<B>(n, nid, 1, 1, n.y[c], lo[1], hi[1]);
            vec_pair<B>(n, nid, 2, 2, n.z[c], lo[2], hi[2]);
            out[c] = (hi[0] - lo[0]) * 0.5f + (hi[1] - lo[1]) * 0.5f + (hi[2] - lo[2]) * 0.5f;
        }
    }
};
template <int B>
__global__ void k_divergence_vec(int nc, FrameK f, Vec3Field n, float* __restrict__ out) {
    k_divergence_vec_f<B>::run(blockIdx.x, gridDim.x, nc, f, n, out);
}

// SparseField::lerp (SparseField.h:150-173) of one scalar field at p, background bg.
template <int B>
__device__ __forceinline__ float lerp_scalar(const FrameK& f, const float* __restrict__ s, float bg, float x, float y,
                                             float z) {
    const int i = rclamp((int)(x - 0.5f), 0, f.nx - 2);
    const int j = rclamp((int)(y - 0.5f), 0, f.ny - 2);
    const int k = rclamp((int)(z - 0.5f), 0, f.nz - 2);
    const float fx = rclamp(x - ((float)i + 0.5f), 0.f, 1.f);
    const float fy = rclamp(y - ((float)j + 0.5f), 0.f, 1.f);
    const float fz = rclamp(z - ((float)k + 0.5f), 0.f, 1.f);
    float v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int c = cell_index<B>(f, i + (q & 1), j + ((q >> 1) & 1), k + (q >> 2));
        v[q] = c >= 0 ? s[c] : bg;
    }
    const float c00 = v[0] * (1.f - fx) + v[1] * fx;
    const float c01 = v[4] * (1.f - fx) + v[5] * fx;
    const float c10 = v[2] * (1.f - fx) + v[3] * fx;
    const float c11 = v[6] * (1.f - fx) + v[7] * fx;
    const float c0 = c00 * (1.f - fy) + c10 * fy;
    const float c1 = c01 * (1.f - fy) + c11 * fy;
    return c0 * (1.f - fz) + c1 * fz;
}

// vort_kernel (FLIPMethod.cu:375-385): v = max(0, |worldVector(curl lerp)| - 0.2); vort -= vort * dis;
// vort = max(v, vort).  Particles are in physical order, vorticity in the reference order (pid).
template <int B>
struct k_particle_vorticity_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int np, const float* __restrict__ px, const float* __restrict__ py,
                                         const float* __restrict__ pz, const int* __restrict__ pid, FrameK f, Vec3Field curl,
                                         float sx, float sy, float sz, float dis, float* __restrict__ vort) {
        const int q = vbid * WG + threadIdx.x;
        if (q >= np) return;
        const float x = px[q], y = py[q], z = pz[q];
        const float cx = lerp_scalar<B>(f, curl.x, 0.f, x, y, z) * sx;
        const float cy = lerp_scalar<B>(f, curl.y, 0.f, x, y, z) * sy;
        const float cz = lerp_scalar<B>(f, curl.z, 0.f, x, y, z) * sz;
        float v = sqrtf(cx * cx + cy * cy + cz * cz);
        v = rmax(0.f, v - 0.2f);
        const int lg = pid[q];
        float w = vort[lg];
        w -= w * dis;
        vort[lg] = rmax(v, w);
    }
};
template <int B>
__global__ void k_particle_vorticity(int np, const float* __restrict__ px, const float* __restrict__ py,
                                     const float* __restrict__ pz, const int* __restrict__ pid, FrameK f, Vec3Field curl,
                                     float sx, float sy, float sz, float dis, float* __restrict__ vort) {
    k_particle_vorticity_f<B>::run(blockIdx.x, gridDim.x, np, px, py, pz, pid, f, curl, sx, sy, sz, dis, vort);
}

// curv_kernel (FLIPMethod.cu:393-398): the curvature field's lerp (background 10) at each particle.
template <int B>
struct k_particle_curvature_f {
    static constexpr int kLB = 1024;
    static constexpr int kWPE = 1;
    static __device__ __forceinline__ void run(const unsigned vbid, const unsigned vnbid, int np, const float* __restrict__ px, const float* __restrict__ py,
                                         const float* __restrict__ pz, const int* __restrict__ pid, FrameK f,
                                         const float* __restrict__ kfield, float* __restrict__ curv) {
        const int q = vbid * WG + threadIdx.x;
        if (q >= np) return;
        curv[pid[q]] = lerp_scalar<B>(f, kfield, kSurfBackground, px[q], py[q], pz[q]);
    }
};
template <int B>
__global__ void k_particle_curvature(int np, const float* __restrict__ px, const float* __restrict__ py,
                                     const float* __restrict__ pz, const int* __restrict__ pid, FrameK f,
                                     const float* __restrict__ kfield, float* __restrict__ curv) {
    k_particle_curvature_f<B>::run(blockIdx.x, gridDim.x, np, px, py, pz, pid, f, kfield, curv);
}

}  // namespace vare
