// upf_adjoint.hip -- differentiable PCISPH rollout and its adjoint on gfx950.
//
// The step is the static-boundary PCISPH step (CpuParticlePhysicsStaticBoundariesPCISPH,
// src/cpu/UnifiedPhysics.cpp:4220-4238, CPU boundary mode) with a fixed number
// of correction iterations and the kernels in closed form (KernelM4
// src/cpu/WeightingKernel.cpp:75-97, kernelPressureGrad :165-180,
// kernelViscosityLaplacian :190-200) rather than through the 300-entry tables,
// whose step values have no useful derivative:
//   F_i   = -sum_j (v_i - v_j) V0^2 mu L(r_ij) - m g e_y + B(x_i)  (fluid j)
//   with B the wall penalty force (CalculateBoundaryForcePerLiquidParticleHost,
//   src/cpu/UnifiedPhysics.cpp:2032-2069), k_b = 2 F_max / d_f per unit depth
//   within d_f of a face of the collision box;
//   iteration k (Fp = 0 before the first):
//     v*_i = v_i + (F_i + Fp_i) dt/m;  z_i = x_i + v*_i dt;  x*_i = clamp(z_i)
//     rho_i = m W(0) + sum_j [r*_ij < h] m W(r*_ij)
//     p_i  += delta sp(rho_i - rho0),  sp(s) = tau log(1 + exp(s / tau))
//     Fp_i  = -(m^2/rho0^2) sum_j (p_i + p_j) G(r_ij) (x_i - x_j)   (p_wall = 0)
//   v'_i = v_i + (F_i + Fp_i) dt/m;  y_i = x_i + v'_i dt;  x'_i = clamp(y_i).
// sp with tau = 1 kg/m^3 (0.1 % of rho0) is a smooth form of the CPU path's
// max(s, 0) (:5477-5478): a settling fluid keeps many particles at the kink,
// where rounding alone decides the branch and FP32 and FP64 gradients of a
// 64-step rollout differ by O(1).
// The CPU path's velocity reflection at the box (:4000-4044) jumps when a
// particle crosses a face; here the position clamp stays, the reflection is
// dropped and the penalty force pushes particles back, so that the rollout is
// a continuous function of its inputs.
// Particles keep their upload order; each step sorts z-index keys only to
// build the cell table and neighbour lists (ascending sorted order, so every
// sum runs in the order of the CPU restatement).
//
// The backward pass is the exact reverse of these passes with the neighbour
// sets held fixed, in gather form over the symmetric lists (no atomics):
//   dL/dp_i  += -(m^2/rho0^2) sum_j G (x_ij . (gFp_i - gFp_j))
//   dL/dx_i  += -(m^2/rho0^2) sum_j (p_i + p_j) H(x_ij) (gFp_i - gFp_j),
//               H(d) = G I + (G'/r) d d^T
//   dL/drho_i = delta sp'(rho_i - rho0) dL/dp_i
//   dL/dx*_i  = sum_j m W'(r*)/r* x*_ij (dL/drho_i + dL/drho_j)
// and the elementwise reverses of the prediction, the clamps and the
// integration.  Memory: the state every `ckpt` steps is kept; the backward
// pass recomputes one segment at a time, keeping its steps' lists and
// iteration states, then runs the segment in reverse.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "upf.h"
#include "upf_device.h"
#include "upf_internal.h"
#include "upf_kernels.h"
#include "upf_sort.h"

namespace upf {
namespace ad {

constexpr int AB = 256;  // threads per block

struct Kern {  // closed-form kernels and their radial derivatives, FP32
    float h, a, cs, cv;
    __device__ explicit Kern(float sr) : h(sr) {
        a = 2.546479089470325472f / (sr * sr * sr);  // 8 / (pi h^3)
        cs = -45.0f / (3.1415926535897932384626f * sr * sr * sr * sr * sr * sr);
        cv = 45.0f / (3.1415926535897932384626f * sr * sr * sr * sr * sr * sr);
    }
    __device__ float W(float r) const {
        if (!(r < h)) return 0.0f;
        const float s = r / h;
        if (s < 0.5f) return a * (1.0f - 6.0f * s * s + 6.0f * s * s * s);
        const float t = 1.0f - s;
        return a * (2.0f * t * t * t);
    }
    __device__ float dW(float r) const {
        if (!(r < h)) return 0.0f;
        const float s = r / h;
        if (s < 0.5f) return a / h * (-12.0f * s + 18.0f * s * s);
        const float t = 1.0f - s;
        return a / h * (-6.0f * t * t);
    }
    __device__ float G(float r) const {
        if (!(r < h) || r == 0.0f) return 0.0f;
        const float t = h - r;
        return cs * t * t / r;
    }
    __device__ float dG(float r) const {
        if (!(r < h) || r == 0.0f) return 0.0f;
        const float t = h - r;
        return cs * (-2.0f * t / r - t * t / (r * r));
    }
    __device__ float L(float r) const { return r < h ? cv * (h - r) : 0.0f; }
    __device__ float dL(float r) const { return r < h ? -cv : 0.0f; }
};

constexpr float kTau = 1.0f;  // width of the smooth pressure clamp, kg/m^3

__device__ inline float softplus(float s) {
    const float sx = s / kTau;
    return sx > 30.0f ? s : kTau * log1pf(expf(sx));
}
__device__ inline float softplus_grad(float s) { return 1.0f / (1.0f + expf(-s / kTau)); }

struct Consts {
    float m, rho0, dt, q, cvis, cp, delta, g, h;
    float fd, kb;  // wall penalty: depth and stiffness 2 F_max / d_f
    float cmin[3], cmax[3];
};

// wall penalty force on a liquid particle at (x, y, z) and its diagonal derivative
__device__ inline void wall_force(const Consts& C, float x, float y, float z, float* f, float* g) {
    const float c[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        f[a] = 0.0f;
        g[a] = 0.0f;
        if (c[a] < C.cmin[a] + C.fd) { f[a] = f[a] + (C.cmin[a] + C.fd - c[a]) * C.kb; g[a] = g[a] - C.kb; }
        if (c[a] > C.cmax[a] - C.fd) { f[a] = f[a] - (c[a] + C.fd - C.cmax[a]) * C.kb; g[a] = g[a] - C.kb; }
    }
}

__device__ inline bool is_fluid(const float4& x) { return __float_as_int(x.w) == UPF_LIQUID; }
__device__ inline float len3(float dx, float dy, float dz) { return sqrtf(dx * dx + dy * dy + dz * dz); }
__device__ inline float dot3(const float4& a, const float4& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline float ax(const float4& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// ---------------------------------------------------------- neighbour search
__global__ void k_keys(int n, const float4* __restrict__ x, upf_params P, uint32_t* __restrict__ keys, Ctl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = x[i];
    bool ok;
    keys[i] = dkey(p.x, p.y, p.z, P.scales, ok);
    if (!ok) atomicOr(&ctl->err, 1u << UPF_ERR_RANGE);
}

__global__ void k_clear_cells(uint2* __restrict__ cells, uint32_t nc) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) cells[c] = make_uint2(0u, 0u);
}

__global__ void k_cell_bounds(int n, const uint32_t* __restrict__ keys, uint32_t shift3, uint2* __restrict__ cells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = keys[i] >> shift3;
    if (i == 0 || (keys[i - 1] >> shift3) != c) cells[c].x = (uint32_t)i;
    if (i == n - 1 || (keys[i + 1] >> shift3) != c) cells[c].y = (uint32_t)(i + 1);
}

// neighbour ids of every fluid particle: the cells of its query box (the
// reference's BoxQuery) in Morton order, candidates in sorted order, r < h
__global__ void k_lists(int n, const float4* __restrict__ x, const uint32_t* __restrict__ perm,
                        const uint2* __restrict__ cells, upf_params P, uint32_t* __restrict__ nb,
                        int* __restrict__ cnt, Ctl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = x[i];
    if (!is_fluid(xi)) { cnt[i] = 0; return; }
    const float h = P.support_radius, hh = h * h;
    const int cap = P.max_neighbors;
    BoxCells bc;
    bc.init(xi.x, xi.y, xi.z, h, P.scales, P.block_shift);
    int c = 0;
    for (int e = 0; e < bc.count(); ++e) {
        uint32_t q[3];
        if (!bc.cell(e, q)) continue;
        if (q[0] >= P.grid_dim || q[1] >= P.grid_dim || q[2] >= P.grid_dim) continue;
        const uint2 r = cells[dmorton3(q[0], q[1], q[2])];
        for (uint32_t s = r.x; s < r.y; ++s) {
            const int j = (int)perm[s];
            if (j == i) continue;
            const float4 xj = x[j];
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            if (dx * dx + dy * dy + dz * dz < hh) {
                if (c < cap) nb[(size_t)c * n + i] = (uint32_t)j;
                ++c;
            }
        }
    }
    cnt[i] = c < cap ? c : cap;
    if (c > cap) atomicOr(&ctl->err, 1u << UPF_ERR_NBR_OVERFLOW);
}

// --------------------------------------------------------------- forward
__global__ void k_ext(int n, const float4* __restrict__ x, const float4* __restrict__ v, const uint32_t* __restrict__ nb,
                      const int* __restrict__ cnt, Consts C, float4* __restrict__ F, float4* __restrict__ Fp,
                      float* __restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = x[i];
    Fp[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    p[i] = 0.0f;
    if (!is_fluid(xi)) { F[i] = make_float4(0.f, 0.f, 0.f, 0.f); return; }
    const Kern kn(C.h);
    const float4 vi = v[i];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = (int)nb[(size_t)k * n + i];
        const float4 xj = x[j];
        if (!is_fluid(xj)) continue;
        const float r = len3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
        const float4 vj = v[j];
        const float s = C.cvis * kn.L(r);
        fx = fx - (vi.x - vj.x) * s;
        fy = fy - (vi.y - vj.y) * s;
        fz = fz - (vi.z - vj.z) * s;
    }
    fy = fy - C.m * C.g;
    float wf[3], wg[3];
    wall_force(C, xi.x, xi.y, xi.z, wf, wg);
    F[i] = make_float4(fx + wf[0], fy + wf[1], fz + wf[2], 0.0f);
}

__device__ inline float clampc(float c, float mn, float mx) {
    if (c < mn) c = mn;
    if (c > mx) c = mx;
    return c;
}

// prediction of iteration k: z (unclamped) to the tape, x* = clamp(z)
__global__ void k_predict(int n, const float4* __restrict__ x, const float4* __restrict__ v, const float4* __restrict__ F,
                          const float4* __restrict__ Fp, Consts C, float4* __restrict__ xs, float4* __restrict__ z,
                          Ctl* ctl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = x[i];
    if (!is_fluid(xi)) { xs[i] = xi; if (z) z[i] = xi; return; }
    const float4 vi = v[i], f = F[i], fp = Fp[i];
    const float vx = vi.x + (f.x + fp.x) * C.q, vy = vi.y + (f.y + fp.y) * C.q, vz = vi.z + (f.z + fp.z) * C.q;
    const float4 zz = make_float4(xi.x + vx * C.dt, xi.y + vy * C.dt, xi.z + vz * C.dt, xi.w);
    if (z) z[i] = zz;
    xs[i] = make_float4(clampc(zz.x, C.cmin[0], C.cmax[0]), clampc(zz.y, C.cmin[1], C.cmax[1]),
                        clampc(zz.z, C.cmin[2], C.cmax[2]), xi.w);
    if (!(zz.x == zz.x) || !(zz.y == zz.y) || !(zz.z == zz.z)) atomicOr(&ctl->err, 1u << UPF_ERR_NAN);
}

__global__ void k_density(int n, const float4* __restrict__ xs, const uint32_t* __restrict__ nb,
                          const int* __restrict__ cnt, Consts C, float* __restrict__ p, float* __restrict__ p_tape,
                          float* __restrict__ rho_tape) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = xs[i];
    if (!is_fluid(xi)) { if (p_tape) { p_tape[i] = 0.0f; rho_tape[i] = 0.0f; } return; }
    const Kern kn(C.h);
    float r_ = C.m * kn.W(0.0f);
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const float4 xj = xs[nb[(size_t)k * n + i]];
        r_ = r_ + C.m * kn.W(len3(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z));
    }
    const float pi = p[i] + C.delta * softplus(r_ - C.rho0);
    p[i] = pi;
    if (p_tape) { p_tape[i] = pi; rho_tape[i] = r_; }
}

__global__ void k_pforce(int n, const float4* __restrict__ x, const uint32_t* __restrict__ nb,
                         const int* __restrict__ cnt, Consts C, const float* __restrict__ p, float4* __restrict__ Fp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = x[i];
    if (!is_fluid(xi)) return;
    const Kern kn(C.h);
    const float pi = p[i];
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = (int)nb[(size_t)k * n + i];
        const float4 xj = x[j];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float r = len3(dx, dy, dz);
        const float pj = is_fluid(xj) ? p[j] : 0.0f;
        const float s = kn.G(r) * (pi + pj) * C.cp;
        fx = fx - dx * s;
        fy = fy - dy * s;
        fz = fz - dz * s;
    }
    Fp[i] = make_float4(fx, fy, fz, 0.0f);
}

__global__ void k_integrate(int n, float4* __restrict__ x, float4* __restrict__ v, const float4* __restrict__ F,
                            const float4* __restrict__ Fp, Consts C, float4* __restrict__ y_tape) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = x[i];
    if (!is_fluid(xi)) { if (y_tape) y_tape[i] = xi; return; }
    const float4 vi = v[i], f = F[i], fp = Fp[i];
    const float vn[3] = {vi.x + (f.x + fp.x) * C.q, vi.y + (f.y + fp.y) * C.q, vi.z + (f.z + fp.z) * C.q};
    const float xo[3] = {xi.x, xi.y, xi.z};
    float y[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) y[a] = xo[a] + vn[a] * C.dt;
    if (y_tape) y_tape[i] = make_float4(y[0], y[1], y[2], xi.w);
#pragma unroll
    for (int a = 0; a < 3; ++a) y[a] = clampc(y[a], C.cmin[a], C.cmax[a]);
    x[i] = make_float4(y[0], y[1], y[2], xi.w);
    v[i] = make_float4(vn[0], vn[1], vn[2], 0.0f);
}

// -------------------------------------------------------------- backward
__device__ inline bool outside(float c, int a, const Consts& C) { return c < C.cmin[a] || c > C.cmax[a]; }

// reverse of the integration: (gx, gv) = dL/d(x', v') -> nx, nv, gF, gFp
__global__ void kr_integrate(int n, const float4* __restrict__ y, const float4* __restrict__ gx,
                             const float4* __restrict__ gv, Consts C, float4* __restrict__ nx, float4* __restrict__ nv,
                             float4* __restrict__ gF, float4* __restrict__ gFp, float* __restrict__ gp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 yi = y[i];
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    gp[i] = 0.0f;
    if (!is_fluid(yi)) { nx[i] = z4; nv[i] = z4; gF[i] = z4; gFp[i] = z4; return; }
    const float4 a = gx[i], b = gv[i];
    float gy[3], gvn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        gy[k] = outside(ax(yi, k), k, C) ? 0.0f : ax(a, k);
        gvn[k] = ax(b, k);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) gvn[k] = gvn[k] + gy[k] * C.dt;
    nx[i] = make_float4(gy[0], gy[1], gy[2], 0.0f);
    nv[i] = make_float4(gvn[0], gvn[1], gvn[2], 0.0f);
    const float4 f = make_float4(gvn[0] * C.q, gvn[1] * C.q, gvn[2] * C.q, 0.0f);
    gF[i] = f;
    gFp[i] = f;
}

// reverse of the pressure force of one iteration: dL/dp (running) and dL/dx
__global__ void kr_pforce(int n, const float4* __restrict__ x, const uint32_t* __restrict__ nb,
                          const int* __restrict__ cnt, Consts C, const float* __restrict__ p,
                          const float4* __restrict__ gFp, float* __restrict__ gp, float4* __restrict__ nx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = x[i];
    if (!is_fluid(xi)) return;
    const Kern kn(C.h);
    const float4 gi = gFp[i];
    const float pi = p[i];
    float acc = 0.0f, gxx = 0.0f, gxy = 0.0f, gxz = 0.0f;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = (int)nb[(size_t)k * n + i];
        const float4 xj = x[j];
        const float4 d = make_float4(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z, 0.0f);
        const float r = len3(d.x, d.y, d.z);
        const bool fj = is_fluid(xj);
        const float4 gj = fj ? gFp[j] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 w = fj ? make_float4(gi.x - gj.x, gi.y - gj.y, gi.z - gj.z, 0.0f) : gi;
        const float G = kn.G(r);
        const float dw = dot3(d, w);
        acc = acc - C.cp * G * dw;
        const float pp = pi + (fj ? p[j] : 0.0f);
        const float s = r > 0.0f ? kn.dG(r) / r * dw : 0.0f;
        const float t = C.cp * pp;
        gxx = gxx - (w.x * G + d.x * s) * t;
        gxy = gxy - (w.y * G + d.y * s) * t;
        gxz = gxz - (w.z * G + d.z * s) * t;
    }
    gp[i] = gp[i] + acc;
    float4 o = nx[i];
    o.x = o.x + gxx; o.y = o.y + gxy; o.z = o.z + gxz;
    nx[i] = o;
}

// reverse of density, pressure update and prediction of one iteration
__global__ void kr_density(int n, const float4* __restrict__ x, const float4* __restrict__ z,
                           const uint32_t* __restrict__ nb, const int* __restrict__ cnt, Consts C,
                           const float* __restrict__ rho, const float* __restrict__ gp, float4* __restrict__ nx,
                           float4* __restrict__ nv, float4* __restrict__ gF, float4* __restrict__ gFp_prev) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = x[i];
    if (!is_fluid(xi)) return;
    const Kern kn(C.h);
    const float4 zi = z[i];
    const float xsi[3] = {clampc(zi.x, C.cmin[0], C.cmax[0]), clampc(zi.y, C.cmin[1], C.cmax[1]),
                          clampc(zi.z, C.cmin[2], C.cmax[2])};
    const float gri = C.delta * gp[i] * softplus_grad(rho[i] - C.rho0);
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = (int)nb[(size_t)k * n + i];
        const float4 xj = x[j];
        float xsj[3];
        float grj = 0.0f;
        if (is_fluid(xj)) {
            const float4 zj = z[j];
            xsj[0] = clampc(zj.x, C.cmin[0], C.cmax[0]);
            xsj[1] = clampc(zj.y, C.cmin[1], C.cmax[1]);
            xsj[2] = clampc(zj.z, C.cmin[2], C.cmax[2]);
            grj = C.delta * gp[j] * softplus_grad(rho[j] - C.rho0);
        } else {
            xsj[0] = xj.x; xsj[1] = xj.y; xsj[2] = xj.z;
        }
        const float dx = xsi[0] - xsj[0], dy = xsi[1] - xsj[1], dz = xsi[2] - xsj[2];
        const float r = len3(dx, dy, dz);
        if (!(r < kn.h) || r == 0.0f) continue;
        const float gr = gri + grj;
        const float s = C.m * kn.dW(r) / r * gr;
        gx = gx + dx * s;
        gy = gy + dy * s;
        gz = gz + dz * s;
    }
    const float g3[3] = {gx, gy, gz};
    float zb[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) zb[a] = outside(ax(zi, a), a, C) ? 0.0f : g3[a];
    float4 o = nx[i];
    o.x = o.x + zb[0]; o.y = o.y + zb[1]; o.z = o.z + zb[2];
    nx[i] = o;
    const float vs[3] = {zb[0] * C.dt, zb[1] * C.dt, zb[2] * C.dt};
    float4 ov = nv[i];
    ov.x = ov.x + vs[0]; ov.y = ov.y + vs[1]; ov.z = ov.z + vs[2];
    nv[i] = ov;
    float4 of = gF[i];
    of.x = of.x + vs[0] * C.q; of.y = of.y + vs[1] * C.q; of.z = of.z + vs[2] * C.q;
    gF[i] = of;
    gFp_prev[i] = make_float4(vs[0] * C.q, vs[1] * C.q, vs[2] * C.q, 0.0f);
}

// reverse of the viscosity force
__global__ void kr_ext(int n, const float4* __restrict__ x, const float4* __restrict__ v, const uint32_t* __restrict__ nb,
                       const int* __restrict__ cnt, Consts C, const float4* __restrict__ gF, float4* __restrict__ nx,
                       float4* __restrict__ nv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = x[i];
    if (!is_fluid(xi)) return;
    const Kern kn(C.h);
    const float4 vi = v[i], gi = gF[i];
    float gvx = 0.0f, gvy = 0.0f, gvz = 0.0f, gxx = 0.0f, gxy = 0.0f, gxz = 0.0f;
    const int c = cnt[i];
    for (int k = 0; k < c; ++k) {
        const int j = (int)nb[(size_t)k * n + i];
        const float4 xj = x[j];
        if (!is_fluid(xj)) continue;
        const float4 d = make_float4(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z, 0.0f);
        const float r = len3(d.x, d.y, d.z);
        const float4 gj = gF[j];
        const float4 w = make_float4(gi.x - gj.x, gi.y - gj.y, gi.z - gj.z, 0.0f);
        const float sl = C.cvis * kn.L(r);
        gvx = gvx - w.x * sl;
        gvy = gvy - w.y * sl;
        gvz = gvz - w.z * sl;
        if (r > 0.0f) {
            const float4 vj = v[j];
            const float4 dv = make_float4(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, 0.0f);
            const float s = C.cvis * kn.dL(r) / r * dot3(dv, w);
            gxx = gxx - d.x * s;
            gxy = gxy - d.y * s;
            gxz = gxz - d.z * s;
        }
    }
    float wf[3], wg[3];
    wall_force(C, xi.x, xi.y, xi.z, wf, wg);
    gxx = gxx + gi.x * wg[0];
    gxy = gxy + gi.y * wg[1];
    gxz = gxz + gi.z * wg[2];
    float4 ov = nv[i];
    ov.x = ov.x + gvx; ov.y = ov.y + gvy; ov.z = ov.z + gvz;
    nv[i] = ov;
    float4 o = nx[i];
    o.x = o.x + gxx; o.y = o.y + gxy; o.z = o.z + gxz;
    nx[i] = o;
}

}  // namespace ad
}  // namespace upf

using namespace upf;
using namespace upf::ad;

#define ADCHK(x)                                   \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) return UPF_ERR_HIP;  \
    } while (0)

struct upf_adjoint {
    upf_params P;
    int n = 0, steps = 0, K = 0, ckpt = 0, nseg = 0;
    bool have_forward = false;
    hipStream_t st = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float fwd_ms = 0.0f, bwd_ms = 0.0f;
    std::vector<void*> allocs;
    // state and neighbour search
    float4 *x = nullptr, *v = nullptr;
    uint32_t *keys = nullptr, *skeys = nullptr, *perm = nullptr;
    uint2* cells = nullptr;
    uint32_t ncells = 0;
    void* sort_tmp = nullptr;
    Ctl* ctl = nullptr;
    // forward work
    float4 *F = nullptr, *Fp = nullptr, *xs = nullptr;
    float* p = nullptr;
    // checkpoints (x, v) at the start of every segment
    float4 *ck_x = nullptr, *ck_v = nullptr;
    // segment tape: per step x, v, y, lists; per iteration z, p, rho
    float4 *t_x = nullptr, *t_v = nullptr, *t_y = nullptr, *t_z = nullptr;
    uint32_t* t_nb = nullptr;
    int* t_cnt = nullptr;
    float *t_p = nullptr, *t_rho = nullptr;
    // adjoint state
    float4 *gx = nullptr, *gv = nullptr, *nx = nullptr, *nv = nullptr, *gF = nullptr, *gFp = nullptr, *gFp2 = nullptr;
    float* gp = nullptr;

    template <class T> hipError_t alloc(T** ptr, size_t count) {
        hipError_t e = hipMalloc((void**)ptr, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) allocs.push_back((void*)*ptr);
        return e;
    }
    Consts consts() const {
        Consts C;
        C.m = P.initial_mass;
        C.rho0 = P.rest_density;
        C.dt = P.dt;
        C.q = P.dt / P.initial_mass;
        C.cvis = (P.initial_mass / P.rest_density) * (P.initial_mass / P.rest_density) * P.visc_const;
        C.cp = (P.initial_mass * P.initial_mass) / (P.rest_density * P.rest_density);
        C.delta = P.density_error_factor;
        C.g = P.gravity;
        C.h = P.support_radius;
        C.fd = P.force_distance;
        C.kb = 2.0f * P.max_boundary_force / P.force_distance;
        for (int a = 0; a < 3; ++a) {
            C.cmin[a] = P.collision_min[a];
            C.cmax[a] = P.collision_max[a];
        }
        return C;
    }
    dim3 grid() const { return dim3((unsigned)((n + AB - 1) / AB)); }
    size_t nb_stride() const { return (size_t)P.max_neighbors * n; }
};

// lists of the positions in a->x into (nb, cnt)
static void enqueue_lists(upf_adjoint* a, uint32_t* nb, int* cnt) {
    const int n = a->n;
    hipLaunchKernelGGL(k_keys, a->grid(), dim3(AB), 0, a->st, n, a->x, a->P, a->keys, a->ctl);
    (void)radix_sort(a->keys, nullptr, a->skeys, a->perm, n, 30, a->sort_tmp, a->st);
    hipLaunchKernelGGL(k_clear_cells, dim3((a->ncells + AB - 1) / AB), dim3(AB), 0, a->st, a->cells, a->ncells);
    hipLaunchKernelGGL(k_cell_bounds, a->grid(), dim3(AB), 0, a->st, n, a->skeys, 3u * a->P.block_shift, a->cells);
    hipLaunchKernelGGL(k_lists, a->grid(), dim3(AB), 0, a->st, n, a->x, a->perm, a->cells, a->P, nb, cnt, a->ctl);
}

// one forward step of the state (a->x, a->v); with `tape` >= 0 the step's
// lists and iteration states go to tape slot `tape`
static void enqueue_step(upf_adjoint* a, int tape) {
    const int n = a->n, K = a->K;
    const Consts C = a->consts();
    const dim3 g = a->grid();
    const bool rec = tape >= 0;
    uint32_t* nb = rec ? a->t_nb + (size_t)tape * a->nb_stride() : a->t_nb;
    int* cnt = rec ? a->t_cnt + (size_t)tape * n : a->t_cnt;
    if (rec) {
        (void)hipMemcpyAsync(a->t_x + (size_t)tape * n, a->x, (size_t)n * 16, hipMemcpyDeviceToDevice, a->st);
        (void)hipMemcpyAsync(a->t_v + (size_t)tape * n, a->v, (size_t)n * 16, hipMemcpyDeviceToDevice, a->st);
    }
    enqueue_lists(a, nb, cnt);
    hipLaunchKernelGGL(k_ext, g, dim3(AB), 0, a->st, n, a->x, a->v, nb, cnt, C, a->F, a->Fp, a->p);
    for (int k = 0; k < K; ++k) {
        const size_t o = ((size_t)tape * K + k) * n;
        hipLaunchKernelGGL(k_predict, g, dim3(AB), 0, a->st, n, a->x, a->v, a->F, a->Fp, C, a->xs,
                           rec ? a->t_z + o : nullptr, a->ctl);
        hipLaunchKernelGGL(k_density, g, dim3(AB), 0, a->st, n, a->xs, nb, cnt, C, a->p, rec ? a->t_p + o : nullptr,
                           rec ? a->t_rho + o : nullptr);
        hipLaunchKernelGGL(k_pforce, g, dim3(AB), 0, a->st, n, a->x, nb, cnt, C, a->p, a->Fp);
    }
    hipLaunchKernelGGL(k_integrate, g, dim3(AB), 0, a->st, n, a->x, a->v, a->F, a->Fp, C,
                       rec ? a->t_y + (size_t)tape * n : nullptr);
}

// reverse of the step in tape slot `tape`: (gx, gv) -> (gx, gv)
static void enqueue_step_back(upf_adjoint* a, int tape) {
    const int n = a->n, K = a->K;
    const Consts C = a->consts();
    const dim3 g = a->grid();
    const float4* X = a->t_x + (size_t)tape * n;
    const float4* V = a->t_v + (size_t)tape * n;
    const uint32_t* nb = a->t_nb + (size_t)tape * a->nb_stride();
    const int* cnt = a->t_cnt + (size_t)tape * n;
    hipLaunchKernelGGL(kr_integrate, g, dim3(AB), 0, a->st, n, a->t_y + (size_t)tape * n, a->gx, a->gv, C, a->nx, a->nv,
                       a->gF, a->gFp, a->gp);
    for (int k = K - 1; k >= 0; --k) {
        const size_t o = ((size_t)tape * K + k) * n;
        hipLaunchKernelGGL(kr_pforce, g, dim3(AB), 0, a->st, n, X, nb, cnt, C, a->t_p + o, a->gFp, a->gp, a->nx);
        hipLaunchKernelGGL(kr_density, g, dim3(AB), 0, a->st, n, X, a->t_z + o, nb, cnt, C, a->t_rho + o, a->gp, a->nx,
                           a->nv, a->gF, a->gFp2);
        std::swap(a->gFp, a->gFp2);
    }
    hipLaunchKernelGGL(kr_ext, g, dim3(AB), 0, a->st, n, X, V, nb, cnt, C, a->gF, a->nx, a->nv);
    std::swap(a->gx, a->nx);
    std::swap(a->gv, a->nv);
}

static int adj_error(upf_adjoint* a) {
    Ctl h;
    ADCHK(hipMemcpyAsync(&h, a->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, a->st));
    ADCHK(hipStreamSynchronize(a->st));
    const int order[] = {UPF_ERR_NAN, UPF_ERR_RANGE, UPF_ERR_NBR_OVERFLOW};
    for (int c : order)
        if (h.err & (1u << c)) return c;
    return UPF_OK;
}

extern "C" {

int upf_adjoint_destroy(upf_adjoint* a) {
    if (!a) return UPF_ERR_ARG;
    if (a->st) (void)hipStreamSynchronize(a->st);
    for (void* p : a->allocs) (void)hipFree(p);
    for (auto& e : a->ev)
        if (e) (void)hipEventDestroy(e);
    if (a->st) (void)hipStreamDestroy(a->st);
    delete a;
    return UPF_OK;
}

int upf_adjoint_create(const upf_params* p, int n, int steps, int iterations, int checkpoint_every, upf_adjoint** out) {
    if (!p || !out || n <= 0 || steps <= 0 || iterations <= 0 || checkpoint_every <= 0) return UPF_ERR_ARG;
    if (upf::check_params(p) != UPF_OK) return UPF_ERR_ARG;
    if (p->add_wall_weight) return UPF_ERR_ARG;  // the differentiable step has no wall weight
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPF_ERR_NO_DEVICE;
    upf_adjoint* a = new upf_adjoint();
    a->P = *p;
    a->n = n;
    a->steps = steps;
    a->K = iterations;
    a->ckpt = std::min(checkpoint_every, steps);
    a->nseg = (steps + a->ckpt - 1) / a->ckpt;
    a->ncells = p->grid_dim * p->grid_dim * p->grid_dim;
    const size_t T = (size_t)a->ckpt;
    bool ok = hipStreamCreateWithFlags(&a->st, hipStreamNonBlocking) == hipSuccess;
    for (auto& e : a->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    ok = ok && a->alloc(&a->x, n) == hipSuccess && a->alloc(&a->v, n) == hipSuccess &&
         a->alloc(&a->keys, n) == hipSuccess && a->alloc(&a->skeys, n) == hipSuccess &&
         a->alloc(&a->perm, n) == hipSuccess && a->alloc(&a->cells, a->ncells) == hipSuccess &&
         a->alloc(&a->ctl, 1) == hipSuccess && a->alloc(&a->F, n) == hipSuccess && a->alloc(&a->Fp, n) == hipSuccess &&
         a->alloc(&a->xs, n) == hipSuccess && a->alloc(&a->p, n) == hipSuccess &&
         a->alloc(&a->ck_x, (size_t)a->nseg * n) == hipSuccess && a->alloc(&a->ck_v, (size_t)a->nseg * n) == hipSuccess &&
         a->alloc(&a->t_x, T * n) == hipSuccess && a->alloc(&a->t_v, T * n) == hipSuccess &&
         a->alloc(&a->t_y, T * n) == hipSuccess && a->alloc(&a->t_z, T * a->K * n) == hipSuccess &&
         a->alloc(&a->t_nb, T * a->nb_stride()) == hipSuccess && a->alloc(&a->t_cnt, T * n) == hipSuccess &&
         a->alloc(&a->t_p, T * a->K * n) == hipSuccess && a->alloc(&a->t_rho, T * a->K * n) == hipSuccess &&
         a->alloc(&a->gx, n) == hipSuccess && a->alloc(&a->gv, n) == hipSuccess && a->alloc(&a->nx, n) == hipSuccess &&
         a->alloc(&a->nv, n) == hipSuccess && a->alloc(&a->gF, n) == hipSuccess && a->alloc(&a->gFp, n) == hipSuccess &&
         a->alloc(&a->gFp2, n) == hipSuccess && a->alloc(&a->gp, n) == hipSuccess;
    if (ok) {
        void* tmp = nullptr;
        ok = hipMalloc(&tmp, radix_sort_temp_bytes(n)) == hipSuccess;
        if (ok) { a->sort_tmp = tmp; a->allocs.push_back(tmp); }
    }
    if (!ok) {
        upf_adjoint_destroy(a);
        return UPF_ERR_HIP;
    }
    *out = a;
    return UPF_OK;
}

int upf_adjoint_forward(upf_adjoint* a, const float* pos4, const float* vel4, const int* type, float* pos_out4,
                        float* vel_out4) {
    if (!a || !pos4 || !type) return UPF_ERR_ARG;
    const int n = a->n;
    std::vector<float> x((size_t)4 * n), v((size_t)4 * n, 0.0f);
    for (int i = 0; i < n; ++i) {
        if (type[i] != UPF_LIQUID && type[i] != UPF_FROZEN) return UPF_ERR_ARG;
        for (int c = 0; c < 3; ++c) {
            x[4 * i + c] = pos4[4 * i + c];
            if (vel4) v[4 * i + c] = vel4[4 * i + c];
            if (!(a->P.scales[c] * (pos4[4 * i + c] + a->P.support_radius) < 1024.0f) || !(pos4[4 * i + c] >= 0.0f))
                return UPF_ERR_RANGE;
        }
        std::memcpy(&x[4 * i + 3], &type[i], 4);
    }
    ADCHK(hipMemcpyAsync(a->x, x.data(), (size_t)n * 16, hipMemcpyHostToDevice, a->st));
    ADCHK(hipMemcpyAsync(a->v, v.data(), (size_t)n * 16, hipMemcpyHostToDevice, a->st));
    ADCHK(hipMemsetAsync(a->ctl, 0, sizeof(Ctl), a->st));
    ADCHK(hipEventRecord(a->ev[0], a->st));
    for (int s = 0; s < a->steps; ++s) {
        if (s % a->ckpt == 0) {
            const size_t o = (size_t)(s / a->ckpt) * n;
            ADCHK(hipMemcpyAsync(a->ck_x + o, a->x, (size_t)n * 16, hipMemcpyDeviceToDevice, a->st));
            ADCHK(hipMemcpyAsync(a->ck_v + o, a->v, (size_t)n * 16, hipMemcpyDeviceToDevice, a->st));
        }
        enqueue_step(a, -1);
    }
    ADCHK(hipEventRecord(a->ev[1], a->st));
    ADCHK(hipGetLastError());
    int rc = adj_error(a);
    if (rc) return rc;
    ADCHK(hipEventElapsedTime(&a->fwd_ms, a->ev[0], a->ev[1]));
    if (pos_out4 || vel_out4) {
        ADCHK(hipMemcpy(x.data(), a->x, (size_t)n * 16, hipMemcpyDeviceToHost));
        ADCHK(hipMemcpy(v.data(), a->v, (size_t)n * 16, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 4; ++c) {
                if (pos_out4) pos_out4[4 * i + c] = c < 3 ? x[4 * i + c] : 0.0f;
                if (vel_out4) vel_out4[4 * i + c] = c < 3 ? v[4 * i + c] : 0.0f;
            }
    }
    a->have_forward = true;
    return UPF_OK;
}

int upf_adjoint_backward(upf_adjoint* a, const float* dl_dpos4, const float* dl_dvel4, float* dl_dpos0_4,
                         float* dl_dvel0_4) {
    if (!a || !dl_dpos4) return UPF_ERR_ARG;
    if (!a->have_forward) return UPF_ERR_STATE;
    const int n = a->n;
    std::vector<float> g((size_t)4 * n), h((size_t)4 * n, 0.0f);
    for (size_t k = 0; k < g.size(); ++k) {
        g[k] = (k & 3) == 3 ? 0.0f : dl_dpos4[k];
        if (dl_dvel4) h[k] = (k & 3) == 3 ? 0.0f : dl_dvel4[k];
    }
    ADCHK(hipMemcpyAsync(a->gx, g.data(), (size_t)n * 16, hipMemcpyHostToDevice, a->st));
    ADCHK(hipMemcpyAsync(a->gv, h.data(), (size_t)n * 16, hipMemcpyHostToDevice, a->st));
    ADCHK(hipEventRecord(a->ev[2], a->st));
    for (int seg = a->nseg - 1; seg >= 0; --seg) {
        const int s0 = seg * a->ckpt, s1 = std::min(s0 + a->ckpt, a->steps);
        const size_t o = (size_t)seg * n;
        ADCHK(hipMemcpyAsync(a->x, a->ck_x + o, (size_t)n * 16, hipMemcpyDeviceToDevice, a->st));
        ADCHK(hipMemcpyAsync(a->v, a->ck_v + o, (size_t)n * 16, hipMemcpyDeviceToDevice, a->st));
        for (int s = s0; s < s1; ++s) enqueue_step(a, s - s0);      // recompute, taping
        for (int s = s1 - 1; s >= s0; --s) enqueue_step_back(a, s - s0);
    }
    ADCHK(hipEventRecord(a->ev[3], a->st));
    ADCHK(hipGetLastError());
    int rc = adj_error(a);
    if (rc) return rc;
    ADCHK(hipEventElapsedTime(&a->bwd_ms, a->ev[2], a->ev[3]));
    ADCHK(hipMemcpy(g.data(), a->gx, (size_t)n * 16, hipMemcpyDeviceToHost));
    ADCHK(hipMemcpy(h.data(), a->gv, (size_t)n * 16, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < 4; ++c) {
            if (dl_dpos0_4) dl_dpos0_4[4 * i + c] = c < 3 ? g[4 * i + c] : 0.0f;
            if (dl_dvel0_4) dl_dvel0_4[4 * i + c] = c < 3 ? h[4 * i + c] : 0.0f;
        }
    return UPF_OK;
}

int upf_adjoint_timing(upf_adjoint* a, float* forward_ms, float* backward_ms, size_t* device_bytes) {
    if (!a) return UPF_ERR_ARG;
    if (forward_ms) *forward_ms = a->fwd_ms;
    if (backward_ms) *backward_ms = a->bwd_ms;
    if (device_bytes) {
        const size_t n = (size_t)a->n, T = (size_t)a->ckpt, K = (size_t)a->K;
        *device_bytes = n * 16 * (2 + 3 + 2 * a->nseg + 3 * T + T * K + 7) + n * 4 * (4 + T + 2 * T * K) +
                        T * a->nb_stride() * 4 + (size_t)a->ncells * 8 + radix_sort_temp_bytes(a->n);
    }
    return UPF_OK;
}

}  // extern "C"
