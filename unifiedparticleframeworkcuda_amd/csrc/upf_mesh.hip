// upf_mesh.hip -- OBJ model -> particles, the scene-input path of the
// reference's rigid objects: ObjLoader::Load / LoadFaces (src/cpu/ObjLoader.cpp:18-258)
// read the triangles, ObjectLoader::LoadParticles (src/cpu/ObjectLoader.cpp:8-97)
// keeps the lattice points inside the mesh or on a face (Polygon,
// src/cpu/Polygon.cpp).  The reader runs on the host; the point-in-mesh test of
// every lattice point against every face -- the O(points x faces) part -- runs
// on the GPU, one thread per lattice point over faces staged through LDS, and
// the kept points are compacted in the reference's loop order.
//
// Floating point: -ffp-contract=off and correctly rounded division, every
// expression in the reference's evaluation order, so the particle set and its
// positions equal the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "upf.h"

namespace upf {
namespace {

#define MCHK(x)                                     \
    do {                                            \
        if ((x) != hipSuccess) return UPF_ERR_HIP; \
    } while (0)

struct F3 {
    float x, y, z;
};

// ---------------------------------------------------------------- the reader
// Two passes with a 256-byte fgets buffer under while (!feof(fp)): the fgets
// that meets the end of the file returns nothing and leaves the buffer as it
// was, so the reference processes its last line twice (a file ending in a
// face line gets that face twice).  No faces or no vertices: the reference's
// FileError, i.e. an empty model.
int read_obj(const char* path, std::vector<F3>& tri) {
    tri.clear();
    FILE* fp = std::fopen(path, "r");
    if (!fp) return UPF_ERR_IO;
    char line[256];
    std::memset(line, 0, sizeof(line));
    long n_vertex = 0, n_face = 0;
    while (!std::feof(fp)) {
        if (std::fgets(line, sizeof(line), fp) == nullptr) {
            // end of file: `line` keeps the previous line
        }
        if (std::strncmp("vt ", line, 3) == 0 || std::strncmp("vn ", line, 3) == 0) continue;
        if (std::strncmp("v ", line, 2) == 0) ++n_vertex;
        else if (std::strncmp("f ", line, 2) == 0) ++n_face;
    }
    std::fclose(fp);
    if (n_face == 0 || n_vertex == 0) return UPF_OK;
    fp = std::fopen(path, "r");
    if (!fp) return UPF_ERR_IO;
    std::memset(line, 0, sizeof(line));
    std::vector<F3> vertex;
    std::vector<int> corner;
    bool bad = false;
    while (!std::feof(fp)) {
        if (std::fgets(line, sizeof(line), fp) == nullptr) {
        }
        if (std::strncmp("vt ", line, 3) == 0 || std::strncmp("vn ", line, 3) == 0) continue;
        if (std::strncmp("v ", line, 2) == 0) {
            F3 v{0.0f, 0.0f, 0.0f};
            bad |= std::sscanf(line + 2, "%f%f%f", &v.x, &v.y, &v.z) != 3;
            vertex.push_back(v);
        } else if (std::strncmp("f ", line, 2) == 0) {
            // v/vt/vn triples; the vertex index of the third corner is field 7
            int f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            const int got = std::sscanf(line + 2, "%d/%d/%d %d/%d/%d %d/%d/%d", &f[0], &f[1], &f[2], &f[3], &f[4],
                                        &f[5], &f[6], &f[7], &f[8]);
            bad |= got < 7;
            corner.push_back(f[0] - 1);
            corner.push_back(f[3] - 1);
            corner.push_back(f[6] - 1);
        }
    }
    std::fclose(fp);
    if (bad) return UPF_ERR_ARG;  // a line the reference reads into undefined values
    tri.reserve(corner.size());
    for (int c : corner) {
        if (c < 0 || c >= (int)vertex.size()) return UPF_ERR_ARG;
        tri.push_back(vertex[(size_t)c]);
    }
    return UPF_OK;
}

// ----------------------------------------------------------- face constants
// Per face, what the crossing and touching tests read: the corners, the unit
// normal n (Polygon::AddVertex: (v0 - v1) x (v2 - v1), vmml's normalize with
// 1/l in double), d = -(v0 . n), the numerator -(d + o . n) of the segment
// parameter for the common segment start o, the in-plane test's far point
// (box minimum - 1, the box starting at the origin like the reference's), and
// the two coordinates the in-plane test projects on.
struct Face {
    float4 a;  // v0, d
    float4 b;  // v1, -(d + o . n)
    float4 c;  // v2, index of the first projected coordinate (0 or 1)
    float4 n;  // normal, index of the second (1 or 2)
    float4 o;  // in-plane far point
};

inline float dot(const F3& a, const F3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

void face_box(const F3* v, F3& mn, F3& mx) {
    mn = {0.0f, 0.0f, 0.0f};
    mx = {0.0f, 0.0f, 0.0f};
    for (int k = 0; k < 3; ++k) {
        const float c[3] = {v[k].x, v[k].y, v[k].z};
        float* pmn = &mn.x;
        float* pmx = &mx.x;
        for (int i = 0; i < 3; ++i) {
            pmn[i] = std::min(c[i], pmn[i]);
            pmx[i] = std::max(c[i], pmx[i]);
        }
    }
}

F3 face_normal(const F3* v) {
    const F3 a{v[0].x - v[1].x, v[0].y - v[1].y, v[0].z - v[1].z};
    const F3 b{v[2].x - v[1].x, v[2].y - v[1].y, v[2].z - v[1].z};
    F3 n{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    const float l = std::sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    if (l != 0.0f) {
        const float ll = (float)(1.0 / (double)l);
        n.x *= ll;
        n.y *= ll;
        n.z *= ll;
    }
    return n;
}

// ------------------------------------------------------------------ device
__device__ __forceinline__ float pick(const float3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// Polygon::ArePointsOnBothSidesOfTheLine (Polygon.cpp:73-103)
__device__ __forceinline__ bool both_sides(int fi, int si, const float3& ls, const float3& le, const float3& p1,
                                           const float3& p2) {
    const float dxES = pick(le, fi) - pick(ls, fi);
    const float dx1E = pick(p1, fi) - pick(le, fi);
    const float dySE = pick(ls, si) - pick(le, si);
    const float dyE1 = pick(le, si) - pick(p1, si);
    const bool ccw1 = (dySE * dx1E) > (dyE1 * dxES);
    const float dx2E = pick(p2, fi) - pick(le, fi);
    const float dyE2 = pick(le, si) - pick(p2, si);
    const bool ccw2 = (dySE * dx2E) > (dyE2 * dxES);
    return ccw1 != ccw2;
}

// Polygon::IsPointInPlaneInside (Polygon.cpp:118-137)
__device__ __forceinline__ bool in_plane_inside(const Face& f, const float3& pt) {
    const int fi = (int)f.c.w, si = (int)f.n.w;
    const float3 out = make_float3(f.o.x, f.o.y, f.o.z);
    const float3 v[3] = {make_float3(f.a.x, f.a.y, f.a.z), make_float3(f.b.x, f.b.y, f.b.z),
                         make_float3(f.c.x, f.c.y, f.c.z)};
    bool inside = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float3& s = v[k];
        const float3& e = v[k == 2 ? 0 : k + 1];
        if (both_sides(fi, si, out, pt, s, e) && both_sides(fi, si, s, e, out, pt)) inside = !inside;
    }
    return inside;
}

constexpr int MF_BLOCK = 256;
constexpr int MF_FACES = 256;  // faces per LDS stage (20 KB)

// One thread per lattice point (x-major, then y, then z: the reference's loop
// order).  For each face in file order: Polygon::DoesLineIntersect of the
// segment from o to the point toggles the point; otherwise
// Polygon::DoesPointTouch keeps it and ends the scan.  keep[] gets the
// decision, counts[] the kept points of each block.
__global__ __launch_bounds__(MF_BLOCK) void k_mesh_test(const Face* __restrict__ faces, int nf,
                                                        const float* __restrict__ X, const float* __restrict__ Y,
                                                        const float* __restrict__ Z, int nx, int ny, int nz, float3 o,
                                                        uint8_t* __restrict__ keep, int* __restrict__ counts) {
    __shared__ Face sf[MF_FACES];
    const long long npts = (long long)nx * ny * nz;
    const long long idx = (long long)blockIdx.x * MF_BLOCK + threadIdx.x;
    const bool live = idx < npts;
    float3 p = make_float3(0.f, 0.f, 0.f);
    if (live) {
        const int z = (int)(idx % nz);
        const int y = (int)((idx / nz) % ny);
        const int x = (int)(idx / ((long long)ny * nz));
        const int par = x & 1;  // odd x: y and z start half a spacing later
        p = make_float3(X[x], Y[par * ny + y], Z[par * nz + z]);
    }
    // segment o -> p
    const float3 dv = make_float3(p.x - o.x, p.y - o.y, p.z - o.z);
    bool add = false, done = !live;
    for (int f0 = 0; f0 < nf; f0 += MF_FACES) {
        const int m = min(MF_FACES, nf - f0);
        __syncthreads();
        for (int k = threadIdx.x; k < m; k += MF_BLOCK) sf[k] = faces[f0 + k];
        __syncthreads();
        if (!done) {
            for (int k = 0; k < m; ++k) {
                const Face& f = sf[k];
                const float3 n = make_float3(f.n.x, f.n.y, f.n.z);
                bool crosses = false;
                const float den = dv.x * n.x + dv.y * n.y + dv.z * n.z;
                if (!(fabsf(den) < 0.000001f)) {
                    const float dof = f.b.w / den;
                    if (!(dof < 0.0f || dof > 1.0f)) {
                        const float3 q = make_float3(o.x + dv.x * dof, o.y + dv.y * dof, o.z + dv.z * dof);
                        crosses = in_plane_inside(f, q);
                    }
                }
                if (crosses) {
                    add = !add;
                } else {
                    const float eq = (p.x * n.x + p.y * n.y + p.z * n.z) + f.a.w;
                    if (!(fabsf(eq) > 0.000001f) && in_plane_inside(f, p)) {
                        add = true;
                        done = true;
                        break;
                    }
                }
            }
        }
        if (__syncthreads_count(!done) == 0) break;
    }
    if (live) keep[idx] = (uint8_t)add;
    const int c = __syncthreads_count(live && add);
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// exclusive prefix of the block counts (one block)
__global__ __launch_bounds__(1024) void k_mesh_scan(int* __restrict__ counts, int nb, int* __restrict__ total) {
    __shared__ int part[1024];
    const int per = (nb + 1023) / 1024;
    const int beg = threadIdx.x * per, end = min(beg + per, nb);
    int s = 0;
    for (int i = beg; i < end; ++i) s += counts[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int run = threadIdx.x ? part[threadIdx.x - 1] : 0;
    for (int i = beg; i < end; ++i) {
        const int c = counts[i];
        counts[i] = run;
        run += c;
    }
    if (threadIdx.x == 1023) *total = part[1023];
}

// kept points in lattice order: (p - box minimum) * scale
__global__ __launch_bounds__(MF_BLOCK) void k_mesh_emit(const uint8_t* __restrict__ keep,
                                                        const int* __restrict__ offsets, const float* __restrict__ X,
                                                        const float* __restrict__ Y, const float* __restrict__ Z,
                                                        int nx, int ny, int nz, float3 mn, float3 sc,
                                                        float4* __restrict__ out, int cap) {
    __shared__ int wave_base[MF_BLOCK / 64];
    const long long npts = (long long)nx * ny * nz;
    const long long idx = (long long)blockIdx.x * MF_BLOCK + threadIdx.x;
    const bool k = idx < npts && keep[idx];
    const uint64_t m = __ballot(k);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wave_base[w] = __popcll(m);
    __syncthreads();
    int base = offsets[blockIdx.x];
    for (int i = 0; i < w; ++i) base += wave_base[i];
    const uint64_t below = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
    const int slot = base + __popcll(below);
    if (k && slot < cap) {
        const int z = (int)(idx % nz);
        const int y = (int)((idx / nz) % ny);
        const int x = (int)(idx / ((long long)ny * nz));
        const int par = x & 1;
        out[slot] = make_float4((X[x] - mn.x) * sc.x, (Y[par * ny + y] - mn.y) * sc.y, (Z[par * nz + z] - mn.z) * sc.z,
                                0.0f);
    }
}

// number of lattice steps: the reference's `for (int i = 0; i <= steps; ++i)`
long long loop_count(float steps) {
    if (!(steps >= 0.0f)) return 0;
    if (steps >= 16777216.0f) return -1;  // past exact float integers: refused
    long long c = 0;
    while ((float)c <= steps) ++c;
    return c;
}

}  // namespace
}  // namespace upf

using namespace upf;

extern "C" {

int upf_obj_read(const char* path, float* faces9, int cap, int* n_faces) {
    if (!path || !n_faces || (faces9 && cap < 0)) return UPF_ERR_ARG;
    std::vector<F3> tri;
    const int rc = read_obj(path, tri);
    if (rc) return rc;
    const int n = (int)(tri.size() / 3);
    *n_faces = n;
    if (faces9) {
        if (cap < n) return UPF_ERR_RANGE;
        std::memcpy(faces9, tri.data(), tri.size() * sizeof(F3));
    }
    return UPF_OK;
}

int upf_mesh_fill(const float* faces9, int n_faces, const float scale[3], float spacing, float* pos4, int cap,
                  int* n_out) {
    if (!n_out || n_faces < 0 || (n_faces > 0 && !faces9) || !scale || (pos4 && cap < 0)) return UPF_ERR_ARG;
    if (!(spacing > 0.0f) || !(scale[0] > 0.0f) || !(scale[1] > 0.0f) || !(scale[2] > 0.0f)) return UPF_ERR_ARG;
    *n_out = 0;
    if (n_faces == 0) return UPF_OK;
    // faces: the box of each (from the origin) and of the model, constants
    std::vector<Face> faces((size_t)n_faces);
    const F3* tri = reinterpret_cast<const F3*>(faces9);
    F3 mn{}, mx{};
    std::vector<F3> fmn((size_t)n_faces), fmx((size_t)n_faces), fn((size_t)n_faces);
    for (int f = 0; f < n_faces; ++f) {
        face_box(tri + 3 * (size_t)f, fmn[f], fmx[f]);
        fn[f] = face_normal(tri + 3 * (size_t)f);
    }
    mn = fmn[0];
    mx = fmx[0];
    for (int f = 0; f < n_faces; ++f) {
        float* a = &mn.x;
        float* b = &mx.x;
        const float* c = &fmn[f].x;
        const float* d = &fmx[f].x;
        for (int i = 0; i < 3; ++i) {
            a[i] = std::min(a[i], c[i]);
            b[i] = std::max(b[i], d[i]);
        }
    }
    const F3 o{mn.x - spacing, mn.y - spacing, mn.z - spacing};  // the segments' common start
    for (int f = 0; f < n_faces; ++f) {
        const F3* v = tri + 3 * (size_t)f;
        const F3& n = fn[f];
        const float d = -dot(v[0], n);
        const float num = -(d + dot(o, n));
        int fi, si;
        if (n.x == 0) {
            fi = 0;
            si = n.y == 0 ? 1 : 2;
        } else {
            fi = 1;
            si = 2;
        }
        Face& F = faces[(size_t)f];
        F.a = make_float4(v[0].x, v[0].y, v[0].z, d);
        F.b = make_float4(v[1].x, v[1].y, v[1].z, num);
        F.c = make_float4(v[2].x, v[2].y, v[2].z, (float)fi);
        F.n = make_float4(n.x, n.y, n.z, (float)si);
        F.o = make_float4(fmn[f].x - 1.0f, fmn[f].y - 1.0f, fmn[f].z - 1.0f, 0.0f);
    }
    // the lattice: coordinates accumulated step by step as the reference's
    // loops do (odd x: y and z start half a spacing later)
    const float xs = (mx.x - mn.x) * scale[0] / spacing;
    const float ys = (mx.y - mn.y) * scale[1] / spacing;
    const float zs = (mx.z - mn.z) * scale[2] / spacing;
    const long long nx = loop_count(xs), ny = loop_count(ys), nz = loop_count(zs);
    if (nx < 0 || ny < 0 || nz < 0) return UPF_ERR_ARG;
    if (nx == 0 || ny == 0 || nz == 0) return UPF_OK;
    const long long npts = nx * ny * nz;
    if (npts > (1ll << 31) - 1) return UPF_ERR_ARG;
    std::vector<float> X((size_t)nx), Y(2 * (size_t)ny), Z(2 * (size_t)nz);
    {
        float cx = mn.x;
        for (long long x = 0; x < nx; ++x) {
            X[(size_t)x] = cx;
            cx += spacing / scale[0];
        }
        for (int par = 0; par < 2; ++par) {
            float cy = mn.y + par * spacing / (2 * scale[1]);
            for (long long y = 0; y < ny; ++y) {
                Y[(size_t)(par * ny + y)] = cy;
                cy += spacing / scale[1];
            }
            float cz = mn.z + par * spacing / (2 * scale[2]);
            for (long long z = 0; z < nz; ++z) {
                Z[(size_t)(par * nz + z)] = cz;
                cz += spacing / scale[2];
            }
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UPF_ERR_NO_DEVICE;
    const int nb = (int)((npts + MF_BLOCK - 1) / MF_BLOCK);
    Face* d_faces = nullptr;
    float *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
    uint8_t* d_keep = nullptr;
    int *d_counts = nullptr, *d_total = nullptr;
    float4* d_out = nullptr;
    hipStream_t st = nullptr;
    int rc = UPF_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == UPF_OK) rc = UPF_ERR_HIP;
        return rc == UPF_OK;
    };
    int total = 0;
    if (ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && ok(hipMalloc(&d_faces, faces.size() * sizeof(Face))) &&
        ok(hipMalloc(&d_x, X.size() * sizeof(float))) && ok(hipMalloc(&d_y, Y.size() * sizeof(float))) &&
        ok(hipMalloc(&d_z, Z.size() * sizeof(float))) && ok(hipMalloc(&d_keep, (size_t)npts)) &&
        ok(hipMalloc(&d_counts, (size_t)nb * sizeof(int))) && ok(hipMalloc(&d_total, sizeof(int))) &&
        ok(hipMemcpyAsync(d_faces, faces.data(), faces.size() * sizeof(Face), hipMemcpyHostToDevice, st)) &&
        ok(hipMemcpyAsync(d_x, X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice, st)) &&
        ok(hipMemcpyAsync(d_y, Y.data(), Y.size() * sizeof(float), hipMemcpyHostToDevice, st)) &&
        ok(hipMemcpyAsync(d_z, Z.data(), Z.size() * sizeof(float), hipMemcpyHostToDevice, st))) {
        hipLaunchKernelGGL(k_mesh_test, dim3(nb), dim3(MF_BLOCK), 0, st, d_faces, n_faces, d_x, d_y, d_z, (int)nx,
                           (int)ny, (int)nz, make_float3(o.x, o.y, o.z), d_keep, d_counts);
        hipLaunchKernelGGL(k_mesh_scan, dim3(1), dim3(1024), 0, st, d_counts, nb, d_total);
        if (ok(hipGetLastError()) && ok(hipMemcpyAsync(&total, d_total, sizeof(int), hipMemcpyDeviceToHost, st)) &&
            ok(hipStreamSynchronize(st))) {
            *n_out = total;
            if (pos4 && cap < total) rc = UPF_ERR_RANGE;
            if (rc == UPF_OK && pos4 && total > 0 && ok(hipMalloc(&d_out, (size_t)total * sizeof(float4)))) {
                hipLaunchKernelGGL(k_mesh_emit, dim3(nb), dim3(MF_BLOCK), 0, st, d_keep, d_counts, d_x, d_y, d_z,
                                   (int)nx, (int)ny, (int)nz, make_float3(mn.x, mn.y, mn.z),
                                   make_float3(scale[0], scale[1], scale[2]), d_out, total);
                if (ok(hipGetLastError()))
                    ok(hipMemcpyAsync(pos4, d_out, (size_t)total * sizeof(float4), hipMemcpyDeviceToHost, st)) &&
                        ok(hipStreamSynchronize(st));
            }
        }
    }
    if (st) (void)hipStreamSynchronize(st);
    for (void* p : {(void*)d_faces, (void*)d_x, (void*)d_y, (void*)d_z, (void*)d_keep, (void*)d_counts,
                    (void*)d_total, (void*)d_out})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

}  // extern "C"
