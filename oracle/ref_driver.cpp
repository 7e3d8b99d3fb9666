// ref_driver.cpp -- test-only C entry points over the reference's own
// translation units (WeightingKernel.cpp, zIndex.cpp, UnifiedConstants.cpp,
// RigidBody.cpp, Matrix3x3.cpp, UnifiedParticle.cpp, the OBJ model loader
// ObjLoader.cpp, ObjectLoader.cpp, Polygon.cpp, Face.cpp, Mesh.cpp, and the
// particle file writers and reader UnifiedIO.cpp), compiled from their
// place under the reference tree by oracle/Makefile into oracle/_ref/.
// It pins the oracle's restatement of those files; nothing here is shipped.

#include <unistd.h>

#include <climits>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>

// the driver sets / reads private RigidBody members that have no accessor
// (inverted_inertia_local_, inverted_inertia_world_)
#define private public
#include "RigidBody.h"
#undef private
#include "ObjLoader.h"
#include "UnifiedConstants.h"
#include "UnifiedIO.h"
#include "UnifiedParticle.h"
#include "WeightingKernel.h"
#include "zIndex.h"

// globals the reference translation units expect (src/sceneConfigure/global.cpp)
UnifiedConstants* fc = 0;
UnifiedPhysics* myFluid = 0;

namespace {
struct Quiet {  // the reference prints to std::cout in its constructors
    std::streambuf* old;
    std::ostringstream sink;
    Quiet() { old = std::cout.rdbuf(sink.rdbuf()); }
    ~Quiet() { std::cout.rdbuf(old); }
};
}  // namespace

extern "C" {

// UnifiedConstants ctor followed by setParticleSpacing(spacing):
// out = {particleSpacing, particleRadius, globalSupportRadius, forceDistance,
//        initialMass, fluidRestDensity, gravityConst, fluidViscConst,
//        deltaT_pcisph, fluidGasConstantWCSPH, densityErrorFactorParameter,
//        maxDensityErrorAllowed, minLoops, maxLoops, springCoefficient,
//        springCoefficientBoundary, dampingCoefficient, terminalSpeed,
//        surface_tension_gamma, surface_adhesion_beta, scales.x, scales.y, scales.z}
int ref_constants(float spacing, float* out) {
    Quiet q;
    UnifiedConstants c;
    c.setParticleSpacing(spacing);
    float v[] = {c.particleSpacing, c.particleRadius, c.globalSupportRadius, c.forceDistance,
                 c.initialMass, c.fluidRestDensity, c.gravityConst, c.fluidViscConst,
                 c.deltaT_pcisph, c.fluidGasConstantWCSPH, c.densityErrorFactorParameter,
                 c.maxDensityErrorAllowed, (float)c.minLoops, (float)c.maxLoops, c.springCoefficient,
                 c.springCoefficientBoundary, c.dampingCoefficient, c.terminalSpeed,
                 c.surface_tension_gamma, c.surface_adhesion_beta, c.scales.x, c.scales.y, c.scales.z};
    std::memcpy(out, v, sizeof(v));
    return (int)(sizeof(v) / sizeof(float));
}

// WeightingKernel::initLookupTables for support radius h (5 x 301 floats),
// KernelM4(0, h), and the LUT lookups of n distances (one h per process:
// the lookups cache 1/h in a function-local static).
int ref_luts(float h, float* out, float* kernel_self, int n, const float* dist, float* w, float* g, float* v) {
    Quiet q;
    UnifiedConstants c;
    c.globalSupportRadius = h;
    WeightingKernel k(&c);
    std::memcpy(out, k.lut_kernel_m4(), 301 * sizeof(float));
    std::memcpy(out + 301, k.lut_kernel_pressure_grad(), 301 * sizeof(float));
    std::memcpy(out + 602, k.lut_kernel_viscosity_lap(), 301 * sizeof(float));
    std::memcpy(out + 903, k.lut_spline_surface_tension(), 301 * sizeof(float));
    std::memcpy(out + 1204, k.lut_spline_surface_adhesion(), 301 * sizeof(float));
    *kernel_self = k.KernelM4(0.0f, h);
    for (int i = 0; i < n; ++i) {
        w[i] = k.kernelM4Lut(dist[i]);
        g[i] = k.kernelPressureGradLut(dist[i]);
        v[i] = k.kernelViscosityLaplacianLut(dist[i]);
    }
    return 0;
}

// Analytic kernels at n distances: M4, pressure gradient, viscosity laplacian.
int ref_kernels(float h, int n, const float* dist, float* m4, float* pg, float* vl) {
    Quiet q;
    UnifiedConstants c;
    c.globalSupportRadius = h;
    WeightingKernel k(&c);
    for (int i = 0; i < n; ++i) {
        m4[i] = k.KernelM4(dist[i], h);
        pg[i] = k.kernelPressureGrad(dist[i], h);
        vl[i] = k.kernelViscosityLaplacian(dist[i], h);
    }
    return 0;
}

int ref_ztable(unsigned* out) {
    ZIndex z;
    std::memcpy(out, z.z_table(), 1024 * 3 * sizeof(unsigned));
    return 0;
}

// ZIndex::SetBlockSize(size) then BoxQuery(min, max); returns the number of ranges.
int ref_box_query(int size, const unsigned short* mn, const unsigned short* mx, unsigned* out, int cap,
                  unsigned* block_size) {
    Quiet q;
    ZIndex z;
    z.SetBlockSize(size);
    *block_size = z.block_size();
    ZRanges r;
    z.BoxQuery(Vector3ui(mn[0], mn[1], mn[2]), Vector3ui(mx[0], mx[1], mx[2]), r);
    int m = (int)r.size();
    for (int i = 0; i < m && i < cap; ++i) { out[2 * i] = r[i].first; out[2 * i + 1] = r[i].second; }
    return m;
}

// One rigid-body update as CalculateRigidBodyForceTorqueMomenta does it
// (src/cpu/UnifiedPhysics.cpp:3620-3703): apply per-particle forces, torques
// and gravity, UpdateMomenta, PerformStep, UpdateRotationMatrix.
// state (in/out): pos[3] vel[3] angvel[3] P[3] L[3] q[4] R[9] Iinv_local[9] Iinv_world[9] mass
int ref_rigid_step(float* st, int n, const float* ppos, const float* pforce, const float* pmass, float dt,
                   float g, float terminal_speed) {
    RigidBody b;
    b.set_rigidbody_pos(vmml::Vector3f(st[0], st[1], st[2]));
    b.set_velocity(vmml::Vector3f(st[3], st[4], st[5]));
    b.set_angular_velocity(vmml::Vector3f(st[6], st[7], st[8]));
    b.set_linear_momentum(vmml::Vector3f(st[9], st[10], st[11]));
    b.set_angular_momentum(vmml::Vector3f(st[12], st[13], st[14]));
    b.set_quaternion(vmml::Vector4f(st[15], st[16], st[17], st[18]));
    Matrix3x3 R, Il, Iw;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            R.elements[r][c] = st[19 + 3 * r + c];
            Il.elements[r][c] = st[28 + 3 * r + c];
            Iw.elements[r][c] = st[37 + 3 * r + c];
        }
    b.set_rotation_matrix(R);
    b.inverted_inertia_local_ = Il;
    b.inverted_inertia_world_ = Iw;
    b.set_mass(st[46]);
    for (int i = 0; i < n; ++i) {
        UnifiedParticle p(ppos[3 * i], ppos[3 * i + 1], ppos[3 * i + 2], 1000.0, pmass[i], 0);
        vmml::Vector3f f(pforce[3 * i], pforce[3 * i + 1], pforce[3 * i + 2]);
        b.ApplyForce(f);
        b.ApplyTorque(f, p);
        vmml::Vector3f gf(0, -p.particle_mass_ * g, 0);
        b.ApplyForce(gf);
    }
    b.set_center_of_mass(b.rigidbody_pos());
    const float terminal = b.mass() * terminal_speed;
    b.UpdateMomenta(dt, terminal);
    b.PerformStep(dt);
    b.UpdateRotationMatrix();
    vmml::Vector3f x = b.rigidbody_pos(), v = b.velocity(), w = b.angular_velocity(), P = b.linear_momentum(),
                   L = b.angular_momentum();
    vmml::Vector4f qq = b.quaternion();
    for (int a = 0; a < 3; ++a) { st[a] = x[a]; st[3 + a] = v[a]; st[6 + a] = w[a]; st[9 + a] = P[a]; st[12 + a] = L[a]; }
    for (int a = 0; a < 4; ++a) st[15 + a] = qq[a];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            st[19 + 3 * r + c] = b.rotation_matrix().elements[r][c];
            st[37 + 3 * r + c] = b.inverted_inertia_world_.elements[r][c];
        }
    return 0;
}

// ObjLoader::GetInstance().LoadParticles(path, sx, sy, sz, spacing, template)
// (src/cpu/ObjectLoader.cpp:8-97): the particle positions, 3 floats each, up
// to cap of them; returns the number of particles.
int ref_obj_particles(const char* path, float sx, float sy, float sz, float spacing, float* out3, int cap) {
    Quiet q;
    UnifiedParticle templ(0.0f, 0.0f, 0.0f, 1000.0, 1.0, 0);
    std::vector<UnifiedParticle>* v = ObjLoader::GetInstance().LoadParticles(path, sx, sy, sz, spacing, templ);
    const int n = (int)v->size();
    for (int i = 0; i < n && i < cap; ++i)
        for (int a = 0; a < 3; ++a) out3[3 * i + a] = (*v)[i].position_[a];
    delete v;
    return n;
}

// Matrix3x3::CalculateInverse (src/cpu/Matrix3x3.cpp:121-158) on a row-major
// 3x3; returns 1 if it inverted, 0 if it found the matrix singular.
int ref_inverse3(const float* m9, float* out9) {
    Matrix3x3 m, r;
    r.SetZero();
    for (int k = 0; k < 9; ++k) m.elements[k / 3][k % 3] = m9[k];
    const bool ok = Matrix3x3::CalculateInverse(m, r);
    for (int k = 0; k < 9; ++k) out9[k] = r.elements[k / 3][k % 3];
    return ok ? 1 : 0;
}

// One call of a UnifiedIO writer (src/cpu/UnifiedIO.cpp:13-166) with `dir` as
// the working directory: the writers open "K:/SimulationData/output/..."
// relative to it, so dir must hold K:/SimulationData/output/FluidParticles and
// .../RigidParticles.  which: 1 SaveLiquidParticleInfoFull, 2
// SaveRigidParticleInfoFull, 3 SaveParticlePositions (particles built from
// pos4 / type).  count is the count argument of writers 1 and 2.  Returns the
// writer's file counter of this call (its file is particles%04d.txt,
// rigid1Particles%04d.txt or particles%04d.txt under the writer's folder),
// -1 on a bad argument.
int ref_save_text(const char* dir, int which, int n, const float* pos4, const int* type, int count) {
    static UnifiedIO io;
    static int calls[4] = {0, 0, 0, 0};
    if (which < 1 || which > 3) return -1;
    char cwd[PATH_MAX];
    if (!getcwd(cwd, sizeof cwd) || chdir(dir) != 0) return -1;
    {
        Quiet q;
        std::vector<float> p4(pos4, pos4 + 4 * (size_t)n);
        std::vector<int> t(type, type + n);
        if (which == 1) io.SaveLiquidParticleInfoFull(p4.data(), t.data(), n, count);
        else if (which == 2) io.SaveRigidParticleInfoFull(p4.data(), t.data(), n, count);
        else {
            std::vector<UnifiedParticle> ps;
            for (int i = 0; i < n; ++i)
                ps.push_back(UnifiedParticle(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], 1000.0, 1.0, type[i]));
            io.SaveParticlePositions(ps);
        }
    }
    if (chdir(cwd) != 0) return -1;
    return calls[which]++;
}

// ReadInFluidParticleNum + ReadInFluidParticlesPCISPH (src/cpu/UnifiedIO.cpp:168-220)
// into arrays of cap particles (3 floats per position / velocity); returns
// the count the file states.  The file must exist (the reference exits).
int ref_read_restart(const char* path, int cap, float* pos3, float* vel3, int* type, float* cp, float* rho) {
    static UnifiedIO io;
    const int cnt = io.ReadInFluidParticleNum(path);
    if (cnt < 0 || cnt > cap) return cnt;
    io.ReadInFluidParticlesPCISPH(path, pos3, vel3, type, cp, rho);
    return cnt;
}

}  // extern "C"
