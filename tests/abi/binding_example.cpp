// The reference-side binding of the C ABI (include/upf.h) as a standalone C++
// program: the calls UnifiedPhysics::InitCudaVariables and
// UnifiedPhysics::CudaParticlePhysics (src/cpu/UnifiedPhysics.cpp:615-643,
// 4051-4141) would make, compiled with g++ against the header and linked to
// libupf.so.
//
//   binding_example host       parameters, domain, density error factor and a
//                              dam-break scene (no device)
//   binding_example step N     also creates a simulation, uploads, runs N
//                              PCISPH steps and downloads; prints every
//                              position's float bits (hex, one particle per line)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "upf.h"

static int fail(const char* what, int rc) {
    std::fprintf(stderr, "%s failed: %d\n", what, rc);
    return 1;
}

int main(int argc, char** argv) {
    const bool step = argc > 1 && std::strcmp(argv[1], "step") == 0;
    const int nsteps = argc > 2 ? std::atoi(argv[2]) : 1;
    // InitCudaVariables: constants, domain, LUT-derived factor
    upf_params p;
    int rc = upf_params_init(0.01f, &p);
    if (rc != UPF_OK) return fail("upf_params_init", rc);
    p.dt = 0.001f;
    // the scene the reference's Scene::build would hold in particles_
    upf_dam_break_desc d;
    std::memset(&d, 0, sizeof(d));
    d.nx = 12; d.ny = 12; d.nz = 12;
    d.box[0] = 0.24f; d.box[1] = 0.12f; d.box[2] = 0.12f;
    d.jitter_rel = 0.001f; d.wall_rel = 1.0f; d.lattice_rel = 1.0025f; d.seed = 0;
    int nf = 0, nw = 0;
    rc = upf_scene_dam_break(&d, &p, &nf, &nw, nullptr, nullptr);
    if (rc != UPF_OK) return fail("upf_scene_dam_break (count)", rc);
    const int n = nf + nw;
    std::vector<float> pos4(4 * (size_t)n), vel4(4 * (size_t)n, 0.0f);
    std::vector<int> type(n);
    rc = upf_scene_dam_break(&d, &p, &nf, &nw, pos4.data(), type.data());
    if (rc != UPF_OK) return fail("upf_scene_dam_break", rc);
    int maxn = 0;
    rc = upf_compute_density_error_factor(&p, &maxn);
    if (rc != UPF_OK) return fail("upf_compute_density_error_factor", rc);
    std::printf("params spacing %.9g h %.9g mass %.9g delta %.9g block %u grid %u\n", p.spacing, p.support_radius,
                p.initial_mass, p.density_error_factor, p.block_size, p.grid_dim);
    std::printf("scene fluid %d wall %d\n", nf, nw);
    if (!step) return 0;
    upf_sim* sim = nullptr;
    rc = upf_create(&p, n, &sim);
    if (rc != UPF_OK) return fail("upf_create", rc);
    rc = upf_upload_particles(sim, n, pos4.data(), vel4.data(), type.data());
    if (rc != UPF_OK) return fail("upf_upload_particles", rc);
    // CudaParticlePhysics, per frame
    rc = upf_step(sim, nsteps, UPF_PCISPH_STATIC);
    if (rc != UPF_OK) return fail("upf_step", rc);
    rc = upf_download(sim, pos4.data(), vel4.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc != UPF_OK) return fail("upf_download", rc);
    upf_destroy(sim);
    for (int i = 0; i < n; ++i) {
        unsigned b[3];
        std::memcpy(b, &pos4[4 * (size_t)i], 12);
        std::printf("%08x %08x %08x\n", b[0], b[1], b[2]);
    }
    return 0;
}
