/*
 * upf.h -- C ABI of the MI355X (gfx950) per-step particle pipeline.
 *
 * The reference exposes its GPU path as an `extern "C"` block of functions
 * that take a caller-owned `dataPointers&` plus ~35 global constant setters
 * (src/gpu/particles.cuh:430-586).  This header replaces that block with a
 * handle-based ABI: every function takes plain pointers and sizes, returns an
 * int status (never exit()/abort()), and keeps no process-global state, so
 * several simulations may live in one process.  Each entry point names the
 * reference function(s) it replaces.
 *
 * Units and layouts
 *   positions / velocities : float4 per particle (x, y, z, w), w ignored on
 *                            input and 0 on output, SoA like the reference's
 *                            d_pos_zindex / d_vel (src/gpu/particles.cuh:124-126)
 *   type                   : int per particle, the reference's particle types
 *                            (src/cpu/UnifiedParticle.h:14-23):
 *                            0 rigid, 2 liquid, 6 frozen (static boundary)
 *   keys                   : 30-bit z-index (Morton) codes, 10 bits per axis
 *                            (src/cpu/zIndex.cpp:10-24)
 */
#ifndef UPF_H_
#define UPF_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
/* The reference aborts or exits on these conditions; here they are codes.  */
#define UPF_OK                  0
#define UPF_ERR_HIP             1  /* HIP runtime error (CUDA_SAFE_CALL -> exit, src/cpu/cutil.h:761-769) */
#define UPF_ERR_ARG             2  /* invalid argument / unsupported configuration */
#define UPF_ERR_RANGE           3  /* a position maps outside the 1024^3 z-index grid (z_table_ overflow, src/cpu/zIndex.h:53) */
#define UPF_ERR_NAN             4  /* NaN predicted position (abort at src/cpu/UnifiedPhysics.cpp:5446-5452) */
#define UPF_ERR_NBR_OVERFLOW    5  /* a particle has more neighbours than the list capacity */
#define UPF_ERR_DELTA           6  /* density error factor divisor is zero (exit at src/cpu/UnifiedPhysics.cpp:5677-5681) */
#define UPF_ERR_NO_DEVICE       7  /* no HIP device */
#define UPF_ERR_ISOLATED        8  /* boundary particle without boundary neighbours (abort at src/cpu/UnifiedPhysics.cpp:2739-2740) */
#define UPF_ERR_STATE           9  /* call order violated (e.g. step before upload) */
#define UPF_ERR_IO             10  /* file missing or unreadable (ObjectLoader::FileNotFound, src/cpu/ObjectLoader.h:36-46) */

/* ------------------------------------------------------ particle types */
#define UPF_RIGID   0
#define UPF_LIQUID  2
#define UPF_FROZEN  6

/* ----------------------------------------------------------- methods */
/* Step drivers; each replaces one CPU driver and its GPU counterpart.     */
#define UPF_PCISPH_STATIC  0  /* CpuParticlePhysicsStaticBoundariesPCISPH src/cpu/UnifiedPhysics.cpp:4220 /
                                 ParticlePhysicsOnDeviceFluidStaticBoundariesPCISPH src/gpu/particles.cu:2603 */
#define UPF_WCSPH_STATIC   1  /* CpuParticlePhysicsStaticBoundariesSPH src/cpu/UnifiedPhysics.cpp:4180 (physicsType 'w') /
                                 ParticlePhysicsOnDeviceSPH src/gpu/particles.cu:1281 */
#define UPF_PCISPH_RIGID   2  /* CpuParticlePhysicsVersatileCouplingPCISPH src/cpu/UnifiedPhysics.cpp:4991 /
                                 ParticlePhysicsOnDeviceFluidRigidCouplingPCISPH src/gpu/particles.cu:2515 */
#define UPF_SPH_STATIC     3  /* CpuParticlePhysicsStaticBoundariesSPH src/cpu/UnifiedPhysics.cpp:4180 with
                                 physicsType 'o': original SPH, p = max(0, k (rho - rho0))
                                 (CalculatePressureSPH :2479-2489), k = gas_const_sph */
/* On a scene without boundary particles the two SPH methods are also the
 * reference's default CPU driver for 'o' / 'w', CpuParticlePhysicsPureFluidsSPH
 * (src/cpu/UnifiedPhysics.cpp:4252-4330), whose passes differ from the
 * static-boundary driver's only in not testing the particle type.         */

/* Boundary-pressure mode (Akinci boundary pressure force on liquid from
 * boundary particles).  The reference CPU path passes an integer LUT index to
 * kernelPressureGradLut(float) (src/cpu/UnifiedPhysics.cpp:4245), which makes
 * the force identically zero; its GPU path indexes the table directly
 * (src/gpu/particles_kernel.cuh:791-802).                                    */
#define UPF_BOUNDARY_CPU_REF   0  /* as the CPU path computes it (zero force) */
#define UPF_BOUNDARY_AKINCI    1  /* table indexed by the LUT bin, Psi of the boundary neighbour */

/* ------------------------------------------------------------ params */
/* Scalar part of UnifiedConstants (src/sceneConfigure/UnifiedConstants.h:13-219)
 * and of dataPointers (src/gpu/particles.cuh:49-121) that the step reads.
 * Filled by upf_params_init + upf_params_set_domain; fields may be edited
 * afterwards (dt, stiffness, loop limits, coupling coefficients).          */
typedef struct upf_params {
    /* UnifiedConstants::setParticleSpacing (UnifiedConstants.cpp:168-191) */
    float spacing;                 /* particleSpacing (delta)                      */
    float particle_radius;         /* particleRadius = 0.5 * spacing               */
    float support_radius;          /* globalSupportRadius h = 2.00002 * spacing    */
    float force_distance;          /* forceDistance = 0.25 h                       */
    float initial_mass;            /* initialMass = spacing^3 * rest density       */
    float rest_density;            /* fluidRestDensity                             */
    float visc_const;              /* fluidViscConst                               */
    float gravity;                 /* gravityConst                                 */
    float dt;                      /* deltaT of the step                           */
    float gas_const_wcsph;         /* fluidGasConstantWCSPH (Tait stiffness)       */
    float gamma;                   /* Tait exponent (7)                            */
    float max_boundary_force;      /* maxBoundaryForce                             */
    int   add_boundary_force;      /* addBoundaryForce                             */
    /* PCISPH (UnifiedConstants.cpp:89-103,152-155)                         */
    float density_error_factor;    /* densityErrorFactor (delta, already scaled)   */
    float density_error_factor_parameter;
    float max_density_error_allowed; /* percent                                   */
    int   min_loops;
    int   max_loops;
    float kernel_self;             /* KernelM4(0, h)                               */
    /* coupling (UnifiedConstants.cpp:85-87,113-116,158)                    */
    float surface_tension_gamma;
    float surface_adhesion_beta;
    float spring_coefficient;
    float spring_coefficient_boundary;
    float damping_coefficient;
    float terminal_speed;
    int   boundary_pressure_mode;  /* UPF_BOUNDARY_*                               */
    int   renormalize_quaternion;  /* 0: CPU path (no renorm), 1: GPU path          */
    int   rigid_predicted_current; /* 0: CPU path (rigid x* stays at its initial
                                      value), 1: rigid x* = x                       */
    /* z-index domain (UnifiedConstants::setBoundingBoxValues :308-382)     */
    float vbox_min[3];             /* virtual bounding box (origin must be 0)      */
    float vbox_max[3];
    float scales[3];               /* GRID_RESOLUTION / extent                     */
    uint32_t block_size;           /* ZIndex::block_size_ (power of two; must exceed
                                      scales[a] * support_radius by 0.01 % so that a
                                      query box spans at most 3 cells per axis) */
    uint32_t block_shift;          /* log2(block_size)                             */
    uint32_t grid_dim;             /* cells per axis = 1024 / block_size           */
    float collision_min[3];        /* collisionBox                                 */
    float collision_max[3];
    /* neighbour search                                                        */
    float support_radius_sq_limit; /* smallest float T with sqrtf(T) >= h: a pair
                                      is a neighbour iff |x_i-x_j|^2 < T, which is
                                      exactly `sqrtf(d2) < h` of the CPU path      */
    int   max_neighbors;           /* neighbour list capacity per particle         */
    /* wall weight function (addWallWeightFunction, UnifiedConstants.cpp:144;
     * WallWeight / DistanceToWall, src/cpu/UnifiedPhysics.cpp:5926-6060): a
     * density contribution of the container's walls as a function of the
     * distance to the nearest wall face.  WCSPH adds it after the neighbour
     * sum, as the CPU path's CalculateDensityPressure (:2339-2406); PCISPH
     * adds it to the predicted density before the neighbour sum, as the GPU
     * path's ComputePredictedDensityAndPressureBlocksKernelPCISPHWallWeight
     * (src/gpu/particles_kernel.cu:3845-3990, the CPU path's PCISPH has none).
     * Single-scene PCISPH_STATIC and WCSPH_STATIC steps.                    */
    int   add_wall_weight;         /* 0 off (upf_params_init), 1 on                */
    float container_min[3];        /* realBoxContainer (upf_params_set_domain: the */
    float container_max[3];        /* collision box widened by particle_radius)    */
    float gas_const_sph;           /* fluidGasConst of original SPH (UPF_SPH_STATIC),
                                      1000 (UnifiedConstants.cpp:260)              */
} upf_params;

/* ------------------------------------------------------ host helpers */
/* No device is touched by these.                                          */

/* UnifiedConstants ctor + setParticleSpacing + computeGasConstAndTimeStep
 * (src/sceneConfigure/UnifiedConstants.cpp:7-191,241-306), for a given
 * spacing; dt defaults to deltaT_pcisph.  */
int upf_params_init(float spacing, upf_params* out);

/* setBoundingBoxValues (UnifiedConstants.cpp:308-382) with the virtual box and
 * collision box given by the caller, plus the block-size rule of Scene::build /
 * ZIndex::SetBlockSize (src/sceneConfigure/Scene.cpp:29-34, src/cpu/zIndex.cpp:27-40).
 * Returns UPF_ERR_ARG when vbox_min is not the origin or when the block is
 * smaller than the scaled support radius (cells must hold the h-neighbourhood). */
int upf_params_set_domain(upf_params* p, const float vbox_max[3],
                          const float collision_min[3], const float collision_max[3]);

/* WeightingKernel::initLookupTables (src/cpu/WeightingKernel.cpp:303-320):
 * five 301-entry tables (entry 300 = 0) in this order: M4, pressure gradient,
 * viscosity laplacian, surface tension spline, adhesion spline.  out: 5*301 floats. */
int upf_build_luts(const upf_params* p, float* out);

/* ComputeDensityErrorFactor (src/cpu/UnifiedPhysics.cpp:5525-5689): fills
 * p->density_error_factor from the prototype block.  *max_neighbors_out may be NULL. */
int upf_compute_density_error_factor(upf_params* p, int* max_neighbors_out);

/* Seeded dam-break scene (replaces Scene::build / createFluidObject /
 * createBoxBoundaryFrozenParticles, src/sceneConfigure/Scene.cpp:23-138,597-645,890-1042):
 * an nx*ny*nz fluid lattice at (i+1/2)*spacing with uniform jitter of
 * +-jitter_rel*spacing per component in the corner of a box of extent box[3],
 * walls sampled every wall_rel*spacing on the floor and four sides (open top).
 * Call with pos4 == NULL to get the counts.  Also sets the domain of *p.     */
typedef struct upf_dam_break_desc {
    int   nx, ny, nz;
    float box[3];          /* container extent at lattice_rel = 1 (scaled by lattice_rel) */
    float jitter_rel;
    float wall_rel;        /* wall sampling / spacing                                  */
    float lattice_rel;     /* fluid lattice spacing / spacing (1.0025 puts the nearest
                              neighbours mid-bin of the 300-entry kernel table, so the
                              lattice starts at the table kernel's rest density)       */
    uint64_t seed;
} upf_dam_break_desc;
int upf_scene_dam_break(const upf_dam_break_desc* d, upf_params* p,
                        int* n_fluid, int* n_wall, float* pos4, int* type);

/* Counts of the rigid-fluid coupling scene: a fluid block plus n_bodies seeded
 * boxes and spheres sampled on their surface (see upf_scene_rigid).  */
typedef struct upf_rigid_scene_desc {
    int   nx, ny, nz;       /* fluid lattice                                 */
    float box[3];           /* container extent                              */
    int   n_bodies;
    float size_min, size_max;       /* edge / diameter range (m)             */
    float density_min, density_max; /* body density range (kg/m^3)           */
    float jitter_rel;
    float surface_rel;      /* surface sampling / spacing                    */
    float wall_rel;         /* wall sampling / spacing                       */
    float lattice_rel;      /* fluid lattice spacing / spacing               */
    uint64_t seed;
} upf_rigid_scene_desc;

/* Rigid body state, one per body (replaces RigidBody_GPU, src/gpu/particles.cuh:27-44,
 * without its fixed int[15000] index array).  Matrices are row-major 3x3.   */
typedef struct upf_body {
    float pos[3];               /* centre of mass                             */
    float mass;
    float vel[3];
    float pad0;
    float angular_velocity[3];
    float pad1;
    float linear_momentum[3];
    float pad2;
    float angular_momentum[3];
    float pad3;
    float quaternion[4];        /* (s, vx, vy, vz)                            */
    float rotation[9];
    float inv_inertia_local[9];
    float inv_inertia_world[9];
    float force[3];             /* last reduced force  (diagnostic)           */
    float torque[3];            /* last reduced torque (diagnostic)           */
} upf_body;

/* Rigid scene (config of the coupling benchmark; replaces the OBJ-based
 * rigid objects of Scene::build, src/sceneConfigure/Scene.cpp:59-131): the
 * fluid lattice and walls of upf_scene_dam_break plus n_bodies bodies --
 * even ones cubes, odd ones spheres -- with seeded size and density, sampled
 * on their surface every surface_rel * spacing and stacked above the fluid.
 * Call with pos4 == NULL to get the counts.  body_id / rel4 are per
 * particle: -1 / 0 for non-rigid particles; rel4 = initial offset from the
 * body's centre of mass, w = the particle's mass.  Bodies carry the inverse
 * of their solid inertia tensor (the reference computes it into a temporary
 * and keeps a zero one, src/cpu/UnifiedPhysics.cpp:150).  Sets the domain.  */
int upf_scene_rigid(const upf_rigid_scene_desc* d, upf_params* p,
                    int* n_fluid, int* n_rigid, int* n_wall,
                    float* pos4, int* type, int* body_id, float* rel4,
                    upf_body* bodies);

/* ------------------------------------------------------- simulation */
typedef struct upf_sim upf_sim;

/* MallocDeviceArraysParticles + SetDeviceConstants + CreateLut* (src/gpu/particles.cu:100-261).
 * capacity = maximum particle count; the device is the current HIP device.  */
int upf_create(const upf_params* p, int capacity, upf_sim** out);
/* FreeDeviceArrays (src/gpu/particles.cu). */
int upf_destroy(upf_sim* s);
/* Replace the parameters (dt, stiffness, loop limits ...); the domain (block
 * size and shift, grid, scales, support radius, list capacity) must not change. */
int upf_set_params(upf_sim* s, const upf_params* p);

/* CopyZIndicesKVPairHostToDevice + CopyParticleDataHostToDevice
 * (src/gpu/particles.cu:349-501): host buffers are borrowed for the call.
 * vel4 may be NULL (zero velocity).  The upload order defines particle ids
 * 0..n-1.  Computes the z-index keys of the uploaded positions.  Boundary
 * volumes (Psi, CalculateWeightedVolume src/cpu/UnifiedPhysics.cpp:2703-2746)
 * are computed on the first step.                                        */
int upf_upload_particles(upf_sim* s, int n, const float* pos4, const float* vel4,
                         const int* type);
/* CopyRigidBodyDataHostToDevice (src/gpu/particles.cu:563): bodies and the
 * per-particle body id / initial relative position (upload order).  Body
 * ids must lie in [-1, n_bodies): otherwise UPF_ERR_ARG and nothing on the
 * device changes.  Any other failure leaves the simulation without bodies
 * (a rigid step then returns UPF_ERR_STATE) until an upload succeeds.      */
int upf_upload_bodies(upf_sim* s, int n_bodies, const upf_body* bodies,
                      const int* body_id, const float* rel4);

/* One or more full time steps of `method` (UPF_*), equivalent to calling the
 * reference's per-frame driver nsteps times.                              */
int upf_step(upf_sim* s, int nsteps, int method);

/* CopyParticlesCUDA (src/gpu/particles.cu:62-72) in upload (id) order.  Any
 * pointer may be NULL.  pos4/vel4: float4; rho = last predicted density
 * (PCISPH) or density (WCSPH); p = correction pressure (PCISPH) or pressure
 * (WCSPH); force4 = external force of the last step; psi = boundary volume.  */
int upf_download(upf_sim* s, float* pos4, float* vel4, float* rho, float* p,
                 float* force4, float* psi, int* type);
int upf_download_bodies(upf_sim* s, upf_body* bodies);

/* Per-step PCISPH statistics of the last upf_step call (the reference prints
 * them behind printDebuggingInfo, src/cpu/UnifiedPhysics.cpp:5391-5394):
 * iterations K and max predicted density of the last iteration, per step. */
int upf_get_step_stats(upf_sim* s, int max_steps, int* iterations, float* max_rho);

/* ------------------------------------------- stage entry points (tests) */
/* FillZindicesKernel / CalcIndex (src/gpu/particles_kernel.cu:3360,
 * src/cpu/zIndex.h:40-45) on the current positions, then RadixSort +
 * CopySortedParticleValues + BlockGeneration (src/gpu/particles.cu:31-60,904-964):
 * sorts the particle state by key (stable) and builds the cell table.     */
int upf_sort_and_bin(upf_sim* s);
/* Sorted keys, ids (the permutation: sorted slot -> upload id), and the
 * occupied-cell table (cell Morton code, first slot, end slot) in ascending
 * cell order.  *n_cells receives the number of occupied cells; cells may be
 * NULL to only query it.                                                   */
int upf_download_sorted(upf_sim* s, uint32_t* keys, int32_t* ids,
                        uint32_t* cells, int* n_cells);

/* GetNeighbors (src/cpu/UnifiedPhysics.cpp:2242-2295) of every liquid particle
 * on the current state, after sorting it as upf_sort_and_bin does: counts
 * per sorted slot and the lists (neighbour slots, ascending) concatenated
 * into flat (up to cap entries).  Returns the number of entries, or a
 * negative UPF_ERR_* code.                                                 */
int upf_download_neighbors(upf_sim* s, int* counts, int32_t* flat, int cap);

/* Device helpers used by the parity tests: keys of arbitrary positions and
 * a stable sort of arbitrary (key) arrays, both fully on the device.       */
int upf_keys(const upf_params* p, int n, const float* pos4, uint32_t* keys);
int upf_radix_sort(int n, const uint32_t* keys_in, uint32_t* keys_out, int32_t* perm_out);

/* No counterpart in the reference: upf_radix_sort for callers that cannot
 * allocate device memory (code under stream capture, serving loops,
 * harnesses that forbid hipMalloc).  keys_in, keys_out and perm_out are
 * device pointers to n elements each (no aliasing); stream is a hipStream_t
 * (NULL: the default stream); the call enqueues the sort and returns without
 * synchronising.  The sort's working memory is a fixed array of the library
 * (about 9 MB on every device it runs on, held until the process ends),
 * sized for n <= UPF_NOALLOC_SORT_MAX: larger n gives UPF_ERR_ARG.  The call
 * never allocates; calls that overlap on one device share the array and
 * race, so a caller serialises them (one stream, or events).               */
#define UPF_NOALLOC_SORT_MAX (1 << 20)
int upf_radix_sort_noalloc(int n, const uint32_t* keys_in, uint32_t* keys_out, int32_t* perm_out, void* stream);

/* ------------------------------------------------------ batched form */
/* Many independent scenes stepped together (no reference counterpart; the
 * reference holds one simulation per process).  Scenes share the physics
 * params; each brings its own domain (z-index scales, cell grid, collision
 * box) and density error factor in its own params, and its particles occupy
 * a fixed slot range.  One
 * launch per stage serves every scene: per-scene cell tables and loop
 * control, every scene running its own PCISPH loop to its own convergence.
 * Results per scene equal those of a plain simulation of that scene.       */
typedef struct upf_batch upf_batch;
/* counts[k] particles in scene k; max_record_steps bounds the per-step
 * statistics kept on the device.                                          */
int upf_batch_create(int n_scenes, const upf_params* scene_params,
                     const int* counts, int max_record_steps, upf_batch** out);
int upf_batch_destroy(upf_batch* b);
/* Particles of all scenes concatenated scene by scene. */
int upf_batch_upload(upf_batch* b, const float* pos4, const float* vel4, const int* type);
/* nsteps PCISPH static-boundary steps of every scene.  use_graph == 1: each
 * step replays one hipGraph holding the whole step, the correction loop
 * included (max_loops iterations, the convergence test on the device;
 * iterations past every scene's decision exit at once), with no host
 * round trip; errors are checked once at the end.  use_graph == 2: split
 * graphs -- the step's head with the first 4 iterations, then graphs of 4
 * further iterations while the device's decision (read back after each, the
 * previous step's count going ahead unread) asks for more, then the tail;
 * no idle iterations beyond the last chunk.  use_graph == 0: no graphs, the
 * host reads the loop decision back as upf_step does.  All three give the
 * same results bit for bit.                                               */
int upf_batch_step(upf_batch* b, int nsteps, int use_graph);
/* Concatenated upload order, like upf_batch_upload. */
int upf_batch_download(upf_batch* b, float* pos4, float* vel4, float* rho);
/* iterations / max_rho as [step][scene] for up to max_steps steps; returns
 * the number of steps recorded (negative UPF_ERR_* on failure).           */
int upf_batch_get_stats(upf_batch* b, int max_steps, int* iterations, float* max_rho);
/* The batch's simulation handle, for the profiling and download calls.
 * upf_set_params / upf_set_tuning on it apply to the graph replays too: the
 * batch recaptures its step graph when either changed since the capture.  */
int upf_batch_sim(upf_batch* b, upf_sim** sim);

/* ------------------------------------------------- differentiable form */
/* A rollout of `steps` static-boundary PCISPH steps with a fixed number of
 * correction iterations and its adjoint (no reference counterpart).  The
 * step is CpuParticlePhysicsStaticBoundariesPCISPH (src/cpu/UnifiedPhysics.cpp:4220)
 * in the CPU boundary mode with the kernels in closed form (KernelM4,
 * kernelPressureGrad, kernelViscosityLaplacian, src/cpu/WeightingKernel.cpp:75,165,190)
 * instead of their 300-entry tables, the pressure clamp max(rho - rho0, 0)
 * smoothed to tau log(1 + exp((rho - rho0) / tau)), tau = 1 kg/m^3, and the
 * box handled by position clamp + wall penalty force (no velocity
 * reflection), so that the rollout is differentiable.  Particles
 * keep their upload order (types: liquid and frozen).  The forward pass keeps
 * the state every `checkpoint_every` steps; the backward pass recomputes one
 * segment at a time and runs it in reverse.                                */
typedef struct upf_adjoint upf_adjoint;
int upf_adjoint_create(const upf_params* p, int n, int steps, int iterations,
                       int checkpoint_every, upf_adjoint** out);
int upf_adjoint_destroy(upf_adjoint* a);
/* x_final, v_final from x0, v0 (vel4 may be NULL: at rest). */
int upf_adjoint_forward(upf_adjoint* a, const float* pos4, const float* vel4, const int* type,
                        float* pos_out4, float* vel_out4);
/* dL/dx0, dL/dv0 from dL/dx_final (and dL/dv_final, may be NULL) of the
 * last forward call; outputs may be NULL.                                 */
int upf_adjoint_backward(upf_adjoint* a, const float* dl_dpos4, const float* dl_dvel4,
                         float* dl_dpos0_4, float* dl_dvel0_4);
/* Device time of the last forward / backward call (HIP events on the
 * rollout's stream) and the device memory the rollout holds.              */
int upf_adjoint_timing(upf_adjoint* a, float* forward_ms, float* backward_ms, size_t* device_bytes);

/* -------------------------------------------- OBJ model -> particles */
/* The scene-input path of the reference's rigid objects
 * (Scene::CreateFullParticleObject, src/sceneConfigure/Scene.cpp:797-824).
 *
 * upf_obj_read replaces ObjLoader::Load + LoadFaces (src/cpu/ObjLoader.cpp:18-258):
 * the triangles of a Wavefront OBJ file in file order, 9 floats each (three
 * corners), read as the reference reads them -- lines through a 256-byte
 * buffer, the last line of the file read twice (its end-of-file quirk), faces
 * as v/vt/vn triples.  faces9 == NULL: count only.  A missing file is
 * UPF_ERR_IO; a file without faces or vertices has no triangles (the
 * reference's FileError); a face or vertex line the reference would read
 * into undefined values is UPF_ERR_ARG.
 *
 * upf_mesh_fill replaces ObjectLoader::LoadParticles (src/cpu/ObjectLoader.cpp:8-97)
 * with its Polygon tests (src/cpu/Polygon.cpp): the lattice points of the
 * model's box (spacing / scale apart, odd x planes shifted half a step in y
 * and z) that the segment from outside the box reaches through an odd number
 * of faces, or that lie on a face, as (point - box minimum) * scale, in the
 * reference's loop order (x, then y, then z).  The box of every face starts
 * at the origin, as the reference's does.  Runs on the GPU (one thread per
 * lattice point).  pos4 == NULL: count only (n_out); cap < count:
 * UPF_ERR_RANGE with n_out set.  The caller adds the object's offset
 * (CreateFullParticleObject: position += offset).                          */
int upf_obj_read(const char* path, float* faces9, int cap, int* n_faces);
int upf_mesh_fill(const float* faces9, int n_faces, const float scale[3], float spacing, float* pos4, int cap,
                  int* n_out);

/* ----------------------------------------------------- particle state files */
/* upf_save_text writes the simulation's particles in slot order (the order
 * of the last z-index sort, as CopyParticlesCUDA hands them to the writers,
 * src/cpu/UnifiedPhysics.cpp:6176-6178; upload order before the first
 * step), byte for byte in the reference's
 * formats (src/cpu/UnifiedIO.cpp):
 *   UPF_TEXT_LIQUID     SaveLiquidParticleInfoFull (:62-113): count line, then
 *                       per liquid particle its position and type, the other
 *                       fields written as zeros (as the reference does);
 *   UPF_TEXT_RIGID      SaveRigidParticleInfoFull (:115-166): the same for
 *                       rigid and frozen particles;
 *   UPF_TEXT_POSITIONS  SaveParticlePositions (:13-60): every particle,
 *                       20 significant digits;
 *   UPF_TEXT_RESTART    the lines ReadInFluidParticlesPCISPH reads (:196-220):
 *                       position, type, velocity, correction pressure and
 *                       predicted density of the last step, 9 significant
 *                       digits (each float exactly).  Loaded back by
 *                       upf_read_restart + upf_upload_restart, the next steps
 *                       equal the steps of the simulation that wrote it.
 * The count line is the number of particle lines that follow.  The file name
 * is the caller's (the reference numbers files under a fixed directory).  */
#define UPF_TEXT_LIQUID     1
#define UPF_TEXT_RIGID      2
#define UPF_TEXT_POSITIONS  3
#define UPF_TEXT_RESTART    4
int upf_save_text(upf_sim* s, const char* path, int which);
/* The same files from host arrays (n particles, e.g. from upf_download):
 * pos4 float4, type per particle; vel4 / corr_pressure / predicted_density
 * are read for UPF_TEXT_RESTART only.  Host code, no device needed.      */
int upf_write_text(const char* path, int which, int n, const float* pos4, const int* type, const float* vel4,
                   const float* corr_pressure, const float* predicted_density);

/* DumpGPUParamsToFile (src/cpu/UnifiedPhysics.cpp:411-474), the binary
 * cross-check dump of the particles and the block (cell) table: sorts the
 * current state and builds its cell table as the next step's start does
 * (the state and the next steps are unchanged), then appends one record to
 * each file.  particles_path: uint32 n, then n float4 (x, y, z, w) in slot
 * order, w = the z-index of the particle's uploaded position as a float (the
 * reference's w, set at upload and never rewritten).  blocks_path: uint32
 * number of occupied blocks, then per occupied block, by ascending block
 * index (grid_dim^2 * bz + grid_dim * by + bx), uint32 (block index, first
 * slot, particle count).  Native-endian.  Single-scene simulations.     */
int upf_dump_cells(upf_sim* s, const char* particles_path, const char* blocks_path);

/* ReadInFluidParticleNum + ReadInFluidParticlesPCISPH (src/cpu/UnifiedIO.cpp:168-220):
 * the count, then 9 whitespace-separated fields per particle.  All output
 * pointers NULL: count only (n_out).  cap < count: UPF_ERR_RANGE with n_out
 * set.  A missing file is UPF_ERR_IO (the reference exits); fewer particles
 * than the count, or a field that is not a decimal number (int for the
 * type), is UPF_ERR_ARG (the reference stops reading and keeps undefined
 * values).  pos4/vel4 are float4 with w = 0.                             */
int upf_read_restart(const char* path, int cap, int* n_out, float* pos4, float* vel4, int* type,
                     float* corr_pressure, float* predicted_density);

/* Scene::readFluidObjectFromFile (src/sceneConfigure/Scene.cpp:648-697)
 * followed by the upload: as upf_upload_particles, with each particle's
 * correction pressure and predicted density of the step before (the
 * UnifiedParticle constructor that takes them, src/cpu/UnifiedParticle.cpp:29-45;
 * the pressure of two steps back, which that constructor leaves unset, is 0).
 * Liquid and frozen particles (rigid bodies are not part of the format).  */
int upf_upload_restart(upf_sim* s, int n, const float* pos4, const float* vel4, const int* type,
                       const float* corr_pressure, const float* predicted_density);

/* Rigid bodies from particles: Scene::addAllTo -> SolidifyParticle ->
 * FreezeParticle (src/sceneConfigure/Scene.cpp:149-208,
 * src/cpu/UnifiedPhysics.cpp:810-813,1657-1847) with AddToRigidBody
 * (:825-883), finished as UnifiedPhysics' constructor finishes them
 * (:141-160).  The n particles given are the solid ones (e.g. an OBJ fill of
 * upf_mesh_fill); those closer than the support radius (|x_i - x_j| < h, the
 * CPU path's neighbour test) belong to one body, and a body needs two
 * particles (a lone particle gets body_id -1).  Bodies are numbered by their
 * first particle and accumulate their particles in index order with
 * AddToRigidBody's float arithmetic: mass, centre of mass (pos), velocity
 * (from vel4, NULL = at rest), inertia about the centre, angular momentum;
 * inv_inertia_local = Matrix3x3::CalculateInverse of the inertia (the zero
 * matrix if it is singular, as there); rotation identity, quaternion
 * (1,0,0,0); linear momentum, angular velocity and inv_inertia_world 0, as
 * RigidBody() leaves them until the first rigid step.  rel4 (may be NULL):
 * position - centre of mass (init_relative_pos_), w = the particle's mass.
 * bodies == NULL: body ids and count only; cap_bodies < count: UPF_ERR_RANGE
 * with n_bodies set.  Host code.                                           */
int upf_bodies_from_particles(const upf_params* p, int n, const float* pos4, const float* vel4,
                              const float* mass, int* body_id, float* rel4, upf_body* bodies,
                              int cap_bodies, int* n_bodies);

/* ------------------------------------------------------ launch tuning */
/* Launch shapes that change speed only, never results (no reference
 * counterpart; the reference fixes them at compile time).  Defaults
 * {256, 8192, 4, 0, 4, 1, 8, 0, 4, 1}.  A batch recaptures its step graph when its
 * tuning (set through its upf_batch_sim handle) changed.                   */
typedef struct upf_tuning {
    int loop_block;   /* threads per block of the grid-stride passes: 64, 128, 256     */
    int grid_cap;     /* blocks of the grid-stride passes at most (>= 1)               */
    int nbr_rounds;   /* list build: candidates per lane per chunk, 2, 4 or 8 (x 64)    */
    int nbr_stage;    /* list build: entries per list staged in LDS: 0, 32 or 64       */
    int speculate;    /* host loop: iterations enqueued per read-back of the decision (1 .. 64) */
    int loop_list;    /* correction loop reads: 0 packed 32-bit lists, 1 delta-coded,
                         2 the neighbours' records from LDS through per-chunk unions
                         (static PCISPH; needs loop_block 256)                     */
    int loop_width;   /* correction loop: neighbours whose loads are issued together, 4 or 8 */
    int nbr_tile;     /* list build: 0 cell by cell (nbr_rounds, nbr_stage), or the tile
                         form with 256 / 512 candidates staged in LDS at a time (the
                         coupling step's rigid lists: then one thread per rigid slot) */
    int spec_first;   /* host loop: first enqueue the previous step's iteration count,
                         at most this many past the minimum (0 .. 64)                  */
    int xcd_remap;    /* 1: each XCD (own L2) takes one contiguous eighth of the blocks
                         of the list build and the grid-stride passes; 0: round robin   */
    int nbr_skip;     /* list build: 1 skips, per particle, a round of 64 candidates none
                         of whose cells lies in its query box (pays where cells are much
                         wider than h); 0 tests every round                            */
    int nbr_fuse;     /* 1: the list build and the external-force pass run as one kernel
                         (cell-by-cell build without LDS staging, delta-coded loop lists);
                         0: two kernels                                               */
} upf_tuning;
int upf_set_tuning(upf_sim* s, const upf_tuning* t);
int upf_get_tuning(const upf_sim* s, upf_tuning* t);

/* ------------------------------------------------------ profiling */
/* Per-kernel-class device time measured with HIP events recorded on the
 * simulation's own stream around every launch (the reference's
 * SPH_PROFILING_VERBOSE cudaEvent timers, src/gpu/particles.cu:1460-1690).
 * Classes are numbered 0 .. upf_profile_classes()-1.                      */
int upf_profile_enable(upf_sim* s, int on);
int upf_profile_reset(upf_sim* s);
int upf_profile_classes(void);
const char* upf_profile_name(int cls);
int upf_profile_read(upf_sim* s, int cls, double* total_ms, long long* launches);

/* ------------------------------------------------------ version */
const char* upf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* UPF_H_ */
