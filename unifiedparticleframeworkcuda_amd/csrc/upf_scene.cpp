// upf_scene.cpp -- seeded scene generators (host).  They replace the
// reference's Scene::build / createFluidObject / createBoxBoundaryFrozenParticles
// (src/sceneConfigure/Scene.cpp:23-138,597-645,890-1042) with analytic,
// reproducible inputs: a jittered fluid lattice in a box corner, wall
// particles on the floor and four sides, and surface-sampled rigid boxes and
// spheres.  The domain (virtual z-index box, collision box) is chosen so that
// every particle keys inside the 1024^3 grid and one cell holds the support
// radius.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "upf.h"
#include "upf_internal.h"

namespace upf {

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Domain for a container of extent box[3] whose wall layer starts at `pad`:
// cubic virtual box (isotropic cells) grown until the block holds h.  A cell
// is a power-of-two block of the 1024-unit z-index grid, so that first extent
// can leave cells up to 2 h wide, and the list build tests (cell / h)^3 times
// the candidates it would with cells of h.  When cells come out wider than
// 1.2 h, the extent grows further, to where the next smaller block holds h
// with a small margin (cells of 1.0005 h, eight times as many).
int choose_domain(upf_params* p, const float box[3], float pad) {
    float e = 0.0f;
    for (int a = 0; a < 3; ++a) e = std::max(e, box[a]);
    e += 2.0f * pad + 2.0f * p->spacing;
    const float r = p->particle_radius;
    float cmin[3], cmax[3];
    for (int a = 0; a < 3; ++a) {
        cmin[a] = pad + r;
        cmax[a] = pad + box[a] - r;  // container interior [pad, pad + box]
    }
    auto fit = [&](float e0, upf_params& out) {
        for (int tries = 0; tries < 400; ++tries, e0 *= 1.002f) {
            const float vmax[3] = {e0, e0, e0};
            upf_params q = *p;
            if (upf_params_set_domain(&q, vmax, cmin, cmax) == UPF_OK) { out = q; return true; }
        }
        return false;
    };
    upf_params q;
    if (!fit(e, q)) return UPF_ERR_ARG;
    const float wide = (float)q.block_size / (q.scales[0] * q.support_radius);
    if (wide > 1.2f && q.block_size >= 8) {
        upf_params f;
        const float e1 = 1024.0f * p->support_radius * 1.0005f / (float)(q.block_size / 2);
        if (fit(e1, f) && f.block_size == q.block_size / 2) q = f;
    }
    *p = q;
    return UPF_OK;
}

}  // namespace upf

using namespace upf;

extern "C" {

int upf_scene_dam_break(const upf_dam_break_desc* d, upf_params* p, int* n_fluid, int* n_wall,
                        float* pos4, int* type) {
    if (!d || !p || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || !(d->wall_rel > 0.0f)) return UPF_ERR_ARG;
    const double lat = d->lattice_rel > 0.0f ? d->lattice_rel : 1.0;
    const double dx = p->spacing;
    const double ws = d->wall_rel * dx;
    const float pad = p->support_radius;
    const float box[3] = {(float)(d->box[0] * lat), (float)(d->box[1] * lat), (float)(d->box[2] * lat)};
    int rc = choose_domain(p, box, pad);
    if (rc != UPF_OK) return rc;
    // Container interior [pad, pad + box]; the wall layer lies one particle
    // radius outside it and the collision box one radius inside it, so the
    // first fluid layer is one spacing from the wall layer, as in the
    // reference (walls at h - r, fluid and collision box from h + r:
    // Scene.cpp:41,902,928, UnifiedConstants.cpp:374-375).
    const double r = 0.5 * dx;
    const double o = pad - r;
    const double lx = box[0] + 2.0 * r, ly = box[1] + 2.0 * r, lz = box[2] + 2.0 * r;
    // wall lattice: floor (y = o), sides x = o, x = o+lx, z = o, z = o+lz; open top
    // per axis the wall grid spans the wall extent exactly (the sampling step
    // is stretched to a whole number of steps), so the far walls keep the
    // same gap to the fluid as the near ones
    const int wx = std::max(1, (int)std::floor(lx / ws + 0.5)), wy = std::max(1, (int)std::floor(ly / ws + 0.5)),
              wz = std::max(1, (int)std::floor(lz / ws + 0.5));
    const double sx = lx / wx, sy = ly / wy, sz = lz / wz;
    std::vector<float> walls;
    auto add_wall = [&](int i, int j, int k) {
        walls.push_back((float)(o + i * sx));
        walls.push_back((float)(o + j * sy));
        walls.push_back((float)(o + k * sz));
    };
    for (int i = 0; i <= wx; ++i)
        for (int k = 0; k <= wz; ++k) add_wall(i, 0, k);           // floor
    for (int j = 1; j <= wy; ++j)
        for (int k = 0; k <= wz; ++k) { add_wall(0, j, k); add_wall(wx, j, k); }   // x sides
    for (int j = 1; j <= wy; ++j)
        for (int i = 1; i < wx; ++i) { add_wall(i, j, 0); add_wall(i, j, wz); }    // z sides
    const int nf = d->nx * d->ny * d->nz;
    const int nw = (int)(walls.size() / 3);
    if (n_fluid) *n_fluid = nf;
    if (n_wall) *n_wall = nw;
    if (!pos4) return UPF_OK;
    uint64_t s = d->seed * 0x2545F4914F6CDD1Dull + 1;
    const double jit = d->jitter_rel * dx;
    int w = 0;
    // fluid lattice (y slowest, z fastest, as createFluidObject loops)
    for (int iy = 0; iy < d->ny; ++iy)
        for (int ix = 0; ix < d->nx; ++ix)
            for (int iz = 0; iz < d->nz; ++iz) {
                const double jx = (2.0 * uniform01(s) - 1.0) * jit;
                const double jy = (2.0 * uniform01(s) - 1.0) * jit;
                const double jz = (2.0 * uniform01(s) - 1.0) * jit;
                pos4[4 * w + 0] = (float)(pad + (ix + 0.5) * lat * dx + jx);
                pos4[4 * w + 1] = (float)(pad + (iy + 0.5) * lat * dx + jy);
                pos4[4 * w + 2] = (float)(pad + (iz + 0.5) * lat * dx + jz);
                pos4[4 * w + 3] = 0.0f;
                if (type) type[w] = UPF_LIQUID;
                ++w;
            }
    for (int k = 0; k < nw; ++k) {
        pos4[4 * w + 0] = walls[3 * k];
        pos4[4 * w + 1] = walls[3 * k + 1];
        pos4[4 * w + 2] = walls[3 * k + 2];
        pos4[4 * w + 3] = 0.0f;
        if (type) type[w] = UPF_FROZEN;
        ++w;
    }
    return UPF_OK;
}

}  // extern "C"

namespace {

struct RigidOut {
    std::vector<float> pos, rel;  // xyz per particle; rel w = particle mass
    std::vector<int> body;
};

// points on the surface of a cube of edge a centred at the origin: the cube
// lattice points with at least one coordinate on a face
void sample_cube(double a, double s, std::vector<double>& pts) {
    const int m = std::max(1, (int)std::floor(a / s + 0.5));
    const double st = a / m;
    for (int i = 0; i <= m; ++i)
        for (int j = 0; j <= m; ++j)
            for (int k = 0; k <= m; ++k) {
                if (i != 0 && i != m && j != 0 && j != m && k != 0 && k != m) continue;
                pts.push_back(-0.5 * a + i * st);
                pts.push_back(-0.5 * a + j * st);
                pts.push_back(-0.5 * a + k * st);
            }
}

// Fibonacci lattice on a sphere of radius r (about one point per s^2)
void sample_sphere(double r, double s, std::vector<double>& pts) {
    const int n = std::max(8, (int)std::floor(4.0 * M_PI * r * r / (s * s) + 0.5));
    const double ga = M_PI * (3.0 - std::sqrt(5.0));
    for (int i = 0; i < n; ++i) {
        const double y = 1.0 - 2.0 * (i + 0.5) / n;
        const double rr = std::sqrt(std::max(0.0, 1.0 - y * y));
        const double th = ga * i;
        pts.push_back(r * std::cos(th) * rr);
        pts.push_back(r * y);
        pts.push_back(r * std::sin(th) * rr);
    }
}

}  // namespace

extern "C" int upf_scene_rigid(const upf_rigid_scene_desc* d, upf_params* p, int* n_fluid, int* n_rigid, int* n_wall,
                               float* pos4, int* type, int* body_id, float* rel4, upf_body* bodies) {
    if (!d || !p || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || d->n_bodies < 0 || !(d->surface_rel > 0.0f)) return UPF_ERR_ARG;
    // fluid + walls of the dam break
    upf_dam_break_desc db;
    std::memset(&db, 0, sizeof(db));
    db.nx = d->nx; db.ny = d->ny; db.nz = d->nz;
    for (int a = 0; a < 3; ++a) db.box[a] = d->box[a];
    db.jitter_rel = d->jitter_rel;
    db.wall_rel = d->wall_rel > 0.0f ? d->wall_rel : 1.0f;
    db.lattice_rel = d->lattice_rel > 0.0f ? d->lattice_rel : 1.0f;
    db.seed = d->seed;
    int nf = 0, nw = 0;
    int rc = upf_scene_dam_break(&db, p, &nf, &nw, nullptr, nullptr);
    if (rc != UPF_OK) return rc;
    // bodies: columns 0.2 m apart over the floor, layers 0.2 m apart above the fluid
    const double lat = db.lattice_rel;
    const double pad = p->support_radius;
    const double Lx = d->box[0] * lat, Ly = d->box[1] * lat, Lz = d->box[2] * lat;
    const double fluid_top = pad + d->ny * lat * p->spacing;
    const double slot = 0.2;
    const int cx = std::max(1, (int)std::floor(Lx / slot)), cz = std::max(1, (int)std::floor(Lz / slot));
    uint64_t rs = d->seed * 0x9E3779B97F4A7C15ull + 7;
    RigidOut R;
    std::vector<upf_body> B(d->n_bodies);
    const double s = d->surface_rel * p->spacing;
    for (int b = 0; b < d->n_bodies; ++b) {
        const double size = d->size_min + (d->size_max - d->size_min) * uniform01(rs);
        const double dens = d->density_min + (d->density_max - d->density_min) * uniform01(rs);
        const int col = b % (cx * cz), layer = b / (cx * cz);
        const double jx = (uniform01(rs) - 0.5) * 0.04, jz = (uniform01(rs) - 0.5) * 0.04, jy = uniform01(rs) * 0.03;
        const double ccx = pad + (col % cx + 0.5) * (Lx / cx) + jx;
        const double ccz = pad + (col / cx + 0.5) * (Lz / cz) + jz;
        const double ccy = fluid_top + 0.1 + 0.5 * size + layer * slot + jy;
        if (ccy + 0.5 * size > pad + Ly - p->particle_radius) return UPF_ERR_ARG;  // container too low
        std::vector<double> pts;
        double mass, I;
        if (b % 2 == 0) {
            sample_cube(size, s, pts);
            mass = dens * size * size * size;
            I = mass * size * size / 6.0;
        } else {
            sample_sphere(0.5 * size, s, pts);
            mass = dens * M_PI * size * size * size / 6.0;
            I = 0.4 * mass * 0.25 * size * size;
        }
        const int np = (int)(pts.size() / 3);
        const float pm = (float)(mass / np);
        for (int k = 0; k < np; ++k) {
            R.pos.push_back((float)(ccx + pts[3 * k])); R.pos.push_back((float)(ccy + pts[3 * k + 1]));
            R.pos.push_back((float)(ccz + pts[3 * k + 2]));
            R.rel.push_back((float)pts[3 * k]); R.rel.push_back((float)pts[3 * k + 1]); R.rel.push_back((float)pts[3 * k + 2]);
            R.rel.push_back(pm);
            R.body.push_back(b);
        }
        upf_body& o = B[b];
        std::memset(&o, 0, sizeof(o));
        o.pos[0] = (float)ccx; o.pos[1] = (float)ccy; o.pos[2] = (float)ccz;
        o.mass = (float)mass;
        o.quaternion[0] = 1.0f;
        for (int k = 0; k < 3; ++k) {
            o.rotation[4 * k] = 1.0f;
            o.inv_inertia_local[4 * k] = (float)(1.0 / I);
            o.inv_inertia_world[4 * k] = (float)(1.0 / I);
        }
    }
    const int nr = (int)R.body.size();
    if (n_fluid) *n_fluid = nf;
    if (n_wall) *n_wall = nw;
    if (n_rigid) *n_rigid = nr;
    if (!pos4) return UPF_OK;
    rc = upf_scene_dam_break(&db, p, &nf, &nw, pos4, type);  // fluid, then walls
    if (rc != UPF_OK) return rc;
    const int n0 = nf + nw;
    for (int i = 0; i < n0; ++i) {
        if (body_id) body_id[i] = -1;
        if (rel4) { rel4[4 * i] = rel4[4 * i + 1] = rel4[4 * i + 2] = rel4[4 * i + 3] = 0.0f; }
    }
    for (int k = 0; k < nr; ++k) {
        const int i = n0 + k;
        pos4[4 * i] = R.pos[3 * k]; pos4[4 * i + 1] = R.pos[3 * k + 1]; pos4[4 * i + 2] = R.pos[3 * k + 2];
        pos4[4 * i + 3] = 0.0f;
        if (type) type[i] = UPF_RIGID;
        if (body_id) body_id[i] = R.body[k];
        if (rel4) for (int a = 0; a < 4; ++a) rel4[4 * i + a] = R.rel[4 * k + a];
    }
    if (bodies) std::memcpy(bodies, B.data(), B.size() * sizeof(upf_body));
    return UPF_OK;
}

// ---------------------------------------------------- bodies from particles
namespace {

// Matrix3x3::CalculateInverse (src/cpu/Matrix3x3.cpp:121-158): the adjugate
// over the determinant, expression by expression; false (result untouched)
// when |det| < 1e-20.
bool invert3(const float m[3][3], float out[9]) {
    const float det = m[0][0] * m[1][1] * m[2][2] - m[2][0] * m[1][1] * m[0][2] + m[0][1] * m[1][2] * m[2][0] -
                      m[2][1] * m[1][2] * m[0][0] + m[0][2] * m[1][0] * m[2][1] - m[2][2] * m[1][0] * m[0][1];
    if ((double)det < 1e-20 && (double)det > -1e-20) return false;
    const float f = 1 / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float s[4];
            int e = 0;
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l)
                    if (k != i && l != j) s[e++] = m[k][l];
            float v = s[0] * s[3] - s[2] * s[1];
            v *= ((i + j) & 1) ? -f : f;
            out[3 * j + i] = v;
        }
    return true;
}

struct Accum {  // RigidBody fields that AddToRigidBody changes
    float mass = 0.0f;
    float com[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, L[3] = {0, 0, 0};
    float I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
};

void cross3(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// AddToRigidBody (src/cpu/UnifiedPhysics.cpp:825-883): one particle joins the
// body; centre of mass, inertia about it, angular momentum and velocity are
// updated incrementally in float (vmml's v / s is v * (1 / s)).
void add_particle(Accum& b, const float x[3], const float v[3], float m) {
    const float M = b.mass, newMass = M + m, inv = 1.0f / newMass;
    float c[3], dc[3], dp[3];
    for (int a = 0; a < 3; ++a) c[a] = (b.com[a] * M + x[a] * m) * inv;
    for (int a = 0; a < 3; ++a) dc[a] = b.com[a] - c[a];
    for (int a = 0; a < 3; ++a) dp[a] = x[a] - c[a];
    const float cs0 = dc[0] * dc[0], cs1 = dc[1] * dc[1], cs2 = dc[2] * dc[2];
    const float ps0 = dp[0] * dp[0], ps1 = dp[1] * dp[1], ps2 = dp[2] * dp[2];
    b.I[0][0] += M * (cs1 + cs2) + m * (ps1 + ps2);
    b.I[1][1] += M * (cs0 + cs2) + m * (ps0 + ps2);
    b.I[2][2] += M * (cs1 + cs0) + m * (ps1 + ps0);
    b.I[0][1] -= M * dc[0] * dc[1] + m * dp[0] * dp[1];
    b.I[0][2] -= M * dc[0] * dc[2] + m * dp[0] * dp[2];
    b.I[1][2] -= M * dc[1] * dc[2] + m * dp[1] * dp[2];
    b.I[1][0] = b.I[0][1];
    b.I[2][0] = b.I[0][2];
    b.I[2][1] = b.I[1][2];
    float mom[3], t[3];
    for (int a = 0; a < 3; ++a) mom[a] = b.vel[a] * M;
    cross3(dc, mom, t);
    for (int a = 0; a < 3; ++a) b.L[a] = b.L[a] + t[a];
    for (int a = 0; a < 3; ++a) mom[a] = v[a] * m;
    cross3(dp, mom, t);
    for (int a = 0; a < 3; ++a) b.L[a] += t[a];
    for (int a = 0; a < 3; ++a) b.vel[a] = (b.vel[a] * M + v[a] * m) * inv;
    for (int a = 0; a < 3; ++a) b.com[a] = c[a];
    b.mass = newMass;
}

int find_root(std::vector<int>& parent, int i) {
    while (parent[i] != i) {
        parent[i] = parent[parent[i]];
        i = parent[i];
    }
    return i;
}

}  // namespace

extern "C" int upf_bodies_from_particles(const upf_params* p, int n, const float* pos4, const float* vel4,
                                         const float* mass, int* body_id, float* rel4, upf_body* bodies,
                                         int cap_bodies, int* n_bodies) {
    if (!p || n < 0 || (n > 0 && (!pos4 || !mass)) || !body_id || !n_bodies || cap_bodies < 0) return UPF_ERR_ARG;
    const float T = p->support_radius_sq_limit;
    const double cell = (double)p->support_radius;
    if (!(T > 0.0f) || !(cell > 0.0)) return UPF_ERR_ARG;
    for (int i = 0; i < 3 * n; ++i)  // finite positions, cell indices well inside 64 bits
        if (!(std::fabs(pos4[4 * (i / 3) + i % 3] / cell) < 1e15)) return UPF_ERR_ARG;
    // neighbours (|x_i - x_j| < h, the CPU path's lists) through a hash grid
    // of h-sized cells; components by union-find
    std::vector<int> parent(n);
    for (int i = 0; i < n; ++i) parent[i] = i;
    struct Key {
        long long x, y, z;
    };
    auto key_of = [&](int i) {
        return Key{(long long)std::floor(pos4[4 * i] / cell), (long long)std::floor(pos4[4 * i + 1] / cell),
                   (long long)std::floor(pos4[4 * i + 2] / cell)};
    };
    auto hash = [](long long x, long long y, long long z) {
        return (uint64_t)(x * 73856093LL) ^ (uint64_t)(y * 19349663LL) ^ (uint64_t)(z * 83492791LL);
    };
    std::vector<std::pair<uint64_t, int>> cells((size_t)n);
    for (int i = 0; i < n; ++i) {
        const Key k = key_of(i);
        cells[i] = {hash(k.x, k.y, k.z), i};
    }
    std::sort(cells.begin(), cells.end());
    for (int i = 0; i < n; ++i) {
        const Key k = key_of(i);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const uint64_t h = hash(k.x + dx, k.y + dy, k.z + dz);
                    auto it = std::lower_bound(cells.begin(), cells.end(), std::make_pair(h, -1));
                    for (; it != cells.end() && it->first == h; ++it) {
                        const int j = it->second;
                        if (j <= i) continue;
                        const float ex = pos4[4 * i] - pos4[4 * j], ey = pos4[4 * i + 1] - pos4[4 * j + 1],
                                    ez = pos4[4 * i + 2] - pos4[4 * j + 2];
                        if (!(ex * ex + ey * ey + ez * ez < T)) continue;
                        const int a = find_root(parent, i), b = find_root(parent, j);
                        if (a != b) parent[std::max(a, b)] = std::min(a, b);
                    }
                }
    }
    // bodies numbered by their first particle; a particle alone forms none
    std::vector<int> size(n, 0), index(n, -1);
    for (int i = 0; i < n; ++i) ++size[find_root(parent, i)];
    int nb = 0;
    for (int i = 0; i < n; ++i) {
        const int r = find_root(parent, i);
        if (size[r] < 2) {
            body_id[i] = -1;
            continue;
        }
        if (index[r] < 0) index[r] = nb++;
        body_id[i] = index[r];
    }
    *n_bodies = nb;
    if (!bodies) return UPF_OK;
    if (nb > cap_bodies) return UPF_ERR_RANGE;
    std::vector<Accum> acc(nb);
    const float zero[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        if (body_id[i] >= 0) add_particle(acc[body_id[i]], &pos4[4 * i], vel4 ? &vel4[4 * i] : zero, mass[i]);
    for (int b = 0; b < nb; ++b) {
        upf_body& o = bodies[b];
        std::memset(&o, 0, sizeof(o));  // RigidBody(): momenta, angular velocity, forces 0
        const Accum& A = acc[b];
        for (int a = 0; a < 3; ++a) {
            o.pos[a] = A.com[a];
            o.vel[a] = A.vel[a];
            o.angular_momentum[a] = A.L[a];
            o.rotation[4 * a] = 1.0f;
        }
        o.mass = A.mass;
        o.quaternion[0] = 1.0f;
        // UnifiedPhysics' constructor inverts the inertia once the scene is
        // built (:150); a singular one leaves the zero matrix
        invert3(A.I, o.inv_inertia_local);
    }
    if (rel4)
        for (int i = 0; i < n; ++i) {  // init_relative_pos_ = position - centre of mass (:158)
            const int b = body_id[i];
            for (int a = 0; a < 3; ++a) rel4[4 * i + a] = b >= 0 ? pos4[4 * i + a] - acc[b].com[a] : 0.0f;
            rel4[4 * i + 3] = mass[i];
        }
    return UPF_OK;
}
