// oracle_capi.cpp - C ABI of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// Loaded by the tests, the smoke check and the CPU-baseline leg of the
// benchmark; the engine never links it.  `precision` selects float (4) or
// double (8) arithmetic for the whole substep.
#include <memory>
#include <new>

#include "mpm_oracle.hpp"

using namespace mpm_oracle;

namespace {

struct Handle {
    int precision;
    std::unique_ptr<State<float>> f;
    std::unique_ptr<State<double>> d;
};

template <class R>
void load(State<R> &s, const mpm_consts &c, const mpm_particles_view &p) {
    const size_t n = p.count;
    s.consts = c;
    s.time = p.time;
    s.sort_map.resize(n);
    s.reverse_sort_map.resize(n);
    for (size_t i = 0; i < n; ++i) s.sort_map[i] = s.reverse_sort_map[i] = (uint32_t)i;
    s.flags.assign(p.flags, p.flags + n);
    s.parameters.assign(p.parameters, p.parameters + n);
    s.elastic_energies.resize(n);
    s.collider_bits.resize(n);
    s.positions.resize(n);
    s.initial_positions.resize(n);
    s.velocities.resize(n);
    s.position_gradients.resize(n);
    s.velocity_gradients.resize(n);
    s.plastic_jp.resize(n);
    for (size_t i = 0; i < n; ++i) {
        s.elastic_energies[i] = p.elastic_energies ? R(p.elastic_energies[i]) : R(0);
        s.collider_bits[i] = p.collider_bits ? p.collider_bits[i] : 0u;
        s.plastic_jp[i] = p.plastic_jp ? R(p.plastic_jp[i]) : R(1);
        for (int a = 0; a < 3; ++a) {
            s.positions[i][a] = R(p.positions[3 * i + a]);
            s.initial_positions[i][a] = p.initial_positions ? R(p.initial_positions[3 * i + a]) : R(0);
            s.velocities[i][a] = R(p.velocities[3 * i + a]);
        }
        // Column-major [[f32;3];3]: element (row r, col c) at 3*c + r.
        for (int col = 0; col < 3; ++col)
            for (int r = 0; r < 3; ++r) {
                s.position_gradients[i].m[r][col] = R(p.position_gradients[9 * i + 3 * col + r]);
                s.velocity_gradients[i].m[r][col] = R(p.velocity_gradients[9 * i + 3 * col + r]);
            }
    }
}

template <class R>
void store(const State<R> &s, mpm_particles_view &o) {
    // to_io_state, rust/crates/cpu/src/cpu_state.rs:71-100: permute by reverse_sort_map.
    const size_t n = s.positions.size();
    o.time = s.time;
    for (size_t orig = 0; orig < n; ++orig) {
        size_t i = s.reverse_sort_map[orig];
        if (o.flags) o.flags[orig] = s.flags[i];
        if (o.parameters) o.parameters[orig] = s.parameters[i];
        if (o.elastic_energies) o.elastic_energies[orig] = (float)s.elastic_energies[i];
        if (o.collider_bits) o.collider_bits[orig] = s.collider_bits[i];
        if (o.plastic_jp) o.plastic_jp[orig] = (float)s.plastic_jp[i];
        for (int a = 0; a < 3; ++a) {
            if (o.positions) o.positions[3 * orig + a] = (float)s.positions[i][a];
            if (o.velocities) o.velocities[3 * orig + a] = (float)s.velocities[i][a];
            if (o.initial_positions) o.initial_positions[3 * orig + a] = (float)s.initial_positions[i][a];
        }
        for (int col = 0; col < 3; ++col)
            for (int r = 0; r < 3; ++r) {
                if (o.position_gradients) o.position_gradients[9 * orig + 3 * col + r] = (float)s.position_gradients[i].m[r][col];
                if (o.velocity_gradients) o.velocity_gradients[9 * orig + 3 * col + r] = (float)s.velocity_gradients[i].m[r][col];
            }
    }
}

template <class R>
void store_f64(const State<R> &s, double *x, double *v, double *F, double *C) {
    const size_t n = s.positions.size();
    for (size_t orig = 0; orig < n; ++orig) {
        size_t i = s.reverse_sort_map[orig];
        for (int a = 0; a < 3; ++a) {
            if (x) x[3 * orig + a] = (double)s.positions[i][a];
            if (v) v[3 * orig + a] = (double)s.velocities[i][a];
        }
        for (int col = 0; col < 3; ++col)
            for (int r = 0; r < 3; ++r) {
                if (F) F[9 * orig + 3 * col + r] = (double)s.position_gradients[i].m[r][col];
                if (C) C[9 * orig + 3 * col + r] = (double)s.velocity_gradients[i].m[r][col];
            }
    }
}

void copy_input(FrameInputCopy &dst, const mpm_frame_input *in, size_t n, const MeshTopology &mesh) {
    dst = FrameInputCopy();
    if (!in) return;
    const size_t nv = mesh.num_vertices, nt = mesh.num_triangles();
    if (nt > 0 && in->vertex_positions_a && in->triangle_frictions_a && in->triangle_dampings_a) {
        dst.vertices_a.assign(in->vertex_positions_a, in->vertex_positions_a + 3 * nv);
        dst.frictions_a.assign(in->triangle_frictions_a, in->triangle_frictions_a + nt);
        dst.dampings_a.assign(in->triangle_dampings_a, in->triangle_dampings_a + nt);
        if (in->has_b && in->vertex_positions_b) {
            dst.vertices_b.assign(in->vertex_positions_b, in->vertex_positions_b + 3 * nv);
            dst.frictions_b.assign(in->triangle_frictions_b, in->triangle_frictions_b + nt);
            dst.dampings_b.assign(in->triangle_dampings_b, in->triangle_dampings_b + nt);
        }
    }
    dst.frame = in->frame;
    for (int a = 0; a < 3; ++a) {
        dst.gravity_a[a] = in->gravity_a[a];
        dst.gravity_b[a] = in->gravity_b[a];
    }
    dst.has_b = in->has_b != 0;
    if (in->particle_flags_a && in->goal_positions_a) {
        dst.flags_a.assign(in->particle_flags_a, in->particle_flags_a + n);
        dst.goals_a.assign(in->goal_positions_a, in->goal_positions_a + 3 * n);
        if (dst.has_b) {
            dst.flags_b.assign(in->particle_flags_b, in->particle_flags_b + n);
            dst.goals_b.assign(in->goal_positions_b, in->goal_positions_b + 3 * n);
        }
    }
}

template <class R>
int64_t read_grid(const State<R> &s, int32_t *ids, float *masses, float *vel, int64_t cap) {
    int64_t n = (int64_t)s.keys.size();
    for (int64_t k = 0; k < n && k < cap; ++k) {
        for (int a = 0; a < 3; ++a) {
            if (ids) ids[3 * k + a] = s.keys[k].id[a];
            if (vel) vel[3 * k + a] = (float)s.node_velocities[k][a];
        }
        if (masses) masses[k] = (float)s.masses[k];
    }
    return n;
}

template <class R>
void read_grid_f64(const State<R> &s, double *masses, double *vel) {
    for (size_t k = 0; k < s.keys.size(); ++k) {
        if (masses) masses[k] = (double)s.masses[k];
        for (int a = 0; a < 3; ++a)
            if (vel) vel[3 * k + a] = (double)s.node_velocities[k][a];
    }
}

template <class R>
void stress_impl(const uint32_t *flags, const mpm_particle_parameters *params, const double *F, const double *jp,
                 double hardening, double *tau, double *P, double *energy, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        M3<R> f;
        for (int col = 0; col < 3; ++col)
            for (int r = 0; r < 3; ++r) f.m[r][col] = R(F[9 * i + 3 * col + r]);
        R j = jp ? R(jp[i]) : R(1);
        M3<R> p = first_piola<R>(flags[i], params[i], j, R(hardening), f);
        M3<R> t = p * f.transpose();
        R e = 0;
        if (flags[i] & MPM_IS_FLUID) {
            e = energy_inviscid<R>(params[i].bulk_modulus, params[i].exponent, f);
        } else if (flags[i] & (MPM_FIXED_COROTATED | MPM_SNOW)) {
            R hfac = (flags[i] & MPM_SNOW) ? std::exp(R(hardening) * (R(1) - j)) : R(1);
            kirchhoff_fixed_corotated<R>(R(params[i].mu) * hfac, R(params[i].lambda) * hfac, f, e);
        } else if (!energy_neo_hookean<R>(params[i].mu, params[i].lambda, f, e)) {
            e = std::numeric_limits<R>::quiet_NaN();
        }
        for (int col = 0; col < 3; ++col)
            for (int r = 0; r < 3; ++r) {
                if (tau) tau[9 * i + 3 * col + r] = (double)t.m[r][col];
                if (P) P[9 * i + 3 * col + r] = (double)p.m[r][col];
            }
        if (energy) energy[i] = (double)e;
    }
}

template <class R>
void svd_impl(const double *F, double *U, double *s, double *V, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        M3<R> f, u, v;
        V3<R> sv;
        for (int col = 0; col < 3; ++col)
            for (int r = 0; r < 3; ++r) f.m[r][col] = R(F[9 * i + 3 * col + r]);
        svd3(f, u, sv, v);
        for (int col = 0; col < 3; ++col)
            for (int r = 0; r < 3; ++r) {
                U[9 * i + 3 * col + r] = (double)u.m[r][col];
                V[9 * i + 3 * col + r] = (double)v.m[r][col];
            }
        for (int a = 0; a < 3; ++a) s[3 * i + a] = (double)sv[a];
    }
}

template <class R> M3<R> mat_in(const double *a) {
    M3<R> f;
    for (int col = 0; col < 3; ++col)
        for (int r = 0; r < 3; ++r) f.m[r][col] = R(a[3 * col + r]);
    return f;
}
template <class R> void mat_out(const M3<R> &f, double *a) {
    for (int col = 0; col < 3; ++col)
        for (int r = 0; r < 3; ++r) a[3 * col + r] = (double)f.m[r][col];
}

template <class R>
void plasticity_impl(int kind, const double *F, double mu, double lambda, double alpha, double *out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        const M3<R> f = mat_in<R>(F + 9 * i);
        mat_out<R>(kind == 0 ? sand_project<R>(f, R(mu), R(lambda), R(alpha)) : fluid_project<R>(f), out + 9 * i);
    }
}

}  // namespace

extern "C" {

// The advance's plasticity projection of n deformation gradients (column-major):
// kind 0 = sand (Drucker-Prager with mu, lambda, alpha), 1 = fluid.
void oracle_plasticity(int precision, int kind, const double *F, double mu, double lambda, double alpha, double *out,
                       uint64_t n) {
    if (precision == 8) plasticity_impl<double>(kind, F, mu, lambda, alpha, out, n);
    else plasticity_impl<float>(kind, F, mu, lambda, alpha, out, n);
}

// cauchy_stress_general_viscosity of n velocity gradients (column-major), FP64.
void oracle_cauchy_viscosity(double dynamic, double bulk, const double *C, double *out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) mat_out<double>(cauchy_viscosity<double>(dynamic, bulk, mat_in<double>(C + 9 * i)), out + 9 * i);
}

// distance_to_triangle of n points p to triangles (a, b, c) with unit normals n, FP64.
void oracle_distance_to_triangle(const double *p, const double *tri, double *out, uint64_t count) {
    for (uint64_t i = 0; i < count; ++i) {
        const double *t = tri + 12 * i;
        out[i] = distance_to_triangle<double>(V3<double>(p[3 * i], p[3 * i + 1], p[3 * i + 2]), V3<double>(t[0], t[1], t[2]),
                                              V3<double>(t[3], t[4], t[5]), V3<double>(t[6], t[7], t[8]),
                                              V3<double>(t[9], t[10], t[11]));
    }
}

// collider_bits::compatible.
int oracle_collider_compatible(uint32_t a, uint32_t b) { return collider_compatible(a, b) ? 1 : 0; }

// The sphere collider's slip on n node velocities v at node positions xi, FP64.
void oracle_sphere_slip(const double *xi, const double *center, double radius, const double *v, double *out,
                        uint64_t n) {
    const V3<double> c(center[0], center[1], center[2]);
    for (uint64_t i = 0; i < n; ++i) {
        const V3<double> r = sphere_slip<double>(V3<double>(xi[3 * i], xi[3 * i + 1], xi[3 * i + 2]), c, radius,
                                                 V3<double>(v[3 * i], v[3 * i + 1], v[3 * i + 2]));
        for (int a = 0; a < 3; ++a) out[3 * i + a] = r[a];
    }
}

void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_max_threads(void) { return omp_get_max_threads(); }

void *oracle_create(int precision, const mpm_consts *consts, const mpm_particles_view *particles) {
    if (!consts || !particles) return nullptr;
    Handle *h = new (std::nothrow) Handle();
    if (!h) return nullptr;
    h->precision = precision;
    if (precision == 8) {
        h->d.reset(new State<double>());
        load(*h->d, *consts, *particles);
    } else {
        h->f.reset(new State<float>());
        load(*h->f, *consts, *particles);
    }
    return h;
}

void oracle_destroy(void *handle) { delete static_cast<Handle *>(handle); }

// The collider mesh (mpm_set_collider_mesh); NULL removes it.
int oracle_set_collider_mesh(void *handle, const mpm_collider_mesh *mesh) {
    Handle *h = static_cast<Handle *>(handle);
    MeshTopology m;
    if (mesh && mesh->num_triangles > 0) {
        for (uint32_t t = 0; t < mesh->num_triangles; ++t)
            if (mesh->triangle_colliders[t] >= 16) return MPM_ERR_INVALID_ARGUMENT;
        m = make_topology(mesh->num_vertices, mesh->triangles, mesh->triangle_colliders, mesh->num_triangles);
    }
    if (h->precision == 8) h->d->mesh = m;
    else h->f->mesh = m;
    return MPM_OK;
}

// Adaptive time steps with the given max_time_step (0: the fixed step).
void oracle_set_adaptive(void *handle, int adaptive, float max_time_step) {
    Handle *h = static_cast<Handle *>(handle);
    if (h->precision == 8) {
        h->d->adaptive = adaptive != 0;
        h->d->ats.max_time_step = max_time_step;
    } else {
        h->f->adaptive = adaptive != 0;
        h->f->ats.max_time_step = max_time_step;
    }
}

double oracle_time(void *handle) {
    Handle *h = static_cast<Handle *>(handle);
    return h->precision == 8 ? h->d->time : h->f->time;
}

int oracle_step_to(void *handle, double target_time, const mpm_frame_input *input, const volatile int32_t *cancel,
                   mpm_status *status) {
    Handle *h = static_cast<Handle *>(handle);
    uint64_t substeps = 0;
    int r;
    try {
        if (h->precision == 8) {
            copy_input(h->d->input, input, h->d->positions.size(), h->d->mesh);
            r = h->d->step_to(target_time, cancel, substeps);
        } else {
            copy_input(h->f->input, input, h->f->positions.size(), h->f->mesh);
            r = h->f->step_to(target_time, cancel, substeps);
        }
    } catch (...) {
        r = MPM_ERR_INVALID_ARGUMENT;
    }
    if (status) {
        status->result = r;
        status->substeps = substeps;
        status->time = oracle_time(handle);
        status->device_status = r == MPM_ERR_ENERGY ? MPM_STATUS_ENERGY_ERROR : 0;
        status->active_blocks = 0;
        status->active_nodes = 0;
        status->svd_fp64_fallbacks = 0;
    }
    return r;
}

// Runs InterpolateInput .. ScatterMomentum of the next substep (the P2G half).
int oracle_run_to_scatter(void *handle, const mpm_frame_input *input) {
    Handle *h = static_cast<Handle *>(handle);
    try {
        if (h->precision == 8) {
            copy_input(h->d->input, input, h->d->positions.size(), h->d->mesh);
            return h->d->run_to_scatter();
        }
        copy_input(h->f->input, input, h->f->positions.size(), h->f->mesh);
        return h->f->run_to_scatter();
    } catch (...) {
        return MPM_ERR_INVALID_ARGUMENT;
    }
}

// Completes a substep started by oracle_run_to_scatter.
int oracle_run_from_meld(void *handle) {
    Handle *h = static_cast<Handle *>(handle);
    try {
        return h->precision == 8 ? h->d->run_from_meld() : h->f->run_from_meld();
    } catch (...) {
        return MPM_ERR_INVALID_ARGUMENT;
    }
}

int oracle_read_particles(void *handle, mpm_particles_view *out) {
    Handle *h = static_cast<Handle *>(handle);
    if (h->precision == 8) store(*h->d, *out);
    else store(*h->f, *out);
    return MPM_OK;
}

int oracle_read_particles_f64(void *handle, double *x, double *v, double *F, double *C) {
    Handle *h = static_cast<Handle *>(handle);
    if (h->precision == 8) store_f64(*h->d, x, v, F, C);
    else store_f64(*h->f, x, v, F, C);
    return MPM_OK;
}

// Grid nodes in node-id order: after ScatterMomentum the velocities hold
// momentum, after MeldGrid they hold velocity.  Returns the node count.
int64_t oracle_read_grid(void *handle, int32_t *ids, float *masses, float *vel, int64_t capacity) {
    Handle *h = static_cast<Handle *>(handle);
    return h->precision == 8 ? read_grid(*h->d, ids, masses, vel, capacity) : read_grid(*h->f, ids, masses, vel, capacity);
}

int oracle_read_grid_f64(void *handle, double *masses, double *vel) {
    Handle *h = static_cast<Handle *>(handle);
    if (h->precision == 8) read_grid_f64(*h->d, masses, vel);
    else read_grid_f64(*h->f, masses, vel);
    return MPM_OK;
}

void oracle_stress(int precision, const uint32_t *flags, const mpm_particle_parameters *params, const double *F,
                   const double *jp, double hardening, double *tau, double *P, double *energy, uint64_t n) {
    if (precision == 8) stress_impl<double>(flags, params, F, jp, hardening, tau, P, energy, n);
    else stress_impl<float>(flags, params, F, jp, hardening, tau, P, energy, n);
}

void oracle_svd(int precision, const double *F, double *U, double *s, double *V, uint64_t n) {
    if (precision == 8) svd_impl<double>(F, U, s, V, n);
    else svd_impl<float>(F, U, s, V, n);
}

double oracle_kernel_quadratic(double x) { return kernel_quadratic<double>(x); }

double oracle_lame_mu(double E, double nu) { return lame_mu<double>(E, nu); }
double oracle_lame_lambda(double E, double nu) { return lame_lambda<double>(E, nu); }

}  // extern "C"
