// The C ABI (include/rahman_hip.h): status codes, a thread-local last error, and thin wrappers
// over rh::Engine<3> / rh::Engine<2> and the gpu-prims launchers.
#include <exception>
#include <new>
#include <string>

#include <cstring>
#include "batch.h"
#include "engine.h"

namespace rh {
void kick_convert(const float* x, int32_t* out, uint64_t n, hipStream_t s);
void stochastic_draws(uint64_t seed, double t, double mass, double dt, uint32_t id0, uint64_t count, uint32_t clock, int32_t* out, hipStream_t s);
void check_division(float r2_lo, float r2_hi, float sigma2, float six_eps4, uint64_t* checked, uint64_t* mismatches, float* first_bad, hipStream_t s);
}

struct rh_engine3 {
    rh::Engine<3>* e;
};
struct rh_engine2 {
    rh::Engine<2>* e;
};
struct rh_batch {
    rh::Batch* b;
};

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
    try {
        f();
        return RH_OK;
    } catch (const rh::Error& err) {
        g_last_error = err.what();
        std::string m = err.what();
        if (m.find("HIP error") != std::string::npos) return RH_ERR_HIP;
        if (m.find("allocation") != std::string::npos) return RH_ERR_NOMEM;
        if (m.find("not supported") != std::string::npos) return RH_ERR_UNSUPPORTED;
        return RH_ERR_INVALID;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return RH_ERR_NOMEM;
    } catch (const std::exception& err) {
        g_last_error = err.what();
        return RH_ERR_INVALID;
    } catch (...) {
        g_last_error = "unknown error";
        return RH_ERR_INVALID;
    }
}

rh::SystemDesc desc3(const rh_system3* s) {
    RH_REQUIRE(s != nullptr, "null system");
    rh::SystemDesc d{};
    d.dim = 3;
    for (int a = 0; a < 3; a++) {
        RH_REQUIRE(s->box_len[a] > 0, "box lengths must be positive");
        d.box_len[a] = s->box_len[a];
        d.walled[a] = s->walled[a];
        for (int side = 0; side < 2; side++) {
            RH_REQUIRE(s->wall_kind[a][side] <= RH_WALL_THERMAL, "unknown wall kind");
            d.wall_kind[a][side] = s->wall_kind[a][side];
            d.wall_t[a][side] = s->wall_t[a][side];
            if (s->walled[a] && s->wall_kind[a][side] == RH_WALL_THERMAL)
                RH_REQUIRE(s->wall_t[a][side] >= 0.0, "heat-bath wall temperature must be >= 0");
        }
        d.gravity[a] = s->gravity[a];
    }
    d.seed = s->seed;
    d.has_bath = s->has_bath ? 1u : 0u;
    if (d.has_bath) {
        RH_REQUIRE(s->bath_t >= 0.0 && s->bath_rate >= 0.0, "bath temperature and rate must be >= 0");
        d.bath_species = s->bath_species;
        d.bath_t = s->bath_t;
        d.bath_rate = s->bath_rate;
    }
    d.dt = s->dt;
    d.nspecies = s->nspecies;
    d.pinned = s->pinned;
    for (int k = 0; k < 4; k++) d.masses[k] = k < (int)s->nspecies ? s->masses[k] : 1.0;
    for (int k = 0; k < 16; k++) d.pairs[k] = s->pairs[k];
    return d;
}

// md_core::System with its boundaries, checked as System's builders check them
// (system.rs:204-330).
rh::SystemDesc desc2(const rh_system2* s) {
    RH_REQUIRE(s != nullptr, "null system");
    rh::SystemDesc d{};
    d.dim = 2;
    for (int a = 0; a < 2; a++) {
        RH_REQUIRE(s->box_len[a] > 0, "box lengths must be positive");
        d.box_len[a] = s->box_len[a];
        d.walled[a] = s->walled[a] ? 1u : 0u;
        for (int side = 0; side < 2; side++) {
            const uint32_t kind = s->walled[a] ? s->wall_kind[a][side] : (uint32_t)RH_WALL_HARD;
            RH_REQUIRE(kind <= RH_WALL_THERMAL, "unknown wall kind");
            d.wall_kind[a][side] = kind;
            d.wall_t[a][side] = s->wall_t[a][side];
            if (kind == RH_WALL_THERMAL) RH_REQUIRE(s->wall_t[a][side] >= 0.0, "heat-bath wall temperature must be >= 0");
        }
        d.gravity[a] = s->gravity[a];
        if (s->compartments[a] != 0) {
            const uint32_t shift = s->compartments[a];
            RH_REQUIRE(s->walled[a], "compartments need walls on their axis");
            RH_REQUIRE(shift >= 32 && shift < 52, "compartment side must be 2^k sigma, 0 <= k < 20");
            RH_REQUIRE(s->box_len[a] % (int64_t(1) << shift) == 0, "box must be a whole number of compartments");
            RH_REQUIRE((s->box_len[a] >> shift) <= RH_MAX_COMPARTMENTS, "at most 256 compartments along an axis");
            d.compartments[a] = shift;
        }
    }
    d.box_len[2] = 1;
    d.dt = s->dt;
    d.nspecies = s->nspecies;
    d.pinned = s->pinned;
    d.seed = s->seed;
    for (int k = 0; k < 4; k++) d.masses[k] = k < (int)s->nspecies ? s->masses[k] : 1.0;
    for (int k = 0; k < 16; k++) d.pairs[k] = s->pairs[k];
    d.diathermal = s->diathermal ? 1u : 0u;
    if (s->has_bath) {
        RH_REQUIRE(s->bath_t >= 0.0 && s->bath_rate >= 0.0, "bath temperature and rate must be >= 0");
        RH_REQUIRE(s->bath_axis < 2, "thermostat region axis must be 0 or 1");
        d.has_bath = 1u;
        d.bath_species = s->bath_species;
        d.bath_t = s->bath_t;
        d.bath_rate = s->bath_rate;
        d.bath_axis = s->bath_axis;
        d.bath_drift = s->bath_drift;
        d.bath_region[0] = s->bath_region[0];
        d.bath_region[1] = s->bath_region[1];
    }
    if (s->has_membrane) {
        RH_REQUIRE(s->membrane_axis < 2 && s->walled[s->membrane_axis], "a membrane needs walls on its axis");
        RH_REQUIRE(s->pore_shift == 0 || (s->pore_shift >= 1 && s->pore_shift < 62), "pore period out of range");
        d.has_membrane = 1u;
        d.membrane_axis = s->membrane_axis;
        d.membrane_blocks = s->membrane_blocks;
        d.pore_shift = s->pore_shift;
        d.membrane_position = s->membrane_position;
        d.pore_width = s->pore_width;
    }
    if (s->has_piston) {
        RH_REQUIRE(s->walled[1], "a piston needs walls on y");
        RH_REQUIRE(s->piston_start <= s->piston_max && s->piston_min < s->piston_start, "piston needs min < start <= max");
        if (s->piston_free) RH_REQUIRE(s->piston_mass > 0.0, "a free piston needs a positive mass");
        d.has_piston = 1u;
        d.piston_free = s->piston_free ? 1u : 0u;
        d.piston_speed = s->piston_speed;
        d.piston_mass = s->piston_mass;
        d.piston_force = s->piston_force;
        d.piston_start = s->piston_start;
        d.piston_min = s->piston_min;
        d.piston_max = s->piston_max;
    }
    RH_REQUIRE(s->nbond_types <= 4, "at most 4 bond types");
    d.nbond_types = s->nbond_types;
    for (uint32_t t = 0; t < 4; t++) {
        d.bond_fene[t] = t < s->nbond_types && s->bond_fene[t] ? 1u : 0u;
        d.bond_k[t] = t < s->nbond_types ? s->bond_k[t] : 0.0;
        d.bond_r0[t] = t < s->nbond_types ? s->bond_r0[t] : 0.0;
    }
    if (s->nbonds) {
        RH_REQUIRE(s->bonds != nullptr, "null bond list");
        RH_REQUIRE(s->nbonds < (uint64_t(1) << 30), "too many bonds");
        d.bonds.assign(s->bonds, s->bonds + 3 * s->nbonds);
        for (uint64_t b = 0; b < s->nbonds; b++) RH_REQUIRE(d.bonds[3 * b + 2] < s->nbond_types, "bond type out of range");
    }
    return d;
}

rh_mode default_mode3(const rh_system3* s, uint64_t n) {
    rh_mode m;
    rh_verlet_mode3(s, n, 0.4, &m);
    return m;
}

rh_mode default_mode2(const rh_system2* s, uint64_t n) {
    rh_mode m;
    rh_verlet_mode2(s, n, 0.4, &m);
    return m;
}

template <class H>
auto* eng(H* h) {
    RH_REQUIRE(h != nullptr && h->e != nullptr, "null engine");
    return h->e;
}

rh::Batch* eng(rh_batch* h) {
    RH_REQUIRE(h != nullptr && h->b != nullptr, "null batch");
    return h->b;
}

}  // namespace

extern "C" {

const char* rh_last_error(void) { return g_last_error.c_str(); }
const char* rh_version(void) { return "rahman-hip 0.1 (gfx950)"; }

int rh_device_count(int* count) {
    return guarded([&] { RH_HIP(hipGetDeviceCount(count)); });
}

int rh_engine3_create(const rh_system3* sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, const rh_mode* mode, rh_engine3** out) {
    return guarded([&] {
        RH_REQUIRE(out && pos && vel && species, "null argument");
        *out = nullptr;
        rh_mode m = mode ? *mode : default_mode3(sys, n);
        auto* h = new rh_engine3{nullptr};
        try {
            h->e = new rh::Engine<3>(desc3(sys), n, pos, vel, species, m);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int rh_engine3_bytes_needed(const rh_system3* sys, uint64_t n, const rh_mode* mode, uint64_t* bytes) {
    return guarded([&] {
        RH_REQUIRE(bytes != nullptr, "null argument");
        rh_mode m = mode ? *mode : default_mode3(sys, n);
        *bytes = rh::Engine<3>(desc3(sys), n, m).planned_bytes();
    });
}
int rh_engine3_device_bytes(const rh_engine3* e, uint64_t* bytes) {
    return guarded([&] {
        RH_REQUIRE(e && bytes, "null argument");
        *bytes = e->e->alloc_bytes();
    });
}

int rh_engine3_destroy(rh_engine3* e) {
    return guarded([&] {
        if (!e) return;
        delete e->e;
        delete e;
    });
}

int rh_engine3_n(const rh_engine3* e, uint64_t* n) {
    return guarded([&] { *n = eng(e)->n(); });
}
int rh_engine3_steps(const rh_engine3* e, int64_t* steps) {
    return guarded([&] { *steps = eng(e)->steps(); });
}
int rh_engine3_step(rh_engine3* e, uint64_t steps) {
    return guarded([&] { eng(e)->step(steps); });
}
int rh_engine3_reverse(rh_engine3* e) {
    return guarded([&] { eng(e)->reverse(); });
}
int rh_engine3_update_system(rh_engine3* e, const rh_system3* sys) {
    return guarded([&] { eng(e)->update_system(desc3(sys)); });
}
int rh_engine3_kick_region(rh_engine3* e, const double lo[3], const double hi[3], double scale, const double add[3]) {
    return guarded([&] { eng(e)->kick_region(lo, hi, scale, add); });
}
int rh_engine3_set_species(rh_engine3* e, const uint32_t* ids, uint64_t count, uint32_t species) {
    return guarded([&] { eng(e)->set_species(ids, count, species); });
}
int rh_engine3_set_species_velocity(rh_engine3* e, uint32_t species, const double v[3]) {
    return guarded([&] { eng(e)->set_species_velocity(species, v); });
}
int rh_engine3_upload(rh_engine3* e, const int64_t* pos, const int32_t* vel, const uint32_t* species) {
    return guarded([&] { eng(e)->upload(pos, vel, species); });
}
int rh_engine3_read_state(rh_engine3* e, int64_t* pos, int32_t* vel, uint32_t* species) {
    return guarded([&] { eng(e)->read_state(pos, vel, species); });
}
int rh_engine3_kicks(rh_engine3* e, int32_t* kicks, double* potential) {
    return guarded([&] { eng(e)->kicks(kicks, potential); });
}
int rh_engine3_read_kinetic_total(rh_engine3* e, double* ke) {
    return guarded([&] { *ke = eng(e)->kinetic_total(); });
}
int rh_engine3_measure_kinetic(rh_engine3* e) {
    return guarded([&] { eng(e)->measure_kinetic(); });
}
int rh_engine3_read_kinetic(rh_engine3* e, float* ke, uint32_t* species) {
    return guarded([&] { eng(e)->read_kinetic(ke, species); });
}
int rh_engine3_read_potential(rh_engine3* e, double* pe) {
    return guarded([&] { *pe = eng(e)->potential(); });
}
int rh_engine3_read_ledger(rh_engine3* e, rh_ledger3* out) {
    return guarded([&] {
        RH_REQUIRE(out != nullptr, "null ledger");
        eng(e)->read_ledger(out);
    });
}
int rh_engine3_list_stats(rh_engine3* e, rh_list_stats* out) {
    return guarded([&] { *out = eng(e)->list_stats(); });
}
int rh_engine3_set_sort_interval(rh_engine3* e, uint32_t interval) {
    return guarded([&] { eng(e)->set_sort_interval(interval); });
}
int rh_engine3_tune(rh_engine3* e, int variant, uint32_t block) {
    return guarded([&] { eng(e)->tune(variant, block); });
}
int rh_engine3_stream(rh_engine3* e, void** stream) {
    return guarded([&] { *stream = (void*)eng(e)->stream(); });
}
int rh_engine3_set_timing(rh_engine3* e, int on) {
    return guarded([&] { eng(e)->set_timing(on != 0); });
}
int rh_engine3_kernel_ms(rh_engine3* e, double* force_ms, uint64_t* force_launches, double* rebuild_ms, uint64_t* rebuilds) {
    return guarded([&] { eng(e)->kernel_ms(force_ms, force_launches, rebuild_ms, rebuilds); });
}

int rh_engine2_create(const rh_system2* sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, const rh_mode* mode, rh_engine2** out) {
    return guarded([&] {
        RH_REQUIRE(out && pos && vel && species, "null argument");
        *out = nullptr;
        rh_mode m = mode ? *mode : default_mode2(sys, n);
        auto* h = new rh_engine2{nullptr};
        try {
            h->e = new rh::Engine<2>(desc2(sys), n, pos, vel, species, m);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}
int rh_engine2_bytes_needed(const rh_system2* sys, uint64_t n, const rh_mode* mode, uint64_t* bytes) {
    return guarded([&] {
        RH_REQUIRE(bytes != nullptr, "null argument");
        rh_mode m = mode ? *mode : default_mode2(sys, n);
        *bytes = rh::Engine<2>(desc2(sys), n, m).planned_bytes();
    });
}
int rh_engine2_device_bytes(const rh_engine2* e, uint64_t* bytes) {
    return guarded([&] {
        RH_REQUIRE(e && bytes, "null argument");
        *bytes = e->e->alloc_bytes();
    });
}
int rh_engine2_destroy(rh_engine2* e) {
    return guarded([&] {
        if (!e) return;
        delete e->e;
        delete e;
    });
}
int rh_engine2_step(rh_engine2* e, uint64_t steps) {
    return guarded([&] { eng(e)->step(steps); });
}
int rh_engine2_reverse(rh_engine2* e) {
    return guarded([&] { eng(e)->reverse(); });
}
int rh_engine2_upload(rh_engine2* e, const int64_t* pos, const int32_t* vel, const uint32_t* species) {
    return guarded([&] { eng(e)->upload(pos, vel, species); });
}
int rh_engine2_read_state(rh_engine2* e, int64_t* pos, int32_t* vel, uint32_t* species) {
    return guarded([&] { eng(e)->read_state(pos, vel, species); });
}
int rh_engine2_kicks(rh_engine2* e, int32_t* kicks, double* potential) {
    return guarded([&] { eng(e)->kicks(kicks, potential); });
}
int rh_engine2_read_kinetic_total(rh_engine2* e, double* ke) {
    return guarded([&] { *ke = eng(e)->kinetic_total(); });
}
int rh_engine2_compute_forces(rh_engine2* e) {
    return guarded([&] { eng(e)->compute_forces(); });
}
int rh_engine2_list_stats(rh_engine2* e, rh_list_stats* out) {
    return guarded([&] { *out = eng(e)->list_stats(); });
}
int rh_engine2_set_arith(rh_engine2* e, int arith) {
    return guarded([&] { eng(e)->set_arith(arith); });
}
int rh_engine2_tune(rh_engine2* e, int variant, uint32_t block) {
    return guarded([&] { eng(e)->tune(variant, block); });
}
int rh_engine2_stream(rh_engine2* e, void** stream) {
    return guarded([&] { *stream = (void*)eng(e)->stream(); });
}
int rh_engine2_n(const rh_engine2* e, uint64_t* n) {
    return guarded([&] { *n = eng(e)->n(); });
}
int rh_engine2_steps(const rh_engine2* e, int64_t* steps) {
    return guarded([&] { *steps = eng(e)->steps(); });
}
int rh_engine2_update_system(rh_engine2* e, const rh_system2* sys) {
    return guarded([&] { eng(e)->update_system(desc2(sys)); });
}
int rh_engine2_set_sort_interval(rh_engine2* e, uint32_t interval) {
    return guarded([&] { eng(e)->set_sort_interval(interval); });
}
int rh_engine2_sort_now(rh_engine2* e) {
    return guarded([&] { eng(e)->sort_now(); });
}
int rh_engine2_read_potential(rh_engine2* e, double* pe) {
    return guarded([&] { *pe = eng(e)->potential(); });
}
int rh_engine2_read_positions(rh_engine2* e, int64_t* pos) {
    return guarded([&] {
        RH_REQUIRE(pos != nullptr, "null argument");
        eng(e)->read_positions(pos);
    });
}
int rh_engine2_kick_region(rh_engine2* e, const double lo[2], const double hi[2], double scale, const double add[2]) {
    return guarded([&] {
        RH_REQUIRE(lo && hi && add, "null argument");
        const double l3[3] = {lo[0], lo[1], 0.0}, h3[3] = {hi[0], hi[1], 0.0}, a3[3] = {add[0], add[1], 0.0};
        eng(e)->kick_region(l3, h3, scale, a3);
    });
}
int rh_engine2_set_species(rh_engine2* e, const uint32_t* ids, uint64_t count, uint32_t species) {
    return guarded([&] { eng(e)->set_species(ids, count, species); });
}
int rh_engine2_set_species_velocity(rh_engine2* e, uint32_t species, const double v[2]) {
    return guarded([&] {
        RH_REQUIRE(v != nullptr, "null argument");
        const double v3[3] = {v[0], v[1], 0.0};
        eng(e)->set_species_velocity(species, v3);
    });
}
int rh_engine2_read_ledger(rh_engine2* e, rh_ledger2* out) {
    return guarded([&] {
        RH_REQUIRE(out != nullptr, "null ledger");
        rh_ledger3 l;
        eng(e)->read_ledger(&l);
        static_assert(sizeof(rh_ledger2) == sizeof(rh_ledger3), "ledger layouts");
        std::memcpy(out, &l, sizeof l);
    });
}
int rh_engine2_read_compartment_ledger(rh_engine2* e, uint32_t per_wall, double* energy, double* impulse, uint64_t* hits) {
    return guarded([&] { eng(e)->read_compartment_ledger(per_wall, energy, impulse, hits); });
}
int rh_engine2_read_piston(rh_engine2* e, double out[4]) {
    return guarded([&] {
        RH_REQUIRE(out != nullptr, "null argument");
        eng(e)->read_piston(out);
    });
}
int rh_engine2_set_piston_speed(rh_engine2* e, double u) {
    return guarded([&] { eng(e)->set_piston_speed(u); });
}
int rh_engine2_flip_bit(rh_engine2* e, uint32_t particle, uint32_t axis, uint32_t bit) {
    return guarded([&] { eng(e)->flip_bit(particle, axis, bit); });
}
int rh_engine2_flip_species_bit(rh_engine2* e, uint32_t species, uint32_t axis, uint32_t bit) {
    return guarded([&] { eng(e)->flip_species_bit(species, axis, bit); });
}
int rh_engine2_add_bond(rh_engine2* e, uint32_t a, uint32_t b, uint32_t type, int* changed) {
    return guarded([&] {
        const bool c = eng(e)->add_bond(a, b, type);
        if (changed) *changed = c ? 1 : 0;
    });
}
int rh_engine2_remove_bond(rh_engine2* e, uint32_t a, uint32_t b, int* changed) {
    return guarded([&] {
        const bool c = eng(e)->remove_bond(a, b);
        if (changed) *changed = c ? 1 : 0;
    });
}

// ---- batched engine ----

int rh_batch_create(uint64_t budget_bytes, uint32_t max_systems, rh_batch** out) {
    return guarded([&] {
        RH_REQUIRE(out, "null argument");
        *out = nullptr;
        auto* h = new rh_batch{nullptr};
        try {
            h->b = new rh::Batch(budget_bytes, max_systems);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}
int rh_batch_destroy(rh_batch* b) {
    return guarded([&] {
        if (!b) return;
        delete b->b;
        delete b;
    });
}
int rh_batch_add3(rh_batch* b, const rh_system3* sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, double skin, uint64_t steps, int32_t* id) {
    return guarded([&] {
        RH_REQUIRE(pos && vel && species && id, "null argument");
        *id = eng(b)->add(desc3(sys), n, pos, vel, species, skin, steps);
    });
}
int rh_batch_add2(rh_batch* b, const rh_system2* sys, uint64_t n, const int64_t* pos, const int32_t* vel, const uint32_t* species, double skin, uint64_t steps, int32_t* id) {
    return guarded([&] {
        RH_REQUIRE(pos && vel && species && id, "null argument");
        *id = eng(b)->add(desc2(sys), n, pos, vel, species, skin, steps);
    });
}
int rh_batch_remove(rh_batch* b, int32_t id) {
    return guarded([&] { eng(b)->remove(id); });
}
int rh_batch_set_steps(rh_batch* b, int32_t id, uint64_t steps) {
    return guarded([&] { eng(b)->set_steps(id, steps); });
}
int rh_batch_run(rh_batch* b, uint64_t iterations) {
    return guarded([&] { eng(b)->run(iterations); });
}
int rh_batch_sync(rh_batch* b) {
    return guarded([&] { eng(b)->sync(); });
}
int rh_batch_poll(rh_batch* b, uint64_t* remaining, uint32_t count) {
    return guarded([&] { eng(b)->poll(remaining, count); });
}
int rh_batch_read_state(rh_batch* b, int32_t id, int64_t* pos, int32_t* vel, uint32_t* species) {
    return guarded([&] { eng(b)->read_state(id, pos, vel, species); });
}
int rh_batch_list_stats(rh_batch* b, int32_t id, rh_list_stats* out) {
    return guarded([&] { *out = eng(b)->stats(id); });
}
int rh_batch_read_ledger(rh_batch* b, int32_t id, rh_ledger3* out) {
    return guarded([&] {
        RH_REQUIRE(out, "null output");
        std::memset(out, 0, sizeof *out);
        eng(b)->read_ledger(id, out->energy, out->impulse, out->hits);
    });
}
int rh_batch_set_graph_chunk(rh_batch* b, uint32_t chunk) {
    return guarded([&] { eng(b)->set_graph_chunk(chunk); });
}
int rh_batch_set_tuning(rh_batch* b, uint32_t dim, uint32_t size_class, uint32_t variant, double cell_factor, uint32_t block) {
    return guarded([&] { eng(b)->set_tuning((int)dim, size_class, variant, cell_factor, block); });
}
int rh_batch_stream(rh_batch* b, void** stream) {
    return guarded([&] { *stream = (void*)eng(b)->stream(); });
}
int rh_batch_capacity(rh_batch* b, uint64_t* slots, uint64_t* list_entries, uint64_t* cells, uint32_t* active) {
    return guarded([&] {
        eng(b)->budget(slots, list_entries, cells);
        if (active) *active = eng(b)->active();
    });
}
int rh_batch_bytes_needed3(rh_batch* b, const rh_system3* sys, uint64_t n, double skin, uint64_t* bytes) {
    return guarded([&] { *bytes = eng(b)->bytes_needed(3, n, desc3(sys), skin); });
}
int rh_batch_bytes_needed2(rh_batch* b, const rh_system2* sys, uint64_t n, double skin, uint64_t* bytes) {
    return guarded([&] { *bytes = eng(b)->bytes_needed(2, n, desc2(sys), skin); });
}

// ---- gpu-prims ----

int rh_prims_exclusive_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream) {
    return guarded([&] {
        hipStream_t s = (hipStream_t)stream;
        void* temp = nullptr;
        RH_HIP(hipMalloc(&temp, rh::scan_temp_bytes(n)));
        try {
            rh::exclusive_scan_u32(d_in, d_out, n, temp, nullptr, s);
            RH_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipFree(temp);
            throw;
        }
        RH_HIP(hipFree(temp));
    });
}

int rh_prims_radix_sort_u32(uint32_t* d_keys, uint32_t* d_vals, uint64_t n, uint32_t key_bits, void* stream) {
    return guarded([&] {
        RH_REQUIRE(key_bits >= 1 && key_bits <= 32, "key_bits must be 1..32");
        if (n == 0) return;
        hipStream_t s = (hipStream_t)stream;
        rh::RadixSort rs;
        rs.reserve(n);
        RH_HIP(hipMemcpyAsync(rs.keys[0], d_keys, n * 4, hipMemcpyDeviceToDevice, s));
        RH_HIP(hipMemcpyAsync(rs.vals[0], d_vals, n * 4, hipMemcpyDeviceToDevice, s));
        int r = rs.sort(n, key_bits, nullptr, s);
        RH_HIP(hipMemcpyAsync(d_keys, rs.keys[r], n * 4, hipMemcpyDeviceToDevice, s));
        RH_HIP(hipMemcpyAsync(d_vals, rs.vals[r], n * 4, hipMemcpyDeviceToDevice, s));
        RH_HIP(hipStreamSynchronize(s));
    });
}

int rh_prims_radix_sort_u32_noalloc(uint32_t* d_keys, uint32_t* d_vals, uint64_t n, uint32_t key_bits, void* stream) {
    return guarded([&] { rh::radix_sort_u32_fixed(d_keys, d_vals, n, key_bits, (hipStream_t)stream); });
}

int rh_prims_reduce_sum_f64(const double* d_in, double* d_out, uint64_t n, void* stream) {
    return guarded([&] {
        hipStream_t s = (hipStream_t)stream;
        void* temp = nullptr;
        RH_HIP(hipMalloc(&temp, rh::reduce_temp_bytes(n)));
        try {
            rh::reduce_sum_f64(d_in, d_out, n, temp, s);
            RH_HIP(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipFree(temp);
            throw;
        }
        RH_HIP(hipFree(temp));
    });
}

int rh_prims_threefry(uint32_t* d_out, uint64_t n, uint32_t ctr0, uint32_t ctr1, uint32_t key0, uint32_t key1, void* stream) {
    return guarded([&] {
        rh::threefry_fill(d_out, n, ctr0, ctr1, key0, key1, (hipStream_t)stream);
        RH_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

int rh_prims_atomic_add_i64(const int64_t* d_vals, uint64_t n, int64_t* d_acc, void* stream) {
    return guarded([&] {
        rh::atomic_add_i64(d_vals, n, d_acc, (hipStream_t)stream);
        RH_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

int rh_pair_division_floor(double eps, double sigma, double rc, float* r2_fast) {
    return guarded([&] {
        float rc2 = (float)(rc * rc), sigma2 = (float)(sigma * sigma), six_eps4 = 6.0f * (float)(4.0 * eps);
        *r2_fast = rh::fast_division_floor(rc2, sigma2, six_eps4);
    });
}

int rh_prims_check_pair_division(double eps, double sigma, double rc, uint64_t* checked, uint64_t* mismatches, float* first_bad, void* stream) {
    return guarded([&] {
        float rc2 = (float)(rc * rc), sigma2 = (float)(sigma * sigma), six_eps4 = 6.0f * (float)(4.0 * eps);
        float lo = rh::fast_division_floor(rc2, sigma2, six_eps4);
        RH_REQUIRE(lo < rc2, "this pair never takes the fast division");
        rh::check_division(lo, rc2, sigma2, six_eps4, checked, mismatches, first_bad, (hipStream_t)stream);
    });
}

int rh_prims_stochastic_draws(uint64_t seed, double t, double mass, double dt, uint32_t id0, uint64_t count, uint32_t clock, int32_t* d_out, void* stream) {
    return guarded([&] {
        RH_REQUIRE(count < (uint64_t(1) << 32), "count must fit 32 bits");
        rh::stochastic_draws(seed, t, mass, dt, id0, count, clock, d_out, (hipStream_t)stream);
        RH_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

int rh_prims_kick_convert(const float* d_x, int32_t* d_out, uint64_t n, void* stream) {
    return guarded([&] {
        rh::kick_convert(d_x, d_out, n, (hipStream_t)stream);
        RH_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

}  // extern "C"
