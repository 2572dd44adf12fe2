//! md-hip: the MI355X engine behind md::GpuEngine3 and md::GpuEngine.
//!
//! The engine is hand-written HIP for gfx950 in librahman_hip.so, reached through its C ABI
//! (include/rahman_hip.h).  The types here keep the method names, arguments and meanings of
//! crates/md/src/engine3.rs:307-648 and crates/md/src/engine.rs:567-1098; md-core's System3 /
//! State3 / System / State and md's ForceMode / Arith cross the boundary as plain structs and
//! flat arrays in particle-id order.  A caller switches by replacing
//! `md::GpuEngine3::new(&prims, sys, &state)` with `md_hip::GpuEngine3::new(sys, &state)` (no
//! wgpu Prims).  Readbacks return ready futures: the C calls have finished when they return.
//! Contract violations panic with the library's message, as the reference's asserts do.

use md::engine3::Ledger3;
use md::{Arith, ForceMode};
use md_core::system::{BondType, CompartmentWalls, PistonDrive, Wall};
use md_core::three::{State3, System3};
use md_core::walls::{CompartmentLedger, Ledger};
use md_core::{State, System};
use std::ffi::CStr;
use std::future::{ready, Future};
use std::os::raw::{c_char, c_int};

// ---- the C ABI (include/rahman_hip.h) ---------------------------------------------------------

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct RhPair {
    pub eps: f64,
    pub sigma: f64,
    pub rc: f64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct RhSystem3 {
    pub box_len: [i64; 3],
    pub dt: f64,
    pub nspecies: u32,
    pub pinned: u32,
    pub masses: [f64; 4],
    pub pairs: [RhPair; 16],
    pub walled: [u32; 3],
    pub wall_kind: [[u32; 2]; 3],
    pub wall_t: [[f64; 2]; 3],
    pub gravity: [f64; 3],
    pub seed: u64,
    pub has_bath: u32,
    pub bath_species: u32,
    pub bath_t: f64,
    pub bath_rate: f64,
}

#[repr(C)]
#[derive(Clone, Copy)]
pub struct RhSystem2 {
    pub box_len: [i64; 2],
    pub dt: f64,
    pub nspecies: u32,
    pub pinned: u32,
    pub masses: [f64; 4],
    pub pairs: [RhPair; 16],
    pub walled: [u32; 2],
    pub gravity: [f64; 2],
    pub seed: u64,
    pub wall_kind: [[u32; 2]; 2],
    pub wall_t: [[f64; 2]; 2],
    pub compartments: [u32; 2],
    pub diathermal: u32,
    pub has_bath: u32,
    pub bath_species: u32,
    pub bath_axis: u32,
    pub bath_t: f64,
    pub bath_rate: f64,
    pub bath_drift: f64,
    pub bath_region: [i64; 2],
    pub has_membrane: u32,
    pub membrane_axis: u32,
    pub membrane_blocks: u32,
    pub pore_shift: u32,
    pub membrane_position: i64,
    pub pore_width: i64,
    pub has_piston: u32,
    pub piston_free: u32,
    pub piston_speed: f64,
    pub piston_mass: f64,
    pub piston_force: f64,
    pub piston_start: f64,
    pub piston_min: f64,
    pub piston_max: f64,
    pub nbond_types: u32,
    pub bond_fene: [u32; 4],
    pub bond_k: [f64; 4],
    pub bond_r0: [f64; 4],
    pub nbonds: u64,
    pub bonds: *const u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct RhMode {
    pub kind: u32,
    pub cap: u32,
    pub skin: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct RhLedger {
    pub energy: [f64; 7],
    pub impulse: [f64; 7],
    pub hits: [u64; 7],
}

#[repr(C)]
#[derive(Clone, Copy, Default)]
pub struct RhListStats {
    pub rebuilds: u32,
    pub overflowed: u32,
    pub cap: u32,
    pub max_count: u32,
    pub steps_done: u64,
    pub launches_skipped: u64,
}

#[repr(C)]
pub struct RhEngine3 {
    _opaque: [u8; 0],
}

#[repr(C)]
pub struct RhEngine2 {
    _opaque: [u8; 0],
}

#[link(name = "rahman_hip")]
unsafe extern "C" {
    fn rh_last_error() -> *const c_char;
    fn rh_verlet_mode3(sys: *const RhSystem3, n: u64, skin: f64, out: *mut RhMode) -> c_int;
    fn rh_verlet_mode2(sys: *const RhSystem2, n: u64, skin: f64, out: *mut RhMode) -> c_int;

    fn rh_engine3_create(sys: *const RhSystem3, n: u64, pos: *const i64, vel: *const i32, species: *const u32, mode: *const RhMode, out: *mut *mut RhEngine3) -> c_int;
    fn rh_engine3_destroy(e: *mut RhEngine3) -> c_int;
    fn rh_engine3_step(e: *mut RhEngine3, steps: u64) -> c_int;
    fn rh_engine3_reverse(e: *mut RhEngine3) -> c_int;
    fn rh_engine3_update_system(e: *mut RhEngine3, sys: *const RhSystem3) -> c_int;
    fn rh_engine3_kick_region(e: *mut RhEngine3, lo: *const f64, hi: *const f64, scale: f64, add: *const f64) -> c_int;
    fn rh_engine3_set_species(e: *mut RhEngine3, ids: *const u32, count: u64, species: u32) -> c_int;
    fn rh_engine3_set_species_velocity(e: *mut RhEngine3, species: u32, v: *const f64) -> c_int;
    fn rh_engine3_upload(e: *mut RhEngine3, pos: *const i64, vel: *const i32, species: *const u32) -> c_int;
    fn rh_engine3_read_state(e: *mut RhEngine3, pos: *mut i64, vel: *mut i32, species: *mut u32) -> c_int;
    fn rh_engine3_kicks(e: *mut RhEngine3, kicks: *mut i32, potential: *mut f64) -> c_int;
    fn rh_engine3_measure_kinetic(e: *mut RhEngine3) -> c_int;
    fn rh_engine3_read_kinetic(e: *mut RhEngine3, ke: *mut f32, species: *mut u32) -> c_int;
    fn rh_engine3_read_kinetic_total(e: *mut RhEngine3, ke: *mut f64) -> c_int;
    fn rh_engine3_read_potential(e: *mut RhEngine3, pe: *mut f64) -> c_int;
    fn rh_engine3_read_ledger(e: *mut RhEngine3, out: *mut RhLedger) -> c_int;
    fn rh_engine3_list_stats(e: *mut RhEngine3, out: *mut RhListStats) -> c_int;
    fn rh_engine3_set_sort_interval(e: *mut RhEngine3, interval: u32) -> c_int;

    fn rh_engine2_create(sys: *const RhSystem2, n: u64, pos: *const i64, vel: *const i32, species: *const u32, mode: *const RhMode, out: *mut *mut RhEngine2) -> c_int;
    fn rh_engine2_destroy(e: *mut RhEngine2) -> c_int;
    fn rh_engine2_step(e: *mut RhEngine2, steps: u64) -> c_int;
    fn rh_engine2_compute_forces(e: *mut RhEngine2) -> c_int;
    fn rh_engine2_reverse(e: *mut RhEngine2) -> c_int;
    fn rh_engine2_update_system(e: *mut RhEngine2, sys: *const RhSystem2) -> c_int;
    fn rh_engine2_set_arith(e: *mut RhEngine2, arith: c_int) -> c_int;
    fn rh_engine2_set_sort_interval(e: *mut RhEngine2, interval: u32) -> c_int;
    fn rh_engine2_sort_now(e: *mut RhEngine2) -> c_int;
    fn rh_engine2_upload(e: *mut RhEngine2, pos: *const i64, vel: *const i32, species: *const u32) -> c_int;
    fn rh_engine2_read_state(e: *mut RhEngine2, pos: *mut i64, vel: *mut i32, species: *mut u32) -> c_int;
    fn rh_engine2_read_positions(e: *mut RhEngine2, pos: *mut i64) -> c_int;
    fn rh_engine2_kicks(e: *mut RhEngine2, kicks: *mut i32, potential: *mut f64) -> c_int;
    fn rh_engine2_list_stats(e: *mut RhEngine2, out: *mut RhListStats) -> c_int;
    fn rh_engine2_kick_region(e: *mut RhEngine2, lo: *const f64, hi: *const f64, scale: f64, add: *const f64) -> c_int;
    fn rh_engine2_set_species(e: *mut RhEngine2, ids: *const u32, count: u64, species: u32) -> c_int;
    fn rh_engine2_set_species_velocity(e: *mut RhEngine2, species: u32, v: *const f64) -> c_int;
    fn rh_engine2_read_ledger(e: *mut RhEngine2, out: *mut RhLedger) -> c_int;
    fn rh_engine2_read_compartment_ledger(e: *mut RhEngine2, per_wall: u32, energy: *mut f64, impulse: *mut f64, hits: *mut u64) -> c_int;
    fn rh_engine2_read_piston(e: *mut RhEngine2, out: *mut f64) -> c_int;
    fn rh_engine2_set_piston_speed(e: *mut RhEngine2, u: f64) -> c_int;
    fn rh_engine2_flip_bit(e: *mut RhEngine2, particle: u32, axis: u32, bit: u32) -> c_int;
    fn rh_engine2_flip_species_bit(e: *mut RhEngine2, species: u32, axis: u32, bit: u32) -> c_int;
    fn rh_engine2_add_bond(e: *mut RhEngine2, a: u32, b: u32, ty: u32, changed: *mut c_int) -> c_int;
    fn rh_engine2_remove_bond(e: *mut RhEngine2, a: u32, b: u32, changed: *mut c_int) -> c_int;
}

/// A failed call panics with the library's message (the reference asserts on the same
/// contract violations: engine3.rs:343,372,537-541).
fn check(rc: c_int) {
    if rc != 0 {
        let msg = unsafe { CStr::from_ptr(rh_last_error()) }.to_string_lossy().into_owned();
        panic!("rahman-hip ({rc}): {msg}");
    }
}

// ---- conversions at the boundary ------------------------------------------------------------

/// md-core states as flat arrays in particle-id order (the ABI's layout) and back.
pub trait FlatState: Sized {
    fn flat(&self) -> (Vec<i64>, Vec<i32>, Vec<u32>);
    fn from_flat(pos: &[i64], vel: &[i32], species: &[u32]) -> Self;
}

impl FlatState for State3 {
    fn flat(&self) -> (Vec<i64>, Vec<i32>, Vec<u32>) {
        (self.pos.iter().flatten().copied().collect(), self.vel.iter().flatten().copied().collect(), self.species.clone())
    }
    fn from_flat(pos: &[i64], vel: &[i32], species: &[u32]) -> Self {
        State3 {
            pos: pos.chunks_exact(3).map(|p| [p[0], p[1], p[2]]).collect(),
            vel: vel.chunks_exact(3).map(|v| [v[0], v[1], v[2]]).collect(),
            species: species.to_vec(),
        }
    }
}

impl FlatState for State {
    fn flat(&self) -> (Vec<i64>, Vec<i32>, Vec<u32>) {
        (self.pos.iter().flatten().copied().collect(), self.vel.iter().flatten().copied().collect(), self.species.clone())
    }
    fn from_flat(pos: &[i64], vel: &[i32], species: &[u32]) -> Self {
        State {
            pos: pos.chunks_exact(2).map(|p| [p[0], p[1]]).collect(),
            vel: vel.chunks_exact(2).map(|v| [v[0], v[1]]).collect(),
            species: species.to_vec(),
        }
    }
}

/// System3 / System keep pairs row-major `[a * ns + b]`; the ABI a fixed 4 x 4 table.
fn pair_table(pairs: &[md_core::Pair], ns: usize) -> [RhPair; 16] {
    let mut out = [RhPair::default(); 16];
    for a in 0..ns {
        for b in 0..ns {
            let p = pairs[a * ns + b];
            out[a * 4 + b] = RhPair { eps: p.eps, sigma: p.sigma, rc: p.rc };
        }
    }
    out
}

fn wall_code(w: Wall) -> (u32, f64) {
    match w {
        Wall::Hard => (0, 0.0),
        Wall::Soft => (1, 0.0),
        Wall::Thermal(t) => (2, t),
    }
}

fn masses4(m: &[f64]) -> [f64; 4] {
    let mut out = [1.0; 4];
    out[..m.len()].copy_from_slice(m);
    out
}

/// md_core::three::System3 (three.rs:13-31) as rh_system3.
pub fn system3_to_c(sys: &System3) -> RhSystem3 {
    let ns = sys.masses.len();
    let (mut walled, mut wall_kind, mut wall_t) = ([0u32; 3], [[0u32; 2]; 3], [[0.0; 2]; 3]);
    for a in 0..3 {
        if let Some(w) = sys.walls[a] {
            walled[a] = 1;
            for side in 0..2 {
                (wall_kind[a][side], wall_t[a][side]) = wall_code(w[side]);
            }
        }
    }
    let (has_bath, bath_species, bath_t, bath_rate) = match sys.thermostat {
        Some(b) => (1, b.species, b.t, b.rate),
        None => (0, 0, 0.0, 0.0),
    };
    RhSystem3 {
        box_len: sys.box_len,
        dt: sys.dt,
        nspecies: ns as u32,
        pinned: sys.pinned,
        masses: masses4(&sys.masses),
        pairs: pair_table(&sys.pairs, ns),
        walled,
        wall_kind,
        wall_t,
        gravity: sys.gravity,
        seed: sys.seed,
        has_bath,
        bath_species,
        bath_t,
        bath_rate,
    }
}

/// md_core::System (system.rs:26-61) as rh_system2.  The bond triples are borrowed: the
/// returned struct points into `sys.bonds`, which must outlive the call it is passed to.
pub fn system2_to_c(sys: &System) -> RhSystem2 {
    let ns = sys.masses.len();
    let (mut walled, mut wall_kind, mut wall_t) = ([0u32; 2], [[0u32; 2]; 2], [[0.0; 2]; 2]);
    for a in 0..2 {
        if let Some(w) = sys.walls[a] {
            walled[a] = 1;
            for side in 0..2 {
                (wall_kind[a][side], wall_t[a][side]) = wall_code(w[side]);
            }
        }
    }
    let mut c = RhSystem2 {
        box_len: sys.box_len,
        dt: sys.dt,
        nspecies: ns as u32,
        pinned: sys.pinned,
        masses: masses4(&sys.masses),
        pairs: pair_table(&sys.pairs, ns),
        walled,
        gravity: sys.gravity,
        seed: sys.seed,
        wall_kind,
        wall_t,
        compartments: sys.compartments,
        diathermal: (sys.compartment_walls == CompartmentWalls::Diathermal) as u32,
        has_bath: 0,
        bath_species: 0,
        bath_axis: 0,
        bath_t: 0.0,
        bath_rate: 0.0,
        bath_drift: 0.0,
        bath_region: [i64::MIN, i64::MAX],
        has_membrane: 0,
        membrane_axis: 0,
        membrane_blocks: 0,
        pore_shift: 0,
        membrane_position: 0,
        pore_width: 0,
        has_piston: 0,
        piston_free: 0,
        piston_speed: 0.0,
        piston_mass: 0.0,
        piston_force: 0.0,
        piston_start: 0.0,
        piston_min: 0.0,
        piston_max: 0.0,
        nbond_types: sys.bond_types.len() as u32,
        bond_fene: [0; 4],
        bond_k: [0.0; 4],
        bond_r0: [0.0; 4],
        nbonds: sys.bonds.len() as u64,
        bonds: sys.bonds.as_ptr() as *const u32,
    };
    if let Some(t) = sys.thermostat {
        (c.has_bath, c.bath_species, c.bath_axis, c.bath_t, c.bath_rate, c.bath_drift, c.bath_region) = (1, t.species, t.axis as u32, t.t, t.rate, t.drift, t.region);
    }
    if let Some(m) = sys.membrane {
        (c.has_membrane, c.membrane_axis, c.membrane_blocks, c.pore_shift, c.membrane_position, c.pore_width) = (1, m.axis as u32, m.blocks, m.pore_shift, m.position, m.pore_width);
    }
    if let Some(p) = sys.piston {
        c.has_piston = 1;
        match p.drive {
            PistonDrive::Speed(u) => c.piston_speed = u,
            PistonDrive::Free { mass, force } => (c.piston_free, c.piston_mass, c.piston_force) = (1, mass, force),
        }
        (c.piston_start, c.piston_min, c.piston_max) = (p.start, p.min, p.max);
    }
    for (k, b) in sys.bond_types.iter().enumerate() {
        match *b {
            BondType::Harmonic { k: kk, r0 } => (c.bond_fene[k], c.bond_k[k], c.bond_r0[k]) = (0, kk, r0),
            BondType::Fene { k: kk, r0 } => (c.bond_fene[k], c.bond_k[k], c.bond_r0[k]) = (1, kk, r0),
        }
    }
    c
}

fn mode_to_c(mode: ForceMode) -> RhMode {
    match mode {
        ForceMode::CellList => RhMode { kind: 0, cap: 0, skin: 0.0 },
        ForceMode::Verlet { skin, cap } => RhMode { kind: 1, cap, skin },
    }
}

/// engine3.rs:61-68's list capacity without the 128 MiB WebGPU binding clamp (this engine has
/// no binding limit; a list that outgrows its capacity is regrown on the device's signal).
pub fn verlet3_with_skin(sys: &System3, n: usize, skin: f64) -> ForceMode {
    let mut m = RhMode::default();
    check(unsafe { rh_verlet_mode3(&system3_to_c(sys), n as u64, skin, &mut m) });
    ForceMode::Verlet { skin: m.skin, cap: m.cap }
}

/// ForceMode::verlet_with_skin (engine.rs:129-141), unclamped.
pub fn verlet_with_skin(sys: &System, n: usize, skin: f64) -> ForceMode {
    let mut m = RhMode::default();
    check(unsafe { rh_verlet_mode2(&system2_to_c(sys), n as u64, skin, &mut m) });
    ForceMode::Verlet { skin: m.skin, cap: m.cap }
}

fn ledger(l: &RhLedger) -> Ledger {
    Ledger { energy: l.energy, impulse: l.impulse, hits: l.hits }
}

// ---- the 3D engine: md::GpuEngine3 ----------------------------------------------------------

/// md::GpuEngine3 (engine3.rs:307-648) on the MI355X.
pub struct GpuEngine3 {
    h: *mut RhEngine3,
    n: usize,
    pub sys: System3,
    pub mode: ForceMode,
    /// Reorder every this many steps in cell-list mode (0 = never); Verlet mode reorders at
    /// every list rebuild (engine3.rs:327-329, 493-495).
    pub sort_interval: u32,
    sent_interval: u32,
    /// Net steps taken (decreases while running reversed).
    pub steps: i64,
}

impl GpuEngine3 {
    /// With Verlet lists, skin 0.4 (GpuEngine3::new, engine3.rs:337-340).
    pub fn new(sys: System3, state: &State3) -> Self {
        let mode = verlet3_with_skin(&sys, state.n(), 0.4);
        Self::with_mode(sys, state, mode)
    }

    pub fn with_mode(sys: System3, state: &State3, mode: ForceMode) -> Self {
        let (pos, vel, species) = state.flat();
        let mut h = std::ptr::null_mut();
        check(unsafe { rh_engine3_create(&system3_to_c(&sys), state.n() as u64, pos.as_ptr(), vel.as_ptr(), species.as_ptr(), &mode_to_c(mode), &mut h) });
        GpuEngine3 { h, n: state.n(), sys, mode, sort_interval: 100, sent_interval: 100, steps: 0 }
    }

    pub fn n(&self) -> usize {
        self.n
    }

    /// Advance `steps` steps.
    pub fn step(&mut self, steps: usize) {
        if self.sort_interval != self.sent_interval {
            check(unsafe { rh_engine3_set_sort_interval(self.h, self.sort_interval) });
            self.sent_interval = self.sort_interval;
        }
        check(unsafe { rh_engine3_step(self.h, steps as u64) });
        self.steps += steps as i64;
    }

    /// Time reversal, `v <- -(v + kick(x) + g)`.
    pub fn reverse(&mut self) {
        check(unsafe { rh_engine3_reverse(self.h) });
        self.steps = -self.steps;
    }

    pub fn update_system(&mut self, sys: System3) {
        check(unsafe { rh_engine3_update_system(self.h, &system3_to_c(&sys)) });
        self.sys = sys;
    }

    pub fn kick_region(&mut self, lo: [f64; 3], hi: [f64; 3], scale: f64, add: [f64; 3]) {
        check(unsafe { rh_engine3_kick_region(self.h, lo.as_ptr(), hi.as_ptr(), scale, add.as_ptr()) });
    }

    pub fn set_species(&mut self, ids: &[u32], species: u32) {
        check(unsafe { rh_engine3_set_species(self.h, ids.as_ptr(), ids.len() as u64, species) });
    }

    pub fn set_species_velocity(&mut self, species: u32, v: [f64; 3]) {
        check(unsafe { rh_engine3_set_species_velocity(self.h, species, v.as_ptr()) });
    }

    pub fn upload(&mut self, state: &State3) {
        assert_eq!(state.n(), self.n);
        let (pos, vel, species) = state.flat();
        check(unsafe { rh_engine3_upload(self.h, pos.as_ptr(), vel.as_ptr(), species.as_ptr()) });
        self.steps = 0;
    }

    pub fn read_state(&self) -> impl Future<Output = State3> + 'static {
        let (mut pos, mut vel, mut species) = (vec![0i64; 3 * self.n], vec![0i32; 3 * self.n], vec![0u32; self.n]);
        check(unsafe { rh_engine3_read_state(self.h, pos.as_mut_ptr(), vel.as_mut_ptr(), species.as_mut_ptr()) });
        ready(State3::from_flat(&pos, &vel, &species))
    }

    /// Integer kicks at the current positions and the total potential energy.
    pub fn kicks(&self) -> impl Future<Output = (Vec<[i32; 3]>, f64)> + 'static {
        let mut k = vec![0i32; 3 * self.n];
        let mut pe = 0.0;
        check(unsafe { rh_engine3_kicks(self.h, k.as_mut_ptr(), &mut pe) });
        ready((k.chunks_exact(3).map(|c| [c[0], c[1], c[2]]).collect(), pe))
    }

    pub fn measure_kinetic(&self) {
        check(unsafe { rh_engine3_measure_kinetic(self.h) });
    }

    /// Kinetic energies and species in id order (measures now).
    pub fn read_kinetic(&self) -> impl Future<Output = (Vec<f32>, Vec<u32>)> + 'static {
        let (mut ke, mut sp) = (vec![0f32; self.n], vec![0u32; self.n]);
        check(unsafe { rh_engine3_read_kinetic(self.h, ke.as_mut_ptr(), sp.as_mut_ptr()) });
        ready((ke, sp))
    }

    pub fn read_kinetic_total(&self) -> impl Future<Output = f64> + 'static {
        let mut v = 0.0;
        check(unsafe { rh_engine3_read_kinetic_total(self.h, &mut v) });
        ready(v)
    }

    pub fn read_potential(&self) -> impl Future<Output = f64> + 'static {
        let mut v = 0.0;
        check(unsafe { rh_engine3_read_potential(self.h, &mut v) });
        ready(v)
    }

    pub fn read_ledger(&self) -> impl Future<Output = Ledger3> + 'static {
        let mut l = RhLedger::default();
        check(unsafe { rh_engine3_read_ledger(self.h, &mut l) });
        ready(Ledger3 { energy: l.energy, impulse: l.impulse, hits: l.hits })
    }

    /// (neighbour-list rebuilds so far, whether a list ever overflowed): buf.read_list_stats
    /// (engine3.rs:272-279).
    pub fn read_list_stats(&self) -> impl Future<Output = (u32, bool)> + 'static {
        let mut s = RhListStats::default();
        check(unsafe { rh_engine3_list_stats(self.h, &mut s) });
        ready((s.rebuilds, s.overflowed != 0))
    }
}

impl Drop for GpuEngine3 {
    fn drop(&mut self) {
        unsafe {
            rh_engine3_destroy(self.h);
        }
    }
}

// ---- the 2D engine: md::GpuEngine -----------------------------------------------------------

/// md::GpuEngine (engine.rs:567-1098) on the MI355X, with md_core::System's boundaries.
pub struct GpuEngine {
    h: *mut RhEngine2,
    n: usize,
    pub sys: System,
    pub mode: ForceMode,
    pub arith: Arith,
    pub sort_interval: u32,
    sent_interval: u32,
    pub steps: i64,
}

impl GpuEngine {
    /// The force mode chosen by size (ForceMode::auto, engine.rs:119-121).
    pub fn new(sys: System, state: &State) -> Self {
        let mode = if state.n() < ForceMode::VERLET_MIN_N { ForceMode::CellList } else { verlet_with_skin(&sys, state.n(), 0.4) };
        Self::with_mode(sys, state, mode)
    }

    pub fn with_mode(sys: System, state: &State, mode: ForceMode) -> Self {
        Self::with_options(sys, state, mode, Arith::F32)
    }

    pub fn with_options(sys: System, state: &State, mode: ForceMode, arith: Arith) -> Self {
        let (pos, vel, species) = state.flat();
        let c = system2_to_c(&sys);
        let mut h = std::ptr::null_mut();
        check(unsafe { rh_engine2_create(&c, state.n() as u64, pos.as_ptr(), vel.as_ptr(), species.as_ptr(), &mode_to_c(mode), &mut h) });
        if matches!(arith, Arith::Int) {
            check(unsafe { rh_engine2_set_arith(h, 1) });
        }
        GpuEngine { h, n: state.n(), sys, mode, arith, sort_interval: 100, sent_interval: 100, steps: 0 }
    }

    pub fn n(&self) -> u32 {
        self.n as u32
    }

    pub fn step(&mut self, steps: usize) {
        if self.sort_interval != self.sent_interval {
            check(unsafe { rh_engine2_set_sort_interval(self.h, self.sort_interval) });
            self.sent_interval = self.sort_interval;
        }
        check(unsafe { rh_engine2_step(self.h, steps as u64) });
        self.steps += steps as i64;
    }

    /// Evaluate forces at the current positions without moving (for `kicks`).
    pub fn compute_forces(&self) {
        check(unsafe { rh_engine2_compute_forces(self.h) });
    }

    pub fn reverse(&mut self) {
        check(unsafe { rh_engine2_reverse(self.h) });
        self.steps = -self.steps;
    }

    pub fn update_system(&mut self, sys: System) {
        let c = system2_to_c(&sys);
        check(unsafe { rh_engine2_update_system(self.h, &c) });
        self.sys = sys;
    }

    pub fn set_piston_speed(&mut self, u: f64) {
        check(unsafe { rh_engine2_set_piston_speed(self.h, u) });
    }

    pub fn sort_now(&mut self) {
        check(unsafe { rh_engine2_sort_now(self.h) });
    }

    pub fn read_ledger(&self) -> impl Future<Output = Ledger> + 'static {
        let mut l = RhLedger::default();
        check(unsafe { rh_engine2_read_ledger(self.h, &mut l) });
        ready(ledger(&l))
    }

    /// The ledger per wall and per compartment along it (Buffers::read_compartment_ledger,
    /// engine.rs:455-470).
    pub fn read_compartment_ledger(&self, per_wall: usize) -> impl Future<Output = CompartmentLedger> + 'static {
        let (mut e, mut i, mut h) = (vec![0.0; 7 * per_wall], vec![0.0; 7 * per_wall], vec![0u64; 7 * per_wall]);
        check(unsafe { rh_engine2_read_compartment_ledger(self.h, per_wall as u32, e.as_mut_ptr(), i.as_mut_ptr(), h.as_mut_ptr()) });
        ready(CompartmentLedger {
            energy: e.chunks(per_wall.max(1)).map(|c| c.to_vec()).collect(),
            impulse: i.chunks(per_wall.max(1)).map(|c| c.to_vec()).collect(),
            hits: h.chunks(per_wall.max(1)).map(|c| c.to_vec()).collect(),
        })
    }

    /// (height sigma, velocity sigma/tau, gas force eps/sigma, work done by the gas eps).
    pub fn read_piston(&self) -> impl Future<Output = [f64; 4]> + 'static {
        let mut p = [0.0; 4];
        check(unsafe { rh_engine2_read_piston(self.h, p.as_mut_ptr()) });
        ready(p)
    }

    /// XOR bit `bit` of coordinate `axis` of particle `particle` (a flip that would leave the box
    /// is refused with a panic).
    pub fn flip_bit(&mut self, particle: u32, axis: usize, bit: u32) {
        check(unsafe { rh_engine2_flip_bit(self.h, particle, axis as u32, bit) });
    }

    pub fn flip_species_bit(&mut self, species: u32, axis: usize, bit: u32) {
        check(unsafe { rh_engine2_flip_species_bit(self.h, species, axis as u32, bit) });
    }

    pub fn remove_bond(&mut self, a: u32, b: u32) -> bool {
        let mut changed = 0;
        check(unsafe { rh_engine2_remove_bond(self.h, a, b, &mut changed) });
        changed != 0
    }

    pub fn add_bond(&mut self, a: u32, b: u32, ty: u32) -> bool {
        let mut changed = 0;
        check(unsafe { rh_engine2_add_bond(self.h, a, b, ty, &mut changed) });
        changed != 0
    }

    pub fn kick_region(&mut self, lo: [f64; 2], hi: [f64; 2], scale: f64, add: [f64; 2]) {
        check(unsafe { rh_engine2_kick_region(self.h, lo.as_ptr(), hi.as_ptr(), scale, add.as_ptr()) });
    }

    pub fn set_species(&mut self, ids: &[u32], species: u32) {
        check(unsafe { rh_engine2_set_species(self.h, ids.as_ptr(), ids.len() as u64, species) });
    }

    pub fn set_species_velocity(&mut self, species: u32, v: [f64; 2]) {
        check(unsafe { rh_engine2_set_species_velocity(self.h, species, v.as_ptr()) });
    }

    pub fn upload(&mut self, state: &State) {
        assert_eq!(state.n(), self.n);
        let (pos, vel, species) = state.flat();
        check(unsafe { rh_engine2_upload(self.h, pos.as_ptr(), vel.as_ptr(), species.as_ptr()) });
        self.steps = 0;
    }

    pub fn read_state(&self) -> impl Future<Output = State> + 'static {
        let (mut pos, mut vel, mut species) = (vec![0i64; 2 * self.n], vec![0i32; 2 * self.n], vec![0u32; self.n]);
        check(unsafe { rh_engine2_read_state(self.h, pos.as_mut_ptr(), vel.as_mut_ptr(), species.as_mut_ptr()) });
        ready(State::from_flat(&pos, &vel, &species))
    }

    pub fn read_positions(&self) -> impl Future<Output = Vec<[i64; 2]>> + 'static {
        let mut pos = vec![0i64; 2 * self.n];
        check(unsafe { rh_engine2_read_positions(self.h, pos.as_mut_ptr()) });
        ready(pos.chunks_exact(2).map(|p| [p[0], p[1]]).collect())
    }

    pub fn kicks(&self) -> impl Future<Output = (Vec<[i32; 2]>, f64)> + 'static {
        let mut k = vec![0i32; 2 * self.n];
        let mut pe = 0.0;
        check(unsafe { rh_engine2_kicks(self.h, k.as_mut_ptr(), &mut pe) });
        ready((k.chunks_exact(2).map(|c| [c[0], c[1]]).collect(), pe))
    }

    pub fn read_list_stats(&self) -> impl Future<Output = (u32, bool)> + 'static {
        let mut s = RhListStats::default();
        check(unsafe { rh_engine2_list_stats(self.h, &mut s) });
        ready((s.rebuilds, s.overflowed != 0))
    }
}

impl Drop for GpuEngine {
    fn drop(&mut self) {
        unsafe {
            rh_engine2_destroy(self.h);
        }
    }
}
