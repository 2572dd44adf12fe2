"""Total energy over 10,000 steps of the 3D liquid on the plain engine: the 262,144-atom FCC melt
at density 0.8442 and T 0.72 (rc 2.5, skin 0.3, dt 0.005), with energies from the oracle's f64
evaluation (md-ref's f64-force path) of GPU states, as crates/md/tests/three.rs:95-111 evaluates
them.  Relative drift and spread stay within 1e-4 once the lattice has melted.  At small N the
GPU's f32-force trajectory is also set beside the oracle's own f64-force trajectory from the
same start: both conserve energy to 1e-4 and their energies agree to 1e-4."""
import numpy as np
import pytest

import micrographia_amd as rh
from oracle_twin import energy, oracle_system
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-4  # relative total energy over 10k steps


def liquid(n, t, seed=1):
    sys = rh.System3.new(rh.cube_for(n, 0.8442), 0.005, rh.Pair.lj(2.5), [1.0])
    pos = rh.fcc(n, sys.box_len, rh.fcc_cells(n, sys.box_len), 0.05, seed)
    sp = np.zeros(n, np.uint32)
    return sys, rh.State3(pos, rh.maxwell(sys, sp, t, seed), sp)


def test_energy_within_1e4_over_10k_steps_at_262144_atoms():
    n = 262144
    sys, st = liquid(n, 0.72)
    gpu = rh.GpuEngine3(sys, st, rh.verlet3_with_skin(sys, n, 0.3))
    series = []
    for k in range(21):
        if k:
            gpu.step(500)
        s = gpu.read_state()
        ke, pe = energy(sys, s)
        series.append(ke + pe)
        print(f"step {500 * k}: E/N {(ke + pe) / n:.7f}", flush=True)
    e = np.array(series)
    ref = e[4]  # from step 2000, after the lattice has melted
    rel = np.abs(e[4:] - ref) / abs(ref)
    print("max relative deviation from step 2000:", rel.max())
    assert rel.max() < TOL
    assert abs(e[-1] - ref) / abs(ref) < TOL
    assert gpu.read_list_stats().rebuilds > 50 and not gpu.read_list_stats().overflowed


def test_f32_and_f64_force_trajectories_conserve_the_same_energy():
    n = 256
    sys, st = liquid(n, 0.72, seed=3)
    gpu = rh.GpuEngine3(sys, st, rh.verlet3_with_skin(sys, n, 0.3))
    o = oracle_system(sys)
    pos, vel, sp = st.pos.copy(), st.vel.copy(), st.species.copy()
    eg, eo = [], []
    for k in range(11):
        if k:
            gpu.step(1000)
            O.step3(o, pos, vel, sp, 1000, 64)  # the f64-force twin
        eg.append(sum(energy(sys, gpu.read_state())))
        eo.append(sum(O.energy3(o, pos, vel, sp, 64)))
    eg, eo = np.array(eg), np.array(eo)
    print("GPU (f32 forces):", eg / n)
    print("oracle f64 forces:", eo / n)
    ref = eo[2]
    assert np.max(np.abs(eg[2:] - eg[2])) / abs(ref) < TOL
    assert np.max(np.abs(eo[2:] - eo[2])) / abs(ref) < TOL
    assert abs(eg[2:].mean() - eo[2:].mean()) / abs(ref) < TOL
