"""Runs a product-side system and state through the oracle (test infrastructure only): the
oracle's image of a System3 / System, and a state advanced by the oracle's IntEngine3<f32> /
IntEngine<f32> (or <f64>), for the GPU tests to compare with."""
import micrographia_amd as rh
from oracle import oracle as O


def oracle_system(sys):
    c = sys.to_c()
    if isinstance(sys, rh.System3):
        o = O.System3()
        for a in range(3):
            o.box_len[a] = c.box_len[a]
            o.walled[a] = c.walled[a]
            o.gravity[a] = c.gravity[a]
            for s in range(2):
                o.wall_kind[a][s] = c.wall_kind[a][s]
                o.wall_t[a][s] = c.wall_t[a][s]
        o.seed = c.seed
        o.has_bath = c.has_bath
        o.bath_species = c.bath_species
        o.bath_t = c.bath_t
        o.bath_rate = c.bath_rate
    else:
        # field by field: the wall kinds and temperatures, the seed and every boundary travel too
        # (the bond array stays alive on the copy)
        import ctypes as C

        o = O.System2()
        for name, _ in O.System2._fields_:
            src, dst = getattr(type(c), name), getattr(O.System2, name)
            assert src.size == dst.size, name
            C.memmove(C.addressof(o) + dst.offset, C.addressof(c) + src.offset, src.size)
        o._keep = c
        return o
    o.dt = c.dt
    o.nspecies = c.nspecies
    o.pinned = c.pinned
    for k in range(4):
        o.masses[k] = c.masses[k]
    for k in range(16):
        o.pairs[k] = O.Pair(c.pairs[k].eps, c.pairs[k].sigma, c.pairs[k].rc)
    return o


def oracle_state(sys, st, steps, prec=32):
    """The state after `steps` oracle steps (a copy; `st` is left alone)."""
    pos, vel, sp = st.pos.copy(), st.vel.copy(), st.species.copy()
    o = oracle_system(sys)
    if isinstance(sys, rh.System3):
        O.step3(o, pos, vel, sp, steps, prec)
        return rh.State3(pos, vel, sp)
    O.step2(o, pos, vel, sp, steps, prec)
    return rh.State(pos, vel, sp)


def energy(sys, st):
    """(kinetic, potential) of a state by the oracle's f64 evaluation (md-ref's f64-force path)."""
    o = oracle_system(sys)
    if isinstance(sys, rh.System3):
        return O.energy3(o, st.pos, st.vel, st.species, 64)
    return O.energy2(o, st.pos, st.vel, st.species, 64)
