"""Energy-test and unbinding oracles (test infrastructure only).

* ``G_CONST``: halo_gas.py:459-460 / halo_properties.py:345-346, (km/s)^2 Mpc / Msun,
  evaluated with the same two Python float operations so the bits agree.
* ``energy_test``: E = 1/2 |v - v_cm|^2 - factor * G * S, bound iff E <= 0
  (halo_gas.py:456-476 with factor 2; halo_properties.py:342-359 with factor_v**2).
  FP64, with the operation order the GPU kernels follow:
  kin = 0.5 * ((dvx*dvx + dvy*dvy) + dvz*dvz); E = kin - (factor*G) * S.
* ``unbind_iterative``: the iterated star test.  The reference evaluates the test once
  (halo_properties.py:282-361, never called by the driver); iterating it until no
  particle leaves the bound set is this build's extension.  Per halo: start with every
  member bound; each pass recomputes v_cm over the bound set (correctly rounded FP64 sums,
  CM_velocity halo_properties.py:45-60) and S over the bound set (FP64 direct sum on
  FP32-cast inputs, particle_subroutines.f90:466-514), and removes members with E > 0.
* ``rps`` and ``most_bound_particle`` restate halo_gas.py:285-492 and :498-634.  Their
  random subsampling above BRUTE_FORCE_LIM (np.random.choice, unseeded in the reference)
  is taken from ``np.random`` here in the same call order, so a seeded global RNG gives
  the same samples as the reference would draw.
"""
import math

import numpy as np

from .native import pot_sum_f64, pot_sum_f32

_g = 4.3 * 1e-3
_g *= 1e-6
G_CONST = _g


def binding_energy(total_mass, total_x, total_y, total_z, test_x, test_y, test_z,
                   precision="f64"):
    """Positive sum S_i = sum_j m_j / r_ij (halo_gas.py:164-187 semantics: FP32 casts,
    empty test set -> empty float64 array)."""
    if len(test_x) == 0:
        return np.array([])
    fn = pot_sum_f64 if precision == "f64" else pot_sum_f32
    return fn(total_mass, total_x, total_y, total_z, test_x, test_y, test_z)


def kinetic(vx, vy, vz, vcm):
    dvx = np.asarray(vx, np.float64) - vcm[0]
    dvy = np.asarray(vy, np.float64) - vcm[1]
    dvz = np.asarray(vz, np.float64) - vcm[2]
    return 0.5 * ((dvx * dvx + dvy * dvy) + dvz * dvz)


def energy_test(S, vx, vy, vz, vcm, factor):
    """Returns (E, bound) in FP64; bound iff E <= 0."""
    cg = float(factor) * G_CONST
    E = kinetic(vx, vy, vz, vcm) - cg * np.asarray(S, np.float64)
    return E, E <= 0.0


def _vcm(m, vx, vy, vz):
    M = math.fsum(m.tolist())
    if not M > 0:
        return (0.0, 0.0, 0.0)
    return tuple(math.fsum((m * v).tolist()) / M for v in (vx, vy, vz))


def unbind_iterative(offsets, members, x, y, z, vx, vy, vz, mass, factor, max_iter=100):
    """Iterated energy test per halo.  ``members`` lists particle indices halo by halo
    (CSR with ``offsets``).  Returns (bound mask over ``members``, iterations per halo,
    final S (FP64) over ``members`` (NaN for members removed before the last pass))."""
    offsets = np.asarray(offsets, np.int64)
    members = np.asarray(members, np.int64)
    x32 = np.asarray(x, np.float32)
    y32 = np.asarray(y, np.float32)
    z32 = np.asarray(z, np.float32)
    m32 = np.asarray(mass, np.float32)
    m64 = np.asarray(mass, np.float64)
    vx = np.asarray(vx, np.float64)
    vy = np.asarray(vy, np.float64)
    vz = np.asarray(vz, np.float64)
    H = len(offsets) - 1
    bound = np.ones(len(members), dtype=bool)
    S_out = np.full(len(members), np.nan)
    iters = np.zeros(H, dtype=np.int64)
    for h in range(H):
        lo, hi = offsets[h], offsets[h + 1]
        idx = members[lo:hi]
        b = np.ones(hi - lo, dtype=bool)
        for it in range(max_iter):
            sel = idx[b]
            if sel.size == 0:
                break
            iters[h] = it + 1
            vcm = _vcm(m64[sel], vx[sel], vy[sel], vz[sel])
            S = pot_sum_f64(m32[sel], x32[sel], y32[sel], z32[sel],
                            x32[sel], y32[sel], z32[sel])
            _, ok = energy_test(S, vx[sel], vy[sel], vz[sel], vcm, factor)
            pos = np.nonzero(b)[0]
            S_out[lo + pos] = S
            if ok.all():
                break
            b[pos[~ok]] = False
        bound[lo:hi] = b
    return bound, iters, S_out


def rps(gas_x, gas_y, gas_z, gas_vx, gas_vy, gas_vz, gas_mass, gas_temp,
        dm_x, dm_y, dm_z, dm_mass, st_x, st_y, st_z, st_mass, vx, vy, vz,
        BRUTE_FORCE_LIM, mass_dm_part, num_dm_species, pot_fn=None, return_energy=False):
    """halo_gas.py:285-492 restated; ``pot_fn`` defaults to the FP32 restatement of the
    Fortran routine (what the reference calls).  With ``return_energy`` the per-cell E
    (the array the reference thresholds at :470-474) is returned after the four values."""
    if pot_fn is None:
        def pot_fn(m, x, y, z, tx, ty, tz):
            if len(tx) == 0:
                return np.array([])
            return pot_sum_f32(m, x, y, z, tx, ty, tz)

    ngas = len(gas_x)
    be = np.zeros((ngas,), dtype=np.float32)

    def sampled(mm, xx, yy, zz, n):
        ns = np.max([BRUTE_FORCE_LIM, int(0.01 * n)])
        s = np.random.choice(np.arange(n), ns, replace=True)
        return pot_fn(mm[s], xx[s], yy[s], zz[s], gas_x, gas_y, gas_z) * n / ns

    if ngas > BRUTE_FORCE_LIM:
        be += sampled(gas_mass, gas_x, gas_y, gas_z, ngas)
    elif ngas > 0:
        be += pot_fn(gas_mass, gas_x, gas_y, gas_z, gas_x, gas_y, gas_z)

    if num_dm_species > 1:
        mand = dm_mass >= 0.9 * (mass_dm_part / 8)
        if np.count_nonzero(mand) > 0:
            be += pot_fn(dm_mass[mand], dm_x[mand], dm_y[mand], dm_z[mand],
                         gas_x, gas_y, gas_z)
        oth = np.logical_not(mand)
        ox, oy, oz, om = dm_x[oth], dm_y[oth], dm_z[oth], dm_mass[oth]
        n = len(ox)
        if n > BRUTE_FORCE_LIM:
            be += sampled(om, ox, oy, oz, n)
        elif n > 0:
            be += pot_fn(om, ox, oy, oz, gas_x, gas_y, gas_z)
    else:
        n = len(dm_x)
        if n > BRUTE_FORCE_LIM:
            be += sampled(dm_mass, dm_x, dm_y, dm_z, n)
        elif n > 0:
            be += pot_fn(dm_mass, dm_x, dm_y, dm_z, gas_x, gas_y, gas_z)

    nst = len(st_x)
    if nst > BRUTE_FORCE_LIM:
        be += sampled(st_mass, st_x, st_y, st_z, nst)
    elif nst > 0:
        be += pot_fn(st_mass, st_x, st_y, st_z, gas_x, gas_y, gas_z)

    be = -be
    be *= G_CONST
    be *= 2.0
    gas_v2 = 0.5 * ((gas_vx - vx) ** 2 + (gas_vy - vy) ** 2 + (gas_vz - vz) ** 2)
    E = gas_v2 + be
    unbound = E > 0.0
    bound = E <= 0.0
    cold = gas_temp < 5 * 1e4
    hot = gas_temp >= 5 * 1e4
    total = np.sum(gas_mass)
    cold_bound = np.sum(gas_mass[cold * bound])
    frac = cold_bound / total if total != 0.0 else 0.0
    out = (total, frac, np.sum(gas_mass[unbound * cold]), np.sum(gas_mass[unbound * hot]))
    return out + (E,) if return_energy else out


def most_bound_particle(gas_x, gas_y, gas_z, gas_mass, dm_x, dm_y, dm_z, dm_mass,
                        st_x, st_y, st_z, st_mass, st_oripa, BRUTE_FORCE_LIM,
                        mass_dm_part, pot_fn=None, return_energy=False):
    """halo_gas.py:498-634 restated (argmin of -S over the stars).  With ``return_energy``
    the per-star -S (the array of the argmin, :632) is returned after the four values."""
    if pot_fn is None:
        def pot_fn(m, x, y, z, tx, ty, tz):
            if len(tx) == 0:
                return np.array([])
            return pot_sum_f32(m, x, y, z, tx, ty, tz)

    nst = len(st_x)
    be = np.zeros((nst,), dtype=np.float32)

    def sampled(mm, xx, yy, zz, n):
        ns = np.max([BRUTE_FORCE_LIM, int(0.01 * n)])
        s = np.random.choice(np.arange(n), ns, replace=True)
        return pot_fn(mm[s], xx[s], yy[s], zz[s], st_x, st_y, st_z) * n / ns

    ngas = len(gas_x)
    if ngas > BRUTE_FORCE_LIM:
        be += sampled(gas_mass, gas_x, gas_y, gas_z, ngas)
    elif ngas > 0:
        be += pot_fn(gas_mass, gas_x, gas_y, gas_z, st_x, st_y, st_z)

    mand = dm_mass >= 0.9 * (mass_dm_part / 8)
    if np.count_nonzero(mand) > 0:
        be += pot_fn(dm_mass[mand], dm_x[mand], dm_y[mand], dm_z[mand], st_x, st_y, st_z)
    oth = np.logical_not(mand)
    ox, oy, oz, om = dm_x[oth], dm_y[oth], dm_z[oth], dm_mass[oth]
    n = len(ox)
    if n > BRUTE_FORCE_LIM:
        be += sampled(om, ox, oy, oz, n)
    elif n > 0:
        be += pot_fn(om, ox, oy, oz, st_x, st_y, st_z)

    if nst > BRUTE_FORCE_LIM:
        be += sampled(st_mass, st_x, st_y, st_z, nst)
    elif nst > 0:
        be += pot_fn(st_mass, st_x, st_y, st_z, st_x, st_y, st_z)

    be = -be
    k = np.argmin(be)
    out = (st_x[k], st_y[k], st_z[k], st_oripa[k])
    return out + (be,) if return_energy else out
