"""Seeded synthetic particle sets for the five halo-finding workloads.

Units follow the reference (positions Mpc, velocities km/s, masses Msun; G in
(km/s)^2 Mpc / Msun as at halo_gas.py:459-460).  Every generator takes a seed and
returns plain numpy arrays; nothing here touches the GPU.

Workloads (sizes are the full ones; tests use the ``scale`` knobs to shrink them):
  plummer_with_background  one Plummer sphere (20,000 stars, a = 5 kpc, 1e5 Msun each,
                           FP64) plus 4,000 uniform background particles in a 200 kpc box
                           moving at 3x the local escape speed.
  clumps                   1000 Plummer clumps in a 40 Mpc box, sizes p(N) ~ N^-alpha on
                           [100, 20000] (alpha = 1.385 gives ~2M stars; alpha = 2 is the
                           literal slope and gives ~0.53M), 10% uniform field particles,
                           FP32 positions/velocities, log-uniform masses 1e4-1e6 Msun.
  gas_haloes               500 NFW gas haloes (c in [5, 15]), cell counts log-uniform in
                           [1e3, 2e5], bulk + Gaussian velocities, log-normal temperature.
  bridged_pairs            4000 pairs of Plummer spheres 1-3 scale radii apart joined by a
                           uniform cylinder holding 5% of the pair's particles.
  hernquist_haloes         200 Hernquist haloes of 5e4-5e5 particles with off-centre seeds.
"""
import numpy as np

_g = 4.3 * 1e-3
_g *= 1e-6
G = _g  # (km/s)^2 Mpc / Msun


def _unit_vectors(rng, n):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v


def plummer_positions(rng, n, a):
    u = rng.uniform(1e-12, 1.0, size=n)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    return r[:, None] * _unit_vectors(rng, n)


def plummer_sigma1d(M, a):
    return np.sqrt(np.pi * G * M / (32.0 * a))


def linking_length(volume, n, b=0.2):
    """b times the mean interparticle spacing (V/N)^(1/3)."""
    return b * (volume / n) ** (1.0 / 3.0)


def plummer_with_background(seed=20261203, n_sphere=20000, n_bg=4000, a=0.005,
                            m=1e5, box=0.2):
    rng = np.random.default_rng(seed)
    M = n_sphere * m
    pos = plummer_positions(rng, n_sphere, a)
    sig = plummer_sigma1d(M, a)
    vel = rng.normal(scale=sig, size=(n_sphere, 3))
    bpos = rng.uniform(-box / 2, box / 2, size=(n_bg, 3))
    r = np.linalg.norm(bpos, axis=1)
    vesc = np.sqrt(2.0 * G * M / np.sqrt(r * r + a * a))
    bvel = (3.0 * vesc)[:, None] * _unit_vectors(rng, n_bg)
    x = np.concatenate([pos, bpos]).astype(np.float64)
    v = np.concatenate([vel, bvel]).astype(np.float64)
    mass = np.full(n_sphere + n_bg, m, dtype=np.float64)
    ll = linking_length(box ** 3, n_sphere + n_bg)
    return dict(x=x[:, 0].copy(), y=x[:, 1].copy(), z=x[:, 2].copy(),
                vx=v[:, 0].copy(), vy=v[:, 1].copy(), vz=v[:, 2].copy(), mass=mass,
                ll=ll, box=box, n_sphere=n_sphere)


def _powerlaw_sizes(rng, n, lo, hi, alpha):
    u = rng.uniform(size=n)
    if abs(alpha - 1.0) < 1e-12:
        return np.floor(lo * (hi / lo) ** u).astype(np.int64)
    k = 1.0 - alpha
    return np.floor((lo ** k + u * (hi ** k - lo ** k)) ** (1.0 / k)).astype(np.int64)


def clumps(seed=20261204, n_clumps=1000, nmin=100, nmax=20000, alpha=1.385,
           box=40.0, field_frac=0.10, scale=1.0, bulk_sigma=300.0, field_sigma=300.0):
    """Clumped star particle set.  ``scale`` < 1 shrinks the clump count and box volume
    together (density kept) for test-sized instances."""
    rng = np.random.default_rng(seed)
    nc = max(1, int(round(n_clumps * scale)))
    L = box * scale ** (1.0 / 3.0)
    sizes = _powerlaw_sizes(rng, nc, nmin, nmax, alpha)
    n_cl = int(sizes.sum())
    n_field = int(round(field_frac / (1.0 - field_frac) * n_cl))
    centres = rng.uniform(-L / 2, L / 2, size=(nc, 3))
    bulk = rng.normal(scale=bulk_sigma, size=(nc, 3))
    xs, vs, ms = [], [], []
    for k in range(nc):
        n = int(sizes[k])
        a = 0.005 * (n / 2e4) ** (1.0 / 3.0)
        m = np.exp(rng.uniform(np.log(1e4), np.log(1e6), size=n))
        pos = plummer_positions(rng, n, a) + centres[k]
        sig = plummer_sigma1d(m.sum(), a)
        vel = rng.normal(scale=sig, size=(n, 3)) + bulk[k]
        xs.append(pos)
        vs.append(vel)
        ms.append(m)
    xs.append(rng.uniform(-L / 2, L / 2, size=(n_field, 3)))
    vs.append(rng.normal(scale=field_sigma, size=(n_field, 3)))
    ms.append(np.exp(rng.uniform(np.log(1e4), np.log(1e6), size=n_field)))
    x = np.concatenate(xs)
    x = (x + L / 2) % L - L / 2  # periodic placement (FoF itself is non-periodic)
    v = np.concatenate(vs)
    mass = np.concatenate(ms)
    perm = rng.permutation(len(mass))  # particle order carries no clump information
    x, v, mass = x[perm], v[perm], mass[perm]
    ll = linking_length(L ** 3, len(mass))
    f32 = np.float32
    return dict(x=x[:, 0].astype(f32), y=x[:, 1].astype(f32), z=x[:, 2].astype(f32),
                vx=v[:, 0].astype(f32), vy=v[:, 1].astype(f32), vz=v[:, 2].astype(f32),
                mass=mass.astype(f32), ll=ll, box=L, sizes=sizes, n_field=n_field)


def _nfw_m(x):
    return np.log1p(x) - x / (1.0 + x)


def nfw_radii(rng, n, c, rs):
    """Inverse of the enclosed-mass profile m(x) = ln(1+x) - x/(1+x) on [0, c] by
    bisection (monotone), r = rs * x."""
    target = rng.uniform(size=n) * _nfw_m(c)
    lo = np.zeros(n)
    hi = np.full(n, float(c))
    for _ in range(45):  # bracket below 1e-12 of c
        mid = 0.5 * (lo + hi)
        less = _nfw_m(mid) < target
        lo = np.where(less, mid, lo)
        hi = np.where(less, hi, mid)
    return rs * 0.5 * (lo + hi)


def gas_haloes(seed=20261205, n_haloes=500, nmin=1000, nmax=200000, scale=1.0,
               box=40.0, m_cell=1e5, dm_per_gas=0.0, st_per_gas=0.0):
    """Batched gas-cell haloes.  Returns CSR-ordered arrays with ``offsets``.  With
    dm_per_gas / st_per_gas > 0, also NFW dark-matter (5 m_cell each) and stellar
    (0.1 m_cell) particles per halo, in their own CSR arrays (``dm``, ``stars``:
    (x, y, z, mass, offsets)), drawn after all the gas so the gas is unchanged."""
    rng = np.random.default_rng(seed)
    nh = max(1, int(round(n_haloes * scale)))
    sizes = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=nh))).astype(np.int64)
    offsets = np.zeros(nh + 1, np.int64)
    offsets[1:] = np.cumsum(sizes)
    xs, vs, ms, ts, bulks = [], [], [], [], []
    centres = rng.uniform(-box / 2, box / 2, size=(nh, 3))
    for k in range(nh):
        n = int(sizes[k])
        mass = rng.uniform(0.5, 1.5, size=n) * m_cell
        Mtot = mass.sum() * 6.0  # gas is a sixth of the halo's mass
        rvir = (G * Mtot / 100.0 / 1e4) ** (1.0 / 3.0) * 0.05 + 0.02
        c = rng.uniform(5.0, 15.0)
        r = nfw_radii(rng, n, c, rvir / c)
        pos = r[:, None] * _unit_vectors(rng, n) + centres[k]
        vc = np.sqrt(G * Mtot / rvir)
        bulk = rng.normal(scale=200.0, size=3)
        vel = bulk + rng.normal(scale=vc / np.sqrt(3.0), size=(n, 3))
        temp = 10.0 ** rng.normal(4.0, 0.5, size=n)
        xs.append(pos)
        vs.append(vel)
        ms.append(mass)
        ts.append(temp)
        bulks.append(bulk)
    x = np.concatenate(xs)
    v = np.concatenate(vs)
    f32 = np.float32
    out = dict(x=x[:, 0].astype(f32), y=x[:, 1].astype(f32), z=x[:, 2].astype(f32),
               vx=v[:, 0].astype(f32), vy=v[:, 1].astype(f32), vz=v[:, 2].astype(f32),
               mass=np.concatenate(ms).astype(f32), temp=np.concatenate(ts).astype(f32),
               offsets=offsets, bulk=np.array(bulks), sizes=sizes)
    for key, per, mfac in (("dm", dm_per_gas, 5.0), ("stars", st_per_gas, 0.1)):
        if per <= 0:
            continue
        cnt = np.maximum(1, (sizes * per).astype(np.int64))
        off = np.zeros(nh + 1, np.int64)
        off[1:] = np.cumsum(cnt)
        ps = []
        for k in range(nh):
            n = int(cnt[k])
            c = rng.uniform(5.0, 15.0)
            rv = 0.05 * (sizes[k] / 1e4) ** (1.0 / 3.0) + 0.02
            ps.append(nfw_radii(rng, n, c, rv / c)[:, None] * _unit_vectors(rng, n)
                      + centres[k])
        p = np.concatenate(ps)
        out[key] = (p[:, 0].astype(f32), p[:, 1].astype(f32), p[:, 2].astype(f32),
                    np.full(len(p), mfac * m_cell, f32), off)
    return out


def bridged_pairs(seed=20261206, n_pairs=4000, n_per_pair=2000, bridge_frac=0.05,
                  box=40.0, scale=1.0):
    """Pairs of Plummer spheres joined by uniform cylinders.  The linking length is
    0.35 a (a = sphere scale radius), which the bridge's axial density percolates.
    Velocities (from a second generator, so positions do not depend on them): each
    sphere moves along the pair axis at -+v_rel/2 plus its virial dispersion, with
    v_rel = f * sqrt(2 G M_pair / sep), f uniform in [0.5, 2], so some pairs are bound
    and some fly apart; bridge particles move with the linear interpolation between the
    two spheres."""
    rng = np.random.default_rng(seed)
    npairs = max(1, int(round(n_pairs * scale)))
    L = box * scale ** (1.0 / 3.0)
    nb = int(round(bridge_frac * n_per_pair))
    ns = (n_per_pair - nb) // 2
    a = 0.005 * (ns / 2e4) ** (1.0 / 3.0)
    xs, pair_id, geo = [], [], []
    centres = rng.uniform(-L / 2, L / 2, size=(npairs, 3))
    for k in range(npairs):
        sep = rng.uniform(1.0, 3.0) * a
        axis = _unit_vectors(rng, 1)[0]
        c1 = centres[k] - 0.5 * sep * axis
        c2 = centres[k] + 0.5 * sep * axis
        p1 = plummer_positions(rng, ns, a) + c1
        p2 = plummer_positions(rng, ns, a) + c2
        # uniform cylinder of radius 0.25 a from c1 to c2
        t = rng.uniform(0.0, 1.0, size=nb)
        rr = 0.25 * a * np.sqrt(rng.uniform(size=nb))
        ph = rng.uniform(0, 2 * np.pi, size=nb)
        e1 = np.cross(axis, [1.0, 0.0, 0.0])
        if np.linalg.norm(e1) < 1e-6:
            e1 = np.cross(axis, [0.0, 1.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(axis, e1)
        pb = (c1 + (t * sep)[:, None] * axis + (rr * np.cos(ph))[:, None] * e1
              + (rr * np.sin(ph))[:, None] * e2)
        xs += [p1, p2, pb]
        pair_id.append(np.full(2 * ns + nb, k))
        geo.append((axis, sep, t))
    x = np.concatenate(xs)
    pid = np.concatenate(pair_id)
    perm = rng.permutation(len(x))
    x, pid = x[perm], pid[perm]
    # velocities, in generation order, then the same permutation
    vr = np.random.default_rng(seed + 7919)
    m_p = 1e5
    sig = plummer_sigma1d(ns * m_p, a)
    vs = []
    for k in range(npairs):
        axis, sep, t = geo[k]
        vrel = vr.uniform(0.5, 2.0) * np.sqrt(2.0 * G * (2 * ns + nb) * m_p / sep)
        v1 = -0.5 * vrel * axis
        v2 = 0.5 * vrel * axis
        vs += [v1 + vr.normal(scale=sig, size=(ns, 3)), v2 + vr.normal(scale=sig, size=(ns, 3)),
               v1 + t[:, None] * (v2 - v1) + vr.normal(scale=sig, size=(nb, 3))]
    v = np.concatenate(vs)[perm]
    f32 = np.float32
    n = len(x)
    return dict(x=x[:, 0].astype(f32), y=x[:, 1].astype(f32), z=x[:, 2].astype(f32),
                vx=v[:, 0].astype(f32), vy=v[:, 1].astype(f32), vz=v[:, 2].astype(f32),
                mass=np.full(n, m_p, f32), ll=0.35 * a, pair_id=pid, a=a, box=L)


def hernquist_positions(rng, n, a):
    u = rng.uniform(1e-12, 1.0 - 1e-9, size=n)
    s = np.sqrt(u)
    r = a * s / (1.0 - s)
    return r[:, None] * _unit_vectors(rng, n)


def hernquist_haloes(seed=20261207, n_haloes=200, nmin=50000, nmax=500000, scale=1.0,
                     nscale=1.0, box=40.0):
    """Hernquist haloes (CSR order) with a seed centre offset by 0.1-0.5 a.  ``nscale``
    shrinks every halo's particle count (tests)."""
    rng = np.random.default_rng(seed)
    nh = max(1, int(round(n_haloes * scale)))
    sizes = np.floor(np.exp(rng.uniform(np.log(nmin * nscale), np.log(nmax * nscale),
                                        size=nh))).astype(np.int64)
    offsets = np.zeros(nh + 1, np.int64)
    offsets[1:] = np.cumsum(sizes)
    xs, seeds, centres, scales = [], [], [], []
    cen = rng.uniform(-box / 2, box / 2, size=(nh, 3))
    for k in range(nh):
        n = int(sizes[k])
        a = 0.01 * (n / 1e5) ** (1.0 / 3.0)
        xs.append(hernquist_positions(rng, n, a) + cen[k])
        off = rng.uniform(0.1, 0.5) * a * _unit_vectors(rng, 1)[0]
        seeds.append(cen[k] + off)
        centres.append(cen[k])
        scales.append(a)
    x = np.concatenate(xs)
    f32 = np.float32
    n = len(x)
    return dict(x=x[:, 0].astype(f32), y=x[:, 1].astype(f32), z=x[:, 2].astype(f32),
                mass=np.full(n, 1e5, f32), offsets=offsets, seeds=np.array(seeds),
                centres=np.array(centres), a=np.array(scales), sizes=sizes)


def dark_matter_around(centres, seed=20261208, per_halo=4000, n_field=2000000, box=40.0,
                       a=0.02):
    """Dark-matter particles: a Plummer sphere of scale ``a`` (Mpc) and ``per_halo``
    particles around each centre plus a uniform field, FP32, masses 1e-12 code units."""
    rng = np.random.default_rng(seed)
    centres = np.asarray(centres, np.float64)
    parts = [plummer_positions(rng, per_halo, a) + c for c in centres]
    parts.append(rng.uniform(-box / 2, box / 2, size=(n_field, 3)))
    x = np.concatenate(parts)
    x = x[rng.permutation(len(x))].astype(np.float32)
    m = rng.uniform(0.8e-12, 1.2e-12, len(x)).astype(np.float32)
    return dict(x=x[:, 0].copy(), y=x[:, 1].copy(), z=x[:, 2].copy(), mass=m)


def amr_around(centres, seed=20261209, L=40.0, ncoarse=128, levels=(6, 7), cells=16,
               refined=0.15, overlapped=0.05):
    """A MASCLET-like AMR snapshot: for every centre one patch per level in ``levels``
    of ``cells`` cells per axis centred on it (parent-cell aligned), FP32 gas fields
    (delta lognormal, velocities in c = 1 units, temperatures 1e4-1e7 K) and random
    refined / overlapped flags.  Returns the AmrGrid constructor arguments (level-0 base
    patch included, empty)."""
    rng = np.random.default_rng(seed)
    base = L / ncoarse
    lev, nx, rx, ry, rz = [0], [1], [0.0], [0.0], [0.0]
    fields = [[np.zeros((1, 1, 1), np.float32)] for _ in range(5)] + \
        [[np.zeros((1, 1, 1), bool)] for _ in range(2)]
    for c in np.asarray(centres, np.float64):
        for level in levels:
            res = base / 2 ** level
            corner = np.floor((c - cells * res / 2) / (2 * res)) * 2 * res + res
            lev.append(level)
            nx.append(cells)
            rx.append(corner[0])
            ry.append(corner[1])
            rz.append(corner[2])
            shp = (cells, cells, cells)
            fields[0].append(np.exp(rng.normal(1.0, 1.0, shp)).astype(np.float32) - 1.0)
            for k in (1, 2, 3):
                fields[k].append(rng.normal(scale=5e-4, size=shp).astype(np.float32))
            fields[4].append(np.exp(rng.uniform(np.log(1e4), np.log(1e7), shp)).astype(np.float32))
            fields[5].append(rng.random(shp) >= refined)
            fields[6].append(rng.random(shp) >= overlapped)
    n = np.array(nx)
    return dict(L=L, ncoarse=ncoarse, levels=np.array(lev), nx=n, ny=n, nz=n,
                rx=np.array(rx), ry=np.array(ry), rz=np.array(rz), cells=fields)
