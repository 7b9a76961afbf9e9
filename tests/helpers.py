"""Test inputs: small seeded SOP systems built in numpy, independent of the
engine's own generator, plus parameter presets for the configs."""
import numpy as np

from sopgpu import System  # dataclass only


def numpy_chain(n, seed, density=175.0, eh=(0.9, 1.5), bond=3.8, dmin=4.0):
    """Confined self-avoiding walk (numpy), contacts |i-j|>2 and r<8 A."""
    rng = np.random.default_rng(seed)
    R = (3.0 * n * density / (4.0 * np.pi)) ** (1.0 / 3.0)
    pos = np.zeros((n, 3))
    i = 1
    while i < n:
        for _ in range(5000):
            u = rng.uniform(-1, 1)
            ph = rng.uniform(0, 2 * np.pi)
            s = np.sqrt(1 - u * u)
            q = pos[i - 1] + bond * np.array([s * np.cos(ph), s * np.sin(ph), u])
            if q @ q > R * R:
                continue
            if i > 1 and np.min(np.sum((pos[: i - 1] - q) ** 2, axis=1)) < dmin * dmin:
                continue
            pos[i] = q
            i += 1
            break
        else:
            R *= 1.05
    c = np.zeros((n, 4), np.float32)
    c[:, :3] = pos.astype(np.float32)
    cd = c[:, :3].astype(np.float64)
    bonds = np.stack([np.arange(n - 1), np.arange(1, n)], axis=1).astype(np.int32)
    br0 = np.linalg.norm(cd[1:] - cd[:-1], axis=1).astype(np.float32)
    d = np.linalg.norm(cd[:, None, :] - cd[None, :, :], axis=2)
    ii, jj = np.nonzero(np.triu(d < 8.0, k=3))
    nat = np.stack([ii, jj], axis=1).astype(np.int32)
    nr0 = d[ii, jj].astype(np.float32)
    neh = rng.uniform(eh[0], eh[1], size=len(ii)).astype(np.float32)
    return System(c, bonds, br0, nat, nr0, neh)


def c0_params(**kw):
    """Plain-method small config: FENE k=14, R0=2; repulsive sigma 3.8,
    eps 1; cutoffs 20/32 A; rebuilds 10/100; T=0.6 kcal/mol, h=40 ps in
    tau_H units (0.16); force-ramp on the last bead, bead 0 fixed."""
    d = dict(kspring_cov=14.0, R_limit=2.0, a=3.8, el=1.0, pairs_cutoff=20.0, pairlist_cutoff=20.0,
             pairs_threshold=32.0, pairs_freq=10, possiblepairs_freq=100, timestep=0.16, zeta=50.0,
             temperature=0.6, seed=4242, nav=1000, outputtiming=10000, dcdfreq=10000)
    d.update(kw)
    return d


def pulling_params(n, k_trans=0.05, speed_A_per_step=None, **kw):
    d = c0_params(**kw)
    d.update(pulling=1, fixedEnd=0, pulledEnd=n - 1, k_trans=k_trans, pullFreq=d.get("nav", 1000),
             deltax=d.get("deltax", 0.0) or 0.05)
    return d
