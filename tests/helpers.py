"""Seeded synthetic inputs shared by the tests (one generator per quantity:
lens 11, uv 12, noise 13, source 14)."""
import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def annulus_uv(n, r1=2e4, r2=2e6, seed=12):
    g = rng(seed)
    r = np.sqrt(g.uniform(r1 ** 2, r2 ** 2, n))
    ph = g.uniform(0, 2 * np.pi, n)
    return np.stack([r * np.cos(ph), r * np.sin(ph), np.zeros(n)], 1)


def reference_test_uv(seed=5):
    """80 uv points at scales 0..55e5 like nufft_tests.py:47-52."""
    g = rng(seed)
    uv = [g.normal(scale=1e5 * s, size=(4, 3)) for s in np.linspace(0.0, 55.0, 20)]
    return np.array(uv).reshape(-1, 3)


def c1_lens(O):
    return [O.pack_pemd(b=1.0, q=0.8, th=30.0, qh=0.5, x=0.02, y=-0.01, z=0.5),
            O.pack_shear(ss=0.03, sa=45.0, x=0.02, y=-0.01, z=0.5)]


def random_lenses(O, n, seed=11):
    """c2's lens law: b U(0.5,1.5), q U(0.5,0.95), gamma U(1.8,2.2), theta
    U(0,180), centre U(-0.05,0.05), shear U(0,0.05) at U(0,180)."""
    g = rng(seed)
    out = []
    for _ in range(n):
        b = g.uniform(0.5, 1.5)
        q = g.uniform(0.5, 0.95)
        gam = g.uniform(1.8, 2.2)
        th = g.uniform(0, 180)
        x, y = g.uniform(-0.05, 0.05, 2)
        ss, sa = g.uniform(0, 0.05), g.uniform(0, 180)
        out.append([O.pack_pemd(b=b, q=q, th=th, qh=(gam - 1) / 2, x=x, y=y),
                    O.pack_shear(ss=ss, sa=sa, x=x, y=y)])
    return out
