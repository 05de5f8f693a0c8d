"""The learned preconditioner inside the batched source inversion.

* an untrained network (zero last layer) is exactly the lens-structured
  preconditioner M: same iterations, same solutions;
* a network whose parameters are NaN is caught by the guard every
  iteration, and flexible CG still reaches the FP64 true residual with M."""
import numpy as np
import pytest

from test_gpu_solver import RESID_BAR, _oracle_apply, _problem

pytestmark = pytest.mark.gpu


def _random_models(n, seed):
    from lentils_amd import ExternalPotential, GlobalLensModel, PowerLawEllipsoid
    g = np.random.default_rng(seed)
    return [GlobalLensModel([PowerLawEllipsoid(b=g.uniform(0.2, 0.4), f=g.uniform(0.6, 0.95), th=g.uniform(0, 180),
                                               qh=g.uniform(0.4, 0.6), x=0.011, y=-0.007, rc=0.0, z=0.5),
                             ExternalPotential(ss=g.uniform(0, 0.05), sa=g.uniform(0, 180), x=0.011, y=-0.007,
                                               z=0.5)])
            for _ in range(n)]


def _inversion(im, src, vs, uv, data, models):
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    lb = LensBatch(im, [src] * len(models), models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    inv.set_graphs(False)
    return inv, lb, w


def test_untrained_network_is_lens_preconditioner(hip):
    from lentils_amd.learned import PCNet
    im, src, uv, vs, _, data = _problem(nlens=4)
    inv, _, _ = _inversion(im, src, vs, uv, data, _random_models(4, 5))
    xl, itl, rl = inv.solve(rtol=1e-6, maxiter=5000, precond='lens')
    net = PCNet(src.nx, src.ny, batch=4, seed=1)
    inv.set_preconditioner(net)
    xn, itn, rn = inv.solve(rtol=1e-6, maxiter=5000, precond='learned')
    print(f'lens iterations {itl}, untrained learned {itn}')
    assert np.array_equal(itl, itn)
    for a, b in zip(xl, xn):
        assert np.array_equal(a, b)
    # M's direction is a descent direction throughout (a sample may still
    # switch its zero network off at a slow stretch: that changes nothing)
    c, sw = inv.fallbacks()
    print(f'guard fallbacks {c}, switched {sw}')
    assert np.all(c == 0)


def test_nan_network_falls_back_to_lens_direction(hip):
    from lentils_amd.learned import PCNet
    im, src, uv, vs, _, data = _problem(nlens=3)
    inv, _, _ = _inversion(im, src, vs, uv, data, _random_models(3, 6))
    xl, itl, rl = inv.solve(rtol=1e-6, maxiter=5000, precond='lens')
    net = PCNet(src.nx, src.ny, batch=3, params=np.full(PCNet.param_count(), np.nan, dtype=np.float32))
    inv.set_preconditioner(net)
    xn, itn, rn = inv.solve(rtol=1e-6, maxiter=5000, precond='learned')
    c, sw = inv.fallbacks()
    print(f'lens iterations {itl}; NaN network {itn}, guard fallbacks {c}, switched {sw}')
    assert np.all(rn <= 1e-6)
    # the guard takes m in every iteration (and every restart)
    assert np.all(c >= itn)
    assert np.array_equal(itl, itn)


def test_trained_network_against_plain_cg_and_m(hip, oracle):
    """Trained on the GPU on the solver's own vectors (right-hand sides of
    blob sources and residuals of truncated M-preconditioned solves) over
    random lenses, the network lowers its training loss below M's own (zero
    network); on held-out lenses the learned flexible CG reaches the FP64
    true residual (confirmed by the oracle's operator) in fewer iterations
    than plain CG.  Its iterations against M alone are printed: on these
    problems M takes fewer (the correction's BF16 error costs more in the
    last decades of the residual than it gains in the first)."""
    from lentils_amd.learned import PCNet, blob_sources, solver_residuals
    O = oracle
    nl = 8
    im, src, uv, vs, _, data = _problem(nlens=nl)
    net = PCNet(src.nx, src.ny, batch=nl, seed=2)
    g = np.random.default_rng(31)
    first = last = None
    for tb in range(3):
        inv, _, _ = _inversion(im, src, vs, uv, data, _random_models(nl, 100 + tb))
        xs = blob_sources(src, nl * 8, g, width=(0.02, 0.08), spread=0.15)
        draws = [list(xs[d * nl:(d + 1) * nl]) for d in range(8)]
        vecs = solver_residuals(inv, draws, iterations=(0, 2, 6, 12), precond='lens')
        if tb == 0:
            first = inv.train_preconditioner(PCNet(src.nx, src.ny, batch=nl, seed=2), vecs, epochs=0)[0]
        h = inv.train_preconditioner(net, vecs, epochs=4, lr=2e-3, lr_final=2e-4, seed=tb)
        last = inv.train_preconditioner(net, vecs, epochs=0)[0]
        print(f'training batch {tb}: loss per epoch {np.round(h, 4)}; after: {last:.4f}')
    print(f'loss of M alone on the first batch {first:.4f}, trained network on the last {last:.4f}')
    assert last < first
    models = _random_models(nl, 7)
    inv, lb, w = _inversion(im, src, vs, uv, data, models)
    x0, it0, r0 = inv.solve(rtol=1e-6, maxiter=5000)
    xl, itl, rl = inv.solve(rtol=1e-6, maxiter=5000, precond='lens')
    inv.set_preconditioner(net)
    inv.set_flexible_depth(4)
    xn, itn, rn = inv.solve(rtol=1e-6, maxiter=5000, precond='learned')
    c, sw = inv.fallbacks()
    print(f'held-out lenses: plain CG {it0} (median {np.median(it0)}), M alone {itl} (median {np.median(itl)}), '
          f'learned {itn} (median {np.median(itn)}); guard fallbacks {c}, switched {sw}')
    assert np.all(rn <= 1e-6)
    assert np.median(itn) < np.median(it0)
    for k in (0, nl - 1):
        x = xn[k].reshape(-1)
        Ax, (L, Cm, geo) = _oracle_apply(O, im, src, uv, lb.csr(k), x, w, 1e2)
        b = L.T @ O.nufft_adjoint((data * w).reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
        assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= RESID_BAR
