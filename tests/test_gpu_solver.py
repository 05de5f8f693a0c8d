"""GPU source inversion: operator parity with the oracle and CG convergence.

* the solver's A x (lens matrix, exact N^T C N, curvature prior) equals the
  oracle's composition of the same operators (max-relative 1e-12);
* the precomputed grid stencil and the per-visibility explicit pass agree;
* CG reaches the FP64 true relative residual it reports, and the oracle's
  operator confirms it.  The BCB operator's tests are in test_gpu_bcb.py."""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import annulus_uv, rng

pytestmark = pytest.mark.gpu

# the oracle's FP64 operator must confirm the 1e-6 true relative residual the
# GPU reports; the two operators agree to ~1e-12 of |A x|, which moves a
# residual of 1e-6 |b| by ~1e-6 of itself
RESID_BAR = 1e-6 * (1 + 1e-6)


def _problem(nv=30000, n_img=64, n_src=32, nlens=3, seed=71):
    from lentils_amd import ExternalPotential, GlobalLensModel, ImageSpace, PowerLawEllipsoid, VisibilitySpace
    g = rng(seed)
    im = ImageSpace(shape=(n_img, n_img), bounds=[(-0.64, 0.64), (-0.64, 0.64)])
    src = ImageSpace(shape=(n_src, n_src), bounds=[(-0.32, 0.32), (-0.32, 0.32)])
    uv = annulus_uv(nv, r2=1.5e6, seed=seed)
    vs = VisibilitySpace(channels=[1.0], uvcoords=uv)
    models = [GlobalLensModel([PowerLawEllipsoid(b=g.uniform(0.2, 0.4), f=g.uniform(0.6, 0.95), th=g.uniform(0, 180),
                                                 qh=g.uniform(0.45, 0.55), x=0.011, y=-0.007, rc=0.0, z=0.5),
                               ExternalPotential(ss=0.02, sa=g.uniform(0, 180), x=0.011, y=-0.007, z=0.5)])
              for _ in range(nlens)]
    data = (g.normal(size=(nv, 2)) * 1e-3).view(np.complex128).reshape(-1)
    return im, src, uv, vs, models, data


def _oracle_apply(O, im, src, uv, csr, x, w, lam):
    """Oracle composition L^T N^T C N L + lam Cc^T Cc with the GPU's own lens
    matrix (lens-matrix parity is tested separately; the reduction quirk
    makes rare entries depend on last-bit deflection differences)."""
    og = O.Grid2D(im.nx, im.ny, im.bounds)
    osr = O.Grid2D(src.nx, src.ny, src.bounds)
    rp, ci, va = csr
    L = sp.csr_matrix((va, ci, rp), shape=(im.size, src.size))
    crp, cci, cva = O.curv_matrix(osr)
    Cm = sp.csr_matrix((cva, cci, crp), shape=(src.size, src.size))
    geo = O.NUFFTGeometry(og, pad=2, wsup=3, beta=14.04)
    vis = O.nufft_forward((L @ x).reshape(im.nx, im.ny), uv, [1.0], geo, nthreads=8).reshape(-1) * w
    back = O.nufft_adjoint(vis.reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
    return L.T @ back + lam * (Cm.T @ (Cm @ x)), (L, Cm, geo)


@pytest.mark.parametrize('operator', ['nufft', 'nufft_explicit'])
def test_solver_operator_matches_oracle(hip, oracle, operator):
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem()
    lb = LensBatch(im, [src] * len(models), models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2, operator=operator)
    xs = [rng(72 + k).normal(size=src.size) for k in range(len(models))]
    ys = inv.apply(xs)
    for k, m in enumerate(models):
        ref, _ = _oracle_apply(O, im, src, uv, lb.csr(k), xs[k], w, 1e2)
        assert O.max_relative_error(ys[k].reshape(-1), ref) < 1e-12


def test_operator_rows_of_1024(hip, oracle):
    """512^2 image (padded rows of 1024, the register-fed row transforms):
    the FP64 operator against the oracle's, the FP32 operator against the
    FP64 one."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem(n_img=512, nlens=2)
    lb = LensBatch(im, [src] * len(models), models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    xs = [rng(76 + k).normal(size=src.size) for k in range(len(models))]
    ys = inv.apply(xs)
    y32 = inv.apply(xs, precision='f32')
    for k in range(len(models)):
        ref, _ = _oracle_apply(O, im, src, uv, lb.csr(k), xs[k], w, 1e2)
        err = O.max_relative_error(ys[k].reshape(-1), ref)
        err32 = O.max_relative_error(y32[k].reshape(-1), ys[k].reshape(-1))
        print(f'sample {k}: FP64 vs oracle {err:.2e}, FP32 vs FP64 {err32:.2e}')
        assert err < 1e-12
        assert err32 < 1e-5


def test_row_forms_of_1024(hip, oracle):
    """Rows padded to 1024: the register-fed row transforms give the
    256-thread LDS form's bits in FP64 and in FP32; both precisions against
    the oracle's operator."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem(n_img=512, nlens=3)
    lb = LensBatch(im, [src] * len(models), models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    xs = [rng(81 + k).normal(size=src.size) for k in range(len(models))]
    out = {}
    for form in (0, 1):
        plan.set_row_form(form)
        inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
        out[form] = (np.stack([y.reshape(-1) for y in inv.apply(xs)]),
                     np.stack([y.reshape(-1) for y in inv.apply(xs, precision='f32')]))
    plan.set_row_form(1)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    for k in range(len(models)):
        ref, _ = _oracle_apply(O, im, src, uv, lb.csr(k), xs[k], w, 1e2)
        err = O.max_relative_error(out[1][0][k], ref)
        e32 = O.max_relative_error(out[1][1][k], ref)
        print(f'sample {k}: FP64 vs oracle {err:.2e}, FP32 vs oracle {e32:.2e}')
        assert err < 1e-12
        assert e32 < 1e-5


def test_column_forms_of_1024(hip, oracle):
    """Padded grid of 1024 x 1024: the own column transforms against rocFFT's
    (rounding apart), both precisions against the oracle's operator."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem(n_img=512, nlens=3, seed=83)
    lb = LensBatch(im, [src] * len(models), models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    xs = [rng(84 + k).normal(size=src.size) for k in range(len(models))]
    out = {}
    for form in (0, 1):
        plan.set_col_form(form)
        inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
        out[form] = (np.stack([y.reshape(-1) for y in inv.apply(xs)]),
                     np.stack([y.reshape(-1) for y in inv.apply(xs, precision='f32')]))
    plan.set_col_form(1)
    for k in range(len(models)):
        ref, _ = _oracle_apply(O, im, src, uv, lb.csr(k), xs[k], w, 1e2)
        d64 = O.max_relative_error(out[1][0][k], out[0][0][k])
        err = O.max_relative_error(out[1][0][k], ref)
        e32 = O.max_relative_error(out[1][1][k], ref)
        print(f'sample {k}: FP64 own vs rocFFT columns {d64:.2e}, own vs oracle {err:.2e}, FP32 own vs oracle '
              f'{e32:.2e}')
        assert d64 < 1e-13
        assert err < 1e-12
        assert e32 < 1e-5


def test_stencil_shapes_same_bits(hip, oracle):
    """Every samples-per-workgroup choice of the normal stencil gives the
    same operator bits; one of them against the oracle's operator."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem()
    lb = LensBatch(im, [src] * len(models), models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    xs = [rng(74 + k).normal(size=src.size) for k in range(len(models))]
    base = None
    for nb in (1, 2, 4, 8):
        plan.set_stencil_shape(nb)
        inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
        ys = np.stack([y.reshape(-1) for y in inv.apply(xs)])
        if base is None:
            base = ys
            ref, _ = _oracle_apply(O, im, src, uv, lb.csr(0), xs[0], w, 1e2)
            assert O.max_relative_error(ys[0], ref) < 1e-12
        assert np.array_equal(ys, base), nb
    plan.set_stencil_shape(0)


def test_stencil_matrix_core_form(hip, oracle):
    """The FP64 matrix-core form of the normal stencil (v_mfma_f64_16x16x4f64
    over 4 x 4 tiles of rows) equals the vector form to rounding and the
    oracle's operator to 1e-12, gives the same bits for every number of
    sample groups per workgroup (19 samples: groups of 8, 8 and 3), and a CG
    solve with it (converged samples leaving the launches) reaches the FP64
    true residual that the oracle's operator confirms."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    n = 19
    im, src, uv, vs, models, data = _problem(nlens=n, seed=77)
    lb = LensBatch(im, [src] * n, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    xs = [rng(90 + k).normal(size=src.size) for k in range(n)]
    plan.set_stencil_shape(2)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    yv = np.stack([y.reshape(-1) for y in inv.apply(xs)])
    base = None
    for g in (-1, -2, -4, -8):
        plan.set_stencil_shape(g)
        inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
        ym = np.stack([y.reshape(-1) for y in inv.apply(xs)])
        err = O.max_relative_error(ym, yv)
        print(f'matrix-core form, {-g} groups per workgroup: max relative difference to the vector form {err:.2e}')
        assert err < 1e-13
        if base is None:
            base = ym
            for k in (0, n - 1):
                ref, _ = _oracle_apply(O, im, src, uv, lb.csr(k), xs[k], w, 1e2)
                assert O.max_relative_error(ym[k], ref) < 1e-12
        assert np.array_equal(ym, base), g
    plan.set_stencil_shape(-2)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    x, its, res = inv.solve(rtol=1e-6, maxiter=20000)
    print(f'CG with the matrix-core stencil: iterations {its}')
    assert np.all(res <= 1e-6)
    k = int(np.argmax(its))
    Ax, (L, Cm, geo) = _oracle_apply(O, im, src, uv, lb.csr(k), x[k].reshape(-1), w, 1e2)
    b = L.T @ O.nufft_adjoint((data * w).reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
    assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= RESID_BAR
    plan.set_stencil_shape(0)


def test_stencil_matrix_core_form_f32(hip, oracle):
    """The precision form's FP32 stencil on the matrix cores
    (v_mfma_f32_16x16x4f32, its own result layout): the FP32 operator with it
    agrees with the FP32 vector form and with the FP64 operator to FP32
    rounding for every number of sample groups per workgroup (19 samples:
    groups of 8, 8 and 3), and an FP32-operator CG solve with it reaches the
    FP64 true residual that the oracle's operator confirms."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    n = 19
    im, src, uv, vs, models, data = _problem(nlens=n, seed=79)
    lb = LensBatch(im, [src] * n, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    xs = [rng(110 + k).normal(size=src.size) for k in range(n)]
    plan.set_stencil_shape(4)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    y64 = np.stack([y.reshape(-1) for y in inv.apply(xs)])
    yv32 = np.stack([y.reshape(-1) for y in inv.apply(xs, precision='f32')])
    assert O.max_relative_error(yv32, y64) < 1e-5
    base = None
    for g in (-1, -2, -4):
        plan.set_stencil_shape(g)
        inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
        ym = np.stack([y.reshape(-1) for y in inv.apply(xs, precision='f32')])
        e64, ev = O.max_relative_error(ym, y64), O.max_relative_error(ym, yv32)
        print(f'FP32 matrix-core form, {-g} groups per workgroup: vs FP64 {e64:.2e}, vs FP32 vector form {ev:.2e}')
        assert e64 < 1e-5 and ev < 1e-5
        if base is None:
            base = ym
        assert np.array_equal(ym, base), g
    plan.set_stencil_shape(-2)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    x, its, res = inv.solve(rtol=1e-6, maxiter=20000, precision='f32', refine_every=20)
    print(f'FP32-operator CG with the matrix-core stencil: iterations {its}')
    assert np.all(res <= 1e-6)
    k = int(np.argmax(its))
    Ax, (L, Cm, geo) = _oracle_apply(O, im, src, uv, lb.csr(k), x[k].reshape(-1), w, 1e2)
    b = L.T @ O.nufft_adjoint((data * w).reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
    assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= RESID_BAR
    plan.set_stencil_shape(0)


def test_cg_converges_and_oracle_confirms(hip, oracle):
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem(nlens=2)
    lb = LensBatch(im, [src] * 2, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    xs, its, res = inv.solve(rtol=1e-6, maxiter=20000)
    assert np.all(res <= 1e-6), (its, res)
    for k, m in enumerate(models):
        x = xs[k].reshape(-1)
        Ax, (L, Cm, geo) = _oracle_apply(O, im, src, uv, lb.csr(k), x, w, 1e2)
        b = L.T @ O.nufft_adjoint((data * w).reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
        assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= RESID_BAR


def test_many_samples_cg(hip, oracle):
    """40 samples (one-workgroup-per-sample dots, sample-grid products as
    samples converge): every sample reaches the FP64 true residual and the
    oracle's operator confirms it for three of them."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    n = 40
    im, src, uv, vs, models, data = _problem(nlens=n, seed=77)
    lb = LensBatch(im, [src] * n, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    xs, its, res = inv.solve(rtol=1e-6, maxiter=20000)
    assert np.all(res <= 1e-6), (its, res)
    assert its.min() < its.max()  # samples leave the active set at different iterations
    for k in (0, 17, n - 1):
        x = xs[k].reshape(-1)
        Ax, (L, Cm, geo) = _oracle_apply(O, im, src, uv, lb.csr(k), x, w, 1e2)
        b = L.T @ O.nufft_adjoint((data * w).reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
        assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= RESID_BAR


def test_per_sample_dirty_images(hip, oracle):
    """Each sample with its own observation: rhs_s = L_s^T D^T C d_s, given
    through per-sample dirty images; a single dirty image equals passing the
    data themselves."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem(nlens=3)
    lb = LensBatch(im, [src] * 3, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    og = O.Grid2D(im.nx, im.ny, im.bounds)
    geo = O.NUFFTGeometry(og, pad=2, wsup=3, beta=14.04)
    datas = [data * (k + 1) + (rng(80 + k).normal(size=(uv.shape[0], 2)) * 1e-3).view(np.complex128).reshape(-1)
             for k in range(3)]
    dirty = [plan.adjoint(w * d) for d in datas]
    for k in range(3):
        ref = O.nufft_adjoint((w * datas[k]).reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
        assert O.max_relative_error(dirty[k].reshape(-1), ref) < 1e-12
    inv = BatchedSourceInversion(lb, plan, weights=w, dirty_images=dirty, lambda_curvature=1e2)
    rhs = inv.rhs()
    for k in range(3):
        rp, ci, va = lb.csr(k)
        L = sp.csr_matrix((va, ci, rp), shape=(im.size, src.size))
        assert O.max_relative_error(rhs[k].reshape(-1), L.T @ dirty[k].reshape(-1)) < 1e-12
    one = BatchedSourceInversion(lb, plan, weights=w, dirty_images=[dirty[0]], lambda_curvature=1e2).rhs()
    direct = BatchedSourceInversion(lb, plan, datas[0], weights=w, lambda_curvature=1e2).rhs()
    for k in range(3):
        assert O.max_relative_error(one[k].reshape(-1), direct[k].reshape(-1)) < 1e-12
    with pytest.raises(ValueError):
        BatchedSourceInversion(lb, plan, weights=w, dirty_images=dirty[:2])


def test_resident_inputs(hip, oracle):
    """Weights resident in the plan (with the normal stencil built from them
    once) and dirty images in device memory give the same right-hand sides,
    operator and solutions as the host-input solver; a second solver on the
    same plan reuses them; device-input misuse is rejected."""
    from lentils_amd.solvers import BatchedSourceInversion, DeviceArray, LensBatch, NUFFTPlan
    im, src, uv, vs, models, data = _problem(nlens=3)
    lb = LensBatch(im, [src] * 3, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = rng(90).uniform(0.5e6, 1.5e6, size=uv.shape[0])
    datas = [data * (k + 1) for k in range(3)]
    dirty = [plan.adjoint(w * d) for d in datas]
    host = BatchedSourceInversion(lb, plan, weights=w, dirty_images=dirty, lambda_curvature=1e2)
    with pytest.raises(ValueError):
        BatchedSourceInversion(lb, plan, dirty_images=dirty, lambda_curvature=1e2)  # no weights anywhere
    plan.set_weights(w)
    dd = DeviceArray(np.stack(dirty))
    assert np.array_equal(dd.numpy(), np.stack(dirty))
    res = BatchedSourceInversion(lb, plan, dirty_images=dd, lambda_curvature=1e2)
    for a, b in zip(host.rhs(), res.rhs()):
        assert np.array_equal(a, b)
    g = rng(91)
    xs = [g.normal(size=src.shape) for _ in range(3)]
    for a, b in zip(host.apply(xs), res.apply(xs)):
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a))
    x0, it0, r0 = host.solve(rtol=1e-6, maxiter=20000)
    x1, it1, r1 = res.solve(rtol=1e-6, maxiter=20000)
    assert np.all(r1 <= 1e-6) and np.all(np.abs(np.asarray(it0) - np.asarray(it1)) <= 2)
    for a, b in zip(x0, x1):
        assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(a))
    again = BatchedSourceInversion(lb, plan, dirty_images=dd, lambda_curvature=1e2)
    x2, it2, r2 = again.solve(rtol=1e-6, maxiter=20000)
    assert np.array_equal(it1, it2)
    for a, b in zip(x1, x2):
        assert np.array_equal(a, b)  # run-to-run identical
    with pytest.raises(ValueError):
        BatchedSourceInversion(lb, plan, weights=w, dirty_images=dd, lambda_curvature=1e2)
    with pytest.raises(ValueError):
        BatchedSourceInversion(lb, plan, dirty_images=DeviceArray(np.stack(dirty[:2])), lambda_curvature=1e2)


@pytest.mark.parametrize('precond', ['none', 'jacobi', 'lens'])
def test_graph_blocks_match_launches(hip, precond):
    """The FP64 CG with its iteration blocks replayed as HIP graphs (many
    replays per solve, across launch widths as samples converge) gives the
    same solutions and iteration counts, bit for bit, as one-by-one launches,
    and again on a second solve with the same solver."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    im, src, uv, vs, models, data = _problem(nlens=11)
    lb = LensBatch(im, [src] * 11, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    dirty = [plan.adjoint(w * data * (1 + 0.3 * k)) for k in range(11)]
    inv = BatchedSourceInversion(lb, plan, weights=w, dirty_images=dirty, lambda_curvature=1e2)
    inv.set_graphs(False)
    x0, it0, r0 = inv.solve(rtol=1e-6, maxiter=20000, precond=precond)
    inv.set_graphs(True)
    for _ in range(2):
        x1, it1, r1 = inv.solve(rtol=1e-6, maxiter=20000, precond=precond)
        assert np.array_equal(np.asarray(it0), np.asarray(it1))
        assert np.array_equal(np.asarray(r0), np.asarray(r1))
        for a, b in zip(x0, x1):
            assert np.array_equal(a, b)
    assert np.all(np.asarray(r1) <= 1e-6)


def test_precision_form(hip, oracle):
    """FP32 operator inside FP64 CG with true-residual refresh: the FP32
    operator agrees with the FP64 one to FP32 rounding, every sample still
    reaches an FP64 true residual of 1e-6 (checked with the oracle's
    operator), and the solutions agree with the FP64 solve."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem(nlens=3)
    lb = LensBatch(im, [src] * 3, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    xs = [rng(90 + k).normal(size=src.size) for k in range(3)]
    y64 = inv.apply(xs)
    y32 = inv.apply(xs, precision='f32')
    for k in range(3):
        err = O.max_relative_error(y32[k].reshape(-1), y64[k].reshape(-1))
        print(f'sample {k}: FP32 operator max-relative error {err:.2e}')
        assert err < 1e-5
    # the FP32 operator is scale-free: inputs of 1e15 (a first CG direction
    # with large weights) must not overflow
    big = [v * 1e15 for v in xs]
    yb32, yb64 = inv.apply(big, precision='f32'), inv.apply(big)
    for k in range(3):
        assert O.max_relative_error(yb32[k].reshape(-1), yb64[k].reshape(-1)) < 1e-5
    x64, it64, r64 = inv.solve(rtol=1e-6, maxiter=20000)
    x32, it32, r32 = inv.solve(rtol=1e-6, maxiter=20000, precision='f32', refine_every=20)
    print(f'iterations FP64 {it64}, FP32 form {it32}; true residuals {r32}')
    assert np.all(r32 <= 1e-6)
    for k, m in enumerate(models):
        x = x32[k].reshape(-1)
        Ax, (L, Cm, geo) = _oracle_apply(O, im, src, uv, lb.csr(k), x, w, 1e2)
        b = L.T @ O.nufft_adjoint((data * w).reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
        assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= RESID_BAR
        d = np.linalg.norm(x - x64[k].reshape(-1)) / np.linalg.norm(x64[k])
        print(f'sample {k}: |x32 - x64| / |x64| = {d:.2e}')


def test_lens_preconditioner(hip, oracle):
    """Flexible CG with the lens-structured preconditioner (inner CG steps on
    lambda Cc^T Cc + c L_s^T L_s) reaches the FP64 true residual in fewer
    iterations than plain CG; the oracle's operator confirms the solutions."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem(nlens=3)
    lb = LensBatch(im, [src] * 3, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    x0, it0, r0 = inv.solve(rtol=1e-6, maxiter=20000)
    inv.set_lens_preconditioner(inner_steps=2)
    x1, it1, r1 = inv.solve(rtol=1e-6, maxiter=20000, precond='lens')
    print(f'plain CG iterations {it0}, lens-preconditioned {it1}')
    assert np.all(r1 <= 1e-6)
    assert np.median(it1) < np.median(it0)
    for k, m in enumerate(models):
        x = x1[k].reshape(-1)
        Ax, (L, Cm, geo) = _oracle_apply(O, im, src, uv, lb.csr(k), x, w, 1e2)
        b = L.T @ O.nufft_adjoint((data * w).reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
        assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= RESID_BAR
    # the explicit M against scipy's c L_s^T L_s + lambda Cc^T Cc, for the
    # default row limit and for one that leaves rows out of L_s
    for max_row in (256, 4):
        inv.set_lens_preconditioner(inner_steps=2, max_row=max_row)
        M, c = inv.lens_preconditioner_matrix()
        assert c > 0
        n = src.size
        for k in range(3):
            L, Cm = _oracle_apply(O, im, src, uv, lb.csr(k), np.zeros(n), w, 1e2)[1][:2]
            L = L.copy()
            L.eliminate_zeros()  # the row limit counts nonzero entries
            keep = sp.diags((np.diff(L.indptr) <= max_row).astype(np.float64))
            Ls = keep @ L
            ref = (c * (Ls.T @ Ls) + 1e2 * (Cm.T @ Cm)).toarray()
            got = M[k * n:(k + 1) * n, k * n:(k + 1) * n].toarray()
            assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), (max_row, k)
            if k == 0:
                assert M[:n, n:].nnz == 0  # block-diagonal
        if max_row == 4:
            assert np.sum(np.diff(lb.csr(0)[0]) > 4) > 0  # the limit does remove rows
    x2, it2, r2 = inv.solve(rtol=1e-6, maxiter=20000, precond='lens')
    assert np.all(r2 <= 1e-6)


def test_lens_preconditioner_precision_form(hip, oracle):
    """The lens-structured preconditioner with the FP32 operator (precision
    form) inside the flexible FP64 CG: every sample reaches the FP64 true
    residual 1e-6 that the oracle's operator confirms, in about the FP64
    lens-preconditioned iteration count."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    O = oracle
    im, src, uv, vs, models, data = _problem(nlens=3, seed=83)
    lb = LensBatch(im, [src] * 3, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    w = np.full(uv.shape[0], 1e6)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, lambda_curvature=1e2)
    x64, it64, r64 = inv.solve(rtol=1e-6, maxiter=20000, precond='lens')
    x32, it32, r32 = inv.solve(rtol=1e-6, maxiter=20000, precond='lens', precision='f32', refine_every=20)
    print(f'lens-preconditioned iterations: FP64 operator {it64}, FP32 operator {it32}; residuals {r32}')
    assert np.all(r64 <= 1e-6) and np.all(r32 <= 1e-6)
    assert np.max(it32) <= 1.5 * np.max(it64)
    for k in range(3):
        x = x32[k].reshape(-1)
        Ax, (L, Cm, geo) = _oracle_apply(O, im, src, uv, lb.csr(k), x, w, 1e2)
        b = L.T @ O.nufft_adjoint((data * w).reshape(1, -1), uv, [1.0], geo, nthreads=8).reshape(-1)
        assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= RESID_BAR
