"""GPU dirty beam and blurred covariance operator (RadioDataset,
data_base.py:115-129,148-172) against the oracle's restatement and the
reference's own bars (nufft_tests.py:70-118):

* the dirty beam equals the oracle's to 1e-12 and its peak is the weight sum
  to 1e-6;
* BCB applied to images equals the oracle's BCB to 1e-12;
* at the reference's default kernel support 4, BCB equals the explicit
  N^T C N and the DFT's D^T C D to the reference's 1e-6 (nufft_tests.py:93-97);
* the solver's L^T BCB L equals the oracle's composition to 1e-12;
* RadioDataset on the reference's own uvfits file: the beam-peak invariant
  and BCB against N^T C N."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import annulus_uv, rng

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
TINY = os.path.join(HERE, 'golden', 'j0751_lwmp_nopos_nosub_0.000000_tiny.uvfits')


def _setup(n=48, nv=20000, seed=31):
    from lentils_amd import ImageSpace, VisibilitySpace
    im = ImageSpace(shape=(n, n), bounds=[(-0.64, 0.64), (-0.6, 0.68)])
    uv = annulus_uv(nv, r2=1.5e6, seed=seed)
    w = rng(seed + 1).uniform(0.5e6, 1.5e6, nv)
    return im, VisibilitySpace(channels=[1.0], uvcoords=uv), uv, w


def test_dirty_beam_matches_oracle(hip, oracle):
    from lentils_amd.datasets import dirty_beam
    O = oracle
    im, vs, uv, w = _setup()
    og = O.Grid2D(im.nx, im.ny, im.bounds)
    for wsup, beta in ((4, None), (3, 14.04)):
        beam, peak = dirty_beam(vs, im, w, kernel_support=wsup, kb_beta=beta)
        ref = O.dirty_beam(uv, [1.0], og, w, wsup=wsup, beta=beta, nthreads=8)
        err = O.max_relative_error(beam, ref)
        print(f'W={wsup}: beam vs oracle {err:.2e}, peak/sum(w) - 1 = {peak / w.sum() - 1:.2e}')
        assert err < 1e-12
        assert peak == beam.max()
        assert abs(peak / w.sum() - 1.0) < 1e-6
        assert np.unravel_index(np.argmax(beam), beam.shape) == (im.nx, im.ny)


def test_bcb_apply_matches_oracle(hip, oracle):
    from lentils_amd.datasets import BlurredCovarianceOperator, dirty_beam
    O = oracle
    im, vs, uv, w = _setup()
    og = O.Grid2D(im.nx, im.ny, im.bounds)
    geo = O.NUFFTGeometry(og, pad=2, wsup=4)
    beam, _ = dirty_beam(vs, im, w)
    op = BlurredCovarianceOperator(im, beam)
    xs = rng(33).normal(size=(3, im.nx, im.ny))
    for x in xs:
        y = op * x
        ref = O.bcb_apply(x, beam, geo)
        assert O.max_relative_error(y.reshape(im.nx, im.ny), ref) < 1e-12
    assert op.T is op


def test_bcb_reference_bar(hip, oracle):
    """nufft_tests.py:70-97 at its default kernel support 4: BCB, the
    explicit N^T C N and the DFT's D^T C D agree to 1e-6 on a masked image."""
    from lentils_amd import DFTOperator, NUFFTOperator
    from lentils_amd.datasets import BlurredCovarianceOperator, dirty_beam
    O = oracle
    im, vs, uv, w = _setup(n=32, nv=4000, seed=35)
    mask = np.zeros((im.nx, im.ny), dtype=bool)
    mask[4:28, 6:30] = True
    im.mask = mask
    im._refresh()
    x = rng(36).normal(size=(im.nx, im.ny)) * mask
    nufft = NUFFTOperator(vs, im)  # kernel_support 4, Beatty's beta
    beam, _ = dirty_beam(vs, im, w)
    bcb = BlurredCovarianceOperator(im, beam)
    explicit = nufft.T * (w * (nufft * x).reshape(-1)).reshape(vs.shape)
    convolved = bcb * x
    D = DFTOperator(vs, im)._mat  # two real rows per visibility, masked pixels only
    by_dft = D.T @ (np.repeat(w, 2) * (D @ x.reshape(-1)))
    e1 = O.max_relative_error(convolved, explicit)
    e2 = O.max_relative_error(convolved.reshape(-1) * mask.reshape(-1), by_dft)
    e3 = O.max_relative_error(explicit.reshape(-1) * mask.reshape(-1), by_dft)
    print(f'BCB vs N^T C N {e1:.2e}, BCB vs DFT {e2:.2e}, N^T C N vs DFT {e3:.2e}')
    assert e1 < 1e-6 and e2 < 1e-6 and e3 < 1e-6


def test_solver_bcb_matches_oracle(hip, oracle):
    """The solver's BCB operator L^T BCB L x (+ curvature) against the
    oracle's composition with the same lens matrix, for a plan whose own
    kernel support (3) differs from the beam's reference default (4)."""
    from lentils_amd import ExternalPotential, GlobalLensModel, PowerLawEllipsoid
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    from lentils_amd import ImageSpace
    O = oracle
    im, vs, uv, w = _setup(n=64, nv=30000, seed=37)
    src = ImageSpace(shape=(32, 32), bounds=[(-0.32, 0.32), (-0.3, 0.34)])
    models = [GlobalLensModel([PowerLawEllipsoid(b=0.3 + 0.05 * k, f=0.8, th=30.0 * k, qh=0.5, x=0.011, y=0.03,
                                                 rc=0.0, z=0.5),
                               ExternalPotential(ss=0.02, sa=40.0, x=0.011, y=0.03, z=0.5)]) for k in range(2)]
    lb = LensBatch(im, [src] * 2, models, z_src=2.0)
    plan = NUFFTPlan(vs, im, 2, 3, 14.04)
    data = (rng(38).normal(size=(uv.shape[0], 2)) * 1e-3).view(np.complex128).reshape(-1)
    inv = BatchedSourceInversion(lb, plan, data, weights=w, operator='bcb', lambda_curvature=1e2)
    xs = [rng(39 + k).normal(size=src.size) for k in range(2)]
    ys = inv.apply(xs)
    og = O.Grid2D(im.nx, im.ny, im.bounds)
    geo4 = O.NUFFTGeometry(og, pad=2, wsup=4)
    beam = O.dirty_beam(uv, [1.0], og, w, nthreads=8)
    crp, cci, cva = O.curv_matrix(O.Grid2D(src.nx, src.ny, src.bounds))
    Cm = sp.csr_matrix((cva, cci, crp), shape=(src.size, src.size))
    for k in range(2):
        rp, ci, va = lb.csr(k)
        L = sp.csr_matrix((va, ci, rp), shape=(im.size, src.size))
        ref = L.T @ O.bcb_apply(L @ xs[k], beam, geo4).reshape(-1) + 1e2 * (Cm.T @ (Cm @ xs[k]))
        err = O.max_relative_error(ys[k].reshape(-1), ref)
        print(f'sample {k}: solver BCB vs oracle {err:.2e}')
        assert err < 1e-12


def test_radio_dataset_reference_file(hip, oracle):
    """RadioDataset on the reference's j0751 tiny uvfits file: the beam peak
    is the weight sum (nufft_tests.py:116-118), the dirty image is the NUFFT
    adjoint of C d, and BCB matches N^T C N to 1e-6."""
    from lentils_amd import ImageSpace, RadioDataset
    O = oracle
    nx = 32
    dx = 1.5 / 1024
    ext = 0.5 * dx * nx
    im = ImageSpace(shape=(nx, nx), bounds=[(-ext, ext), (-ext, ext)])
    ds = RadioDataset(TINY, image_space=im)
    cvec = ds.covariance_operator.T * ds.space.new_vector(1.0 + 0.0j)
    wtsum = np.sum(cvec).real
    assert abs(np.max(ds.dirty_beam) / wtsum - 1.0) < 1e-6
    dimg = ds.dirty_image
    ref = ds.nufft_operator.T * (ds.weights * ds.data)
    assert O.max_relative_error(dimg, ref) < 1e-12
    x = rng(41).normal(size=im.shape)
    explicit = ds.nufft_operator.T * ds.covariance_operator * (ds.nufft_operator * x)
    assert O.max_relative_error(ds.blurred_covariance_operator * x, explicit) < 1e-6
