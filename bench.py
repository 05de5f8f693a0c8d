"""Benchmark: source solves/s of the plain FP64 GPU method on the 64-lens
workload (512^2 image at 0.01"/px, 128^2 source, 1e7 visibilities, 1024 x 513
grid, KB W = 6 full width (kernel_support 3), beta = 14.04, curvature prior
lambda = 1e2, FP64 CG to a true relative residual of 1e-6).

One step = the whole hot path for one batch of 64 lens-parameter samples:
deflect the stencil lattices, rasterize the 64 CSR lens matrices (and their
transposes), form the dirty image and right-hand sides, and run the batched
CG to convergence.  Inputs (uv set, data) are resident in HBM before timing.

Prints one JSON line (see the "roofline" and "cpu_baseline" fields).
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NVIS = 10_000_000
NLENS = 64


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def gen_lenses(n, seed=11):
    from lentils_amd import ExternalPotential, GlobalLensModel, PowerLawEllipsoid
    g = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        b = g.uniform(0.5, 1.5)
        q = g.uniform(0.5, 0.95)
        gam = g.uniform(1.8, 2.2)
        th = g.uniform(0, 180)
        x, y = g.uniform(-0.05, 0.05, 2)
        ss, sa = g.uniform(0, 0.05), g.uniform(0, 180)
        out.append(GlobalLensModel([PowerLawEllipsoid(b=b, f=q, th=th, qh=(gam - 1) / 2, x=x, y=y, rc=0.0, z=0.5),
                                    ExternalPotential(ss=ss, sa=sa, x=x, y=y, z=0.5)]))
    return out


def gen_uv(n, seed=12):
    g = np.random.Generator(np.random.PCG64(seed))
    r = np.sqrt(g.uniform(2e4 ** 2, 2e6 ** 2, n))
    ph = g.uniform(0, 2 * np.pi, n)
    return np.stack([r * np.cos(ph), r * np.sin(ph), np.zeros(n)], 1)


def blob_source(src, seed=14):
    g = np.random.Generator(np.random.PCG64(seed))
    p = src.points
    img = np.zeros(p.shape[:2])
    for _ in range(4):
        c = g.uniform(-0.3, 0.3, 2)
        s = g.uniform(0.05, 0.2)
        a = g.uniform(0.5, 1.0)
        img += a * np.exp(-np.sum((p - c) ** 2, axis=-1) / (2 * s * s))
    return img


GRAPHS = True  # CG iteration blocks as captured HIP graphs (bench --no-graphs turns them off)


class Workload:
    """Inputs of one step, resident before timing.  Default: every lens
    sample has its own observation (its own Gaussian-blob source lensed by
    that sample's lens, complex64 visibilities + noise on the shared uv
    set), entering the solver through its dirty image.  shared_data: one
    observation (lensed by the first sample) for all 64 samples."""

    def __init__(self, nvis=NVIS, nlens=NLENS, shared_data=False, src_sizes=None):
        from lentils_amd import ImageSpace, VisibilitySpace
        from lentils_amd import backend as B
        from lentils_amd.solvers import DeviceArray, LensBatch, NUFFTPlan
        self.B = B
        self.shared_data = shared_data
        self.im = ImageSpace(shape=(512, 512), bounds=[(-2.56, 2.56), (-2.56, 2.56)])
        self.src = ImageSpace(shape=(128, 128), bounds=[(-0.64, 0.64), (-0.64, 0.64)])
        # per-sample source grids (all 128^2 unless uneven sizes are given)
        self.srcs = [self.src if n == 128 else ImageSpace(shape=(n, n), bounds=[(-0.64, 0.64), (-0.64, 0.64)])
                     for n in (src_sizes if src_sizes is not None else [128] * nlens)]
        self.nvis = nvis
        self.uvw = gen_uv(nvis)
        self.vs = VisibilitySpace(channels=[1.0], uvcoords=self.uvw)
        self.lenses = gen_lenses(nlens)
        self.kernel_support, self.kb_beta = 3, 14.04
        t = time.perf_counter()
        self.plan = NUFFTPlan(self.vs, self.im, pad_factor=2, kernel_support=3, kb_beta=14.04, precision='f64')
        log(f'NUFFT plan (uv binning) {time.perf_counter() - t:.2f}s')
        self.weights = np.full(nvis, 1e6)
        self.lambda_curv = 1e2
        g = np.random.Generator(np.random.PCG64(13))
        nobs = 1 if shared_data else nlens
        truth = LensBatch(self.im, self.srcs[:nobs], self.lenses[:nobs], z_src=2.0)
        self.data = None
        self.dirty = []
        for s in range(nobs):
            rp, ci, va = truth.csr(s)
            x = np.ascontiguousarray(blob_source(self.srcs[s], seed=14 + s).reshape(-1))
            img = np.zeros(self.im.size)
            B.check(B.lib().lt_csr_spmv_host(self.im.size, self.srcs[s].size, B.ptr(rp), B.ptr(ci), B.ptr(va),
                                             B.ptr(x), B.ptr(img)))
            vis = self.plan.forward(img).astype(np.complex64)  # complex64 visibilities
            vis.real += g.standard_normal(nvis, dtype=np.float32) * np.float32(1e-3)
            vis.imag += g.standard_normal(nvis, dtype=np.float32) * np.float32(1e-3)
            d = vis.astype(np.complex128)
            if shared_data:
                self.data = d
            self.dirty.append(self.plan.adjoint(self.weights * d))
        del truth
        # resident inputs: the weights (and the normal stencil built from
        # them) in the plan, the dirty images D^T C d_s in device memory
        t = time.perf_counter()
        self.plan.set_weights(self.weights)
        self.ddirty = DeviceArray(np.stack(self.dirty))
        log(f'resident weights + stencil + dirty images {time.perf_counter() - t:.2f}s')

    def step(self, rtol=1e-6, maxiter=20000, operator='nufft', precond='none', precision='f64', net=None):
        """One pass of the hot path: lens matrices of the batch, solver set-up
        and batched CG, all from device-resident inputs (net: the trained
        network of precond='learned')."""
        from lentils_amd.solvers import BatchedSourceInversion, LensBatch
        lb = LensBatch(self.im, self.srcs, self.lenses, z_src=2.0)
        if operator == 'bcb':  # the dirty-beam operator derives its beam from host weights
            inv = BatchedSourceInversion(lb, self.plan, weights=self.weights, dirty_images=self.dirty,
                                         lambda_curvature=self.lambda_curv, operator=operator)
        else:
            inv = BatchedSourceInversion(lb, self.plan, dirty_images=self.ddirty,
                                         lambda_curvature=self.lambda_curv, operator=operator)
        inv.set_graphs(GRAPHS)
        if precond == 'learned':
            inv.set_preconditioner(net)
            inv.set_flexible_depth(4)
        x, its, res = inv.solve(rtol=rtol, maxiter=maxiter, precond=precond, precision=precision)
        if precond == 'learned':
            self.last_fallbacks = inv.fallbacks()
        return its, res


def degrid_roofline(wl, reps=20):
    """FP32 degridder on the same uv set: algorithmic bytes per launch =
    16 B per visibility (u, v in FP32; complex64 out) + 8 B per grid cell
    (grid read once), over the event-timed mean launch duration."""
    from lentils_amd.solvers import NUFFTPlan
    p32 = NUFFTPlan(wl.vs, wl.im, pad_factor=2, kernel_support=3, kb_beta=14.04, precision='f32')
    ms = p32.time(0, reps)
    nu, half = 1024, 513
    bytes_ = 16.0 * NVIS + 8.0 * nu * half
    gbs = bytes_ / (ms * 1e-3) / 1e9
    ms64 = wl.plan.time(0, reps)
    return {'ms': ms, 'gbs': gbs, 'bytes': bytes_, 'gvis_s': NVIS / (ms * 1e-3) / 1e9,
            'gvis_s_f64': NVIS / (ms64 * 1e-3) / 1e9}


FP64_PEAK_TFLOPS = 78.6


def stencil_roofline(wl, nb=NLENS, reps=20):
    """The timed step's dominant kernel: the normal stencil (band form, the
    plan's resident stencil, nb samples, the autotuned form: the FP64
    matrix-core kernel k_stencil_mfma or the vector kernel k_stencil_apply),
    event-timed on the library stream.  Algorithmic work per launch
    (lt_nufft_stencil_time): 8 B per FP64 coefficient (169 per stencil row),
    16 B per sample per row written and per distinct band value read; 4 flops
    per row, offset and sample.  Its intensity (~15 flop/B) is above the FP64
    ridge (78.6 TFLOP/s / 8 TB/s), so FP64 throughput is the bound."""
    B = wl.B
    ms, by, fl = C.c_double(0), C.c_double(0), C.c_double(0)
    rows, win = C.c_long(0), C.c_long(0)
    B.check(B.lib().lt_nufft_stencil_time(wl.plan._handle, nb, reps, C.byref(ms), C.byref(by), C.byref(fl),
                                          C.byref(rows), C.byref(win)))
    form = C.c_int(0)
    B.check(B.lib().lt_nufft_stencil_form(wl.plan._handle, C.byref(form)))
    t = ms.value * 1e-3
    if form.value < 0:
        kernel = (f'k_stencil_mfma<{-form.value}> (FP64 normal stencil N^T C N on v_mfma_f64_16x16x4f64, band form, '
                  f'{nb} samples per launch)')
    else:
        kernel = f'k_stencil_apply (FP64 normal stencil N^T C N, vector form, band form, {nb} samples per launch)'
    return {'ms': ms.value, 'bytes': by.value, 'flops': fl.value, 'rows': rows.value, 'win_cells': win.value,
            'tflops': fl.value / t / 1e12, 'gbs': by.value / t / 1e9, 'form': form.value, 'kernel': kernel}


def spmv_roofline(wl, reps=20):
    """The lens-matrix products of the plain workload's solver at the full
    batch, event-timed on the library stream (lt_solver_time_products):
    L x over the nonempty rows and L^T y (with its long-row chunks).
    Algorithmic bytes per product: (4 + 8) per nonzero (int32 column, FP64
    value) + 8 per row + 8 per column (y written, x read once)."""
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch
    B = wl.B
    lb = LensBatch(wl.im, wl.srcs, wl.lenses, z_src=2.0)
    inv = BatchedSourceInversion(lb, wl.plan, dirty_images=wl.ddirty, lambda_curvature=wl.lambda_curv)
    ml, mt = C.c_double(0), C.c_double(0)
    shl, sht = (C.c_long * 3)(), (C.c_long * 3)()
    B.check(B.lib().lt_solver_time_products(inv._handle, reps, C.byref(ml), C.byref(mt), shl, sht))
    out = {}
    for name, ms, sh in (('L', ml.value, shl), ('LT', mt.value, sht)):
        nrow, ncol, nnz = sh[0], sh[1], sh[2]
        by = 12.0 * nnz + 8.0 * nrow + 8.0 * ncol
        out[name] = {'ms': ms, 'bytes': by, 'gbs': by / (ms * 1e-3) / 1e9, 'rows': nrow, 'cols': ncol, 'nnz': nnz}
    return out


def stream_copy_gbs(nbytes=1 << 30, reps=20):
    """Achievable HBM bandwidth: a plain device-to-device streaming copy."""
    from lentils_amd import backend as B
    ms = C.c_double(0)
    B.check(B.lib().lt_stream_copy(nbytes, reps, C.byref(ms)))
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value


def profile_degrid(reps=10):
    """Workload for the counter runs: the FP32 degridder of the plain GPU
    workload's uv set (1e7 visibilities, 1024 x 513 grid, support 3) launched
    `reps` times, then a 256 MiB streaming copy `reps` times (the calibration
    kernel for FETCH_SIZE / WRITE_SIZE)."""
    from lentils_amd import ImageSpace, VisibilitySpace
    from lentils_amd import backend as B
    from lentils_amd.solvers import NUFFTPlan
    B.check(B.lib().lt_init(0))
    im = ImageSpace(shape=(512, 512), bounds=[(-2.56, 2.56), (-2.56, 2.56)])
    vs = VisibilitySpace(channels=[1.0], uvcoords=gen_uv(NVIS))
    p32 = NUFFTPlan(vs, im, pad_factor=2, kernel_support=3, kb_beta=14.04, precision='f32')
    ms = p32.time(0, reps)
    nbytes = 256 << 20
    gbs, cms = stream_copy_gbs(nbytes, reps)
    print(json.dumps({'degrid_ms': ms, 'degrid_launches': reps + 1, 'copy_bytes': nbytes, 'copy_ms': cms,
                      'copy_gbs': gbs, 'copy_launches': reps + 1}), flush=True)


def profile_kernels(reps=10):
    """Workload for the counter runs (rocprofv3 --pmc): one lens batch of the
    plain GPU workload (k_deflect_lattice over 64 lenses and the raster
    kernels), the normal stencil at 64 samples (reps launches), the FP32
    degridder of the 1e7-visibility uv set (reps launches), the learned
    preconditioner's forward pass on 64 128^2 inputs (k_conv), and a 256 MiB
    streaming copy reps times (the calibration kernel for FETCH_SIZE /
    WRITE_SIZE)."""
    from lentils_amd import backend as B
    from lentils_amd.learned import PCNet
    from lentils_amd.solvers import LensBatch
    B.check(B.lib().lt_init(0))
    wl = Workload(nlens=2)
    lb = LensBatch(wl.im, [wl.src] * NLENS, gen_lenses(NLENS), z_src=2.0)
    del lb
    st = stencil_roofline(wl, reps=reps)
    rl = degrid_roofline(wl, reps=reps)
    net = PCNet(128, 128, batch=64, seed=3)
    x0 = np.random.default_rng(4).normal(size=(64, 2, 128, 128)).astype(np.float32)
    for _ in range(reps):
        net.forward(x0)
    nbytes = 256 << 20
    gbs, cms = stream_copy_gbs(nbytes, reps)
    print(json.dumps({'stencil': st, 'stencil_launches': reps + 1, 'degrid_ms': rl['ms'], 'degrid_launches': reps + 1,
                      'copy_bytes': nbytes, 'copy_ms': cms, 'copy_gbs': gbs, 'copy_launches': reps + 1}), flush=True)


def load_traffic(name, key='traffic_bytes_per_launch'):
    """HBM bytes per launch from a committed counter summary under profiles/
    (the entry `key`, else the default one), or None."""
    path = os.path.join(ROOT, 'profiles', name)
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(key, d.get('traffic_bytes_per_launch'))
    except (OSError, ValueError):
        return None


def batching_form(args):
    """Batching workload: 256 lens-parameter samples in one batch, source grids
    of uneven sizes drawn from {32^2, 48^2, 64^2, 96^2, 128^2}, a shared 1e6
    visibility uv set (complex64 data, one observation per sample), 512^2
    image at 0.01", every solve to a 1e-6 FP64 true residual.  Prints one
    JSON line."""
    from lentils_amd import backend as B
    B.check(B.lib().lt_init(0))
    nlens = 256
    sizes = list(np.random.Generator(np.random.PCG64(15)).choice([32, 48, 64, 96, 128], size=nlens))
    t = time.perf_counter()
    wl = Workload(nvis=1_000_000, nlens=nlens, src_sizes=sizes)
    log(f'batching workload setup {time.perf_counter() - t:.1f}s')
    for k in range(args.warmup):
        t = time.perf_counter()
        its, res = wl.step(maxiter=args.maxiter)
        log(f'warmup {k}: {time.perf_counter() - t:.2f}s iters median {np.median(its):.0f}')
    t0 = time.perf_counter()
    all_its, all_res = [], []
    for k in range(args.steps):
        its, res = wl.step(maxiter=args.maxiter)
        all_its.append(its)
        all_res.append(res)
    total = time.perf_counter() - t0
    its, res = np.concatenate(all_its), np.concatenate(all_res)

    def leg(**kw):  # warm-up steps (FP32 plans, tuning, M's build buffers), then one timed step
        for k in range(args.warmup):
            wl.step(maxiter=args.maxiter, **kw)
        t = time.perf_counter()
        out = wl.step(maxiter=args.maxiter, **kw)
        return out, time.perf_counter() - t

    (its32, res32), t32 = leg(precision='f32')
    (itsl, resl), tl = leg(precond='lens')
    (itsp, resp), tp = leg(precond='lens', precision='f32')
    log(f'batching: FP64 {total / args.steps:.2f}s, precision form {t32:.2f}s, lens preconditioner {tl:.2f}s, '
        f'lens preconditioner with the FP32 operator {tp:.2f}s')
    out = {'metric': 'source solves/s at 1e-6 rel resid (batching form)', 'value': nlens * args.steps / total,
           'unit': 'source solves/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
           'ms_per_step': 1e3 * total / args.steps, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
           'dtype': 'f64', 'data': 'synthetic: one observation per lens sample, complex64 vis + noise 1e-3',
           'config': {'workload': 'batching form: 256 lens samples, uneven source grids {32,48,64,96,128}^2, '
                                  '512^2 image @0.01", 1e6 shared vis, 1024x513 grid, W=6, FP64 CG to 1e-6',
                      'source_sizes': {str(n): int(sizes.count(n)) for n in (32, 48, 64, 96, 128)},
                      'parallelism': 'single GPU'},
           'solver': {'iters_median': float(np.median(its)), 'iters_max': int(its.max()),
                      'true_rel_resid_max': float(res.max()), 'all_converged': bool(np.all(res <= 1e-6))},
           'precision_form': {'value': nlens / t32, 'ms_per_step': 1e3 * t32,
                              'iters_median': float(np.median(its32)),
                              'true_rel_resid_max': float(res32.max()),
                              'all_converged': bool(np.all(res32 <= 1e-6))},
           'lens_preconditioner': {'value': nlens / tl, 'ms_per_step': 1e3 * tl, 'iters_median': float(np.median(itsl)),
                                   'iters_max': int(itsl.max()), 'true_rel_resid_max': float(resl.max()),
                                   'all_converged': bool(np.all(resl <= 1e-6)),
                                   'precision_form': {'value': nlens / tp, 'ms_per_step': 1e3 * tp,
                                                      'iters_median': float(np.median(itsp)),
                                                      'iters_max': int(itsp.max()),
                                                      'true_rel_resid_max': float(resp.max()),
                                                      'all_converged': bool(np.all(resp <= 1e-6))}}}
    print(json.dumps(out), flush=True)


def validation_set(wl, nval, seed):
    """Held-out lens samples for model selection: their own lens parameters
    (seed), blob sources and noise on the workload's uv set, entering as
    device-resident dirty images like the workload's own."""
    from lentils_amd.solvers import DeviceArray, LensBatch
    B = wl.B
    lenses = gen_lenses(nval, seed=seed)
    truth = LensBatch(wl.im, [wl.src] * nval, lenses, z_src=2.0)
    g = np.random.Generator(np.random.PCG64(seed + 1))
    dirty = []
    for s in range(nval):
        rp, ci, va = truth.csr(s)
        x = np.ascontiguousarray(blob_source(wl.src, seed=seed + 2 + s).reshape(-1))
        img = np.zeros(wl.im.size)
        B.check(B.lib().lt_csr_spmv_host(wl.im.size, wl.src.size, B.ptr(rp), B.ptr(ci), B.ptr(va), B.ptr(x),
                                         B.ptr(img)))
        vis = wl.plan.forward(img).astype(np.complex64)
        vis.real += g.standard_normal(wl.nvis, dtype=np.float32) * np.float32(1e-3)
        vis.imag += g.standard_normal(wl.nvis, dtype=np.float32) * np.float32(1e-3)
        dirty.append(wl.plan.adjoint(wl.weights * vis.astype(np.complex128)))
    return lenses, DeviceArray(np.stack(dirty))


def learned_form(args):
    """Learned-preconditioner workload: a batch of `nlens` (256) lens samples
    with 128^2 source grids and the 1e7-visibility uv set (one observation per
    sample, as the plain workload), solved to a 1e-6 FP64 true residual by
    plain CG, by the lens-structured preconditioner M alone and by M with the
    network's correction (flexible CG).  The network is first trained on the
    GPU: over `train_batches` batches of random lenses of the same family
    (seeds apart from the evaluation batch) with `draws` random blob-source
    draws each, the solver generates the right-hand sides b = A x and the
    residuals of M-preconditioned solves of those b stopped after 8 and 32
    iterations; `epochs` passes of Adam on the scale-invariant energy loss
    over them.  The network is then selected on held-out lenses: the
    parameters after each training batch, with the correction scaled by 1,
    1/2 or 1/4, and the zero correction (M alone), each timed on a
    validation batch (`val_samples` lens samples of another seed); the
    fastest is kept.  The learned solves use flexible CG keeping 4 directions
    A-orthogonal.  Prints one JSON line."""
    from lentils_amd import backend as B
    from lentils_amd.learned import PCNet, blob_sources, scale_correction, solver_residuals
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch
    B.check(B.lib().lt_init(0))
    nlens = args.learned_samples
    t = time.perf_counter()
    wl = Workload(nlens=nlens)
    log(f'learned workload setup {time.perf_counter() - t:.1f}s')
    net = PCNet(128, 128, batch=nlens, seed=3)
    iters = (0, 8, 32)
    rng = np.random.default_rng(21)
    t_gen = t_train = 0.0
    npairs = 0
    hist = []
    checkpoints = []
    for tb in range(args.train_batches):
        t = time.perf_counter()
        lb = LensBatch(wl.im, wl.srcs, gen_lenses(nlens, seed=1000 + tb), z_src=2.0)
        inv = BatchedSourceInversion(lb, wl.plan, dirty_images=wl.ddirty, lambda_curvature=wl.lambda_curv)
        xs = blob_sources(wl.src, nlens * args.draws, rng)
        draws = [list(xs[d * nlens:(d + 1) * nlens]) for d in range(args.draws)]
        vecs = solver_residuals(inv, draws, iterations=iters, precond='lens')
        t_gen += time.perf_counter() - t
        t = time.perf_counter()
        last = tb == args.train_batches - 1
        h = inv.train_preconditioner(net, vecs, epochs=args.epochs, lr=args.lr,
                                     lr_final=args.lr * (0.03 if last else 0.3), seed=tb)
        t_train += time.perf_counter() - t
        npairs += len(vecs) * nlens
        hist.append([float(v) for v in h])
        log(f'training batch {tb}: {len(vecs)} vectors x {nlens} samples, loss {h[0]:.4f} -> {h[-1]:.4f} '
            f'(generation {t_gen:.1f}s, training {t_train:.1f}s so far)')
        checkpoints.append(net.params)
        del inv, lb, vecs
    # model selection on held-out lenses: solve time of the validation batch
    t = time.perf_counter()
    vlenses, vdirty = validation_set(wl, args.val_samples, seed=3000)
    vlb = LensBatch(wl.im, [wl.src] * args.val_samples, vlenses, z_src=2.0)
    vinv = BatchedSourceInversion(vlb, wl.plan, dirty_images=vdirty, lambda_curvature=wl.lambda_curv)
    vinv.set_flexible_depth(4)
    cands = [(-1, 0.0)] + [(k, a) for k in range(len(checkpoints)) for a in (1.0, 0.5, 0.25)]
    trials = []
    for k, a in cands:
        net.params = scale_correction(checkpoints[max(k, 0)], a)
        vinv.set_preconditioner(net)
        vinv.solve(rtol=1e-6, maxiter=args.maxiter, precond='learned')  # warm-up
        t1 = time.perf_counter()
        _, vits, vres = vinv.solve(rtol=1e-6, maxiter=args.maxiter, precond='learned')
        trials.append({'checkpoint': k, 'scale': a, 'solve_s': time.perf_counter() - t1,
                       'iters_median': float(np.median(vits)), 'iters_max': int(vits.max()),
                       'all_converged': bool(np.all(vres <= 1e-6))})
    ok = [tr for tr in trials if tr['all_converged']] or trials
    best = min(ok, key=lambda tr: tr['solve_s'])
    net.params = scale_correction(checkpoints[max(best['checkpoint'], 0)], best['scale'])
    t_sel = time.perf_counter() - t
    log(f'selection on {args.val_samples} held-out lenses ({t_sel:.1f}s): checkpoint {best["checkpoint"]} '
        f'scale {best["scale"]}: {best["solve_s"]:.2f}s, iters median {best["iters_median"]:.0f}; '
        f'M alone {trials[0]["solve_s"]:.2f}s, iters median {trials[0]["iters_median"]:.0f}')
    del vinv, vlb, vdirty
    results = {}
    for name, pc in (('plain', 'none'), ('lens', 'lens'), ('learned', 'learned')):
        for k in range(args.warmup):
            wl.step(maxiter=args.maxiter, precond=pc, net=net)
        ts, all_its, all_res = [], [], []
        for k in range(args.steps):
            t = time.perf_counter()
            its, res = wl.step(maxiter=args.maxiter, precond=pc, net=net)
            ts.append(time.perf_counter() - t)
            all_its.append(its)
            all_res.append(res)
        its, res = np.concatenate(all_its), np.concatenate(all_res)
        tot = sum(ts)
        results[name] = {'value': nlens * args.steps / tot, 'unit': 'source solves/s',
                         'ms_per_step': 1e3 * tot / args.steps, 'iters_median': float(np.median(its)),
                         'iters_max': int(its.max()), 'true_rel_resid_max': float(res.max()),
                         'all_converged': bool(np.all(res <= 1e-6))}
        if name == 'learned':
            results[name]['guard_fallback_iters'] = int(wl.last_fallbacks[0].sum())
            results[name]['samples_switched_to_M'] = int(wl.last_fallbacks[1].sum())
            results[name]['method'] = ('flexible CG keeping 4 directions A-orthogonal; z = M^-1 r + network '
                                       'correction (checkpoint and scale selected on held-out lenses), M alone for '
                                       'samples whose residual fell < 10x in 100 iterations')
        log(f'{name}: {tot / args.steps:.2f}s/step iters median {np.median(its):.0f} max {its.max()} '
            f'max res {res.max():.2e}')
    out = {'metric': 'source solves/s at 1e-6 rel resid (learned-PCG vs plain CG); Gvis/s degrid',
           'value': results['learned']['value'], 'unit': 'source solves/s', 'n_gpus': 1, 'steps': args.steps,
           'warmup': args.warmup, 'ms_per_step': results['learned']['ms_per_step'], 'higher_is_better': True,
           'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64 (network: bf16 MFMA, f32 accumulate)',
           'data': 'synthetic: seeded lenses and uv annulus; one observation per lens sample, complex64 vis + '
                   'noise 1e-3; training pairs from random lenses and blob sources',
           'config': {'workload': f'learned-preconditioner form: {nlens} lens samples, 512^2 image @0.01", 128^2 '
                                  'source, 1e7 vis, 1024x513 grid, W=6, FP64 flexible CG to 1e-6',
                      'parallelism': 'single GPU'},
           'learned': results['learned'], 'plain_cg': results['plain'], 'lens_preconditioner': results['lens'],
           'speedup_vs_plain': results['learned']['value'] / results['plain']['value'],
           'speedup_vs_lens': results['learned']['value'] / results['lens']['value'],
           'training': {'lens_batches': args.train_batches, 'draws_per_batch': args.draws,
                        'truncation_iters': list(iters), 'pairs': npairs, 'epochs': args.epochs,
                        'pair_visits': npairs * args.epochs, 'lr': args.lr, 'generation_s': t_gen,
                        'training_s': t_train, 'loss_history': hist},
           'selection': {'criterion': f'solve time of a validation batch of {args.val_samples} held-out lens '
                                      'samples (lens seed 3000, own sources and noise); checkpoint -1 with scale 0 '
                                      'is the zero correction (M alone)',
                         'chosen': best, 'trials': trials, 'selection_s': t_sel}}
    print(json.dumps(out), flush=True)


def precision_report(args):
    """Precision-form workload: errors of the FP32 / mixed paths against FP64.
    (1) FP64 deflections against the oracle for q up to 0.999 and radii down
    to 1e-4" (series convergence); (2) FP32 visibilities of a lensed image
    (1024^2 image at 0.005", 5e7 visibilities, 2048 x 1025 grid) against the
    FP64 plan on all visibilities and against the oracle on a subset; (3) the
    precision-form solution against the FP64 solution for 8 lens samples with
    256^2 sources.  Prints one JSON line."""
    from lentils_amd import ImageSpace, VisibilitySpace
    from lentils_amd import backend as B
    from lentils_amd.solvers import BatchedSourceInversion, LensBatch, NUFFTPlan
    from oracle import oracle as O
    B.check(B.lib().lt_init(0))
    out = {'metric': 'errors of the precision form against FP64 (precision workload)'}
    # (1) deflections
    g = np.random.Generator(np.random.PCG64(21))
    worst, rows = 0.0, []
    for q in (0.5, 0.7, 0.9, 0.99, 0.999):
        for gam in (1.8, 2.0, 2.2):
            b = g.uniform(0.5, 1.5)
            lens = [O.pack_pemd(b=b, q=q, th=g.uniform(0, 180), qh=(gam - 1) / 2, x=0.013, y=-0.007),
                    O.pack_shear(ss=0.03, sa=g.uniform(0, 180))]
            r = np.exp(g.uniform(np.log(1e-4), np.log(3.0), 20000))
            ph = g.uniform(0, 2 * np.pi, 20000)
            pts = np.stack([0.013 + r * np.cos(ph), -0.007 + r * np.sin(ph)], 1)
            arr = np.array(lens, dtype=O.MASS_MODEL_DTYPE)
            glm = B.global_lens_model_ctype(2, 1, arr.ctypes.data_as(C.c_void_p))
            got = np.zeros_like(pts)
            B.check(B.lib().lt_deflect_points(C.byref(glm), B.ptr(pts), pts.shape[0], B.ptr(got)))
            ref = O.deflect_points(lens, pts)
            a_g, a_r = pts - got, pts - ref
            e = float(np.max(np.linalg.norm(a_g - a_r, axis=1) / np.linalg.norm(a_r, axis=1)))
            worst = max(worst, e)
            rows.append({'q': q, 'gamma': gam, 'max_rel_err': e})
    out['deflection'] = {'max_rel_err': worst, 'points_per_case': 20000, 'radii_arcsec': [1e-4, 3.0],
                         'cases': rows}
    log(f'deflections: max relative error {worst:.2e}')
    # (2) visibilities
    im = ImageSpace(shape=(1024, 1024), bounds=[(-2.56, 2.56), (-2.56, 2.56)])
    src = ImageSpace(shape=(256, 256), bounds=[(-0.64, 0.64), (-0.64, 0.64)])
    nvis = 50_000_000
    uvw = gen_uv(nvis, seed=16)
    vs = VisibilitySpace(channels=[1.0], uvcoords=uvw)
    lenses = gen_lenses(8, seed=17)
    lb = LensBatch(im, [src] * 8, lenses, z_src=2.0)
    rp, ci, va = lb.csr(0)
    x = np.ascontiguousarray(blob_source(src, seed=18).reshape(-1))
    img = np.zeros(im.size)
    B.check(B.lib().lt_csr_spmv_host(im.size, src.size, B.ptr(rp), B.ptr(ci), B.ptr(va), B.ptr(x), B.ptr(img)))
    p64 = NUFFTPlan(vs, im, pad_factor=2, kernel_support=3, kb_beta=14.04, precision='f64')
    p32 = NUFFTPlan(vs, im, pad_factor=2, kernel_support=3, kb_beta=14.04, precision='f32')
    v64 = p64.forward(img)
    v32 = p32.forward(img)
    e_all = float(np.max(np.abs(v32 - v64)) / np.max(np.abs(v64)))
    k = 200_000
    og = O.Grid2D(1024, 1024, [-2.56, 2.56, -2.56, 2.56])
    geo = O.NUFFTGeometry(og, pad=2, wsup=3, beta=14.04)
    vo = O.nufft_forward(img.reshape(1024, 1024), uvw[:k], [1.0], geo, nthreads=len(os.sched_getaffinity(0)))
    vo = vo.reshape(-1)
    e_or32 = float(np.max(np.abs(v32[:k] - vo)) / np.max(np.abs(vo)))
    e_or64 = float(np.max(np.abs(v64[:k] - vo)) / np.max(np.abs(vo)))
    out['visibilities'] = {'nvis': nvis, 'f32_vs_f64_max_rel_err': e_all, 'oracle_subset': k,
                           'f32_vs_oracle_max_rel_err': e_or32, 'f64_vs_oracle_max_rel_err': e_or64}
    log(f'visibilities: FP32 vs FP64 {e_all:.2e}, FP32 vs oracle {e_or32:.2e}, FP64 vs oracle {e_or64:.2e}')
    del p32, v32
    # (3) solutions: one observation per sample
    gn = np.random.Generator(np.random.PCG64(19))
    w = np.full(nvis, 1e6)
    dirty = []
    for s_ in range(8):
        rp, ci, va = lb.csr(s_)
        xs = np.ascontiguousarray(blob_source(src, seed=30 + s_).reshape(-1))
        im_s = np.zeros(im.size)
        B.check(B.lib().lt_csr_spmv_host(im.size, src.size, B.ptr(rp), B.ptr(ci), B.ptr(va), B.ptr(xs),
                                         B.ptr(im_s)))
        vis = p64.forward(im_s).astype(np.complex64)
        vis.real += gn.standard_normal(nvis, dtype=np.float32) * np.float32(1e-3)
        vis.imag += gn.standard_normal(nvis, dtype=np.float32) * np.float32(1e-3)
        dirty.append(p64.adjoint(w * vis.astype(np.complex128)))
    inv = BatchedSourceInversion(lb, p64, weights=w, dirty_images=dirty, lambda_curvature=1e2)
    t = time.perf_counter()
    x64, it64, r64 = inv.solve(rtol=1e-6, maxiter=args.maxiter)
    t64 = time.perf_counter() - t
    t = time.perf_counter()
    x32, it32, r32 = inv.solve(rtol=1e-6, maxiter=args.maxiter, precision='f32')
    t32 = time.perf_counter() - t
    dif = [float(np.linalg.norm(a - b_) / np.linalg.norm(b_)) for a, b_ in zip(x32, x64)]
    out['solution'] = {'samples': 8, 'f64': {'seconds': t64, 'iters_median': float(np.median(it64)),
                                             'true_rel_resid_max': float(r64.max())},
                       'f32_form': {'seconds': t32, 'iters_median': float(np.median(it32)),
                                    'true_rel_resid_max': float(r32.max())},
                       'rel_diff_f32_vs_f64_max': max(dif)}
    log(f'solutions: FP64 {t64:.2f}s, precision form {t32:.2f}s, max relative difference {max(dif):.2e}')
    print(json.dumps(out), flush=True)


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(wl, runs=5):
    """The reference's radio solve (test_all.py:57-72) on the host for one of
    the 64 lens samples (sample 0), timed as the GPU step is: lens matrix +
    preconditioner + CG to a 1e-6 relative residual per source solve.

    * lens matrix: the reference's own raster.c + deflect.c (oracle/_ref,
      single thread) in a child process (its fragment buffer can overflow on
      the lens-centre pixel, raster.c:204-212); the oracle port if it dies;
    * operator: L^T BCB L + lambda Cc^T Cc, BCB the dirty-beam convolution of
      data_base.py:115-129 with scipy.fft (single thread, as
      fourier_operators.py:19,23 call it);
    * preconditioner: scipy splu of lambda Cc^T Cc + wtsum L_s^T L_s
      (test_all.py:63-65).  L_s is L without its rows of more than 64 entries:
      with the reference's full L the unmasked lens-centre pixels (rows of up
      to ~15k entries) make L^T L nearly dense and SuperLU runs out of memory;
      the reference's own test masks those pixels out;
    * CG: scipy.sparse.linalg.cg(rtol=1e-6).
    Set-up excluded from the timing (as the GPU's resident inputs are): the
    dirty beam (oracle gridder on all cores) and the right-hand side.  One
    warm-up, then the median of `runs` timed runs of each part."""
    from oracle import oracle as O
    import scipy.fft as sfft
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    cores = len(os.sched_getaffinity(0))
    im = O.Grid2D(512, 512, [-2.56, 2.56, -2.56, 2.56])
    src = O.Grid2D(128, 128, [-0.64, 0.64, -0.64, 0.64])
    lm = wl.lenses[0]
    lm.setup_raytracing(2.0)
    kind_raster = 'port'
    note = ''
    t_ras = []
    if O.ref_lib() is not None:
        import subprocess
        code = ('import sys, time\n'
                'import numpy as np\n'
                'sys.path.insert(0, %r)\n'
                'from oracle import oracle as O\n'
                'L = np.frombuffer(bytes.fromhex(sys.argv[1]), dtype=O.MASS_MODEL_DTYPE)\n'
                'im = O.Grid2D(512, 512, [-2.56, 2.56, -2.56, 2.56])\n'
                'src = O.Grid2D(128, 128, [-0.64, 0.64, -0.64, 0.64])\n'
                'for _ in range(%d):\n'
                '    t = time.perf_counter()\n'
                '    O.ref_lens_matrix_csr(im, src, L)\n'
                '    print(time.perf_counter() - t, flush=True)\n') % (ROOT, runs + 1)
        r = subprocess.run([sys.executable, '-c', code, lm.lenses.tobytes().hex()], capture_output=True, text=True)
        if r.returncode == 0:
            kind_raster = 'reference'
            t_ras = [float(v) for v in r.stdout.split()][1:]
        else:
            note = (f'; the reference raster.c build exited with status {r.returncode} on this lens (fragment-buffer '
                    f'overflow at the lens-centre pixel), so the oracle port was timed')
    rp, ci, va, _ = O.lens_matrix_csr(im, src, lm.lenses)
    if not t_ras:
        for k in range(runs + 1):
            t0 = time.perf_counter()
            O.lens_matrix_csr(im, src, lm.lenses)
            if k:
                t_ras.append(time.perf_counter() - t0)
    L = sp.csr_matrix((va, ci, rp), shape=(512 * 512, 128 * 128))
    crp, cci, cva = O.curv_matrix(src)
    Cm = sp.csr_matrix((cva, cci, crp), shape=(128 * 128, 128 * 128))
    R = wl.lambda_curv * (Cm.T @ Cm)
    # set-up: the dirty beam with the reference's default NUFFT, the BCB multiplier, the right-hand side
    t0 = time.perf_counter()
    beam = O.dirty_beam(wl.uvw, [1.0], im, wl.weights, nthreads=cores)
    geo = O.NUFFTGeometry(im, pad=2, wsup=4)
    dbf = sfft.rfft2(np.roll(beam, [geo.nxp // 2, geo.nyp // 2], axis=[0, 1])) / beam.size
    rhs = L.T @ wl.dirty[0].reshape(-1)
    wtsum = float(np.sum(wl.weights))
    t_setup = time.perf_counter() - t0

    def bcb(x):
        p = np.zeros((geo.nxp, geo.nyp))
        p[geo.padx:geo.padx + 512, geo.pady:geo.pady + 512] = x.reshape(512, 512)
        q = sfft.irfft2(dbf * sfft.rfft2(p), s=p.shape, norm='forward')
        return q[geo.padx:geo.padx + 512, geo.pady:geo.pady + 512].reshape(-1)

    A = spl.LinearOperator((src.nx * src.ny,) * 2, matvec=lambda x: L.T @ bcb(L @ x) + R @ x)
    keep = (np.diff(L.indptr) <= 64).astype(np.float64)
    Ls = sp.diags(keep) @ L
    t_solve, iters, resid = [], [], []
    for k in range(runs + 1):
        t0 = time.perf_counter()
        lu = spl.splu((R + wtsum * (Ls.T @ Ls)).tocsc())
        M = spl.LinearOperator(A.shape, matvec=lu.solve)
        n = [0]
        x, info = spl.cg(A, rhs, rtol=1e-6, atol=0.0, M=M, maxiter=5000,
                         callback=lambda xk: n.__setitem__(0, n[0] + 1))
        dt = time.perf_counter() - t0
        if k:
            t_solve.append(dt)
            iters.append(n[0])
            resid.append(float(np.linalg.norm(rhs - A @ x) / np.linalg.norm(rhs)))
    solve_s = float(np.median(t_ras) + np.median(t_solve))
    return {'value': 1.0 / solve_s, 'unit': 'source solves/s', 'cores': 1, 'kind': 'port',
            'sample': (f'1 of the 64 lens samples (sample 0), the reference\'s radio solve structure '
                       f'(test_all.py:57-72): lens matrix {np.median(t_ras):.2f}s ('
                       f'{"reference raster.c/deflect.c via oracle/_ref" if kind_raster == "reference" else "oracle port"}'
                       f', 1 thread) + splu(R + wtsum L_s^T L_s) and scipy CG with the BCB operator (scipy.fft, '
                       f'1 thread) to rtol 1e-6: {np.median(t_solve):.2f}s, {int(np.median(iters))} CG iterations, '
                       f'true relative residual {max(resid):.2e}; medians of {runs} runs after a warm-up; set-up '
                       f'excluded: dirty beam + BCB multiplier + rhs {t_setup:.1f}s on {cores} threads; '
                       f'L_s = L without rows of more than 64 entries (the full L^T L does not fit SuperLU); '
                       f'CPU: {_cpu_model()}{note}'),
            'cg_iters': int(np.median(iters)), 'true_rel_resid_max': max(resid),
            'lens_matrix_s': float(np.median(t_ras)), 'solve_s': float(np.median(t_solve)), 'setup_s': t_setup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=1)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--operator', default='nufft', choices=['nufft', 'bcb'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--maxiter', type=int, default=20000)
    ap.add_argument('--precond', default='none', choices=['none', 'jacobi', 'lens'])
    ap.add_argument('--precision-report', action='store_true',
                    help='errors of the FP32 paths against FP64 on the precision workload (1024^2 image, 5e7 vis)')
    ap.add_argument('--batching', action='store_true',
                    help='the batching workload (256 samples, uneven source grids, 1e6 visibilities) instead')
    ap.add_argument('--profile-degrid', action='store_true',
                    help='only run the degridder and the calibration copy (for rocprofv3 --pmc)')
    ap.add_argument('--profile-kernels', action='store_true',
                    help='only run the hot kernels once each batch and the calibration copy (for rocprofv3 --pmc)')
    ap.add_argument('--no-graphs', action='store_true',
                    help='launch the CG kernels one by one instead of as captured HIP graphs (for kernel tracing)')
    ap.add_argument('--learned', action='store_true',
                    help='the learned-preconditioner workload (256 samples, network trained first) instead')
    ap.add_argument('--learned-samples', type=int, default=256)
    ap.add_argument('--train-batches', type=int, default=6)
    ap.add_argument('--draws', type=int, default=8)
    ap.add_argument('--epochs', type=int, default=6)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--val-samples', type=int, default=32)
    ap.add_argument('--shared-data', action='store_true',
                    help='one observation for all samples (most samples then fit the data with the wrong lens)')
    args = ap.parse_args()
    global GRAPHS
    GRAPHS = not args.no_graphs
    if args.profile_degrid:
        profile_degrid()
        return
    if args.profile_kernels:
        profile_kernels()
        return
    if args.batching:
        batching_form(args)
        return
    if args.learned:
        learned_form(args)
        return
    if args.precision_report:
        precision_report(args)
        return
    if args.gpus != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path (every solve of a batch is independent; '
              'see DESIGN.md for how batches would shard). Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)
    from lentils_amd import backend as B
    B.check(B.lib().lt_init(0))
    t = time.perf_counter()
    wl = Workload(shared_data=args.shared_data)
    log(f'workload setup {time.perf_counter() - t:.1f}s')
    for k in range(args.warmup):
        t = time.perf_counter()
        its, res = wl.step(operator=args.operator, maxiter=args.maxiter, precond=args.precond)
        log(f'warmup {k}: {time.perf_counter() - t:.2f}s iters median {np.median(its):.0f} max res {res.max():.2e}')
    times = []
    all_its, all_res = [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        its, res = wl.step(operator=args.operator, maxiter=args.maxiter, precond=args.precond)
        times.append(time.perf_counter() - t)
        all_its.append(its)
        all_res.append(res)
        log(f'step {k}: {times[-1]:.2f}s iters median {np.median(its):.0f} max {its.max()} max res {res.max():.2e}')
    total = time.perf_counter() - t0
    its = np.concatenate(all_its)
    res = np.concatenate(all_res)
    ms_per_step = 1e3 * total / args.steps
    value = NLENS * args.steps / total
    # precision form: the same step with the FP32 operator inside the FP64 CG
    # (FP64 true residual refreshed every 20 iterations, final FP64 check)
    pf = None
    if args.operator == 'nufft':
        for k in range(args.warmup):  # FP32 rocFFT plans and kernels are first built here
            wl.step(operator=args.operator, maxiter=args.maxiter, precond=args.precond, precision='f32')
        t = time.perf_counter()
        its32, res32 = wl.step(operator=args.operator, maxiter=args.maxiter, precond=args.precond, precision='f32')
        t32 = time.perf_counter() - t
        pf = {'value': NLENS / t32, 'unit': 'source solves/s', 'ms_per_step': 1e3 * t32,
              'iters_median': float(np.median(its32)), 'iters_max': int(its32.max()),
              'true_rel_resid_max': float(res32.max()), 'all_converged': bool(np.all(res32 <= 1e-6)),
              'dtype': 'f32 operator (SpMV, rocFFT, stencil), f64 CG vectors and inner products'}
        log(f'precision form: {t32:.2f}s iters median {np.median(its32):.0f} max res {res32.max():.2e}')
    # the reference's preconditioner structure (lambda Cc^T Cc + c L_s^T L_s,
    # three inner CG steps on it per outer iteration) on the same workload
    lp = None
    if args.operator == 'nufft' and args.precond == 'none':
        for k in range(args.warmup):
            wl.step(operator=args.operator, maxiter=args.maxiter, precond='lens')
        t = time.perf_counter()
        itsl, resl = wl.step(operator=args.operator, maxiter=args.maxiter, precond='lens')
        tl = time.perf_counter() - t
        lp = {'value': NLENS / tl, 'unit': 'source solves/s', 'ms_per_step': 1e3 * tl,
              'iters_median': float(np.median(itsl)), 'iters_max': int(itsl.max()),
              'true_rel_resid_max': float(resl.max()), 'all_converged': bool(np.all(resl <= 1e-6)),
              'method': 'flexible CG, preconditioner: 3 Jacobi-PCG steps on the explicit lambda Cc^T Cc + c L_s^T L_s (L_s: rows of L with <= 64 nonzero entries)'}
        log(f'lens preconditioner: {tl:.2f}s iters median {np.median(itsl):.0f} max res {resl.max():.2e}')
        # both: the lens-structured preconditioner with the FP32 operator
        for k in range(args.warmup):
            wl.step(operator=args.operator, maxiter=args.maxiter, precond='lens', precision='f32')
        t = time.perf_counter()
        itsp, resp = wl.step(operator=args.operator, maxiter=args.maxiter, precond='lens', precision='f32')
        tp = time.perf_counter() - t
        lp['precision_form'] = {'value': NLENS / tp, 'unit': 'source solves/s', 'ms_per_step': 1e3 * tp,
                                'iters_median': float(np.median(itsp)), 'iters_max': int(itsp.max()),
                                'true_rel_resid_max': float(resp.max()), 'all_converged': bool(np.all(resp <= 1e-6)),
                                'dtype': 'f32 operator inside the flexible FP64 CG'}
        log(f'lens preconditioner, FP32 operator: {tp:.2f}s iters median {np.median(itsp):.0f} max res {resp.max():.2e}')
    st = stencil_roofline(wl)
    log(f'stencil (form {st["form"]}) {st["ms"] * 1e3:.1f} us/launch at {NLENS} samples: {st["tflops"]:.2f} TFLOP/s, {st["gbs"]:.0f} GB/s '
        f'({st["rows"]} rows, {st["win_cells"]} window cells)')
    rl = degrid_roofline(wl)
    log(f'degrid f32 {rl["ms"] * 1e3:.1f} us/launch, {rl["gbs"]:.0f} GB/s, {rl["gvis_s"]:.2f} Gvis/s')
    copy_gbs, _ = stream_copy_gbs()
    log(f'streaming copy {copy_gbs:.0f} GB/s')
    sp = spmv_roofline(wl)
    log(f'L^T product {sp["LT"]["ms"] * 1e3:.1f} us ({sp["LT"]["gbs"]:.0f} GB/s), L product {sp["L"]["ms"] * 1e3:.1f} us '
        f'({sp["L"]["gbs"]:.0f} GB/s)')
    out = {
        'metric': 'source solves/s at 1e-6 rel resid (learned-PCG vs plain CG); Gvis/s degrid',
        'value': value, 'unit': 'source solves/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': ms_per_step, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f64',
        'data': ('synthetic: seeded lenses and uv annulus; ' +
                 ('one observation (blob source lensed by sample 0) shared by all samples'
                  if args.shared_data else 'one observation per lens sample (its own blob source lensed by its '
                  'own lens), complex64 vis + noise 1e-3')),
        'config': {'workload': 'plain GPU method: 64 lens samples, 512^2 image @0.01", 128^2 source, 1e7 vis, '
                               '1024x513 grid, W=6 (support 3), beta=14.04, lambda_curv=1e2, FP64 CG to 1e-6',
                   'operator': args.operator,
                   'precond': 'none (plain CG)' if args.precond == 'none' else args.precond, 'lens_samples_per_step': NLENS,
                   'nvis': NVIS, 'parallelism': 'single GPU'},
        'solver': {'iters_median': float(np.median(its)), 'iters_max': int(its.max()),
                   'true_rel_resid_max': float(res.max()), 'all_converged': bool(np.all(res <= 1e-6))},
        'precision_form': pf,
        'lens_preconditioner': lp,
        'degrid_gvis_s': rl['gvis_s'], 'degrid_gvis_s_f64': rl['gvis_s_f64'],
        'roofline': {'kernel': st['kernel'],
                     'bound': 'mfma', 'achieved': st['tflops'], 'peak': FP64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': st['tflops'] / FP64_PEAK_TFLOPS,
                     'traffic': load_traffic('stencil_c2_pmc.json',
                                             'traffic_bytes_per_launch_2_groups_per_workgroup' if st['form'] == -2
                                             else 'traffic_bytes_per_launch'),
                     'algorithmic_flops': st['flops'], 'algorithmic_bytes': st['bytes'],
                     'achieved_gbs': st['gbs'], 'us_per_launch': st['ms'] * 1e3,
                     'compute_unit': 'FP64 matrix cores (MI355X FP64 matrix and vector peaks are both 78.6 '
                                     'TFLOP/s); intensity above the FP64 ridge, so FP64 throughput bounds it; the '
                                     'matrix form also multiplies the zeros of each 4x4 tile of rows against its '
                                     '16x16 window (executed MFMA work 1.73x the algorithmic flops, counted)'},
        'roofline_spmv': {'kernel': 'L^T y (k_spmv_stream + long-row chunks, FP64, all 64 samples)', 'bound': 'hbm',
                          'achieved': sp['LT']['gbs'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                          'frac': sp['LT']['gbs'] / HBM_PEAK_GBS, 'traffic': None,
                          'algorithmic_bytes': sp['LT']['bytes'], 'us_per_launch': sp['LT']['ms'] * 1e3,
                          'nnz': sp['LT']['nnz'],
                          'L': {'us': sp['L']['ms'] * 1e3, 'achieved_gbs': sp['L']['gbs'],
                                'frac': sp['L']['gbs'] / HBM_PEAK_GBS, 'algorithmic_bytes': sp['L']['bytes'],
                                'nnz': sp['L']['nnz']}},
        'roofline_degrid': {'kernel': 'k_degrid32<float,3> (FP32 degridder)', 'bound': 'hbm', 'achieved': rl['gbs'],
                            'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': rl['gbs'] / HBM_PEAK_GBS,
                            'traffic': load_traffic('degrid_f32_c2_pmc.json'), 'algorithmic_bytes': rl['bytes'],
                            'us_per_launch': rl['ms'] * 1e3, 'stream_copy_gbs': copy_gbs},
    }
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        out['cpu_baseline'] = cpu_baseline(wl)
        log(f'cpu baseline {time.perf_counter() - t:.1f}s')
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
