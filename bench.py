"""Benchmark: Scattering2D J=3 L=8 on 256x256 float32 images, batch 64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

One step = one forward of the fused HIP pipeline over the batch (inputs
resident in HBM; plan creation, filter upload and workspace allocation happen
before the timed region).  Prints one JSON line:
  value            images/s over the K timed steps
  roofline         dominant kernel: algorithmic bytes per launch (stage-boundary
                   byte model) / its mean launch time from HIP events recorded
                   around every launch during the timed steps, vs 8 TB/s;
                   .isolated: the same launch timed after the timed region
                   with every chunk on one stream (nothing beside it);
                   .measured_copy_GBps: a plain 1 GiB device copy's rate
  pipeline_roofline the same model over the whole forward
  cpu_baseline     the CPU oracle (op-for-op restatement of the reference's torch
                   CPU path) on a bounded sample of the same workload
Only one GPU is used; --gpus N > 1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = 'images/s Scattering2D J=3 L=8 256² fp32 (%HBM roofline); recon time to 1e-3'
J, L, SIDE, BATCH, SEED = 3, 8, 256, 64, 2


TRAFFIC_PROFILE = os.path.join(ROOT, 'profiles', 'scat2d_j3_256_hbm_traffic.md')


def measured_traffic_per_node_image():
    """HBM bytes per (image, node) of the dominant launch, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith('traffic_bytes_per_image_per_node:'):
                return int(line.split(':', 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(budget_s=15.0, max_images=2048):
    from oracle import scat2d
    from kymatio_amd import workloads
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    torch.set_num_threads(cores)
    sc = scat2d.Scattering2DTorchCPU(J, (SIDE, SIDE))
    x = torch.from_numpy(workloads.grating_images(16, SIDE, seed=SEED))
    n, t_total = 0, 0.0
    with torch.no_grad():
        sc(x[:1])  # warm-up
        while t_total < budget_s and n < max_images:
            t0 = time.perf_counter()
            sc(x)
            t_total += time.perf_counter() - t0
            n += x.shape[0]
    model = 'unknown'
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return {'value': n / t_total, 'unit': 'images/s', 'cores': cores, 'kind': 'port',
            'sample': '%d images of the same 256x256 J=3 workload (batches of 16), torch CPU FP32 '
                      'restatement of the reference torch backend, %.1f s, %s' % (n, t_total, model)}


def _line(metric, value, unit, steps, warmup, ms, dtype, config, model_bytes_per_step, kernel, cpu):
    """One JSON line of the other workloads: the roofline is the whole step's
    stage-boundary model bytes over its wall time (one graph replay or one
    forward / forward+backward call; no per-launch events there)."""
    achieved = model_bytes_per_step / (ms * 1e-3)
    return {'metric': metric, 'value': value, 'unit': unit, 'n_gpus': 1, 'steps': steps, 'warmup': warmup,
            'ms_per_step': ms, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': dtype,
            'data': 'synthetic', 'config': config,
            'roofline': {'bound': 'hbm', 'achieved': achieved / 1e9, 'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
                         'frac': achieved / HBM_PEAK, 'traffic': None, 'kernel': kernel,
                         'model_bytes_per_step': model_bytes_per_step},
            'cpu_baseline': cpu}


def _timed(fn, steps, warmup):
    for _ in range(max(1, warmup)):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def _cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def bench_mixed(args, n_images=256, seed=5):
    """256 square images, side uniform over {64, 128, 256, 512}, pixels
    U[0,1), J = log2(side) - 3 (8x8 outputs): one HIP-graph replay per step."""
    from kymatio_amd import workloads
    from kymatio_amd.scattering2d.mixed import MixedScattering2D, default_J
    dev = torch.device('cuda:0')
    imgs = workloads.mixed_size_images(n_images, seed=seed)
    groups = {}
    for im in imgs:
        groups.setdefault(im.shape[0], []).append(im)
    groups = {s: torch.from_numpy(np.stack(v)).to(dev) for s, v in sorted(groups.items())}
    mx = MixedScattering2D(device=dev, autotune=not args.no_autotune)
    with torch.no_grad():
        dt = _timed(lambda: mx.forward_grouped(groups), args.steps, args.warmup)
    model = 0
    for side, x in groups.items():
        plan = mx.module(side)._plan(dev, torch.float32)
        model += x.shape[0] * sum(l['model_bytes_per_image'] for l in plan.launches())
    cpu = None
    if not args.no_cpu_baseline:
        # oracle per size on a bounded sample, weighted by this batch's size counts
        from oracle import scat2d
        torch.set_num_threads(_cores())
        per = {}
        for side, x in groups.items():
            sc = scat2d.Scattering2DTorchCPU(default_J(side), (side, side))
            xs = x[:2].cpu()
            with torch.no_grad():
                sc(xs[:1])
                t0 = time.perf_counter()
                sc(xs)
                per[side] = (time.perf_counter() - t0) / xs.shape[0]
        t_batch = sum(per[s] * x.shape[0] for s, x in groups.items())
        cpu = {'value': n_images / t_batch, 'unit': 'images/s', 'cores': _cores(), 'kind': 'port',
               'sample': '2 images of each side (%s), oracle torch CPU FP32, time per image weighted by the batch '
                         'counts %s' % (sorted(per), {s: int(x.shape[0]) for s, x in groups.items()})}
    return _line('images/s MixedScattering2D sides {64,128,256,512} J=log2(side)-3 L=8 fp32', n_images / dt,
                 'images/s', args.steps, args.warmup, dt * 1e3, 'f32',
                 {'workload': 'mixed-size batch of %d images (sides uniform over 64/128/256/512, pixels U[0,1), '
                              'seed %d), one graph replay per step' % (n_images, seed),
                  'counts': {int(s): int(x.shape[0]) for s, x in groups.items()}},
                 model, 'graph replay (all sizes)', cpu)


def bench_scat1d(args, batch=256, seed=3):
    """Scattering1D J=6 Q=8 on 256 chirp signals of length 4096."""
    from kymatio_amd import workloads
    from kymatio_amd.torch import Scattering1D
    dev = torch.device('cuda:0')
    J, Q, N = 6, 8, 4096
    x = torch.from_numpy(workloads.chirp_signals(batch, N, seed=seed)).to(dev)
    sc = Scattering1D(J, N, Q).to(dev)
    with torch.no_grad():
        out = sc(x)
        dt = _timed(lambda: sc(x), args.steps, args.warmup)
    model_per_signal = 10913280  # stage-boundary model at J=6 Q=8 N=4096 (write + read of every stage tensor)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import scat1d
        torch.set_num_threads(_cores())
        ref = scat1d.Scattering1DTorchCPU(J, N, Q)
        xs = x[:64].cpu()
        with torch.no_grad():
            ref(xs[:4])
            t0 = time.perf_counter()
            ref(xs)
            t = time.perf_counter() - t0
        cpu = {'value': xs.shape[0] / t, 'unit': 'signals/s', 'cores': _cores(), 'kind': 'port',
               'sample': '64 of the same chirp signals, oracle torch CPU FP32 (%.1f s)' % t}
    return _line('signals/s Scattering1D J=6 Q=8 N=4096 fp32', batch / dt, 'signals/s', args.steps, args.warmup,
                 dt * 1e3, 'f32', {'workload': 'Scattering1D J=6 Q=8 on %d chirp signals of length %d (seed %d); '
                                               'output %s' % (batch, N, seed, tuple(out.shape)), 'batch': batch},
                 model_per_signal * batch, 'whole forward', cpu)


def bench_grad2d(args, batch=32, seed=4):
    """Scattering2D J=4 L=8 on 32 images of 128x128: forward with the saved
    fields plus the hand-written backward of sum((S(x) - S(y))^2)."""
    from kymatio_amd import workloads
    from kymatio_amd.torch import Scattering2D
    dev = torch.device('cuda:0')
    x = torch.from_numpy(workloads.normal_images(batch, 128, seed=seed)).to(dev)
    y = torch.from_numpy(workloads.normal_images(batch, 128, seed=seed + 1)).to(dev)
    sc = Scattering2D(4, (128, 128)).to(dev)
    with torch.no_grad():
        sy = sc(y)

    def step():
        xr = x.clone().requires_grad_(True)
        loss = ((sc(xr) - sy) ** 2).sum()
        loss.backward()
        return xr.grad

    dt = _timed(step, args.steps, args.warmup)
    model_per_image = 87259776  # forward + backward stage-boundary model at J=4 128x128
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import scat2d
        torch.set_num_threads(_cores())
        ref = scat2d.Scattering2DTorchCPU(4, (128, 128))
        xs, ys = x[:8].cpu(), y[:8].cpu()
        with torch.no_grad():
            rys = ref(ys)
        t0 = time.perf_counter()
        xr = xs.clone().requires_grad_(True)
        ((ref(xr) - rys) ** 2).sum().backward()
        t = time.perf_counter() - t0
        cpu = {'value': xs.shape[0] / t, 'unit': 'images/s', 'cores': _cores(), 'kind': 'port',
               'sample': '8 of the same images, oracle torch CPU FP32 forward + autograd backward (%.1f s)' % t}
    return _line('images/s Scattering2D J=4 L=8 128x128 fp32 forward+backward', batch / dt, 'images/s', args.steps,
                 args.warmup, dt * 1e3, 'f32',
                 {'workload': 'sum((S(x) - S(y))^2) and its gradient in x, J=4 L=8, %d N(0,1) images of 128x128 '
                              '(seed %d)' % (batch, seed), 'batch': batch},
                 model_per_image * batch, 'forward with saved fields + backward', cpu)


def bench_jtfs(args, batch=64, seed=6):
    """TimeFrequencyScattering J=6 J_fr=2 Q=8 on 64 chirps of length 4096
    (the path runs op by op over the 1D HIP kernels; no byte model and no CPU
    restatement exist for it, so roofline and cpu_baseline are null)."""
    from kymatio_amd import workloads
    from kymatio_amd.torch import TimeFrequencyScattering
    dev = torch.device('cuda:0')
    x = torch.from_numpy(workloads.chirp_signals(batch, 4096, seed=seed)).to(dev)
    sc = TimeFrequencyScattering(J=6, J_fr=2, Q=8, shape=4096).to(dev)
    with torch.no_grad():
        out = sc(x)
        dt = _timed(lambda: sc(x), args.steps, args.warmup)
    return {'metric': 'signals/s TimeFrequencyScattering J=6 J_fr=2 Q=8 N=4096 fp32', 'value': batch / dt,
            'unit': 'signals/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': dt * 1e3,
            'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
            'config': {'workload': 'JTFS J=6 J_fr=2 Q=8 on %d chirp signals of length 4096 (seed %d); output %s'
                                   % (batch, seed, tuple(out.shape)), 'batch': batch},
            'roofline': None, 'cpu_baseline': None}


def bench_recon(args, batch=256, seed=99, tol=1e-3, time_limit=20.0):
    """Reconstruction from scattering coefficients, Scattering1D J=6 Q=8
    N=4096, 256 unseen chirp signals: wall time until the relative
    scattering error ||S(y) - S(x)|| / ||S(x)|| of the batch (un-squared)
    reaches tol, from the trained decoder's output (decoder forward included)
    and from N(0, 1) noise, by the same descent through the hand-written
    backward for both starts, each stopped at time_limit seconds: per-signal
    L-BFGS (the fastest method measured here for either start) and Adam with
    step 0.1.  value is the warm start's L-BFGS time to tol (null if it did
    not get there); the line carries each run's time to the intermediate
    levels and the plain / warm ratios."""
    from kymatio_amd.learned.evaluate import compare, load_decoder
    from kymatio_amd.learned.train import CHECKPOINT
    from kymatio_amd.torch import Scattering1D
    dev = torch.device('cuda:0')
    sc = Scattering1D(6, 4096, 8).to(dev)
    dec = load_decoder(CHECKPOINT, dev)
    runs = {}
    for method, lr in (('lbfgs', None), ('adam', 0.1)):
        res = compare(sc, dec, n=batch, seed=seed, tol=tol, method=method, lr=lr, time_limit=time_limit)
        for k in ('warm', 'plain'):
            res[k].pop('curve', None)
        runs[method] = {k: res[k] for k in ('warm', 'plain', 'speedup', 'speedup_by_level')}
        runs[method]['speedup_at_least'] = res.get('speedup_at_least')
        start_rel, t_dec = res['start_rel'], res['decoder_forward_s']
    w = runs['lbfgs']['warm']
    return {'metric': 'recon time to 1e-3: seconds until ||S(y)-S(x)||/||S(x)|| <= 1e-3, learned warm start, '
                      'Scattering1D J=6 Q=8 N=4096, batch 256',
            'value': w['seconds'] if w['reached'] else None, 'unit': 's', 'n_gpus': 1, 'steps': 1, 'warmup': 0,
            'ms_per_step': None, 'higher_is_better': False, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
            'data': 'synthetic',
            'config': {'workload': '%d unseen chirp signals (seed %d) of length 4096: 1-8 linear chirps + 0.05 '
                                   'N(0,1) noise; descent through the HIP Scattering1D backward, %g s limit per '
                                   'start and method' % (batch, seed, time_limit),
                       'batch': batch, 'tol': tol, 'loss_definition': 'relative scattering error of the batch, '
                                                                      '||S(y)-S(x)||_F / ||S(x)||_F (un-squared)'},
            'recon': {'start_rel': start_rel, 'decoder_forward_s': t_dec, 'runs': runs},
            'roofline': None, 'cpu_baseline': None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--chunk', type=int, default=0, help='images per launch chunk (0 = default)')
    ap.add_argument('--no-autotune', action='store_true', help='use the heuristic plan choices')
    ap.add_argument('--tune', default='', metavar='V,V,...',
                    help='fixed plan choice (the autotuner fields, comma separated) instead of autotuning')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='BYTES',
                    help='only run a plain device-to-device copy of BYTES (calibrates HBM counters)')
    ap.add_argument('--workload', default='scat2d', choices=('scat2d', 'mixed', 'scat1d', 'grad2d', 'jtfs', 'recon'),
                    help='scat2d: the headline (default); mixed / scat1d / grad2d: the other parts of the path '
                         '(one JSON line each, same keys)')
    args = ap.parse_args()
    if args.copy_calibration:
        n = args.copy_calibration
        src = torch.ones(n, dtype=torch.uint8, device='cuda')
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes': n}))
        return
    if args.gpus != 1:
        print('bench.py measures one GPU only (the scattering path shards over the batch with no '
              'exchange between GPUs, so one GPU\'s throughput is the per-GPU figure). Run with --gpus 1.',
              file=sys.stderr)
        sys.exit(2)
    if args.workload != 'scat2d':
        res = {'mixed': bench_mixed, 'scat1d': bench_scat1d, 'grad2d': bench_grad2d,
               'jtfs': bench_jtfs, 'recon': bench_recon}[args.workload](args)
        print(json.dumps(res))
        return

    from kymatio_amd import workloads
    from kymatio_amd.torch import Scattering2D

    dev = torch.device('cuda:0')
    x = torch.from_numpy(workloads.grating_images(BATCH, SIDE, seed=SEED)).to(dev)
    sc = Scattering2D(J, (SIDE, SIDE), L=L).to(dev)
    tuning = None
    if args.tune:
        from kymatio_amd.scattering2d import autotune
        tuning = tuple(int(v) for v in args.tune.split(','))
        autotune._apply(sc, tuning)
    elif not args.no_autotune:
        # plan choices timed on this device before the timed region (excluded, like plan creation)
        from kymatio_amd.scattering2d import autotune
        tuning, _ = autotune.tune(sc, BATCH)
    if args.chunk:
        sc.images_per_chunk = args.chunk
    with torch.no_grad():
        for _ in range(max(1, args.warmup)):
            out = sc(x)
        torch.cuda.synchronize()
        plan = sc._plan(dev, torch.float32)
        launches = plan.launches()
        # timed region: K forwards, HIP events around every launch
        plan.profile(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = sc(x)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        ms, cnt = plan.profile_read()
        plan.profile(False)
    assert out.shape == (BATCH, 217, 32, 32)
    value = BATCH * args.steps / elapsed

    # per launch: model bytes per launch = model bytes per image * images of the launch
    chunk = sc._chunk(plan, BATCH)
    # images per launch: chunks are split over the plan's streams (tune field 5, 0 = 2) while each gets
    # one, the chunk count rounded up to a multiple of the streams (as the library's chunk_streams)
    streams = (sc.tune[5] if len(sc.tune or ()) > 5 else 0) or 2
    while streams > 1 and (chunk < streams or BATCH <= chunk // streams * (streams - 1)):
        streams -= 1
    per_launch = chunk
    if streams > 1:
        n = -(-BATCH // (chunk // streams))
        n = -(-n // streams) * streams
        per_launch = -(-BATCH // n)
    per_kernel = []
    for i, l in enumerate(launches):
        if cnt[i] == 0:
            continue
        mean_ms = ms[i] / cnt[i]
        imgs = BATCH * args.steps / cnt[i]  # images per launch
        bytes_per_launch = l['model_bytes_per_image'] * imgs
        per_kernel.append({'launch': i, 'kind': l['kind'], 'res': l['res'], 'nodes': l['nodes'],
                           'total_ms': ms[i], 'launches': cnt[i], 'mean_us': 1e3 * mean_ms,
                           'model_bytes_per_launch': bytes_per_launch,
                           'gbps': bytes_per_launch / (mean_ms * 1e-3) / 1e9})
    dom = max(per_kernel, key=lambda r: r['total_ms'])
    achieved = dom['model_bytes_per_launch'] / (dom['mean_us'] * 1e-6)
    isolated = isolated_launch(sc, x, dev, dom['launch'], per_launch)
    isolated['frac'] = dom['model_bytes_per_launch'] / (isolated['mean_launch_us'] * 1e-6) / HBM_PEAK
    per_node_image = measured_traffic_per_node_image()
    traffic = None
    if per_node_image is not None and dom['kind'] == 'lds_node' and dom['res'] == 1:
        traffic = per_node_image * dom['nodes'] * BATCH * args.steps / dom['launches']
    total_model = sum(l['model_bytes_per_image'] for l in launches) * BATCH
    gpu_ms = sum(r['total_ms'] for r in per_kernel) / args.steps
    res = {
        'metric': METRIC, 'value': value, 'unit': 'images/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': 1e3 * elapsed / args.steps, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'Scattering2D J=3 L=8 max_order=2 on 256x256 float32 images: N(0,1) noise '
                               'plus 3 random gratings; output (64, 217, 32, 32)',
                   'batch': BATCH, 'J': J, 'L': L, 'shape': [SIDE, SIDE], 'images_per_chunk': chunk,
                   'streams': streams, 'images_per_launch': per_launch,
                   'seed': SEED,
                   'plan_tuning': None if tuning is None else
                   dict(zip(autotune.FIELDS, tuple(tuning) + (0,) * (len(autotune.FIELDS) - len(tuning))))},
        'roofline': {'bound': 'hbm', 'achieved': achieved / 1e9, 'peak': HBM_PEAK / 1e9, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK, 'traffic': traffic,
                     'traffic_source': 'profiles/scat2d_j3_256_hbm_traffic.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, '
                                       'copy-calibrated), bytes per launch' if traffic is not None else None,
                     'kernel': '%s (res %d, %d nodes/image)' % (dom['kind'], dom['res'], dom['nodes']),
                     'mean_launch_us': dom['mean_us'], 'model_bytes_per_launch': dom['model_bytes_per_launch'],
                     'isolated': isolated},
        'pipeline_roofline': {'model_bytes_per_image': total_model / BATCH,
                              'achieved_GBps': total_model / (elapsed / args.steps) / 1e9,
                              'frac': total_model / (elapsed / args.steps) / HBM_PEAK,
                              'gpu_busy_ms_per_step': gpu_ms},
        'kernels': [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()}
                    for r in sorted(per_kernel, key=lambda r: -r['total_ms'])[:8]],
    }
    res['roofline']['measured_copy_GBps'] = copy_bandwidth(dev)
    res['parity'] = parity_block(dev)
    if not args.no_cpu_baseline:
        res['cpu_baseline'] = cpu_baseline()
    print(json.dumps(res))


def parity_block(dev):
    """Parity of the headline path, measured after the timed region: the FP32
    HIP forward on the four reference-written 256x256 J=3 images
    (tests/golden: the reference's torch-backend outputs) and, per output
    order, against the FP64 HIP plan (spectral leaves; tested against the FP64
    NumPy restatement to 1e-12)."""
    from kymatio_amd.torch import Scattering2D
    gdir = os.path.join(ROOT, 'tests', 'golden')
    try:
        a = np.load(os.path.join(gdir, 'golden_scat2d.npz'))
        b = np.load(os.path.join(gdir, 'golden_headline2d.npz'))
    except OSError:
        return None
    x = np.concatenate([a['cfg2_x'], b['x']])
    ref = np.concatenate([a['cfg2_Sx'], b['Sx']]).astype(np.float64)
    with torch.no_grad():
        out = Scattering2D(J, (SIDE, SIDE), L=L).to(dev)(torch.from_numpy(x).to(dev)).cpu().double().numpy()
        out64 = Scattering2D(J, (SIDE, SIDE), L=L).double().to(dev)(
            torch.from_numpy(x.astype(np.float64)).to(dev)).cpu().numpy()

    def chan_err(o, r):
        o, r = o.reshape(-1, o.shape[-2] * o.shape[-1]), r.reshape(-1, r.shape[-2] * r.shape[-1])
        scale = np.abs(r).max(axis=1)
        scale[scale == 0] = 1.0
        return float((np.abs(o - r).max(axis=1) / scale).max())

    n1 = J * L
    orders = np.array([0] + [1] * n1 + [2] * (out.shape[1] - 1 - n1))
    return {'images': int(x.shape[0]),
            'reference_per_channel_max_rel': chan_err(out, ref),
            'reference_max_rel_of_peak': float(np.abs(out - ref).max() / np.abs(ref).max()),
            'reference_allclose_1e-5_1e-8_fraction': float(np.isclose(out, ref, rtol=1e-5, atol=1e-8).mean()),
            'fp64_per_channel_max_rel_by_order': [chan_err(out[:, orders == o], out64[:, orders == o])
                                                  for o in (0, 1, 2)],
            'bar': 'per output channel max|ours - ref| <= 1e-5 max|ref|'}


def isolated_launch(sc, x, dev, index, per_launch, steps=5):
    """Mean duration of launch `index` with the same images per launch but
    every chunk on one stream, so no other launch shares the GPU with it
    (measured after the timed region; the timed run overlaps two streams)."""
    saved = (sc.tune, sc.images_per_chunk)
    tune = list(sc.tune or [])
    tune += [0] * (6 - len(tune))
    tune[5] = 1
    sc.tune, sc.images_per_chunk = tune, per_launch
    try:
        with torch.no_grad():
            sc(x)
            plan = sc._plan(dev, torch.float32)
            torch.cuda.synchronize()
            plan.profile(True)
            for _ in range(steps):
                sc(x)
            torch.cuda.synchronize()
            ms, cnt = plan.profile_read()
            plan.profile(False)
    finally:
        sc.tune, sc.images_per_chunk = saved
    return {'mean_launch_us': 1e3 * ms[index] / cnt[index], 'streams': 1, 'images_per_launch': per_launch}


def copy_bandwidth(dev, nbytes=1 << 30, reps=5):
    """Best device-to-device copy rate (bytes read + written per second, GB/s)
    of a plain nbytes copy on dev: the attainable HBM rate beside the 8 TB/s
    peak.  Measured after the timed region."""
    src = torch.empty(nbytes, dtype=torch.uint8, device=dev).fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = float('inf')
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    del src, dst
    return 2 * nbytes / best / 1e9


if __name__ == '__main__':
    main()
