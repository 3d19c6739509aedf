"""Benchmark: Mueller voxel-steps/s of generate_tissue + run_simulation on one GPU.

Workload (default): two fiber bundles crossing at 90 deg, 2 x 2000 fibers with
0.5 um segments, in 120 x 120 x 60 um at 0.2 um voxels (600 x 600 x 300),
layers [(0.75, 0, 5, 'b'), (1.0, -0.004, 1, 'p')], 5 tilts {(0), (5.5 deg,
phi = 0/90/180/270)} x 9 polarizer rotations, Slerp interpolation.

One step = generate_tissue (labels + optical axes) + run_simulation of all 5
tilts, with the fiber model already resident in HBM (`value`); `api_path`
times the same step through the Simpli API, with and without the model's
flatten + upload, and once through host arrays.  The unit of work is one
Mueller voxel-step = (tilt, ray, z-step, rotation); a step does
9 * (360000 * 300 + 4 * 360000 * 301) = 4.873e9 of them.

Prints one JSON line.  The CPU baseline is the FP64 OpenMP oracle (a port of
the reference's C++ core) timed on the same workload (by default the whole
volume, median of 3 steps after a warm-up; --cpu-fraction bounds it to a
y-slab).  The `serving`
object reports p50 / p99 latency of a mixed-size request stream.
"""

import argparse
import json
import os
import sys
import time
import warnings

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ROT = np.linspace(0, np.pi, 9, endpoint=False)
TILTS = np.deg2rad([(0, 0), (5.5, 0), (5.5, 90), (5.5, 180), (5.5, 270)])
VOXEL = 0.2
VOI = (120.0, 120.0, 60.0)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_simpli(dim, fbs=None):
    from fastpli_amd.simulation import Simpli
    s = Simpli()
    s.voxel_size = VOXEL
    s.dim = np.array(dim, dtype=int)
    s.dim_origin = [0.0, 0.0, 0.0]
    if fbs is not None:
        s.fiber_bundles = fbs
    s.filter_rotations = ROT
    s.light_intensity = 26000
    s.wavelength = 525
    s.interpolate = 'Slerp'
    return s


def voxel_steps(s):
    return sum(s.count_voxel_steps(t, p) for t, p in TILTS)


TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'hbm_traffic_crossing_600x600x300.txt')


def kernel_sources_hash():
    """sha256 prefix of the kernel sources (fastpli_amd/csrc/*.hip, *.h)."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, 'fastpli_amd', 'csrc')
    for name in sorted(os.listdir(d)):
        if name.endswith(('.hip', '.h')):
            with open(os.path.join(d, name), 'rb') as f:
                h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(kernel, storage):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE + WRITE_SIZE summary.

    The summary holds lines `traffic <kernel> <storage> <bytes> <sources hash>`
    (counter KiB scaled by the factors a 1 GiB device copy gives), measured on
    the kernel sources whose hash it names; None when no line matches these
    sources (the counters are stale).
    """
    src = kernel_sources_hash()
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                w = line.split()
                if (len(w) == 5 and w[0] == 'traffic' and w[1] == kernel and w[2] == storage and
                        w[4] == src):
                    return float(w[3])
    except OSError:
        pass
    return None


def adjoint_bench(reps=5):
    """One vMF bundle (400 fibers, 300 x 300 x 150 @ 0.2 um), 5 tilts x 9 rotations,
    NN and Slerp: the forward march and the adjoint kernel (gradients w.r.t.
    the (dn, mu) table and the optical-axis field of L = sum(g * images)),
    HIP-event times.  The adjoint's algorithmic bytes, NN (one thread per ray):
    per ray-step the 16-byte record read three times (forward, replay,
    backward) and a 40-byte checkpoint written and read once per 8 steps.
    Slerp (the queued adjoint): per ray-step the record read twice (forward,
    replay) and a 40-byte checkpoint written and read once per 4 steps; per
    birefringent step a 32-byte step result written once and read twice
    (birefringent steps taken as the ray-steps times the birefringent share
    of the voxels).  Both: per pixel the float64 gradient images read."""
    import torch
    from fastpli_amd import _lib, synthetic
    from fastpli_amd.simulation import simulate_stack
    dim = (300, 300, 150)
    s = make_simpli(dim, synthetic.single_bundle(seed=0))
    dev = torch.device('cuda', 0)
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        label, vector, props = s.generate_tissue()
    lab = torch.from_numpy(label).to(dev)
    bire = float(((props[label, 0] != 0) & (label != 0)).mean())
    out = {'config': 'one vMF(kappa 10) bundle of 400 fibers, 300x300x150 voxels @0.2um, '
                     '5 tilts x 9 rotations',
           'gradients': 'dn, mu per label; optical axis per voxel',
           'birefringent_share': bire}
    for interp in ('NN', 'Slerp'):
        s.interpolate = interp
        axis = torch.from_numpy(vector).to(dev).requires_grad_(True)
        pr = torch.tensor(props, requires_grad=True)
        g = None
        for k in range(reps + 2):
            if k == 2:
                _lib.profile_enable(True)
            img = simulate_stack(s, lab, axis, pr, TILTS)
            if g is None:
                g = torch.randn_like(img)
            (img * g).sum().backward()
            axis.grad = None
            pr.grad = None
        torch.cuda.synchronize()
        fwd, nf = _lib.profile_read('march')
        adj, na = _lib.profile_read('adjoint')
        _lib.profile_enable(False)
        ray_steps = sum(s.count_voxel_steps(t, p) for t, p in TILTS) / ROT.size
        nbytes = len(TILTS) * dim[0] * dim[1] * 8.0 * ROT.size
        if interp == 'NN':
            nbytes += ray_steps * (3 * 16.0 + 2 * 40.0 / 8)
        else:
            nbytes += ray_steps * (2 * 16.0 + 2 * 40.0 / 4) + ray_steps * bire * 3 * 32.0
        adj_ms = adj / max(na, 1)
        out[interp] = {'forward_march_ms': fwd / max(nf, 1), 'adjoint_ms': adj_ms,
                       'adjoint_bytes': nbytes, 'adjoint_GBs': nbytes / (adj_ms * 1e-3) / 1e9}
    return out


def precision_bench(fbs, dim, reps=10):
    """The march in the three field storages on the crossing-bundle volume
    (600 x 600 x 300): the untilted view with NN (the precision form of the
    plain workload, a pure streaming march) and the 5-tilt Slerp step;
    HIP-event times, the storage's bytes (records read once per tilt + float64
    images written) and GB/s."""
    import torch
    from fastpli_amd import _lib
    nx, ny, nz = dim
    n_vox, n_rho = nx * ny * nz, ROT.size
    out = {}
    for storage in ('f32', 'f16', 'angle16'):
        s = make_simpli(dim, fbs)
        s.field_storage = storage
        with warnings.catch_warnings():
            warnings.simplefilter('ignore')
            _, _, props, fields = s.generate_tissue_fields(with_arrays=False)
        rec = {'f32': 16.0, 'f16': 8.0, 'angle16': 6.0}[storage]
        row = {}
        for name, interp, tilts in (('untilted_nn', 'NN', TILTS[:1]),
                                    ('tilts5_slerp', 'Slerp', TILTS)):
            s.interpolate = interp
            for _ in range(2):
                s.run_simulation_fields(fields, props, tilts)
            torch.cuda.synchronize()
            _lib.profile_enable(True)
            for _ in range(reps):
                s.run_simulation_fields(fields, props, tilts)
            torch.cuda.synchronize()
            ms, n = _lib.profile_read('march')
            _lib.profile_enable(False)
            ms /= max(n, 1)
            nbytes = len(tilts) * (rec * n_vox + 8.0 * n_rho * nx * ny)
            row[name] = {'march_ms': ms, 'bytes': nbytes, 'GBs': nbytes / (ms * 1e-3) / 1e9}
        row['bytes_per_voxel'] = rec
        out[storage] = row
        del fields
    return out


def serving_bench(n_requests, load, seed=0, pool=4):
    """Serving stream: each request a whole simulation of its own fiber model.

    Volumes uniform over 64^2x60, 128^2x120, 256^2x200 and 512^2x300 voxels
    (0.2 um); each request a seeded 1-3-bundle fiber model (synthetic.
    serving_model: radii lognormal(1.0, 0.2) um, a geometry from a pool of
    `pool` seeded models per size) with its layers' dn drawn from U[-0.006,
    -0.001] and mu from U[1, 10] per request; 5 tilts or the untilted view
    alone (1:1), 9 rotations, Slerp.  The server generates every request's
    tissue in its slot and marches it (SimulationServer.submit_model).  Arrivals
    are Poisson at `load` x the rate one request at a time would sustain (mean
    service time of every (size, tilt set) measured alone first).  Latency is
    submit -> completion as the server measures it.
    """
    import torch
    from fastpli_amd import synthetic
    from fastpli_amd.simulation import SimulationServer
    rng = np.random.default_rng(seed)
    sizes = synthetic.SERVE_SIZES
    geoms, simplis = [], []
    for k, dim in enumerate(sizes):
        entries = []
        for p in range(pool):
            fbs = synthetic.serving_model(1000 * seed + 10 * k + p, dim)
            points, fptr, bptr, lptr, _ = fbs.to_flat()
            pinned = torch.from_numpy(points).pin_memory().numpy()
            entries.append((pinned, fptr, bptr, lptr))
            if p == 0:
                simplis.append(make_simpli(list(dim), fbs))
        geoms.append(entries)
    work_of = {(k, nt): sum(simplis[k].count_voxel_steps(t, p) for t, p in TILTS[:nt])
               for k in range(len(sizes)) for nt in (1, 5)}

    def layers(nb):
        scale = np.tile([0.75, 1.0], nb)
        dn = np.zeros(2 * nb)
        dn[1::2] = rng.uniform(-0.006, -0.001, size=nb)
        mu = rng.uniform(1.0, 10.0, size=2 * nb)
        model = np.tile(np.array([ord('b'), ord('p')], dtype=np.uint8), nb)
        return scale, dn, mu, model

    def model_for(k):
        pts, fptr, bptr, lptr = geoms[k][int(rng.integers(0, pool))]
        nb = len(bptr) - 1
        return (pts, fptr, bptr, 2 * np.arange(nb + 1, dtype=np.int64), layers(nb))

    def make_request():
        k = int(rng.integers(0, len(sizes)))
        tilts = TILTS if rng.uniform() < 0.5 else TILTS[:1]
        return k, tilts, model_for(k)

    inflight = 4
    srv = SimulationServer(max_inflight=inflight)
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        # every slot warmed for every (size, tilt set): its buffers, generation
        # workspace and captured graphs exist before anything is timed
        for k in range(len(sizes)):
            for nt in (5, 1):
                ts = [srv.submit_model(simplis[k], TILTS[:nt], model=model_for(k))
                      for _ in range(inflight)]
                for t in ts:
                    srv.wait(t)
        # service time of each (size, tilt set) alone: median of 8 requests
        # (their geometries differ: the pool's models hold 1-3 bundles)
        service = {}
        for k in range(len(sizes)):
            for nt in (5, 1):
                times = []
                for _ in range(8):
                    t = srv.submit_model(simplis[k], TILTS[:nt], model=model_for(k))
                    _, lat = srv.wait(t, return_latency=True)
                    times.append(lat)
                service[(k, nt)] = float(np.median(times))
    mean_service = float(np.mean(list(service.values())))
    rate = load / mean_service
    reqs = [make_request() for _ in range(n_requests)]
    gaps = rng.exponential(1.0 / rate, size=n_requests)
    arrivals = np.cumsum(gaps)
    outstanding, lats, work = [], [], 0
    st0 = srv.stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    i = 0
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')  # interpenetrating bundles: "objects overlap"
        while i < n_requests or outstanding:
            now = time.perf_counter() - t0
            while i < n_requests and arrivals[i] <= now:
                k, tilts, model = reqs[i]
                outstanding.append(srv.submit_model(simplis[k], tilts, model=model))
                work += work_of[(k, len(tilts))]
                i += 1
            # block briefly on the oldest request (the wait releases the
            # server's lock), then collect whatever else has finished
            nxt = arrivals[i] - (time.perf_counter() - t0) if i < n_requests else 1e-3
            if outstanding:
                try:
                    _, lat = srv.wait(outstanding[0], timeout=min(max(nxt, 0.0), 2e-4),
                                      return_latency=True)
                    lats.append(lat)
                    outstanding.pop(0)
                except TimeoutError:
                    pass
            elif nxt > 0:
                time.sleep(min(nxt, 1e-3))
            still = []
            for t in outstanding:
                try:
                    _, lat = srv.wait(t, timeout=0.0, return_latency=True)
                    lats.append(lat)
                except TimeoutError:
                    still.append(t)
            outstanding = still
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lats = np.array(lats) * 1e3
    return {
        'requests': n_requests,
        'p50_ms': float(np.percentile(lats, 50)),
        'p99_ms': float(np.percentile(lats, 99)),
        'mean_ms': float(lats.mean()),
        'max_ms': float(lats.max()),
        'arrival_rate_per_s': rate,
        'offered_load': load,
        'mean_service_alone_ms': mean_service * 1e3,
        'service_alone_ms': {f'{sizes[k][0]}^2x{sizes[k][2]}/{nt}t': v * 1e3
                             for (k, nt), v in service.items()},
        'throughput_requests_per_s': n_requests / wall,
        'throughput_voxel_steps_per_s': work / wall,
        'max_inflight': inflight,
        'tissues_generated': st['generated'] - st0['generated'],
        'graphs_captured': st['graphs'],
        'graph_replays': st['replays'] - st0['replays'],
        'mix': 'each request generates its own tissue: seeded 1-3-bundle fiber model '
               '(radius lognormal(1.0, 0.2) um, dn U[-0.006,-0.001], mu U[1,10] per request; '
               f'geometry from {pool} seeded models per size), sizes 64^2x60 / 128^2x120 / '
               '256^2x200 / 512^2x300 uniform; 5 tilts or 1 (1:1); 9 rotations; Slerp; '
               'Poisson arrivals',
    }


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_threads():
    """Every CPU this process may run on, capped by OMP_NUM_THREADS when set
    (the CPU share a GPU job is given)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get('OMP_NUM_THREADS')
    return min(n, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else n


def cpu_baseline(fbs, fraction, threads, reps=3):
    """Oracle (FP64, OpenMP) on the y-slab [0, fraction * 120 um] of the workload
    (fraction 1: the whole volume): one warm-up, then the median of `reps`
    timed steps.  Like the GPU's timed step, the timed region starts with the
    fiber model flattened and the output arrays allocated."""
    import oracle
    dim = (int(round(VOI[0] / VOXEL)), max(1, int(round(VOI[1] * fraction / VOXEL))),
           int(round(VOI[2] / VOXEL)))
    s = make_simpli(dim, fbs)
    oracle.lib()
    flat = fbs.to_flat()
    label = np.zeros(dim, dtype=np.int32)
    vec = np.zeros(tuple(dim) + (3,), dtype=np.float32)
    images = [np.zeros((dim[0], dim[1], ROT.size)) for _ in TILTS]

    def one():
        t0 = time.perf_counter()
        _, _, props, _ = oracle.generate(flat, dim, np.zeros(3), VOXEL, n_threads=threads,
                                         out=(label, vec))
        for (theta, phi), img in zip(TILTS, images):
            oracle.simulate(label, vec, props, theta, phi, voxel_size=VOXEL,
                            filter_rotations=ROT, interpolate='Slerp', n_threads=threads,
                            out=img)
        return time.perf_counter() - t0

    one()
    runs = [one() for _ in range(reps)]
    dt = float(np.median(runs))
    w = voxel_steps(s)
    return {
        'value': w / dt,
        'unit': 'Mueller voxel-steps/s',
        'cores': threads,
        'kind': 'port',
        'cpu_model': cpu_model(),
        'cpus_available': len(os.sched_getaffinity(0)),
        'runs_s': runs,
        'build': 'g++ -O3 -fopenmp -ffp-contract=off (oracle/Makefile)',
        'sample': (('whole volume' if fraction >= 1 else 'y-slab') +
                   f' {dim[0]}x{dim[1]}x{dim[2]} voxels of the same workload '
                   f'(generate + 5 tilts x 9 rotations, Slerp), {w:.3e} voxel-steps, '
                   f'median of {reps} steps after a warm-up, FP64 OpenMP oracle, fiber model '
                   'flattened and outputs allocated before timing'),
    }


def api_path_bench(fbs, dim, storage, steps, work):
    """The step through the Simpli API with device tensors:
    generate_tissue_fields (label + optical axis + march records) +
    run_simulation_fields (5 tilts).  'cached': the fiber model unchanged
    between steps (flattened and uploaded once); 'model_edited': the model
    counts as edited before every step, so every step flattens it
    (FiberBundles.to_flat) and uploads its points.  Also one step through the
    numpy API (generate_tissue + run_simulation_tilts on host arrays: the
    fields cross PCIe twice)."""
    import torch
    from fastpli_amd.objects._sealed import edited
    s = make_simpli(dim, fbs)
    s.field_storage = storage

    def step():
        _, _, props, fields = s.generate_tissue_fields()
        return s.run_simulation_fields(fields, props, TILTS)

    out = {}
    for name, edit in (('cached', False), ('model_edited', True)):
        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            if edit:
                edited()
            step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        out[name] = {'ms_per_step': dt * 1e3, 'value': work / dt}
    t0 = time.perf_counter()
    label, vector, props = s.generate_tissue()
    s.run_simulation_tilts(label, vector, props, TILTS)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out['numpy_host_arrays'] = {'ms_per_step': dt * 1e3, 'value': work / dt, 'steps': 1}
    out['unit'] = 'Mueller voxel-steps/s'
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--cpu-fraction', type=float, default=1.0,
                    help='y fraction of the volume timed for the CPU baseline')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--field-storage', default='f32', choices=['f32', 'f16', 'angle16'],
                    help='storage of the fields inside the march (precision modes)')
    ap.add_argument('--no-adjoint', action='store_true',
                    help='skip the adjoint timing (vMF bundle forward + reverse march)')
    ap.add_argument('--no-serving', action='store_true',
                    help='skip the mixed-size serving stream (p50 / p99 latency)')
    ap.add_argument('--no-precision', action='store_true',
                    help='skip the march timings of the three field storages')
    ap.add_argument('--no-api-path', action='store_true',
                    help='skip the step timed through the Simpli API')
    ap.add_argument('--serve-requests', type=int, default=1500)
    ap.add_argument('--serve-load', type=float, default=0.7,
                    help='arrival rate as a fraction of the one-at-a-time service rate')
    ap.add_argument('--copy-calibration', action='store_true',
                    help='after timing, copy 1 GiB device to device once (the reference kernel '
                         'that scales FETCH_SIZE / WRITE_SIZE counters to bytes)')
    args = ap.parse_args()

    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (--gpus {args.gpus} requested): the single-GPU '
              'design holds the whole volume on one device and has no multi-GPU path.',
              file=sys.stderr)
        sys.exit(2)

    import torch
    from fastpli_amd import _lib, synthetic

    assert torch.cuda.is_available(), 'no HIP device'
    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)

    log('building fiber model')
    fbs = synthetic.crossing_bundles(seed=0)
    dim = [int(round(v / VOXEL)) for v in VOI]
    s = make_simpli(dim, fbs)
    s.field_storage = args.field_storage
    work = voxel_steps(s)
    nx, ny, nz = dim
    n_vox = nx * ny * nz
    n_rho = ROT.size

    points, fiber_ptr, bundle_ptr, layer_ptr, table = fbs.to_flat()
    points_d = torch.from_numpy(points).to(dev)
    fiber_ptr_d = torch.from_numpy(fiber_ptr).to(dev)
    label = torch.empty(dim, dtype=torch.int32, device=dev)
    vector = torch.empty(dim + [3], dtype=torch.float32, device=dev)
    images = torch.empty((len(TILTS), nx, ny, n_rho), dtype=torch.float64, device=dev)
    setup, _keep = s._setup()
    fields_bytes, _ = _lib.fields_layout(dim, args.field_storage)
    fields_buf = torch.empty(fields_bytes, dtype=torch.uint8, device=dev)
    fields = _lib.fields_struct(fields_buf, dim, args.field_storage)

    def step():
        # generate_tissue: label + optical axis (the API's arrays) and the
        # march's records, written by the voxelizer; run_simulation: 5 tilts
        props, _ = _lib.generate_fields(dim, s.dim_origin, VOXEL, points_d, fiber_ptr_d,
                                        bundle_ptr, layer_ptr, table, label, vector, fields)
        props[0, 1] = s.background_absorption
        _lib.simulate_fields(setup, fields, props, TILTS, images)

    log(f'warmup {args.warmup}')
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    log(f'timing {args.steps} steps')
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    _lib.profile_enable(False)
    ms_step = dt / args.steps * 1e3

    kernels = {}
    alg_bytes = {
        # generation writes int32 label + 3 x f32 vector per voxel
        'voxelize': 16.0 * n_vox,
        # each tilt reads label + vector per voxel once, writes n_rho f64 per pixel
        'march': len(TILTS) * (16.0 * n_vox + 8.0 * n_rho * nx * ny),
    }
    for name in ('voxelize', 'march', 'pack'):
        ms, n = _lib.profile_read(name)
        if name == 'pack':
            if n:
                kernels[name] = {'avg_ms': ms / n, 'launches': n}
            continue
        if n:
            kernels[name] = {'avg_ms': ms / n, 'launches': n,
                             'algorithmic_bytes': alg_bytes[name],
                             'achieved_GBs': alg_bytes[name] / (ms / n * 1e-3) / 1e9}
    dominant = max(alg_bytes, key=lambda k: kernels[k]['avg_ms'])
    kd = kernels[dominant]
    roofline = {
        'bound': 'hbm',
        'achieved': kd['achieved_GBs'],
        'peak': HBM_PEAK_GBS,
        'unit': 'GB/s',
        'frac': kd['achieved_GBs'] / HBM_PEAK_GBS,
        'traffic': measured_traffic(dominant, args.field_storage),
        'kernel': dominant,
    }

    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    adjoint = None
    if not args.no_adjoint:
        log('adjoint: vMF bundle forward + reverse march')
        adjoint = adjoint_bench()

    precision = None
    if not args.no_precision:
        log('precision storages: f32 / f16 / angle16 march')
        precision = precision_bench(fbs, dim)

    serving = None
    if not args.no_serving:
        log(f'serving stream: {args.serve_requests} requests at load {args.serve_load}')
        serving = serving_bench(args.serve_requests, args.serve_load)

    api = None
    if not args.no_api_path:
        log('API path: Simpli.generate_tissue_fields + run_simulation_fields')
        api = api_path_bench(fbs, dim, args.field_storage, max(3, args.steps // 2), work)

    cpu = None
    if not args.no_cpu_baseline:
        threads = cpu_threads()
        log(f'cpu baseline: oracle, {threads} threads, fraction {args.cpu_fraction}')
        cpu = cpu_baseline(fbs, args.cpu_fraction, threads)

    if args.field_storage != 'f32':
        # the march reads the packed storage: 8 (f16) or 6 (angle16) B/voxel
        kernels['march']['storage_bytes'] = len(TILTS) * (
            {'f16': 8.0, 'angle16': 6.0}[args.field_storage] * n_vox + 8.0 * n_rho * nx * ny)
    out = {
        'metric': 'Mueller voxel-steps/s (5 tilts x 9 rot), generate_tissue + run_simulation',
        'value': work / (dt / args.steps),
        'unit': 'Mueller voxel-steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {
            'workload': 'crossing bundles 2x2000 fibers, 600x600x300 voxels @0.2um, '
                        '5 tilts x 9 rotations, Slerp',
            'voxels': n_vox,
            'voxel_steps_per_step': work,
            'tilts': len(TILTS),
            'rotations': n_rho,
            'interpolate': 'Slerp',
            'parallelism': 'single GPU',
            'field_storage': args.field_storage,
        },
        'roofline': roofline,
        # the step's algorithmic bytes (voxelizer writes + march reads and
        # writes) over its wall time, against the HBM peak
        'whole_step': {'algorithmic_bytes': alg_bytes['voxelize'] + alg_bytes['march'],
                       'achieved_GBs': (alg_bytes['voxelize'] + alg_bytes['march']) /
                       (ms_step * 1e-3) / 1e9,
                       'frac': (alg_bytes['voxelize'] + alg_bytes['march']) /
                       (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        'api_path': api,
        'kernels': kernels,
        'cpu_baseline': cpu,
        'serving': serving,
        'adjoint': adjoint,
        'precision': precision,
    }
    if cpu:
        out['speedup_vs_cpu_baseline'] = out['value'] / cpu['value']
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
