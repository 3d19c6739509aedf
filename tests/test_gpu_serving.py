"""Serving loop: results equal the one-shot calls, alone and in a mix.

Tolerance: exact (the server runs the same generation, pack and march kernels
with the same parameters; nothing depends on what else is in flight).  The
fiber-model requests of the serving stream (seeded 1-3-bundle models at the
stream's volume sizes) are also checked against the FP64 oracle through the
one-shot path they equal: labels and axes bit-exact, intensities within 1e-4
relative.
"""

import os
import sys
import warnings

import numpy as np
import pytest

import oracle
from fastpli_amd import synthetic
from fastpli_amd.simulation import Simpli, SimulationServer
from tests.conftest import GOLDEN_ROT, GOLDEN_TILTS
from tests.test_gpu_simulate import assert_intensity_close

pytestmark = pytest.mark.gpu

TILTS5 = np.deg2rad([(0, 0), (5.5, 0), (5.5, 90), (5.5, 180), (5.5, 270)])
ROT9 = np.linspace(0, np.pi, 9, endpoint=False)
THREADS = min(16, os.cpu_count() or 1)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _simpli(voxel, interp='Slerp', rot=GOLDEN_ROT, storage='f32'):
    s = Simpli()
    s.voxel_size = voxel
    s.filter_rotations = rot
    s.light_intensity = 26000
    s.wavelength = 525
    s.interpolate = interp
    s.field_storage = storage
    return s


@pytest.fixture(scope='module')
def tissues(golden):
    import torch
    out = {'cube': (golden['tissue'], golden['optical_axis'], golden['tissue_properties'], 5.0)}
    s = Simpli()
    s.voxel_size = 0.2
    s.dim = [60, 50, 40]
    s.fiber_bundles = synthetic.crossing_bundles(seed=5, n_fibers=120, size_xy=12.0, size_z=8.0)
    with pytest.warns(UserWarning):
        lab, vec, props = s.generate_tissue()
    out['cross'] = (lab, vec, props, 0.2)
    torch.cuda.synchronize()
    return out


def _eager(tissue, axis, props, voxel, tilts, interp='Slerp', rot=GOLDEN_ROT, storage='f32'):
    s = _simpli(voxel, interp, rot, storage)
    s.dim = np.array(tissue.shape)
    return np.asarray(s.run_simulation_tilts(tissue, axis, props, tilts))


def test_serving_alone_equals_run_simulation(tissues):
    _say('serving: single request vs run_simulation_tilts')
    lab, vec, props, voxel = tissues['cube']
    ref = _eager(lab, vec, props, voxel, GOLDEN_TILTS)
    with SimulationServer(max_inflight=2) as srv:
        tid = srv.add_tissue(lab, vec)
        t = srv.submit(tid, _simpli(voxel), props, GOLDEN_TILTS)
        img, lat = srv.wait(t, return_latency=True)
        assert np.array_equal(img.cpu().numpy(), ref)
        assert lat > 0
        st = srv.stats()
        assert st['completed'] == 1 and st['graphs'] == 1


def test_serving_mix_equals_alone(tissues):
    """Uneven sizes, tilt sets, interpolations and layer parameters in flight
    together; identical requests replay cached graphs; eager one-shot
    simulations run between submissions and waits."""
    import torch
    _say('serving: mixed stream vs individual results')
    rng = np.random.default_rng(7)
    cases = []
    for i in range(24):
        name = 'cube' if i % 3 == 0 else 'cross'
        lab, vec, props, voxel = tissues[name]
        tilts = TILTS5 if i % 2 == 0 else TILTS5[:1]
        interp = ['Slerp', 'NN', 'Lerp'][i % 3]
        p = np.array(props, dtype=np.float64)
        if i % 4 == 3:  # other layer parameters
            p[1:, 0] *= rng.uniform(0.5, 1.5)
            p[1:, 1] *= rng.uniform(0.5, 1.5)
        cases.append((name, tilts, interp, p))
    refs = [None] * len(cases)
    with SimulationServer(max_inflight=4) as srv:
        ids = {k: srv.add_tissue(v[0], v[1]) for k, v in tissues.items()}
        tickets = []
        for i, (name, tilts, interp, p) in enumerate(cases):
            voxel = tissues[name][3]
            tickets.append(srv.submit(ids[name], _simpli(voxel, interp), p, tilts))
            if i % 5 == 0:  # eager work between submissions
                lab, vec, _, _ = tissues[name]
                refs[i] = _eager(lab, vec, p, voxel, tilts, interp)
        for i, (name, tilts, interp, p) in enumerate(cases):
            img = srv.wait(tickets[i]).cpu().numpy()
            if refs[i] is None:
                lab, vec, _, voxel = tissues[name]
                refs[i] = _eager(lab, vec, p, voxel, tilts, interp)
            assert np.array_equal(img, refs[i]), f'request {i} ({name}, {interp}) differs'
        # replays of one graph, with eager work and synchronisation in between
        name, tilts, interp, p = cases[0]
        for rep in range(4):
            t = srv.submit(ids[name], _simpli(tissues[name][3], interp), p, tilts)
            lab, vec, _, voxel = tissues[name]
            again = _eager(lab, vec, p, voxel, tilts, interp)
            torch.cuda.synchronize()
            assert np.array_equal(srv.wait(t).cpu().numpy(), refs[0])
            assert np.array_equal(again, refs[0])
        st = srv.stats()
        _say(f'serving stats {st}')
        assert st['completed'] == len(cases) + 4
        assert st['graphs'] < st['replays']


def test_serving_precision_storage(tissues):
    _say('serving: f16 storage tissue')
    lab, vec, props, voxel = tissues['cross']
    ref = _eager(lab, vec, props, voxel, TILTS5, storage='f16')
    with SimulationServer(max_inflight=2) as srv:
        tid = srv.add_tissue(lab, vec, field_storage='f16')
        img = srv.wait(srv.submit(tid, _simpli(voxel), props, TILTS5))
        assert np.array_equal(img.cpu().numpy(), ref)


def test_serving_errors(tissues):
    _say('serving: validation errors')
    lab, vec, props, voxel = tissues['cube']
    with SimulationServer(max_inflight=1) as srv:
        tid = srv.add_tissue(lab, vec)
        with pytest.raises(ValueError, match='unknown tissue'):
            srv.submit(tid + 100, _simpli(voxel), props, [(0, 0)])
        with pytest.raises(ValueError, match='label exceed'):
            srv.submit(tid, _simpli(voxel), props[:2], [(0, 0)])
        with pytest.raises(ValueError, match='background'):
            bad = np.array(props)
            bad[0, 0] = 0.1
            srv.submit(tid, _simpli(voxel), bad, [(0, 0)])
        with pytest.raises(ValueError, match='no progress'):
            srv.submit(tid, _simpli(voxel), props, [(np.pi / 2, 0)])
        with pytest.raises(ValueError, match='z-steps per ray'):
            srv.submit(tid, _simpli(voxel), props, [(1.5707963, 0)])
        neg = lab.astype(np.int32).copy()
        neg[0, 0, 0] = -1
        with pytest.raises(ValueError, match='label < 0'):
            srv.add_tissue(neg, vec)
        srv.remove_tissue(tid)
        with pytest.raises(ValueError, match='unknown tissue'):
            srv.submit(tid, _simpli(voxel), props, [(0, 0)])
    # a budget that holds the tissue but not the request's images
    with SimulationServer(hbm_budget=24**3 * 16 + 1000, max_inflight=1) as srv:
        tid = srv.add_tissue(lab, vec)
        t = srv.submit(tid, _simpli(voxel), props, GOLDEN_TILTS)
        with pytest.raises(RuntimeError, match='budget'):
            srv.wait(t)


# ---------------------------------------------------------------- fiber models
def _model_simpli(seed, dim, interp='Slerp', storage='f32', fbs=None):
    s = _simpli(0.2, interp, ROT9, storage)
    s.dim = np.array(dim)
    s.dim_origin = [0.0, 0.0, 0.0]
    s.fiber_bundles = synthetic.serving_model(seed, dim) if fbs is None else fbs
    return s


def _one_shot(s, tilts):
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')  # interpenetrating bundles: "objects overlap"
        tissue, axis, props = s.generate_tissue_device()
        images = s.run_simulation_tilts(tissue, axis, props, tilts)
    return tissue, axis, props, images


def _wait(srv, t):
    with warnings.catch_warnings():
        warnings.simplefilter('ignore')
        return srv.wait(t)


def _busy_gpu(torch, ms=150):
    """Queues about `ms` of matrix products on torch's current stream."""
    a = torch.randn(4096, 4096, device='cuda')
    for _ in range(max(1, ms // 2)):
        a = torch.tanh(a @ a * 1e-3)
    return a


@pytest.mark.parametrize('k', [0, 1])
def test_serving_model_request_against_oracle(k):
    """A served 64^2x60 / 128^2x120 request of the stream: identical to the
    one-shot generate_tissue + run_simulation, which matches the oracle."""
    import torch
    dim = synthetic.SERVE_SIZES[k]
    _say(f'serving: fiber-model request {dim} vs one-shot and oracle')
    s = _model_simpli(10 + k, dim)
    with SimulationServer(max_inflight=2) as srv:
        img = _wait(srv, srv.submit_model(s, TILTS5))
        assert srv.stats()['generated'] == 1
    tissue, axis, props, ref = _one_shot(s, TILTS5)
    assert torch.equal(img, ref)
    label, vector = tissue.cpu().numpy(), axis.cpu().numpy()
    ol, ov, op, _ = oracle.generate(s.fiber_bundles, dim, np.zeros(3), 0.2, n_threads=THREADS)
    assert np.array_equal(label, ol), f'{np.count_nonzero(label != ol)} labels differ'
    assert np.array_equal(vector, ov)
    op[0, 1] = s.background_absorption
    assert np.array_equal(props, op)
    img = img.cpu().numpy()
    for t, (theta, phi) in enumerate(TILTS5):
        ref_o, _ = oracle.simulate(label, vector, props, theta, phi, voxel_size=0.2,
                                   filter_rotations=ROT9, interpolate='Slerp', n_threads=THREADS)
        assert_intensity_close(img[t], ref_o, 26000)


def test_serving_large_model_alone_and_in_mix():
    """A 512^2x300 request gives the same images alone and among smaller
    requests of other sizes, interpolations, storages and tilt sets."""
    import torch
    _say('serving: 512^2x300 request alone and in a mix')
    big = _model_simpli(20, synthetic.SERVE_SIZES[3])
    smalls = []
    for i in range(6):
        sm = _model_simpli(30 + i, synthetic.SERVE_SIZES[i % 3], ['Slerp', 'NN', 'Lerp'][i % 3],
                           ['f32', 'f16', 'f32'][i % 3])
        smalls.append((sm, TILTS5 if i % 2 == 0 else TILTS5[:1]))
    with SimulationServer(max_inflight=4) as srv:
        alone = _wait(srv, srv.submit_model(big, TILTS5)).clone()
        ts = [srv.submit_model(sm, tl) for sm, tl in smalls[:3]]
        tb = srv.submit_model(big, TILTS5)
        ts += [srv.submit_model(sm, tl) for sm, tl in smalls[3:]]
        mixed = _wait(srv, tb)
        assert torch.equal(alone, mixed)
        for (sm, tl), t in zip(smalls, ts):
            assert torch.equal(_wait(srv, t), _one_shot(sm, tl)[3])
        st = srv.stats()
        _say(f'serving stats {st}')
        assert st['generated'] == 8
    assert torch.equal(alone, _one_shot(big, TILTS5)[3])


def test_serving_budget_admits_two_of_four():
    """A budget that holds two requests' slot memory: with the GPU held busy on
    the caller's stream, two requests are in flight and two wait; all four then
    run and equal their one-shot results."""
    import torch
    _say('serving: HBM budget for 2 of 4 requests')
    dim = synthetic.SERVE_SIZES[1]
    geom = synthetic.serving_model(40, dim)
    ss = []
    for i in range(4):  # one geometry (equal footprints), other layer parameters
        fbs = geom.copy()
        fbs.layers = [[(0.75, 0.0, 2.0 + i, 'b'), (1.0, -0.001 - 0.001 * i, 1.0 + 2 * i, 'p')]
                      for _ in range(len(geom))]
        ss.append(_model_simpli(0, dim, fbs=fbs))
    refs = [_one_shot(s, TILTS5)[3] for s in ss]
    with SimulationServer(max_inflight=1) as probe:
        _wait(probe, probe.submit_model(ss[0], TILTS5))
        one = probe.stats()['slot_bytes']
    budget = int(2.6 * one)
    with SimulationServer(hbm_budget=budget, max_inflight=4) as srv:
        _wait(srv, srv.submit_model(ss[0], TILTS5))  # warm one slot
        torch.cuda.synchronize()
        hold = _busy_gpu(torch)
        ts = [srv.submit_model(s, TILTS5) for s in ss]
        seen = []
        for _ in range(2000):
            st = srv.stats()
            seen.append((st['in_flight'], st['pending'], st['slot_bytes']))
            if st['in_flight'] == 2 and st['pending'] == 2:
                break
        assert (2, 2) in [(a, b) for a, b, _ in seen], seen[-5:]
        outs = [_wait(srv, t) for t in ts]
        st = srv.stats()
        assert st['slot_bytes'] <= budget
        assert max(a for a, _, _ in seen) <= 2 and max(c for _, _, c in seen) <= budget
        del hold
    for o, r in zip(outs, refs):
        assert torch.equal(o, r)


def test_serving_output_written_after_callers_work():
    """The caller's pending work on its output buffer (queued on torch's
    stream behind a long computation) runs before the server writes it."""
    import torch
    _say('serving: output buffer with pending work on the caller stream')
    s = _model_simpli(11, synthetic.SERVE_SIZES[0])
    ref = _one_shot(s, TILTS5)[3]
    with SimulationServer(max_inflight=2) as srv:
        out = torch.empty_like(ref)
        torch.cuda.synchronize()
        hold = _busy_gpu(torch)
        out.fill_(-1.0)  # queued behind the products
        got = _wait(srv, srv.submit_model(s, TILTS5, out=out))
        assert got.data_ptr() == out.data_ptr()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        del hold
        # a dropped ticket keeps its output alive until the server is closed
        srv.submit_model(s, TILTS5[:1])
        torch.cuda.synchronize()


def test_serving_model_errors():
    _say('serving: fiber-model request errors')
    s = _model_simpli(12, synthetic.SERVE_SIZES[0])
    with SimulationServer(max_inflight=1) as srv:
        with pytest.raises(ValueError, match='no progress'):
            srv.submit_model(s, [(np.pi / 2, 0)])
        with pytest.raises(ValueError, match='field_storage'):
            s.field_storage = 'f8'
    # a budget below one request's slot memory
    with SimulationServer(hbm_budget=1 << 20, max_inflight=2) as srv:
        t = srv.submit_model(s, TILTS5)
        with pytest.raises(RuntimeError, match='budget'):
            srv.wait(t)
