"""Serving loop: results equal the one-shot run_simulation, alone and in a mix.

Tolerance: exact (the server replays the same pack and march kernels with the
same parameters; nothing depends on what else is in flight).
"""

import sys

import numpy as np
import pytest

from fastpli_amd import synthetic
from fastpli_amd.simulation import Simpli, SimulationServer
from tests.conftest import GOLDEN_ROT, GOLDEN_TILTS

pytestmark = pytest.mark.gpu

TILTS5 = np.deg2rad([(0, 0), (5.5, 0), (5.5, 90), (5.5, 180), (5.5, 270)])


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
