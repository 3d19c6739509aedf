"""Serving many simulation requests on one GPU at once.

`SimulationServer` wraps the library's device-resident request loop
(fpl_server_* in include/fastpli_amd.h).  Two kinds of request:

  * ``submit``: the march of a registered tissue (packed once by
    ``add_tissue`` and resident) under the request's own setup (taken from a
    ``Simpli``), property table and tilt list;
  * ``submit_model``: a whole simulation of a ``Simpli``'s own fiber model,
    volume and setup: its tissue is generated in the request's slot and then
    marched, as ``generate_tissue`` followed by ``run_simulation`` of every
    tilt.

Native slot threads admit requests in arrival order as slots and the HBM
budget allow, run each on the slot's stream (the march as the replay of a
captured hipGraph), and report the latency from submit to completion.  A
request's images equal those of the one-shot calls for the same inputs,
whatever else is in flight.  A request starts after the work already queued
on torch's current stream at submit (its output buffer may still be in use
there); the server keeps the output tensor alive until the request is waited
for or the server is closed.

There is no counterpart in the reference, which runs one simulation per call.

    srv = SimulationServer(max_inflight=4)
    t = srv.submit_model(simpli, [(0, 0), (np.deg2rad(5.5), 0)])
    images = srv.wait(t)            # torch float64 (n_tilts, nx, ny, n_rho) on the GPU
"""

import ctypes
import warnings

import numpy as np

from .. import _lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('fastpli_amd needs a HIP device (torch.cuda.is_available() is False)')
    return torch


class Ticket:
    """Handle of a submitted request."""

    __slots__ = ('id', 'images')

    def __init__(self, ticket_id, images):
        self.id = ticket_id
        self.images = images


STAT_NAMES = ('submitted', 'completed', 'in_flight', 'pending', 'graphs', 'replays', 'generated',
              'slot_bytes', 'tissue_bytes')


class SimulationServer:
    """Concurrent simulation requests (registered tissues or fiber models) on one device."""

    def __init__(self, hbm_budget=None, max_inflight=4, device=None):
        torch = _torch()
        self._device = torch.cuda.current_device() if device is None else int(device)
        budget = -1 if hbm_budget is None else int(hbm_budget)
        handle = ctypes.c_void_p()
        msg = ctypes.create_string_buffer(512)
        rc = _lib.lib().fpl_server_create(self._device, budget, int(max_inflight),
                                          ctypes.byref(handle), msg, 512)
        _lib.check(rc, msg)
        self._h = handle
        self._dims = {}
        self._live = {}  # ticket -> output tensor and host arrays, until waited for

    # ------------------------------------------------------------------ tissues
    def add_tissue(self, tissue, optical_axis, field_storage='f32'):
        """Registers a (tissue, optical_axis) pair; returns its id.

        The fields are copied into the march's packed layout on the device, so
        the arguments may be released afterwards.  Label checks that depend
        only on the tissue (label < 0, the storage's label width) run here.
        """
        torch = _torch()
        if field_storage not in _lib.STORAGE:
            raise ValueError(f'field_storage must be one of {list(_lib.STORAGE)}')
        dev = torch.device('cuda', self._device)
        t = torch.as_tensor(tissue, device=dev).to(torch.int32).contiguous()
        v = torch.as_tensor(optical_axis, device=dev).to(torch.float32).contiguous()
        if t.dim() != 3:
            raise ValueError('tissue.ndim != 3')
        if tuple(v.shape) != tuple(t.shape) + (3,):
            raise ValueError('optical_axis.shape != tissue.shape + (3,)')
        torch.cuda.synchronize(dev)  # the fields are read on the server's own stream
        dim = np.array(t.shape, dtype=np.int64)
        tid = ctypes.c_int64(0)
        msg = ctypes.create_string_buffer(512)
        rc = _lib.lib().fpl_server_add_tissue(self._h, dim.ctypes.data, t.data_ptr(),
                                              v.data_ptr(), _lib.STORAGE[field_storage],
                                              ctypes.byref(tid), msg, 512)
        _lib.check(rc, msg)
        self._dims[tid.value] = tuple(int(d) for d in dim)
        return tid.value

    def remove_tissue(self, tissue_id):
        msg = ctypes.create_string_buffer(512)
        rc = _lib.lib().fpl_server_remove_tissue(self._h, int(tissue_id), msg, 512)
        _lib.check(rc, msg)
        self._dims.pop(int(tissue_id), None)

    # ----------------------------------------------------------------- requests
    def _output(self, shape, out):
        torch = _torch()
        if out is None:
            return torch.empty(shape, dtype=torch.float64,
                               device=torch.device('cuda', self._device))
        if tuple(out.shape) != shape or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError(f'out must be a contiguous float64 tensor of shape {shape}')
        return out

    def submit(self, tissue_id, simpli, tissue_properties, tilts, out=None):
        """Queues the march of a registered tissue under `simpli`'s setup.

        tissue_properties: (n, 2) (dn, mu); tilts: (n_tilts, 2) (theta, phi)
        radian.  `out` may be a float64 device tensor (n_tilts, nx, ny, n_rho).
        Returns a Ticket; the validation errors of run_simulation raise here.
        """
        if int(tissue_id) not in self._dims:
            raise ValueError(f'unknown tissue {tissue_id}')
        props = np.array(tissue_properties, dtype=np.float64)
        if props.ndim != 2:
            raise ValueError('tissue_properties.ndim !=2')
        if props.shape[1] != 2:
            raise ValueError('tissue_properties.shape[1] != 2')
        props = np.ascontiguousarray(props)
        tilts = np.ascontiguousarray(np.array(tilts, dtype=np.float64, ndmin=2).reshape(-1, 2))
        setup, rho = simpli._setup()
        nx, ny, _ = self._dims[int(tissue_id)]
        out = self._output((tilts.shape[0], nx, ny, setup.n_rho), out)
        req = _lib.FplRequest(int(tissue_id), ctypes.pointer(setup), props.ctypes.data,
                              props.shape[0], tilts.ctypes.data, tilts.shape[0], out.data_ptr(),
                              _lib._stream())
        tk = ctypes.c_int64(0)
        msg = ctypes.create_string_buffer(512)
        rc = _lib.lib().fpl_server_submit(self._h, ctypes.byref(req), ctypes.byref(tk), msg, 512)
        _lib.check(rc, msg)
        self._live[tk.value] = out
        return Ticket(tk.value, out)

    def submit_model(self, simpli, tilts, out=None, model=None):
        """Queues a whole simulation of `simpli`: generate_tissue of its fiber
        bundles in its volume, then run_simulation of every tilt under its setup
        (field_storage included).  Returns a Ticket; the images are float64
        (n_tilts, nx, ny, n_rho) on the device.

        `model` may give the fiber model already flattened, as
        ``FiberBundles.to_flat()`` returns it (points may live in pinned host
        memory); it then replaces ``simpli.fiber_bundles``.  The server reads the
        point arrays until the request is done and keeps them alive until then.
        """
        simpli._check_volume_input()
        simpli._check_generation_input()
        dim = simpli._dim_checked()
        setup, rho = simpli._setup()
        tilts = np.ascontiguousarray(np.array(tilts, dtype=np.float64, ndmin=2).reshape(-1, 2))
        if model is None:
            model = simpli.fiber_bundles.to_flat()
        points, fiber_ptr, bundle_ptr, layer_ptr, table = model
        points = np.ascontiguousarray(points, dtype=np.float64)
        fiber_ptr = np.ascontiguousarray(fiber_ptr, dtype=np.int64)
        fbs, keep = _lib._bundles(points, fiber_ptr, bundle_ptr, layer_ptr, table)
        out = self._output((tilts.shape[0], dim[0], dim[1], setup.n_rho), out)
        req = _lib.FplModelRequest(_lib._volume(dim, simpli.dim_origin, simpli.voxel_size), fbs,
                                   float(simpli.background_absorption), ctypes.pointer(setup),
                                   tilts.ctypes.data, tilts.shape[0], out.data_ptr(),
                                   _lib._stream())
        tk = ctypes.c_int64(0)
        msg = ctypes.create_string_buffer(512)
        rc = _lib.lib().fpl_server_submit_model(self._h, ctypes.byref(req), ctypes.byref(tk), msg,
                                                512)
        _lib.check(rc, msg)
        self._live[tk.value] = (out, points, fiber_ptr)
        return Ticket(tk.value, out)

    def wait(self, ticket, timeout=None, return_latency=False):
        """Blocks until the request is done; returns its images (device tensor).

        Raises TimeoutError if `timeout` seconds pass first (the ticket stays
        valid), ValueError / RuntimeError if the request failed.  Warnings of
        the generation and the march are raised as in the one-shot calls.
        """
        warn = ctypes.c_uint32(0)
        lat = ctypes.c_double(0)
        msg = ctypes.create_string_buffer(512)
        rc = _lib.lib().fpl_server_wait(self._h, ticket.id,
                                        -1.0 if timeout is None else float(timeout),
                                        ctypes.byref(warn), ctypes.byref(lat), msg, 512)
        if rc == _lib.FPL_E_PENDING:
            raise TimeoutError(msg.value.decode(errors='replace'))
        self._live.pop(ticket.id, None)
        _lib.check(rc, msg)
        _lib.raise_warnings(warn.value)
        _lib.check_images(warn.value)
        if return_latency:
            return ticket.images, lat.value
        return ticket.images

    def stats(self):
        """dict: submitted, completed, in_flight, pending, graphs, replays,
        generated, slot_bytes, tissue_bytes."""
        s = (ctypes.c_int64 * len(STAT_NAMES))()
        _lib.lib().fpl_server_stats(self._h, s)
        return dict(zip(STAT_NAMES, list(s)))

    def close(self):
        """Finishes the requests in flight, fails pending ones, frees the device memory."""
        if self._h:
            _lib.lib().fpl_server_destroy(self._h)
            self._h = None
        self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover - interpreter shutdown
            warnings.warn('SimulationServer was not closed cleanly', RuntimeWarning)
