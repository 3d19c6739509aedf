"""Serving many run_simulation requests on one GPU at once.

`SimulationServer` wraps the library's device-resident request loop
(fpl_server_* in include/fastpli_amd.h).  Tissues are registered once and stay
packed on the device; every request names a tissue and carries its own
simulation setup (taken from a `Simpli`), layer property table and tilt list.
A native worker thread admits requests into stream slots as slots and HBM
budget free up, runs each as a replay of a captured hipGraph, and reports the
latency from submit to completion.  A request's images equal those of
`Simpli.run_simulation_tilts` for the same inputs, whatever else is in flight.

There is no counterpart in the reference, which runs one simulation per call.

    srv = SimulationServer(max_inflight=4)
    tid = srv.add_tissue(tissue, optical_axis)
    t = srv.submit(tid, simpli, tissue_properties, [(0, 0), (np.deg2rad(5.5), 0)])
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
    """Handle of a submitted request; keeps its output buffer alive."""

    __slots__ = ('id', 'images', '_keep')

    def __init__(self, ticket_id, images, keep):
        self.id = ticket_id
        self.images = images
        self._keep = keep


class SimulationServer:
    """Concurrent run_simulation requests over registered tissues (one device)."""

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
    def submit(self, tissue_id, simpli, tissue_properties, tilts, out=None):
        """Queues the march of a registered tissue under `simpli`'s setup.

        tissue_properties: (n, 2) (dn, mu); tilts: (n_tilts, 2) (theta, phi)
        radian.  `out` may be a float64 device tensor (n_tilts, nx, ny, n_rho).
        Returns a Ticket; the validation errors of run_simulation raise here.
        """
        torch = _torch()
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
        shape = (tilts.shape[0], nx, ny, setup.n_rho)
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device=torch.device('cuda',
                                                                              self._device))
        elif tuple(out.shape) != shape or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError(f'out must be a contiguous float64 tensor of shape {shape}')
        req = _lib.FplRequest(int(tissue_id), ctypes.pointer(setup), props.ctypes.data,
                              props.shape[0], tilts.ctypes.data, tilts.shape[0], out.data_ptr())
        tk = ctypes.c_int64(0)
        msg = ctypes.create_string_buffer(512)
        rc = _lib.lib().fpl_server_submit(self._h, ctypes.byref(req), ctypes.byref(tk), msg, 512)
        _lib.check(rc, msg)
        return Ticket(tk.value, out, (rho, props, tilts))

    def wait(self, ticket, timeout=None, return_latency=False):
        """Blocks until the request is done; returns its images (device tensor).

        Raises TimeoutError if `timeout` seconds pass first (the ticket stays
        valid), ValueError / RuntimeError if the request failed.  Warnings of
        the march are raised as in run_simulation.
        """
        warn = ctypes.c_uint32(0)
        lat = ctypes.c_double(0)
        msg = ctypes.create_string_buffer(512)
        rc = _lib.lib().fpl_server_wait(self._h, ticket.id,
                                        -1.0 if timeout is None else float(timeout),
                                        ctypes.byref(warn), ctypes.byref(lat), msg, 512)
        if rc == _lib.FPL_E_PENDING:
            raise TimeoutError(msg.value.decode(errors='replace'))
        _lib.check(rc, msg)
        _lib.raise_warnings(warn.value)
        if return_latency:
            return ticket.images, lat.value
        return ticket.images

    def stats(self):
        """dict: submitted, completed, in_flight, pending, graphs, replays."""
        s = (ctypes.c_int64 * 6)()
        _lib.lib().fpl_server_stats(self._h, s)
        return dict(zip(('submitted', 'completed', 'in_flight', 'pending', 'graphs', 'replays'),
                        list(s)))

    def close(self):
        """Finishes the requests in flight, fails pending ones, frees the device memory."""
        if self._h:
            _lib.lib().fpl_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover - interpreter shutdown
            warnings.warn('SimulationServer was not closed cleanly', RuntimeWarning)
