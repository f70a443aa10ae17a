"""Continuous-batching likelihood server (host mirror of dust2_server_*).

The reference evaluates one likelihood per call from its MCMC driver
(R/monty.R:188-204, dust_likelihood_monty); here many concurrent callers'
requests share the device: each is admitted when its footprint fits the
HBM budget, stays resident, and advances a few data intervals per tick
together with every other resident request.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, dptr, f64
from .interface import pars_matrix

MODELS = {"sir": 0, "seir16": 1, "sirode": 2}


class Server:
    def __init__(self, hbm_bytes=4 << 30, chunk_intervals=8):
        out = C.c_void_p()
        call("dust2_server_create", int(hbm_bytes), int(chunk_intervals), C.byref(out))
        self.ptr = out
        self.chunk = int(chunk_intervals)

    def __del__(self):
        if getattr(self, "ptr", None) is not None and _lib._lib is not None:
            _lib.lib().dust2_server_free(self.ptr)
            self.ptr = None

    def submit(self, generator, pars, n_particles, time_start, data, dt=1.0, seed=0, chain=None):
        """Queue one likelihood evaluation; returns its id.  data: dict with
        'time' and 'incidence' ((n_time,) or (n_time, n_data)).  chain: None
        for streams seeded from `seed`; an MCMC chain index (>= 0) for
        streams that continue from the chain's previous request, the first
        seeded at long_jump^chain of `seed` (R/monty.R:29-36)."""
        row = pars_matrix(generator, pars, 1)[0].copy()
        time = f64(data["time"]).reshape(-1)
        obs = np.ascontiguousarray(f64(data["incidence"]).reshape(time.size, -1))
        if obs.shape[1] != generator.n_data:
            raise ValueError(f"Expected {generator.n_data} observations per time")
        rid = np.zeros(1, dtype=np.uint64)
        call("dust2_server_submit", self.ptr, MODELS[generator.name], dptr(row), int(n_particles),
             float(time_start), dptr(time), time.size, float(dt), dptr(obs),
             C.c_uint64(int(seed)), C.c_int64(-1 if chain is None else int(chain)),
             _lib.u64ptr(rid))
        return int(rid[0])

    def step(self):
        """One scheduling pass (collect completed launches, admit, launch);
        returns (n_resident, n_queued)."""
        a = np.zeros(1, dtype=np.uintp)
        q = np.zeros(1, dtype=np.uintp)
        call("dust2_server_step", self.ptr, _lib.szptr(a), _lib.szptr(q))
        return int(a[0]), int(q[0])

    def poll(self, max_results=1024):
        """Finished results as a list of (id, log_likelihood, status, message)."""
        ids = np.zeros(max_results, dtype=np.uint64)
        ll = np.zeros(max_results)
        st = np.zeros(max_results, dtype=np.int32)
        n = np.zeros(1, dtype=np.uintp)
        buf = C.create_string_buffer(4096)
        rc = _lib.lib().dust2_server_poll(self.ptr, _lib.u64ptr(ids), dptr(ll),
                                          st.ctypes.data_as(C.POINTER(C.c_int)), max_results,
                                          _lib.szptr(n), buf, len(buf))
        if rc != 0:
            raise _lib.DustError(buf.value.decode())
        msg = buf.value.decode()
        return [(int(ids[i]), float(ll[i]), int(st[i]), msg if st[i] else "")
                for i in range(int(n[0]))]

    def memory(self):
        used = np.zeros(1, dtype=np.uintp)
        cap = np.zeros(1, dtype=np.uintp)
        call("dust2_server_memory", self.ptr, _lib.szptr(used), _lib.szptr(cap))
        return int(used[0]), int(cap[0])

    def drain(self, max_ticks=10**7):
        """Step until every submitted request has finished; returns results."""
        out = []
        for _ in range(max_ticks):
            n_active, n_queued = self.step()
            out.extend(self.poll())
            if n_active == 0 and n_queued == 0:
                while True:
                    more = self.poll()
                    if not more:
                        return out
                    out.extend(more)
        raise RuntimeError("server did not drain")
