"""Serving loop: many independent explicit advection solves resident on
one GPU (dg_server_* in include/dgops.h). Requests are admitted first come,
first served within a device-memory budget; every step() advances all
admitted solves by one LSRK(5,4) time step with one kernel launch per stage
and degree, and finished solves hand their solution back through
take_result()."""
import ctypes

import torch

from . import _lib as L


def serve_request(degree, cells_per_direction, velocity, center, width,
                  final_time=0.5, courant=0.3):
    r = L.DgServeRequest()
    r.degree = degree
    r.cells_per_direction = cells_per_direction
    for d in range(3):
        r.velocity[d] = velocity[d]
        r.center[d] = center[d]
    r.width = width
    r.final_time = final_time
    r.courant = courant
    return r


def request_config(request):
    """(DgConfig, time step, number of steps) of a request: the same solve
    can be run alone with AdvectionOperation and LSRK(5,4)."""
    cfg = L.DgConfig()
    dt = ctypes.c_double()
    n = ctypes.c_int()
    L.check(L.load().dg_serve_request_config(ctypes.byref(request),
                                             ctypes.byref(cfg),
                                             ctypes.byref(dt),
                                             ctypes.byref(n)))
    return cfg, dt.value, n.value


class Server:
    def __init__(self, hbm_budget_bytes, stream=None):
        self._lib = L.load()
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        h = ctypes.c_void_p()
        L.check(self._lib.dg_server_create(int(hbm_budget_bytes),
                                           ctypes.c_void_p(stream),
                                           ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.dg_server_destroy(h)
            self._h = None

    def submit(self, request):
        i = ctypes.c_int64()
        L.check(self._lib.dg_server_submit(self._h, ctypes.byref(request),
                                           ctypes.byref(i)))
        return i.value

    def step(self, max_steps=1):
        """Admit, advance every active solve by up to max_steps steps
        (returning at the first completion); returns (n_active, n_queued,
        n_finished) after the call."""
        a, q, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.check(self._lib.dg_server_step(self._h, int(max_steps),
                                         ctypes.byref(a), ctypes.byref(q),
                                         ctypes.byref(f)))
        return a.value, q.value, f.value

    def info(self, rid):
        st, nd = ctypes.c_int(), ctypes.c_int64()
        done, tot = ctypes.c_int(), ctypes.c_int()
        L.check(self._lib.dg_server_request_info(
            self._h, rid, ctypes.byref(st), ctypes.byref(nd),
            ctypes.byref(done), ctypes.byref(tot)))
        return dict(state=st.value, n_dofs=nd.value, steps_done=done.value,
                    steps_total=tot.value)

    def take_result(self, rid):
        n = self.info(rid)["n_dofs"]
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        L.check(self._lib.dg_server_take_result(
            self._h, rid, ctypes.c_void_p(out.data_ptr())))
        return out

    def memory(self):
        b = ctypes.c_int64()
        L.check(self._lib.dg_server_memory(self._h, ctypes.byref(b)))
        return b.value
