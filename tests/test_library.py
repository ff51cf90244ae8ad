"""CPU checks of the C ABI library: it loads and exports every entry point
that include/dgops.h declares (no GPU calls)."""
import ctypes
import os
import re

from advection_miniapp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dgops.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows all of them
    assert set(names) == set(_lib.SIGNATURES), (
        set(names) ^ set(_lib.SIGNATURES))


def test_config_default_and_validation_without_gpu():
    lib = _lib.load()
    cfg = _lib.DgConfig()
    lib.dg_config_default(ctypes.byref(cfg))
    assert cfg.dim == 3 and cfg.degree == 4 and cfg.flux_alpha == 1.0
    cfg.degree = 11
    h = ctypes.c_void_p()
    # rejected before any device call
    assert lib.dg_create(ctypes.byref(cfg), None, ctypes.byref(h)) == \
        _lib.DG_ERR_UNSUPPORTED
    assert b"degree" in lib.dg_last_error()
    cfg.degree = 4
    cfg.refine = 11  # 2^33 cells: beyond the 32-bit cell index
    assert lib.dg_create(ctypes.byref(cfg), None, ctypes.byref(h)) == \
        _lib.DG_ERR_UNSUPPORTED
    assert b"too large" in lib.dg_last_error()
    # SIPG and Burgers are periodic-only, as in the reference
    cfg.refine = 2
    for kind, ncomp in ((_lib.DG_OP_ADVECTION_DIFFUSION, 1),
                        (_lib.DG_OP_BURGERS, 3)):
        cfg.operator_kind, cfg.n_components = kind, ncomp
        cfg.periodic[1] = 0
        assert lib.dg_create(ctypes.byref(cfg), None, ctypes.byref(h)) == \
            _lib.DG_ERR_UNSUPPORTED
        assert b"periodic meshes only" in lib.dg_last_error()
        cfg.periodic[1] = 1
    assert lib.dg_lsrk_n_stages(_lib.DG_LSRK_5_4) == 5
    assert lib.dg_lsrk_n_stages(_lib.DG_LSRK_9_5) == 9
    assert lib.dg_lsrk_n_stages(17) == -1


def test_null_handle_is_an_error():
    lib = _lib.load()
    assert lib.dg_apply(None, None, None, 0.0) == _lib.DG_ERR_ARGUMENT


def test_time_step_rules():
    """Configuration 1: h_min = 2/32, |a| = |(1, 0.15)|, k = 3, Courant 0.4,
    T = 2 gives 421 steps; the implicit rule C h_min / 2."""
    import numpy as np
    from advection_miniapp_amd import explicit_time_step, implicit_time_step
    dt = explicit_time_step(2 / 32, np.hypot(1.0, 0.15), 3, 0.4, 2.0)
    assert round(2.0 / dt) == 421 and abs(421 * dt - 2.0) < 1e-13
    dt = implicit_time_step(1 / 32, 5.0, 1.0)
    assert round(1.0 / dt) == 13 and dt <= 5.0 / 64


def test_host_index_map_tool_matches_oracle():
    """tests/cpp/index_map (host only, no GPU): cell centres in vector order
    and dof offsets equal the oracle's refine_global mesh."""
    import os
    import subprocess
    import numpy as np
    from helpers import Cfg
    tool = os.path.join(os.path.dirname(__file__), "cpp", "index_map")
    assert os.path.exists(tool), "run __graft_entry__.build() first"
    out = subprocess.run([tool, "3", "4", "2", "0.1", "42"], check=True,
                         capture_output=True, text=True).stdout
    rows = np.array([[float(x) for x in line.split()]
                     for line in out.splitlines() if not line.startswith("#")])
    o = Cfg(dim=3, degree=4, refine=2, perturbation=0.1).oracle_handle()
    v = o.cell_vertices().reshape(o.n_cells, 8, 3)
    assert rows.shape == (o.n_cells, 5)
    assert np.array_equal(rows[:, 0], np.arange(o.n_cells))
    assert np.max(np.abs(rows[:, 1:4] - v.mean(axis=1))) < 1e-15
    assert np.array_equal(rows[:, 4], np.arange(o.n_cells) * 125)


def test_serve_request_config_without_gpu():
    """Request -> operator configuration (host only): 24 cells per direction
    are 3^3 coarse cells refined 3 times; the time step follows the explicit
    rule and divides the end time."""
    import numpy as np
    from advection_miniapp_amd.serving import request_config, serve_request
    r = serve_request(4, 24, [0.6, 0.0, 0.8], [0.5, 0.5, 0.5], 0.1,
                      final_time=0.5, courant=0.3)
    cfg, dt, n = request_config(r)
    assert (cfg.dim, cfg.degree, cfg.refine) == (3, 4, 3)
    assert list(cfg.n_coarse) == [3, 3, 3] and list(cfg.periodic) == [1] * 3
    dt0 = 0.3 * (1 / 24) / 1.0 / 4 ** 1.5
    assert n == int(np.ceil(0.5 / dt0)) and abs(n * dt - 0.5) < 1e-14
    import pytest
    bad = serve_request(4, 12 * 11, [1, 0, 0], [0.5] * 3, 0.1)
    with pytest.raises(_lib.DgError):
        request_config(bad)


def test_failed_serving_request_is_released_with_its_error():
    """Without a GPU every admission fails at its first device call: the
    step reports the error, the request is FAILED and holds no memory, and
    take_result releases it and returns the same error (CPU only; the
    server itself makes no device call before admission)."""
    lib = _lib.load()
    srv = ctypes.c_void_p()
    assert lib.dg_server_create(1 << 30, None, ctypes.byref(srv)) == _lib.DG_OK
    r = _lib.DgServeRequest()
    r.degree, r.cells_per_direction = 3, 8
    r.velocity[0], r.width, r.final_time, r.courant = 1.0, 0.1, 0.1, 0.3
    rid = ctypes.c_int64()
    assert lib.dg_server_submit(srv, ctypes.byref(r), ctypes.byref(rid)) == 0
    a, q, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = lib.dg_server_step(srv, 1, ctypes.byref(a), ctypes.byref(q),
                            ctypes.byref(f))
    assert rc != _lib.DG_OK
    step_error = lib.dg_last_error()
    st, nd = ctypes.c_int(), ctypes.c_int64()
    done, tot = ctypes.c_int(), ctypes.c_int()
    assert lib.dg_server_request_info(srv, rid.value, ctypes.byref(st),
                                      ctypes.byref(nd), ctypes.byref(done),
                                      ctypes.byref(tot)) == 0
    assert st.value == _lib.DG_REQUEST_FAILED
    mem = ctypes.c_int64()
    assert lib.dg_server_memory(srv, ctypes.byref(mem)) == 0
    assert mem.value == 0
    assert lib.dg_server_take_result(srv, rid.value, None) == \
        _lib.DG_ERR_ARGUMENT                       # null destination
    dst = ctypes.c_double()
    assert lib.dg_server_take_result(srv, rid.value,
                                     ctypes.byref(dst)) == rc
    assert lib.dg_last_error().endswith(step_error)
    lib.dg_server_request_info(srv, rid.value, ctypes.byref(st), None, None,
                               None)
    assert st.value == _lib.DG_REQUEST_UNKNOWN
    assert lib.dg_server_destroy(srv) == 0
