"""Launch shapes (persistent grids, graph replay) through the C-ABI: every
tuning value gives bitwise the results of the default launch."""

import ctypes as C

import numpy as np
import pytest

from pamtra_amd import _capi, autotune, workloads

pytestmark = pytest.mark.gpu

SHAPES = [
    {"mie_persistent": 2, "doubling_persistent": 2, "adding_persistent": 2, "graph": 0,
     "pipeline": 0},
    {"mie_persistent": 1, "doubling_persistent": 4, "adding_persistent": 1, "graph": 0,
     "pipeline": 0},
    {"mie_persistent": 0, "doubling_persistent": 0, "adding_persistent": 0, "graph": 1,
     "pipeline": 0},
    {"mie_persistent": 2, "doubling_persistent": 2, "adding_persistent": 2, "graph": 1,
     "pipeline": 0},
]


def _context():
    from pamtra_amd import _lib
    lib = _lib.library()
    h = C.c_void_p()
    assert lib.pam_context_create(C.byref(h), 0) == 0
    return lib, h


def _run(lib, ctx, pam, freqs, precision="fp64"):
    sets = dict(pam.set)
    sets["precision"] = precision
    nml = dict(pam.nmlSet)
    nml["active"] = False
    m = _capi.marshal(nml, sets, pam.df.data, pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f), diagnostics=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    assert st == 0, _capi.results_message(hr)
    return {k: np.array(v) for k, v in hr.arrays.items()}


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_launch_shapes_bitwise(precision):
    lib, ctx = _context()
    try:
        pam, freqs = workloads.config2(ncol=40, seed=31)
        print(f"default launch, {precision}", flush=True)
        ref = _run(lib, ctx, pam, freqs, precision)
        for shape in SHAPES:
            autotune.apply(lib, ctx, shape)
            assert autotune.get_tuning(lib, ctx) == shape
            for rep in range(2):  # with graph = 1 the second call replays
                print(f"shape {shape} call {rep}", flush=True)
                out = _run(lib, ctx, pam, freqs, precision)
                np.testing.assert_array_equal(out["tb"], ref["tb"])
                np.testing.assert_array_equal(out["ndoubles"], ref["ndoubles"])
    finally:
        lib.pam_context_destroy(ctx)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pipelined_call_is_bitwise_the_single_stream_call(precision):
    """Column chunks on two streams (each chunk's DSD and Mie beside the
    previous chunk's phase matrix, doubling and adding) give bitwise the
    single-stream result, with persistent grids too."""
    from pamtra_amd import _lib
    lib, ctx = _context()
    try:
        pam, freqs = workloads.config2(ncol=320, seed=33)
        autotune.apply(lib, ctx, {"pipeline": 0})
        print(f"single stream, {precision}", flush=True)
        ref = _run(lib, ctx, pam, freqs, precision)
        for shape in ({"pipeline": 4}, {"pipeline": 3, "mie_persistent": 2,
                                        "doubling_persistent": 1, "adding_persistent": 2}):
            autotune.apply(lib, ctx, shape)
            print(f"pipelined {shape}", flush=True)
            out = _run(lib, ctx, pam, freqs, precision)
            np.testing.assert_array_equal(out["tb"], ref["tb"])
            np.testing.assert_array_equal(out["ndoubles"], ref["ndoubles"])
            np.testing.assert_array_equal(out["kexttot"], ref["kexttot"])
            ms = (C.c_double * 6)()
            lib.pam_last_stage_times(ctx, ms, 6)
            assert all(v > 0 for v in ms[1:]), list(ms)
    finally:
        lib.pam_context_destroy(ctx)


def test_graph_replay_follows_new_inputs():
    """A replayed graph reads the current inputs; a changed shape is captured
    anew."""
    lib, ctx = _context()
    try:
        a, fa = workloads.config2(ncol=24, seed=41)
        b, fb = workloads.config2(ncol=24, seed=42)
        c, fc = workloads.config2(ncol=30, seed=43)
        ref = [_run(lib, ctx, *x) for x in ((a, fa), (b, fb), (c, fc))]
        autotune.apply(lib, ctx, {"graph": 1})
        for i, x in enumerate(((a, fa), (b, fb), (a, fa), (c, fc), (b, fb))):
            j = [0, 1, 0, 2, 1][i]
            print(f"graph call {i} (workload {j})", flush=True)
            out = _run(lib, ctx, *x)
            np.testing.assert_array_equal(out["tb"], ref[j]["tb"])
    finally:
        lib.pam_context_destroy(ctx)


def test_tuning_keys_and_rejects():
    lib, ctx = _context()
    try:
        keys = lib.pam_tuning_keys().decode().split(",")
        assert keys == ["mie_persistent", "doubling_persistent", "adding_persistent", "graph",
                        "pipeline"]
        assert lib.pam_context_get_tuning(ctx, b"pipeline") == 0  # the default
        assert lib.pam_context_set_tuning(ctx, b"graph", 2) != 0
        assert lib.pam_context_set_tuning(ctx, b"pipeline", 17) != 0
        assert lib.pam_context_set_tuning(ctx, b"nope", 1) != 0
        assert lib.pam_context_get_tuning(ctx, b"nope") == -1
        assert lib.pam_context_set_tuning(ctx, b"mie_persistent", 65) != 0
    finally:
        lib.pam_context_destroy(ctx)
