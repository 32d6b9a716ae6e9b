"""The T-matrix producer on the GPU against the oracle.

- pam_calc_tmatrix (calc_tmatrix for one hydrometeor) against
  oracle.calc_tmatrix: scattering matrix, extinction matrix and emission
  vector within 1e-6 of their largest entry.  The device solves the Gauss
  sums in FP64 and inverts Q in double-double where the reference works in
  binary128; on the oracle this variant reproduces the single-precision
  T matrix bit for bit for size parameters up to about 10 and within 7e-7
  at 12-15 with axis ratios 0.5 or 2.
- The golden rain/snow scenario with T-matrix hydrometeors through pam_run
  against the oracle's whole run: brightness temperatures within 1e-4 K,
  doubling counts identical, hydrometeor extinction within 1e-6.
- The mixed-precision path on the same scenario within 0.05 K.
"""

import ctypes as C

import numpy as np
import pytest

import oracle
from pamtra_amd import _capi
from _scenario import FREQUENCIES, build_pamtra, passive_nml

pytestmark = pytest.mark.gpu

NMU = 16

# (freq, m, phase, dmax, as_ratio, density, canting)
CASES = [
    (35.5e9, 5.5 - 2.8j, 1, [1.0e-3, 2.0e-3], [0.8, 0.8], 1000.0, 0.0),
    (35.5e9, 5.5 - 2.8j, 1, [3.0e-3, 4.5e-3], [0.6, 0.6], 1000.0, 0.0),
    (94.0e9, 3.5 - 1.9j, 1, [4.0e-3, 1.0e-3], [0.7, 1.2], 1000.0, 0.0),
    (89.0e9, 1.78 - 0.003j, -1, [5.0e-3, 2.0e-3], [0.6, 1.4], 200.0, 0.0),
    (183.31e9, 1.78 - 0.004j, -1, [6.0e-3, 3.0e-3], [0.6, 0.6], 100.0, 0.0),
    (10.65e9, 7.0 - 2.5j, 1, [1.0e-3, 0.5e-3], [1.0, 1.0], 1000.0, 0.0),
    (35.5e9, 5.5 - 2.8j, 1, [2.0e-3, 3.0e-3], [0.8, 0.7], 1000.0, 20.0),
    (89.0e9, 1.78 - 0.003j, -1, [4.0e-3, 2.0e-3], [0.6, 1.5], 200.0, 45.0),
    (94.0e9, 3.5 - 1.9j, 1, [3.0e-3, 1.0e-3], [0.7, 0.9], 1000.0, 90.0),
]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _device_calc_tmatrix(f, m, phase, dmax, ar, dens, nd, dd, cant=0.0):
    from pamtra_amd import _lib
    lib = _lib.library()
    ctx = _lib.get_context(0)
    n = len(dmax)
    arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, float), (n,)))
            for a in (dmax, dd, nd, dens, ar, cant)]
    sm, ex, em = np.zeros(2 * NMU * 2 * NMU * 2), np.zeros(4 * NMU), np.zeros(2 * NMU)
    nm = np.zeros(n, np.int32)
    print(f"pam_calc_tmatrix f={f / 1e9:.2f} GHz dmax={dmax} as_ratio={ar} canting={cant}",
          flush=True)
    st = lib.pam_calc_tmatrix(ctx, f, m.real, -m.imag, phase, n, *[_dp(a) for a in arrs],
                              _dp(sm), _dp(ex), _dp(em), nm.ctypes.data_as(C.POINTER(C.c_int32)))
    return (st, sm.reshape((2, NMU, 2, NMU, 2), order="F"), ex.reshape((2, 2, NMU), order="F"),
            em.reshape((2, NMU), order="F"), nm)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("case", CASES)
def test_calc_tmatrix_matches_oracle(case):
    f, m, phase, dmax, ar, dens, cant = case
    nd, dd = [2.0e6, 5.0e5], [1.0e-4, 2.0e-4]
    st, sm, ex, em, nm = _device_calc_tmatrix(f, m, phase, dmax, ar, dens, nd, dd, cant)
    o = oracle.calc_tmatrix(f, m, phase, dmax, dd, nd, dens, ar, cant)
    assert st == 0 and o[0] == 0
    assert np.all(nm >= 4)
    r = [_rel(sm, o[1]), _rel(ex, o[2]), _rel(em, o[3])]
    print(f"nmax {list(nm)}  rel diff smat {r[0]:.2e} ext {r[1]:.2e} emis {r[2]:.2e}", flush=True)
    assert max(r) <= 1e-6


def test_calc_tmatrix_refuses_bad_bins():
    """A bin with positive number density but zero width is fatal, as the
    reference's assertion (tmatrix.f90:163-171)."""
    st = _device_calc_tmatrix(35.5e9, 5.5 - 2.8j, 1, [1e-3], [0.8], 1000.0, [1e6], [0.0])[0]
    assert st != 0


def _scenario(as_ratio, precision="fp64", canting=-99.0):
    pam = build_pamtra()
    pam.df.data["scat_name"][:] = "tmatrix"
    pam.df.data["as_ratio"][:] = as_ratio
    pam.df.data["canting"][:] = canting
    pam.set["precision"] = precision
    return pam


def _gpu(pam, freqs):
    from pamtra_amd import _lib
    lib = _lib.library()
    ctx = _lib.get_context(0)
    m = _capi.marshal(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D, pam.p)
    f, fp = _capi.freqs_array(freqs)
    hr = _capi.alloc_results(m, len(f), diagnostics=True)
    print(f"pam_run tmatrix ncol={m.nx * m.ny} nf={len(f)}", flush=True)
    st = lib.pam_run(ctx, C.byref(m.settings), m.hydros, C.byref(m.hydro4d), C.byref(m.profile),
                     fp, len(f), C.byref(hr.struct))
    out = dict(hr.arrays)
    out["msg"] = _capi.results_message(hr)
    return out, st


def _cpu(pam, freqs):
    out, st, _ = oracle.run(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D, pam.p, freqs,
                            nthreads=8, diagnostics=True)
    return out, st


@pytest.mark.parametrize("as_ratio", [1.0, 0.6])
def test_golden_tmatrix_scenario_vs_oracle(as_ratio):
    pam = _scenario(as_ratio)
    g, st = _gpu(pam, FREQUENCIES)
    c, st2 = _cpu(_scenario(as_ratio), FREQUENCIES)
    assert st == 0 and st2 == 0, g["msg"]
    d = np.abs(g["tb"] - c["tb"])
    print(f"as_ratio {as_ratio}: max |dTB| = {d.max():.3e} K", flush=True)
    assert d.max() <= 1e-4
    np.testing.assert_array_equal(g["ndoubles"], c["ndoubles"])
    np.testing.assert_array_equal(g["nlegen"], c["nlegen"])
    ok = np.isfinite(c["kexttot"])
    np.testing.assert_allclose(g["kexttot"][ok], c["kexttot"][ok], rtol=1e-6, atol=1e-300)


def test_golden_tmatrix_scenario_with_canting_vs_oracle():
    """Canted spheroids (30 degrees): the tilted-frame amplitude path."""
    g, st = _gpu(_scenario(0.6, canting=30.0), FREQUENCIES)
    c, st2 = _cpu(_scenario(0.6, canting=30.0), FREQUENCIES)
    assert st == 0 and st2 == 0, g["msg"]
    d = np.abs(g["tb"] - c["tb"])
    print(f"canting 30: max |dTB| = {d.max():.3e} K", flush=True)
    assert d.max() <= 1e-4
    np.testing.assert_array_equal(g["ndoubles"], c["ndoubles"])


def test_golden_tmatrix_scenario_mixed_precision():
    g, st = _gpu(_scenario(0.6, "mixed"), FREQUENCIES)
    c, st2 = _cpu(_scenario(0.6), FREQUENCIES)
    assert st == 0 and st2 == 0, g["msg"]
    d = np.abs(g["tb"] - c["tb"])
    print(f"mixed: max |dTB| = {d.max():.3e} K", flush=True)
    assert d.max() <= 0.05


def test_tmatrix_mixed_with_mie_is_refused_by_the_library():
    """The C-ABI refuses what the Python marshalling refuses: a call mixing
    'tmatrix' with Legendre-series hydrometeors."""
    from pamtra_amd import _lib
    pam = _scenario(0.6)
    m = _capi.marshal(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D, pam.p)
    m.hydros[0].scat_name = _capi.SCAT_NAMES["mie-sphere"]
    f, fp = _capi.freqs_array(FREQUENCIES)
    hr = _capi.alloc_results(m, len(f))
    st = _lib.library().pam_run(_lib.get_context(0), C.byref(m.settings), m.hydros,
                                C.byref(m.hydro4d), C.byref(m.profile), fp, len(f),
                                C.byref(hr.struct))
    assert st != 0
    assert "tmatrix" in _capi.results_message(hr)
