"""The pyPamtra ASCII profile formats (.lev levels, .lay layers;
reference python/pyPamtra/core.py:339-504 reads them, :651-749 writes
them): a written profile reads back to the values at the file's number
formats, and the forward operator gives the same TB on it."""

import numpy as np
import pytest

import oracle
import pamtra_amd
from pamtra_amd import workloads
from _scenario import passive_nml


def _same_descriptor(pam):
    q = pamtra_amd.pyPamtra()
    for row in pam.df.data:
        q.df.addHydrometeor(tuple(row.tolist()))
    q.nmlSet.update(pam.nmlSet)
    q.set.update(pam.set)
    return q


def _layer_profile(ncol=3, seed=96):
    """config1-like columns given as layer values (the .lay content)."""
    pam, freqs = workloads.config1(ncol=ncol, seed=seed)
    q = _same_descriptor(pam)
    hl, tl, pl, rl = (pam.p[k][:, 0] for k in ("hgt_lev", "temp_lev", "press_lev", "relhum_lev"))
    lay = lambda a: 0.5 * (a[:, 1:] + a[:, :-1])
    q.createProfile(hgt=lay(hl), temp=lay(tl), press=np.sqrt(pl[:, 1:] * pl[:, :-1]),
                    relhum=lay(rl), hydro_q=pam.p["hydro_q"][:, 0], groundtemp=pam.p["groundtemp"][:, 0],
                    sfc_type=np.zeros(ncol, dtype=int), sfc_refl=np.full(ncol, "F"),
                    sfc_salinity=np.full(ncol, 33.0), timestamp=1370908800)
    q.p["hgt_lev"][..., 0] = hl[:, None, 0] if q.p["hgt_lev"].ndim == 3 else hl[:, 0]
    return q, freqs


def _tb(pam, freqs):
    out, st, _ = oracle.run(passive_nml(pam), pam.set, pam.df.data, pam.df.data4D, pam.p, freqs,
                            nthreads=4)
    assert st == 0, out["msg"]
    return out["tb"]


def test_lev_profile_round_trip(tmp_path):
    pam, freqs = workloads.config1(ncol=3, seed=95)
    pam.p["sfc_refl"][:] = "F"  # what lfrac = 0 reads back as
    f = str(tmp_path / "profile.lev")
    pam.writePamtraProfile(f)
    q = _same_descriptor(pam)
    q.readPamtraProfile(f)
    assert q.p["max_nlyrs"] == pam.p["max_nlyrs"] and q._shape3D == pam._shape3D
    np.testing.assert_array_equal(q.p["nlyrs"], pam.p["nlyrs"])
    np.testing.assert_allclose(q.p["obs_height"], pam.p["obs_height"], rtol=1e-6)
    np.testing.assert_allclose(q.p["hgt_lev"], pam.p["hgt_lev"], atol=0.051)
    np.testing.assert_allclose(q.p["temp_lev"], pam.p["temp_lev"], atol=0.0051)
    np.testing.assert_allclose(q.p["press_lev"], pam.p["press_lev"], atol=0.051)
    np.testing.assert_allclose(q.p["relhum_lev"], pam.p["relhum_lev"], atol=5.1e-5)
    np.testing.assert_allclose(q.p["hydro_q"], pam.p["hydro_q"], rtol=5.1e-6, atol=0)
    np.testing.assert_array_equal(q.p["sfc_type"], 0)
    np.testing.assert_array_equal(q.p["sfc_refl"], "F")
    np.testing.assert_allclose(q.p["groundtemp"], pam.p["groundtemp"], rtol=1e-12)
    d = np.max(np.abs(_tb(q, freqs) - _tb(pam, freqs)))
    assert d < 0.02, d  # from the file's 0.01 K / 0.1 m / 0.1 Pa rounding


def test_lay_profile_round_trip(tmp_path):
    pam, freqs = _layer_profile()
    f = str(tmp_path / "profile.lay")
    pam.writePamtraProfile(f)
    q = _same_descriptor(pam)
    q.readPamtraProfile(f)
    for k, tol in (("hgt", 0.051), ("temp", 0.0051), ("press", 0.051), ("relhum", 5.1e-5)):
        np.testing.assert_allclose(q.p[k], pam.p[k], atol=tol)
    assert np.all(np.isnan(q.p["temp_lev"]))  # derived by fillMissing_atmosphere_vars
    d = np.max(np.abs(_tb(q, freqs) - _tb(pam, freqs)))
    assert d < 0.02, d


def _classic_file(path, n_moments):
    """A 2x1-column, 2-layer classic profile; layer z of column c holds
    contents 1e-5 * (10 c + z + 1) * (category + 1)."""
    nh = 4 + n_moments
    lines = ["2013 06 11 1430 2 1 2 1000.0 1000.0"]
    for c in range(2):
        lines += [f"{c + 1} 1", f"{50.0 + c} {7.0 + c} {float(c)} 1.5 -2.5",
                  " ".join(str(0.1 * (k + 1) + c) for k in range(5 + n_moments)),
                  f"{10.0 * c} 101325.0 {288.0 + c} 80.0"]
        for z in range(2):
            qn = [1e-5 * (10 * c + z + 1) * (h + 1) for h in range(nh)]
            if n_moments == 2:
                qn += [1e3 * (10 * c + z + 1) * (h + 1) for h in range(nh)]
            lines.append(" ".join(str(v) for v in [1000.0 * (z + 1) + c, 90000.0 - 10000 * z,
                                                    280.0 - 6 * z, 70.0 - z] + qn))
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("n_moments", [1, 2])
def test_classic_profile_format(tmp_path, n_moments):
    """readClassicPamtraProfile (reference core.py:505-649): header, per
    column indices, position and surface, the paths, the ground level and
    one line per layer; relative humidity given in percent; the categories
    cloud, ice, rain, snow, graupel (and hail) as hydro_q / hydro_n."""
    import calendar
    f = tmp_path / "profile.cla"
    _classic_file(f, n_moments)
    pam = _one_category(3)
    if n_moments == 1:
        pam.readPamtraProfile(str(f))  # any extension but .lev/.lay
    else:
        pam.readClassicPamtraProfile(str(f), n_moments=2)
    p = pam.p
    nh = 4 + n_moments
    assert (p["ngridx"], p["ngridy"], p["max_nlyrs"], p["noutlevels"]) == (2, 1, 2, 2)
    assert p["hydro_q"].shape == (2, 1, 2, nh)
    np.testing.assert_array_equal(p["nlyrs"], 2)
    assert p["unixtime"][0, 0] == calendar.timegm((2013, 6, 11, 14, 30, 0))
    np.testing.assert_array_equal(p["model_i"][:, 0], [1, 2])
    np.testing.assert_array_equal(p["lat"][:, 0], [50.0, 51.0])
    np.testing.assert_array_equal(p["sfc_type"][:, 0], [0, 1])
    np.testing.assert_array_equal(p["wind10v"], -2.5)
    np.testing.assert_allclose(p["iwv"][:, 0], [0.1, 1.1])
    np.testing.assert_allclose(p["gwp"][:, 0], [0.6, 1.6])
    np.testing.assert_allclose(p["hwp"][:, 0], [0.0, 0.0] if n_moments == 1 else [0.7, 1.7])
    np.testing.assert_array_equal(p["hgt_lev"][1, 0], [10.0, 1001.0, 2001.0])
    np.testing.assert_allclose(p["relhum_lev"][0, 0], [0.8, 0.7, 0.69])
    np.testing.assert_array_equal(p["temp_lev"][1, 0], [289.0, 280.0, 274.0])
    for c in range(2):
        for z in range(2):
            np.testing.assert_allclose(p["hydro_q"][c, 0, z],
                                       1e-5 * (10 * c + z + 1) * (np.arange(nh) + 1))
            want_n = 1e3 * (10 * c + z + 1) * (np.arange(nh) + 1) if n_moments == 2 else 0.0
            np.testing.assert_allclose(p["hydro_n"][c, 0, z], want_n)
    assert np.all(np.isnan(p["airturb"]))


def test_observation_heights_become_levels():
    """add_obs_height_to_layer (core.py:1908-1951, rescaleHeights :1161-1288):
    an observation height inside the column becomes a level (one on an
    existing level moves up 1 m), fields are interpolated linearly in
    height and pressure in its logarithm, and the run outputs the level at
    exactly that height."""
    pam, freqs = workloads.config1(ncol=2, seed=97)
    hl0 = pam.p["hgt_lev"].copy()
    tl0, pl0 = pam.p["temp_lev"].copy(), pam.p["press_lev"].copy()
    nx, ny = pam.p["ngridx"], pam.p["ngridy"]
    pam.p["noutlevels"] = 4
    pam._shape3Dout = (nx, ny, 4)
    pam.p["obs_height"] = np.array([[833000.0, 2150.0, 4000.0, 0.0]] * 2).reshape(nx, ny, 4)
    ref = _tb(pam, freqs)
    q = _same_descriptor(pam)
    q.createProfile(**{k: pam.p[k][:, 0] for k in ("hgt_lev", "temp_lev", "press_lev",
                                                   "relhum_lev", "groundtemp")},
                    hydro_q=pam.p["hydro_q"][:, 0], obs_height=pam.p["obs_height"][:, 0],
                    sfc_type=np.zeros(2, dtype=int), sfc_refl=np.full(2, "S"),
                    sfc_salinity=np.full(2, 33.0), timestamp=0)
    q._addObservationLevels()
    assert q.p["max_nlyrs"] == 32  # two heights inserted, top and ground skipped
    hl = q.p["hgt_lev"][0, 0]
    assert 2150.0 in hl and 4001.0 in hl and 4000.0 in hl  # 4000 m is a level: +1 m
    k = int(np.where(hl == 2150.0)[0][0])
    j = int(np.searchsorted(hl0[0, 0], 2150.0))
    w = (2150.0 - hl0[0, 0, j - 1]) / (hl0[0, 0, j] - hl0[0, 0, j - 1])
    np.testing.assert_allclose(q.p["temp_lev"][0, 0, k],
                               tl0[0, 0, j - 1] + w * (tl0[0, 0, j] - tl0[0, 0, j - 1]), rtol=1e-14)
    np.testing.assert_allclose(np.log(q.p["press_lev"][0, 0, k]),
                               np.log(pl0[0, 0, j - 1]) + w * np.log(pl0[0, 0, j] / pl0[0, 0, j - 1]),
                               rtol=1e-13)
    # layers away from the inserted levels keep their mid-height and values
    # (the interpolation hits its knots); the split layers take interpolated
    # contents, so TB changes by up to a few K where a cloud layer is split
    h_old, h_new = pam.p["hgt"][0, 0], q.p["hgt"][0, 0]
    for z in range(len(h_old)):
        hit = np.where(h_new == h_old[z])[0]
        if len(hit):
            np.testing.assert_array_equal(q.p["hydro_q"][0, 0, hit[0]], pam.p["hydro_q"][0, 0, z])
    assert sum(np.any(h_new == h) for h in h_old) >= 27
    q.nmlSet["add_obs_height_to_layer"] = False  # already applied
    got = _tb(q, freqs)
    assert np.all(np.isfinite(got)) and np.all(np.isfinite(ref))


def _one_category(moment_in):
    pam = pamtra_amd.pyPamtra()
    pam.df.addHydrometeor((
        "test_hyd", -99.0, -1, -99.0, -99.0, -99.0, -99.0, -99.0,
        moment_in, 30, "exp", -99.0, -99.0, 8.0e6, -99.0, 1.0e-5, 6.0e-3,
        "mie-sphere", "khvorostyanov01_drops", -99.0,
    ))
    pam = pamtra_amd.importer.createUsStandardProfile(pam, hgt_lev=[0.0, 1000.0, 2000.0])
    for k in ("press", "temp", "relhum"):
        pam.p[k] = 0.5 * (pam.p[k + "_lev"][..., 1:] + pam.p[k + "_lev"][..., :-1])
    return pam


# the moment_in cases of the file format, as the reference's own round-trip
# tests cover them (tests/test_profile_io.py:17-76 of the reference)
@pytest.mark.parametrize("kind", ["lev", "lay"])
@pytest.mark.parametrize("moment_in", [0, 1, 2, 3, 12, 13, 23])
@pytest.mark.parametrize("layer", [0, 1])
def test_every_moment_combination_round_trips(tmp_path, kind, moment_in, layer):
    values = {"hydro_n": 1.23e6, "hydro_reff": 4.56e-4, "hydro_q": 7.89e-4}
    keys = pamtra_amd.pyPamtra._MOMENT_KEYS[moment_in]
    pam = _one_category(moment_in)
    for k in keys:
        pam.p[k][0, 0, layer, 0] = values[k]
    f = str(tmp_path / f"profile.{kind}")
    pam.writePamtraProfile(f)
    back = _one_category(moment_in)
    back.readPamtraProfile(f)
    for k in keys:
        np.testing.assert_allclose(back.p[k][0, 0, layer, 0], values[k], rtol=1e-6)
    for k in set(values) - set(keys):
        assert np.all(np.isnan(back.p[k]))  # not in the file


def _results_from_oracle(pam, freqs, nml):
    """self.r as runPamtra fills it, from the oracle (the writer is host code)."""
    out, st, _ = oracle.run(nml, pam.set, pam.df.data, pam.df.data4D, pam.p, freqs, nthreads=4)
    assert st == 0
    pam.set["freqs"], pam.set["nfreqs"] = list(freqs), len(freqs)
    pam.set["radar_pol"] = nml["radar_polarisation"].split(",")
    pam.set["radar_npol"], pam.set["att_npol"] = len(pam.set["radar_pol"]), 1
    pam.r = {"tb": out["tb"], "angles_deg": out["angles_deg"], "pamtraVersion": "test",
             "pamtraHash": "test", "nmlSettings": nml}
    if nml["active"]:
        pam.r.update(Ze=out["ze"], Att_hydro=out["att_hydro"], Att_atmo=out["att_atmo"],
                     radar_hgt=out["radar_hgt"])
    return pam


@pytest.mark.parametrize("form,magic", [("NETCDF3_CLASSIC", b"CDF\x01"),
                                        ("NETCDF3_64BIT", b"CDF\x02")])
@pytest.mark.parametrize("active", [False, True])
def test_results_netcdf_layout(tmp_path, form, magic, active):
    """writeResultsToNetCDF (reference core.py:2163-2556): the dimensions,
    variables, types, units and fill values of the reference's file, read
    back with an independent netCDF-3 reader."""
    from scipy.io import netcdf_file
    from _scenario import FREQUENCIES, build_pamtra
    pam = build_pamtra()
    nml = dict(pam.nmlSet, active=active, radar_polarisation="NN,HV")
    pam = _results_from_oracle(pam, FREQUENCIES, nml)
    pam.p["iwv"] = np.full(pam._shape2D, 12.5)
    fn = str(tmp_path / "out.nc")
    pam.writeResultsToNetCDF(fn, ncForm=form)
    with open(fn, "rb") as fh:
        assert fh.read(4) == magic
    with netcdf_file(fn, "r", mmap=False) as nc:
        dims = {"grid_x": 1, "grid_y": 1, "frequency": 2, "angles": 32, "outlevels": 2,
                "passive_polarisation": 2}
        names = {"frequency", "grid_x", "grid_y", "angles", "passive_polarisation", "outlevels",
                 "model_i", "model_j", "nlyr", "datatime", "longitude", "latitude", "sfc_type",
                 "tb", "iwv"}
        if active:
            dims.update(heightbins=3, radar_polarisation=2, attenuation_polarisation=1,
                        radar_peak_number=1)
            names |= {"heightbins", "height", "Ze", "Attenuation_Hydrometeors",
                      "Attenuation_Atmosphere"}
        assert dict(nc.dimensions) == dims
        assert set(nc.variables) == names
        tb = nc.variables["tb"]
        assert tb.dimensions == ("grid_x", "grid_y", "outlevels", "angles", "frequency",
                                 "passive_polarisation")
        assert tb.units == b"K" and tb.data.dtype == np.dtype(">f4")
        assert tb._FillValue == np.float32(-9999.0)
        np.testing.assert_array_equal(tb.data, pam.r["tb"].astype(np.float32))
        assert nc.variables["nlyr"].data.dtype == np.dtype(">i4")
        assert nc.variables["nlyr"]._FillValue == -9999
        assert b"".join(nc.variables["passive_polarisation"].data.tolist()) == b"VH"
        np.testing.assert_array_equal(nc.variables["outlevels"].data,
                                      pam.p["obs_height"].astype(np.float32))
        assert nc.variables["datatime"].units == b"seconds since 1970-01-01 00:00:00"
        assert nc.history.startswith(b"Created with pyPamtra (Version: test")
        if active:
            np.testing.assert_array_equal(nc.variables["Ze"].data, pam.r["Ze"].astype(np.float32))
            assert nc.variables["Ze"].units == b"dBz"
            np.testing.assert_array_equal(nc.variables["Attenuation_Hydrometeors"].data,
                                          pam.r["Att_hydro"].astype(np.float32))


def test_results_netcdf_refusals(tmp_path):
    from _scenario import build_pamtra
    pam = build_pamtra()
    with pytest.raises(IOError):
        pam.writeResultsToNetCDF(str(tmp_path / "a.nc"))
    pam.r = {"pamtraVersion": "x"}
    with pytest.raises(NotImplementedError):
        pam.writeResultsToNetCDF(str(tmp_path / "a.nc"), ncForm="NETCDF4")
    with pytest.raises(ValueError):
        pam.writeResultsToNetCDF(str(tmp_path / "a.nc"), ncForm="HDF9")
