"""Puffin-format HDF5 exchange (puffin_amd/h5io.py over libpuffin_h5.so):
reading the files the reference's own 1-D test expects
(test/expected/f1main_expected_*.h5, copied as fixtures), round trips, and
an independent check of what is written with the HDF5 tools' h5dump."""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from f1main_case import load_fixture, module_params
from puffin_amd import h5io
from puffin_amd.params import ModuleParams, scaling, step_size

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FE = os.path.join(GOLDEN, "f1main_expected_electrons.h5")
FF = os.path.join(GOLDEN, "f1main_expected_field.h5")
FI = os.path.join(GOLDEN, "f1main_expected_integrated.h5")
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
KEYS7 = ("x", "y", "z2", "px", "py", "gam", "chi")


# ---------------------------------------------------------------------------
# files the reference wrote
# ---------------------------------------------------------------------------

def test_read_reference_electrons():
    d, info = load_fixture()          # the same file read through h5dump at %.17g
    parts, hi = h5io.read_electrons(FE)
    for c, k in enumerate(KEYS7):
        assert np.array_equal(parts[k], d["electrons"][:, c]), k
    assert (hi.istep, hi.icsteps, hi.iund_cr, hi.iwrite_cr, hi.il) == (31, 30, 2, 2, 1)
    assert hi.zbar_total == pytest.approx(info["zbarTotal"], rel=1e-15)
    assert hi.num_spatial_dims == 1
    assert (hi.nz2, hi.dz2) == (511, info["sLengthOfElmZ2"])
    assert hi.eta == info["eta"] and hi.kappa == info["kappa"] and hi.rho == info["rho"]


def test_read_reference_field_and_integrated():
    d, info = load_fixture()
    ar, ai, hi = h5io.read_field(FF)
    assert np.array_equal(ar, d["aperp"][0]) and np.array_equal(ai, d["aperp"][1])
    assert (hi.nx, hi.ny, hi.nz2, hi.num_spatial_dims, hi.field_mesh) == (1, 1, 511, 1, 0)
    assert hi.icsteps == 30 and hi.step_size == info["sStepSize"]
    assert np.array_equal(h5io.read_dataset(FI, "power"), d["power"])
    assert np.array_equal(h5io.read_dataset(FI, "bunchingFundamental"), d["bunching_fundamental"])
    # the reference's 1-D power is its intensity times transArea (hdf5_puff.f90:203-212)
    inten = h5io.read_dataset(FI, "Intensity")
    np.testing.assert_allclose(inten * info["transArea"], d["power"], rtol=1e-15, atol=0)


def test_reference_scaled_lengths_follow_the_scaling_relations():
    """The relations info_from_module uses, on the reference's own runInfo."""
    _, _, hi = h5io.read_field(FF)
    f = h5io.info_from_module(module_params(), time=hi.time, lambda_w=hi.lambda_w,
                              trans_area=hi.trans_area)
    for k in ("lambda_r", "lg", "lc", "z_total", "trans_area_si"):
        assert getattr(f, k) == pytest.approx(getattr(hi, k), rel=1e-5), k


# ---------------------------------------------------------------------------
# round trips and an independent reader
# ---------------------------------------------------------------------------

def _mp3d():
    eta, kappa, fx, fy = scaling(0.01, 1.0, 100.0, 1.0, 1.0)
    return ModuleParams(nx=5, ny=4, nz2=7, dx=0.3, dy=0.4, dz2=0.05, rho=0.01, eta=eta, kappa=kappa,
                        gamma_r=100.0, aw=1.0, fx=fx, fy=fy, periodic=1)


def _info(mp):
    return h5io.info_from_module(mp, time=0.75, zbar_local=0.25, istep=12, icsteps=40, n_steps=90,
                                 step_size=step_size(mp.rho, 30), lambda_w=0.03, npk_bar=1.5e9,
                                 counters=dict(iund_cr=3, ichic_cr=2, idrift_cr=4, iquad_cr=5,
                                               imodulation_cr=1, il=9, iwrite_cr=7))


def _dump(path, obj, attr=False):
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-w", "0", "-a" if attr else "-d", obj, path],
                         check=True, stdout=subprocess.PIPE, text=True).stdout
    body = out.split("DATA {", 1)[1].split("}", 1)[0]
    body = re.sub(r"\(\d+(,\d+)*\):", " ", body)
    return out, np.array([float(v) for v in re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", body)])


def _header(path):
    return subprocess.run([H5DUMP, "-H", "-A", path], check=True, stdout=subprocess.PIPE, text=True).stdout


def _attr_types(text):
    """{(object, attribute): HDF5 type class} from an h5dump -H -A listing."""
    out, obj, cur = {}, [], None
    for ln in text.splitlines():
        m = re.match(r'\s*(GROUP|DATASET) "([^"]*)"', ln)
        if m:
            obj.append(m.group(2))
            cur = None
        m = re.match(r'\s*ATTRIBUTE "([^"]*)"', ln)
        if m:
            cur = m.group(1)
        m = re.match(r"\s*DATATYPE\s+(\S+)", ln)
        if m and cur:
            out[(obj[-1], cur)] = m.group(1)
            cur = None
    return out


def test_electrons_round_trip_and_h5dump(tmp_path):
    mp = _mp3d()
    rng = np.random.default_rng(4)
    n = 1001
    parts = {k: rng.standard_normal(n) for k in KEYS7}
    info = _info(mp)
    path = str(tmp_path / "beam.h5")
    h5io.write_electrons(path, parts, info)
    back, bi = h5io.read_electrons(path)
    for k in KEYS7:
        assert np.array_equal(back[k], parts[k]), k
    for k, v in info.as_dict().items():
        assert getattr(bi, k) == v, k
    # an independent reader: the (N, 7) table, row-major, one row per particle
    text, vals = _dump(path, "/electrons")
    assert "H5T_IEEE_F64LE" in text and "( 1001, 7 )" in text
    table = vals.reshape(n, 7)
    for c, k in enumerate(KEYS7):
        assert np.array_equal(table[:, c], parts[k]), k
    _, v = _dump(path, "/electrons/zbarTotal", attr=True)
    assert v[0] == 0.75


def test_field_round_trip_and_h5dump(tmp_path):
    mp = _mp3d()
    rng = np.random.default_rng(5)
    nm = mp.nx * mp.ny * mp.nz2
    ar, ai = rng.standard_normal(nm), rng.standard_normal(nm)
    path = str(tmp_path / "field.h5")
    h5io.write_field(path, ar, ai, _info(mp))
    br, bi_, hi = h5io.read_field(path)
    assert np.array_equal(br, ar) and np.array_equal(bi_, ai)
    assert (hi.nx, hi.ny, hi.nz2, hi.num_spatial_dims, hi.field_mesh) == (5, 4, 7, 3, 1)
    text, vals = _dump(path, "/aperp")
    assert "( 2, 7, 4, 5 )" in text               # (component, z2, y, x), x fastest
    a = vals.reshape(2, 7, 4, 5)
    assert np.array_equal(a[0].ravel(), ar) and np.array_equal(a[1].ravel(), ai)
    assert a[0, 2, 1, 3] == ar[3 + 5 * (1 + 4 * 2)]


def test_one_d_files_match_the_reference_layout(tmp_path):
    """Written 1-D files carry every attribute the reference's readers take,
    with the reference's types, and the reference's dataset shapes."""
    d, info = load_fixture()
    mp = module_params()
    f = h5io.info_from_module(mp, time=info["zbarTotal"], istep=31, icsteps=30, n_steps=30,
                              step_size=info["sStepSize"], lambda_w=info["lambda_w"],
                              trans_area=info["transArea"])
    e = {k: d["electrons"][:, c] for c, k in enumerate(KEYS7)}
    pe, pf, pi = (str(tmp_path / n) for n in ("e.h5", "f.h5", "i.h5"))
    h5io.write_electrons(pe, e, f)
    h5io.write_field(pf, d["aperp"][0], d["aperp"][1], f)
    inten = d["power"] / info["transArea"]
    h5io.write_integrated(pi, f, inten, np.vstack([d["bunching_fundamental"]] * 2))
    for mine, ref in ((pe, FE), (pf, FF), (pi, FI)):
        ours, theirs = _attr_types(_header(mine)), _attr_types(_header(ref))
        shared = set(ours) & set(theirs)
        assert len(shared) >= 40
        for key in shared:
            assert ours[key] == theirs[key], key
    # what readH5Beamfile and readH5FieldfileSingleDump read (h5_in.f90:504-519, 775-781)
    need_e = {("electrons", a) for a in ("zbarTotal", "zbarInter", "zbarLocal", "iCsteps", "istep", "iUnd_cr",
                                         "iChic_cr", "iDrift_cr", "iQuad_cr", "iModulation_cr", "iL", "iWrite_cr")}
    need_f = {("aperp", "iCsteps")} | {("runInfo", a) for a in ("sLengthOfElmX", "sLengthOfElmY",
                                                                  "sLengthOfElmZ2", "nX", "nY", "nZ2")}
    assert need_e <= set(_attr_types(_header(pe)))
    assert need_f <= set(_attr_types(_header(pf)))
    assert "( 2, 511 )" in _header(pf) and "( 319, 7 )" in _header(pe)
    np.testing.assert_allclose(h5io.read_dataset(pi, "power"), d["power"], rtol=1e-15)
    assert np.array_equal(h5io.read_dataset(pi, "Intensity"), inten)
    assert np.array_equal(h5io.read_dataset(pi, "bunching2ndHarmonic"), d["bunching_fundamental"])


def test_errors_are_reported(tmp_path):
    with pytest.raises(h5io.H5Error):
        h5io.read_electrons(str(tmp_path / "missing.h5"))
    with pytest.raises(h5io.H5Error):
        h5io.read_field(FE)                       # no aperp dataset
    with pytest.raises(h5io.H5Error):
        h5io.read_dataset(FI, "noSuchDataset")
    st = h5io.lib().puff_h5_read_electrons(os.fsencode(FE), 318, *([None] * 7), None)
    assert st == -3 and b"319" in h5io.lib().puff_h5_last_error()
    mp = _mp3d()
    with pytest.raises(h5io.H5Error):
        h5io.write_field(str(tmp_path / "f.h5"), np.zeros(3), np.zeros(3), _info(mp))
