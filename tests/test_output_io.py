"""Output and restart files (lq_save_output / lq_load_restart), host only.

The reference's outputs (src/loquat_io.cu:240-549) cannot be restarted from:
the DP shear modulus is written under a key the reader does not know, the
.dat name ends in a newline, and format 2 writes no .dat.  These tests pin
that every output here reads back bit for bit through the same loaders as an
input project, and that the .vtu files hold the reference's arrays."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from loquat_amd import abi, cases

INPUT_FIELDS = ["np", "npb", "npm", "domain_size_x", "domain_size_y", "domain_size_z", "domain_min_x",
                "domain_min_y", "domain_min_z", "dr", "hdr", "rho0", "cs", "cfl", "constitutive_model",
                "model_par_1", "model_par_2", "model_par_3", "model_par_4", "model_par_5", "acc_x", "acc_y",
                "acc_z", "time_max", "time", "dt", "output_frequency", "output_number", "density_shepard_filter",
                "stress_shepard_filter", "dynamic_time_step", "integration_method", "viscosity_type",
                "artificial_viscosity_alpha", "artificial_pressure", "output_format"]


def _state(case, seed=0):
    """The case with every field filled with awkward floats."""
    rng = np.random.default_rng(seed)
    n = case.np
    arrs = {k: getattr(case, k).copy() for k in ("pospres", "velrhop", "idep", "str", "stn")}
    arrs["velrhop"][:, :3] = rng.normal(0, 1, (n, 3)).astype(np.float32) / 3
    arrs["velrhop"][:, 3] = (2650 * (1 + rng.normal(0, 1e-3, n))).astype(np.float32)
    arrs["pospres"][:, 3] = rng.normal(0, 1e4, n).astype(np.float32)
    arrs["idep"][:, 1] = rng.uniform(0, 0.1, n).astype(np.float32)
    arrs["idep"][:, 2] = rng.uniform(0.5, 0.9, n).astype(np.float32)
    arrs["str"][:] = rng.normal(0, 1e4, (n, 6)).astype(np.float32)
    arrs["stn"][:] = rng.normal(0, 1e-3, (n, 6)).astype(np.float32)
    posd = case.posd + rng.uniform(-1e-4, 1e-4, case.posd.shape)
    arrs["pospres"][:, :3] = posd.astype(np.float32)
    return arrs, posd


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("model", [1, 2, 3])
def test_output_reads_back_bit_for_bit(tmp_path, model):
    case = cases.small_column_3d(seed=5)
    par = case.par.copy()
    par.output_number = 7
    par.time = 0.123456789
    par.constitutive_model = model
    vw = abi.LqVw(**cases.VW_SAND) if model == 3 else None
    if model == 2:
        (par.model_par_1, par.model_par_2, par.model_par_3, par.model_par_4, par.model_par_5) = (
            -101.2, -962.1, -877.3, 1229.0, 0.0)
    arrs, posd = _state(case)
    abi.save_output(str(tmp_path), par, vw, arrs, posd)
    p2, vw2 = abi.load_parameters(str(tmp_path / "output_7"))
    for f in INPUT_FIELDS:
        if model == 3 and f.startswith("model_par"):
            continue  # model 3 keeps its constants in lq_vw_params
        a, b = getattr(par, f), getattr(p2, f)
        assert (np.float32(a).view(np.uint32) == np.float32(b).view(np.uint32)) if isinstance(a, float) else a == b, f
    if model == 3:
        for f, _ in abi.LqVw._fields_:
            assert np.float32(getattr(vw, f)) == np.float32(getattr(vw2, f)), f
    back = abi.load_particles(str(tmp_path / "7"), p2)
    for k in ("pospres", "velrhop", "str", "stn"):
        assert np.array_equal(_bits(back[k]), _bits(arrs[k])), k
    assert np.array_equal(_bits(back["idep"][:, :2]), _bits(arrs["idep"][:, :2]))
    idep = back["idep"].copy()
    posd2 = abi.load_restart(str(tmp_path / "7"), p2, idep)
    assert np.array_equal(posd2, posd)
    assert np.array_equal(_bits(idep[:, 2]), _bits(arrs["idep"][:, 2]))


def test_reference_restart_defects_are_fixed(tmp_path):
    case = cases.small_column_3d(seed=6)
    par = case.par.copy()
    par.output_number = 3
    par.output_format = 2
    arrs, _ = _state(case)
    abi.save_output(str(tmp_path), par, None, arrs)
    text = (tmp_path / "output_3.xml").read_text()
    assert "<elastic-G " in text and "poisson-G" not in text
    names = sorted(os.listdir(tmp_path))
    assert "3.dat" in names and "3.rst" in names          # format 2 writes the restart too
    assert not any("\n" in n for n in names)
    ids = sorted({int(v) for v in arrs["idep"][:, 0]}, reverse=True)
    assert all(f"{i}_3.vtu" in names for i in ids)


def test_vtu_holds_the_reference_arrays(tmp_path):
    case = cases.small_column_3d(seed=7)
    par = case.par.copy()
    par.output_number = 0
    par.output_format = 2
    arrs, _ = _state(case)
    abi.save_output(str(tmp_path), par, None, arrs)
    for i in sorted({int(v) for v in arrs["idep"][:, 0]}):
        sel = np.nonzero(arrs["idep"][:, 0].astype(int) == i)[0]   # input order inside an object
        root = ET.parse(tmp_path / f"{i}_0.vtu").getroot()
        piece = root.find("UnstructuredGrid/Piece")
        assert int(piece.get("NumberOfPoints")) == len(sel)
        data = {d.get("Name"): np.array(d.text.split(), np.float64) for d in root.iter("DataArray")}
        assert np.array_equal(data["Points"].astype(np.float32).reshape(-1, 3), arrs["pospres"][sel, :3])
        assert np.array_equal(data["Pressure"].astype(np.float32), arrs["pospres"][sel, 3])
        assert np.array_equal(data["Velocity"].astype(np.float32).reshape(-1, 3), arrs["velrhop"][sel, :3])
        assert np.array_equal(data["Density"].astype(np.float32), arrs["velrhop"][sel, 3])
        assert np.array_equal(data["PlasticStrain"].astype(np.float32), arrs["idep"][sel, 1])
        assert np.array_equal(data["Stress"].astype(np.float32).reshape(-1, 6), arrs["str"][sel])
        assert np.array_equal(data["Strain"].astype(np.float32).reshape(-1, 6), arrs["stn"][sel])
        assert np.array_equal(data["offsets"], np.arange(1, len(sel) + 1))


def test_output_errors(tmp_path):
    case = cases.small_column_3d(seed=8)
    arrs, _ = _state(case)
    par = case.par.copy()
    with pytest.raises(abi.LqError) as e:
        abi.save_output(str(tmp_path / "missing_dir"), par, None, arrs)
    assert e.value.status == 4
    par3 = par.copy()
    par3.constitutive_model = 3
    with pytest.raises(abi.LqError) as e:
        abi.save_output(str(tmp_path), par3, None, arrs)
    assert e.value.status == 1
    abi.save_output(str(tmp_path), par, None, arrs)
    other = par.copy()
    other.np = par.np + 1
    with pytest.raises(abi.LqError) as e:
        abi.load_restart(str(tmp_path / str(par.output_number)), other)
    assert e.value.status == 5


# Column order of a .dat row, as the reference reads it (src/loquat_io.cu:226)
# and writes it (src/loquat_io.cu:357): pospres xyzw, velrhop xyzw, idep xy,
# str xx yy zz xy xz yz, stn xx yy zz xy xz yz.
DAT_COLUMNS = ([("pospres", k) for k in range(4)] + [("velrhop", k) for k in range(4)] +
               [("idep", k) for k in range(2)] + [("str", k) for k in range(6)] + [("stn", k) for k in range(6)])


def _one_particle_par():
    par = cases.small_column_3d(seed=5).par.copy()
    par.np, par.npb, par.npm = 1, 0, 1
    par.output_number = 3
    return par


def test_dat_row_read_in_the_reference_column_order(tmp_path):
    """A hand-written row of 22 distinct numbers lands field by field where
    the reference's stream extraction puts it."""
    values = [float(k + 1) + 0.25 for k in range(22)]
    (tmp_path / "p.dat").write_text(" ".join(repr(v) for v in values) + "\n")
    arrs = abi.load_particles(str(tmp_path / "p"), _one_particle_par())
    for v, (field, k) in zip(values, DAT_COLUMNS):
        assert arrs[field][0, k] == np.float32(v), (field, k)
    assert arrs["idep"][0, 2] == 0.0 and arrs["idep"][0, 3] == 0.0   # not in the file


def test_dat_row_written_in_the_reference_column_order(tmp_path):
    """lq_save_output writes one row per particle, 22 space-separated numbers
    in the reference's order; idep.z and .w are not written."""
    par = _one_particle_par()
    arrs = {"pospres": np.array([[1.5, 2.5, 3.5, 4.5]], np.float32),
            "velrhop": np.array([[5.5, 6.5, 7.5, 8.5]], np.float32),
            "idep": np.array([[9.5, 10.5, 99.0, 98.0]], np.float32),
            "str": np.array([[11.5, 12.5, 13.5, 14.5, 15.5, 16.5]], np.float32),
            "stn": np.array([[17.5, 18.5, 19.5, 20.5, 21.5, 22.5]], np.float32)}
    abi.save_output(str(tmp_path), par, None, arrs)
    rows = (tmp_path / "3.dat").read_text().splitlines()
    assert len(rows) == 1
    tok = rows[0].split(" ")
    assert len(tok) == 22
    assert [float(t) for t in tok] == [k + 1.5 for k in range(22)]
