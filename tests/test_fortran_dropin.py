"""The Fortran drop-in layer (ising_amd/fortran): the reference's entry points
called from Fortran through iso_c_binding.

CPU tests: the library and programs are built and stop with the library's
error text when no device is present. GPU tests: the check program's values
against closed forms / the oracle's identities, and the batched PT driver
with the reference's command line and CSV format."""
import csv
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

LIB = Path(__file__).resolve().parents[1] / "ising_amd" / "lib"
needs_fortran = pytest.mark.skipif(not (LIB / "test_dropin").exists(), reason="Fortran layer not built (no flang)")


def run(cmd, env=None, timeout=120):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=e)


def test_binding_declares_every_header_function():
    """ising_gpu_c.f90 gives Fortran callers every entry point of the header."""
    import re
    root = Path(__file__).resolve().parents[1]
    header = (root / "include" / "ising_gpu.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(ising_\w+)\s*\(", header, re.M))
    binding = (root / "ising_amd" / "fortran" / "ising_gpu_c.f90").read_text()
    bound = set(re.findall(r'bind\(C, name="(ising_\w+)"\)', binding))
    assert declared and declared <= bound, sorted(declared - bound)


@needs_fortran
def test_driver_rejects_bad_cli():
    p = run([str(LIB / "pt_gpu"), "1", "2"])
    assert p.returncode != 0
    assert "Jmin Jmax NJ Dmin Dmax ND Bmin Bmax NB outputFilePath" in (p.stdout + p.stderr)


@needs_fortran
def test_dropin_reports_missing_device_or_runs():
    import ctypes as C
    from ising_amd import _capi
    n = C.c_int()
    _capi.load().ising_device_count(C.byref(n))
    if n.value > 0:
        pytest.skip("device present: covered by the gpu tests")
    p = run([str(LIB / "test_dropin")])
    assert p.returncode != 0
    assert "no HIP device" in (p.stdout + p.stderr)


def parse(out):
    vals = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) >= 2:
            try:
                vals[parts[0]] = [float(x) for x in parts[1:]]
            except ValueError:
                pass
    return vals


@pytest.mark.gpu
@needs_fortran
def test_dropin_entry_points_on_gpu():
    p = run([str(LIB / "test_dropin")])
    assert p.returncode == 0, p.stdout + p.stderr
    v = parse(p.stdout)
    assert "done" in p.stdout
    # spins are stored in FP32 on the device: the closed form holds to FP32 rounding
    assert abs(v["spiral_energy"][0] - v["spiral_expect"][0]) < 2e-7 * abs(v["spiral_expect"][0])
    assert abs(v["local_delta"][0] - v["total_delta"][0]) < 1e-5 * (1 + abs(v["total_delta"][0]))
    assert v["max_norm_dev"][0] < 1e-6
    assert v["energy_after"][0] < v["spiral_energy"][0] + 200        # sweeps ran and stayed finite
    assert np.allclose(v["fm_mag"], [0, 0, 1]) and v["fm_winding"][0] == 0.0
    # skyrmion_number_middle: the signed count of compute_skyrmion_distribution
    assert abs(abs(v["sk_winding_middle"][0]) - 1) < 1e-9
    assert v["sk_number"][0] == round(v["sk_winding_middle"][0])
    assert abs(v["sk_layer1"][0] - v["sk_winding_middle"][0]) < 1e-12
    # calculate_winding_number_density of the same skyrmion: sums to the layer's
    # charge, peaks at the plaquette whose corner cell is next to the centre
    # (7.7, 7.8) -> 1-based (8, 8) or its neighbours
    assert abs(v["sk_density_sum"][0] - v["sk_winding_middle"][0]) < 1e-12
    assert 0 < v["sk_density_max"][0] < 1
    assert all(abs(c - 7.5) <= 1.5 for c in v["sk_density_peak"])
    # hyper-kagome mesh through the same entry points
    assert v["hk_atoms"][0] == 12 * 3 * 3 * 2
    assert abs(v["hk_fm_energy"][0] - v["hk_fm_energy"][1]) < 1e-12 * abs(v["hk_fm_energy"][1])
    assert v["hk_fm_winding"][0] == 0.0 and v["hk_norm_dev"][0] < 1e-6


@pytest.mark.gpu
@needs_fortran
def test_dropin_demagnetisation_on_gpu(tmp_path, orc):
    """calculate_demagnetisation_field, TotalHeisenbergEnergy and
    calculateHeisenbergEnergy with demag = .true., and Metropolis_mcs's field
    schedule, against the oracle's restatement (bcc 4 x 3 x 3)."""
    p = run([str(LIB / "test_dropin"), str(tmp_path)])
    assert p.returncode == 0, p.stdout + p.stderr
    v = parse(p.stdout)
    s = np.loadtxt(tmp_path / "demag_in.txt").T.astype(np.float32).astype(np.float64)
    H = np.loadtxt(tmp_path / "demag_field.txt").T
    L = orc.lattice(4, 3, 3, cell="bcc")
    want = orc.demag_field(L, np.ascontiguousarray(s))
    assert np.max(np.abs(H - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))
    assert np.max(np.abs(want)) > 1e-3
    term = -0.5 * orc.MU0_REF * np.sum(s * want)
    assert abs(v["demag_energy_term"][0] - term) < 1e-6 * abs(term) + 1e-12
    assert v["demag_array_vs_field"][0] < 1e-13
    # local energies with the field argument H = (2e5, -1e5, 3e5) differ by
    # -mu0 S.H of atom 5 (1-based); FP32 local field: 1e-5 of its scale
    sp = np.array([0.0, 0.6, -0.8])
    hb = np.array([2e5, -1e5, 3e5])
    d_old = v["demag_local"][0] - v["nodemag_local"][0]
    d_new = v["demag_local"][1] - v["nodemag_local"][1]
    assert abs(d_old - (-orc.MU0_REF * s[:, 4] @ hb)) < 1e-5 * 10
    assert abs(d_new - (-orc.MU0_REF * sp @ hb)) < 1e-5 * 10
    assert abs(d_new - d_old) > 0.1
    assert v["demag_after_sweeps"][0] < 1e-13
    assert v["demag_norm_dev"][0] < 1e-6


@pytest.mark.gpu
@needs_fortran
def test_pt_driver_reference_cli_and_csv(tmp_path):
    out = tmp_path / "run"
    env = {"ISING_PT_ITERATIONS": "4", "ISING_PT_SWEEPS_PER_SWAP": "10", "ISING_PT_RELAX": "20",
           "ISING_PT_INTERP_SCALE": "2"}
    p = run([str(LIB / "pt_gpu"), "-1.0", "1.0", "2", "0.5", "0.5", "1", "0.1", "0.1", "1", str(out)], env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    files = sorted(out.glob("data_*.csv"))
    assert len(files) == 2 * 10                      # 2 parameter slots x 10 temperatures
    rows = list(csv.reader(files[0].open()))
    assert rows[0] == ["J", "D", "B", "T", "winding_number_middle", "skyrmion_number_middle",
                       "winding_number_spread", "sx", "sy", "sz"]
    vals = [float(x) for x in rows[1]]
    assert vals[0] in (-1.0, 0.0) and vals[1] == 0.5 and vals[2] == 0.1
    assert 400.0 <= vals[3] <= 20000.0 and abs(vals[7] ** 2 + vals[8] ** 2 + vals[9] ** 2) <= 1.0001
    spins = sorted((tmp_path / "run_spins").glob("spins_*.csv"))
    assert len(spins) == 20
    data = np.loadtxt(spins[0], delimiter=",", skiprows=1)
    # bcc cells of edge 2 (src/unit_cells.f90:21): body atoms at +1 on every axis
    assert data.shape == (30 * 30 * 6 * 2, 6)
    assert np.allclose(np.linalg.norm(data[:, 3:], axis=1), 1.0, atol=1e-6)
    assert np.allclose(data[0::2, :3] % 2.0, 0.0) and np.allclose(data[1::2, :3] % 2.0, 1.0)
    assert np.allclose(data[1::2, :3] - data[0::2, :3], 1.0)
    # write_winding_number_density: x,y,z,Winding_Density per plaquette, layer
    # sums equal to the layer charge (middle layer in the data file, F0.4)
    dens = sorted((tmp_path / "run_density").glob("density_*.csv"))
    assert len(dens) == 20 and [d.name[8:] for d in dens] == [f.name[5:] for f in files]
    assert dens[0].read_text().splitlines()[0] == "x,y,z,Winding_Density"
    for df, f in zip(dens, files):
        d = np.loadtxt(df, delimiter=",", skiprows=1)
        assert d.shape == (30 * 30 * 6, 4)
        assert np.allclose(d[:, :3] / 2.0, np.rint(d[:, :3] / 2.0), atol=1e-6)
        assert d[:, 2].min() == pytest.approx(2.0) and d[:, 2].max() == pytest.approx(6 * 2.0)
        mid = d[np.isclose(d[:, 2], 3 * 2.0), 3].sum()
        row = [float(x) for x in list(csv.reader(f.open()))[1]]
        assert abs(mid - row[4]) < 1e-4 + 1e-6 * 900
    # write_winding_number_density_interpolated: every layer on the 2x finer
    # grid (fine indices times the cell edge); each layer's total is an
    # integer (the charge of the interpolated field); skyrmion_number_middle
    # is a count (an integer field of the CSV)
    interp = sorted((tmp_path / "run_density_interpolated").glob("density_interp_*.csv"))
    assert len(interp) == 20 and [d.name[15:] for d in interp] == [f.name[5:] for f in files]
    for df, f in zip(interp[:3], files[:3]):
        d = np.loadtxt(df, delimiter=",", skiprows=1)
        assert d.shape == (60 * 60 * 6, 4)
        assert d[:, 0].max() == pytest.approx(60 * 2.0) and d[:, 1].max() == pytest.approx(60 * 2.0)
        for z in range(1, 7):
            tot = d[np.isclose(d[:, 2], 2.0 * z), 3].sum()
            assert abs(tot - round(tot)) < 1e-4 + 1e-8 * 3600, (z, tot)
        row = list(csv.reader(f.open()))[1]
        assert row[5].lstrip("-").isdigit()
