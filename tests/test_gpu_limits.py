"""Size limits and empty calls.

* A replica of 2^24 sites or more takes the site number's full 32-bit
  multiplies (site_number switches from 24-bit multiplies there): a
  simple-cubic 256^3 replica (exactly 2^24 sites) and a hyper-kagome
  replica of 112^3 cells (16,859,136 sites) replay one sweep on the oracle
  with the same Philox counters. At this size decisions within FP32 error
  of the threshold are expected on about one site per million, and each
  changes the fields of its neighbours in the second colour: the bar is
  2e-5 N sites (a systematic error would reach most of N).
* Lattices of 2^31 sites or more are refused before any allocation.
* Zero sweeps and zero single-site energy changes are no-ops.
"""
import numpy as np
import pytest

from gpu_util import engine, olat
from ising_amd import Hamiltonian, IsingError, Lattice
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

HAM = Hamiltonian(J=1.0, D=0.5, dmi="bulk", Bz=0.2, Kz=0.05)


@pytest.mark.parametrize("lat", [Lattice(256, 256, 256), Lattice(112, 112, 112, cell="hyperkagome")],
                         ids=["sc_256", "hk_112"])
def test_replica_of_2_24_sites_or_more(lat):
    assert lat.sites >= 1 << 24
    beta = 1.4
    with engine(lat, HAM, [beta], seed=17) as e:
        e.init_random()
        s0 = e.get_spins(0)
        e.sweep(1)
        s1 = e.get_spins(0)
        en = e.energy()[0]
    lo, ho = olat(lat), orc.ham_from(HAM)
    orc.set_threads(16)
    assert np.max(np.abs(s0 - orc.init_random(lo, 17, 0))) < 2e-6
    s = s0.copy()
    orc.sweeps(lo, ho, s, beta, 0.5, 17, 0, 0, 1)
    diff = np.max(np.abs(s - s1), axis=0)
    bad = int(np.sum(diff > 1e-3))
    assert bad <= 2e-5 * lat.sites, bad
    assert np.median(diff) < 1e-6
    assert abs(en - orc.total_energy(lo, ho, s1)) < 1e-11 * lat.sites


@pytest.mark.parametrize("lat", [Lattice(2048, 1024, 1024), Lattice(600, 600, 600, cell="hyperkagome")],
                         ids=["sc_2_31", "hk_2_31"])
def test_lattice_of_2_31_sites_refused(lat):
    assert lat.sites >= 1 << 31
    with pytest.raises(IsingError) as ex:
        engine(lat, HAM, [1.0])
    assert ex.value.code == -2


def test_empty_calls():
    with engine(Lattice(8, 8, 8), HAM, [1.0], seed=3) as e:
        e.init_random()
        s = e.get_spins(0)
        e.sweep(0)
        assert np.array_equal(e.get_spins(0), s)
        out = e.delta_energy(0, np.zeros(0, np.int32), np.zeros((0, 3)))
        assert out.shape == (0,)
