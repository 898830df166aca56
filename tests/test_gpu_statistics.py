"""Statistical agreement of GPU chains with independent FP64 oracle chains:
<E>, <|M|>, C_v and the middle layer's <Q> within 3 sigma.

The reference's own chains are not exact Metropolis (snapshot sweeps with
repeated picks, src/energyMin.f90:287-360) and its hot-path modules do not
build here, so the chains compared against are the oracle's exact
checkerboard Metropolis (oracle/ising_oracle.c) started from another seed,
which makes the two statistically independent. Error bars of means are
corrected by 2 tau_int (Sokal window, c = 5); C_v = beta^2 var(E) / N with a
jackknife error over 20 blocks. Seeds are fixed, so the outcome is
deterministic.

Workloads: the thin film of the skyrmion configuration at reduced size
(12x12x4, periodic in-plane, open z, J = 1, bulk D = 0.6, Bz = 0.3) at
T = 1.0 and 1.4, where the layer charge fluctuates (below T ~ 0.8 a chain
stays in one topological sector for longer than either run), a 6^3 bcc
lattice with anisotropy at T = 2.4, and a 4x4x2 hyper-kagome lattice
(three-colour sweeps) with the reference's interfacial DMI at T = 0.8 and
1.6."""
import numpy as np
import pytest

from gpu_util import olat
from ising_amd import Engine, Hamiltonian, Ladder, Lattice
from ising_amd.analysis import consistent, mean_error, specific_heat, specific_heat_error
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

FILM = (Lattice(12, 12, 4, (True, True, False)), Hamiltonian(J=1.0, D=0.6, dmi="bulk", Bz=0.3), (1.0, 1.4))
BCC = (Lattice(6, 6, 6, cell="bcc"), Hamiltonian(J=1.0, D=0.3, dmi="bulk", Bz=0.1, Kz=0.05), (2.4,))
HK = (Lattice(4, 4, 2, cell="hyperkagome"), Hamiltonian(J=1.0, D=0.4, dmi="interfacial_ref", Bz=0.2, Kz=0.05),
      (0.8, 1.6))
EVERY = 2


def gpu_series(lat, ham, temps, n_meas, seed=5):
    betas = [1.0 / t for t in temps]
    zmid = lat.lz // 2
    with Engine([Ladder(lat, ham, betas, 0.5)], seed=seed) as e:
        e.init_random()
        e.pt_run(n_sweeps=4000, swap_every=0, measure_every=0)
        o = e.pt_run(n_sweeps=EVERY * n_meas, swap_every=0, measure_every=EVERY, winding_every=EVERY,
                     winding_layer=zmid)
    out = []
    for t in range(len(temps)):
        out.append((o["energy"][:, t], np.linalg.norm(o["magnetisation"][:, t], axis=1),
                    o["winding"][:, t].astype(float)))
    return out


def oracle_series(lat, ham, temp, n_meas, seed, slot):
    lo, ho = olat(lat), orc.ham_from(ham)
    beta = 1.0 / temp
    s = orc.init_random(lo, seed, slot)
    orc.sweeps(lo, ho, s, beta, 0.5, seed, slot, 0, 2000)
    k = 2000
    E, M, Q = np.empty(n_meas), np.empty(n_meas), np.empty(n_meas)
    for i in range(n_meas):
        orc.sweeps(lo, ho, s, beta, 0.5, seed, slot, k, EVERY)
        k += EVERY
        E[i] = orc.total_energy(lo, ho, s)
        M[i] = np.linalg.norm(orc.magnetisation(lo, s))
        Q[i] = orc.winding_layer(lo, s, lat.lz // 2)[0]
    return E, M, Q


@pytest.mark.parametrize("case", ["film", "bcc", "hyperkagome"])
def test_observables_match_oracle_chains(case):
    lat, ham, temps = {"film": FILM, "bcc": BCC, "hyperkagome": HK}[case]
    orc.set_threads(1)
    gpu = gpu_series(lat, ham, temps, 40000)
    n = lat.sites
    for t, temp in enumerate(temps):
        Eg, Mg, Qg = gpu[t]
        Eo, Mo, Qo = oracle_series(lat, ham, temp, 10000, seed=99, slot=t)
        beta = 1.0 / temp
        report = []
        for name, a, b in (("E/N", Eg / n, Eo / n), ("|M|", Mg, Mo), ("Q", Qg, Qo)):
            ma, ea, _ = mean_error(a)
            mb, eb, _ = mean_error(b)
            report.append((name, ma, ea, mb, eb))
            assert consistent(ma, ea, mb, eb, 3.0), (case, temp, report)
        ca, cea = specific_heat(Eg, beta, n), specific_heat_error(Eg, beta, n)
        cb, ceb = specific_heat(Eo, beta, n), specific_heat_error(Eo, beta, n)
        assert consistent(ca, cea, cb, ceb, 3.0), (case, temp, "C_v", ca, cea, cb, ceb)
        # the comparison resolves the observables: error bars small against the values
        _, me, ee, _, _ = report[0]
        assert ee < 1e-2 * abs(me) and cea < 0.2 * ca
