import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def small_workload(orc):
    """tryptic-2k-100: 2,000 proteins, 100 spectra at 2+, P rule, 7-30 aa."""
    from hfx_amd import synth

    w = synth.WORKLOADS["tryptic-2k-100"]
    prot = synth.make_proteins(w["proteins"], seed=w["seed"])
    mass = orc.mass_table(w["static"])
    r, tc, tn = orc.digest(prot.seq, prot.seg, mass, rule=1, missed=w["missed"],
                           p_rule=w["p_rule"], min_len=w["min_len"], max_len=w["max_len"])
    sp = synth.make_spectra(prot, r, tc, tn, w["spectra"], seed=w["seed"] + 100,
                            charges=w["charges"], charge_p=w["charge_p"], static_mods=w["static"],
                            double_for_3plus=w["double3"])
    return dict(w=w, prot=prot, mass=mass, r=r, tc=tc, tn=tn, sp=sp)
