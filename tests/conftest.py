import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def pytest_collection_modifyitems(config, items):
    # Print the test id before each GPU test so a run that dies names its case.
    pass


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def peg51():
    from oracle.emp_oracle import load_rv_files
    return load_rv_files([os.path.join(GOLDEN, "datafiles/51Peg/RV/51peg.vels")])


def make_sim(folder, k, parameterisation=1, moav=None, jitter=True, acceleration=0,
             conditions=(), semantics="reference", sinusoid=0, magnetic_cycle=0, switch_SA=False, use_c=False):
    """A Simulation with the blocks autorun(k, k) would add, conditions applied."""
    from astroemperor_amd import Simulation
    sim = Simulation()
    sim.switch_SA = switch_SA
    sim.read_loc = GOLDEN + "/"
    sim.load_data(folder)
    sim.set_engine("reddemcee")
    sim.keplerian_parameterisation = parameterisation
    sim.use_c = use_c
    sim.acceleration = acceleration
    sim.switch_jitter = jitter
    sim.semantics = semantics
    sim.sinusoid = sinusoid
    sim.magnetic_cycle = magnetic_cycle
    if moav:
        sim.moav = moav
    sim._autorun_add_RV_ins()
    for _ in range(k):
        sim.add_keplerian()
    for c in conditions:
        sim.add_condition(list(c))
    sim.apply_conditions()
    sim.update_model()
    return sim


def oracle_for(model):
    from oracle.emp_oracle import OracleModel
    d = model.data
    return OracleModel(model.spec(), d.X, d.Y, d.YERR, d.FLAG, SAI=d.SA)
