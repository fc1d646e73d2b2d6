"""The oracle against its committed fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py): a change in the oracle's links, q, flow steps or
multigrid shows up here, on the CPU. The GPU path is compared with the same fixtures
in test_gpu_golden.py."""
import os

import numpy as np
import pytest

from tests.golden import make_golden

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_links_and_q_reproduce_exactly(golden):
    for k, v in make_golden.make_links().items():
        np.testing.assert_array_equal(v, golden[k], err_msg=k)


def test_flow_steps_reproduce(golden):
    for k, v in make_golden.make_steps().items():
        np.testing.assert_allclose(v, golden[k], rtol=1e-6, atol=1e-9, err_msg=k)


def test_multigrid_reproduces(golden):
    got = make_golden.make_mg()
    np.testing.assert_array_equal(got["mg_r"], golden["mg_r"])
    np.testing.assert_allclose(got["mg_z32"], golden["mg_z32"], rtol=1e-6, atol=1e-6 * np.abs(golden["mg_z32"]).max())
    np.testing.assert_allclose(got["mg_z64"], golden["mg_z64"], rtol=1e-12, atol=1e-12 * np.abs(golden["mg_z64"]).max())
    np.testing.assert_allclose(got["mg_x64"], golden["mg_x64"], rtol=1e-10, atol=1e-10 * np.abs(golden["mg_x64"]).max())


def test_fixture_physics(golden):
    """What the fixtures must satisfy whatever wrote them: the first step conserves the
    unit-density start up to the open boundaries (the startup transient then lets the
    inlet pressurise the channel), the Zou/He inlet plane carries u = 0.05 on its fluid
    nodes and 0 on the walls, and the FP32 V-cycle stays close to the FP64 one."""
    n = 64 * 32 * 32
    assert abs(float(golden["c1_s1_mass"]) - n) / n < 1e-6
    masses = [float(golden[f"c1_s{s}_mass"]) for s in make_golden.STEPS]
    assert masses == sorted(masses)
    inlet = golden["c1_s100_u"][0]  # x = 0, z = 16, all y
    assert inlet[0] == 0.0 and inlet[-1] == 0.0
    np.testing.assert_allclose(inlet[1:-1], 0.05, atol=1e-7)
    z32, z64 = golden["mg_z32"], golden["mg_z64"]
    assert np.abs(z32 - z64).max() / np.abs(z64).max() < 1e-5
