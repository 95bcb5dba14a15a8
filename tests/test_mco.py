"""The .mco reader (tissuemc_amd/mco.py) on the MCML output file the reference's own parser test
holds (tests/golden/mcml_fixture.mco), with the values that test expects
(tests/python/test_mcml_compare.py:51-57), plus its rejection of other files."""

from pathlib import Path

import pytest

from tissuemc_amd import mco

FIXTURE = Path(__file__).resolve().parent / "golden" / "mcml_fixture.mco"


def test_reads_results_and_echo():
    run = mco.read_mco(FIXTURE.read_text())
    assert (run["rsp"], run["rd"], run["absorbed"], run["tt"]) == (0.0278, 0.25, 0.6, 0.1222)
    e = run["echo"]
    assert e["photons"] == 250000
    assert (e["nz"], e["nr"], e["na"]) == (40, 50, 1)
    assert e["layers"][1] == [1.4, 2.0, 10.0, 0.0, 0.2]
    assert e["n_below"] == 1.0 and e["n_above"] == 1.0
    assert e["output_name"] == "run000.mco"
    assert list(run["A_l"]) == [0.3, 0.3]


def test_rejects_other_files():
    with pytest.raises(ValueError):
        mco.read_mco("this is not an MCML output file\n")
