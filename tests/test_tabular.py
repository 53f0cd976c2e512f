"""The drop-in CLI's stdout through the results pipeline's line contract
(d.make_tabular:27-44): a stdout captured from the CLI on an MI355X
(`fastscape_mi355x --preset=benchmark 1000 120 out.dem 123`, committed under
tests/golden/) yields the reference's columns -- the algorithm, the seven
stage timers, Overall and the total calculation time -- plus the metadata
columns (throughput, HBM roofline fraction, depth, step paths), with
consistent values.  The GPU test of the CLI applies the same check to a
fresh run."""
import os

import pytest

import tabular
from conftest import GOLDEN

FIXTURE = os.path.join(GOLDEN, "cli_stdout_benchmark_1000_120.txt")


def test_cli_stdout_fixture_parses_into_the_reference_columns():
    lines = open(FIXTURE).read().splitlines()
    datum = tabular.row(lines)
    tabular.check_cli_row(datum, 121)
    assert datum["Random_seed"] == "123" and datum["Preset"] == "benchmark"
    assert datum["hash"] == "git_hash"


def test_contract_rejects_unknown_tags():
    with pytest.raises(ValueError):
        tabular.row(["A FastScape", "q unknown = 1"])
    assert tabular.row(["A Fast Scape X", "t Step1: Initialize = 5 microseconds",
                        "p Step = 20", "m a b = 7"]) == \
        {"Algorithm": "Fast_Scape_X", "Step1_Initialize": "5", "a_b": "7"}
