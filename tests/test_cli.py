"""Command line (CPU): flag handling and output formats, as the reference's
tests/test_cli.py:80-290 check them; the PC classes are replaced by stubs."""
import logging
from pathlib import Path

import networkx as nx
import numpy as np
import pytest

import gpucsl_amd.cli.cli as cli
from gpucsl_amd.cli.cli_io import ResultFiles, csv_path, dataset_stem, load_correlation, save_run, sepset_lines
from gpucsl_amd.cli.cli_util import GpucslError


class _Inner:
    def __init__(self):
        g = nx.DiGraph()
        g.add_edges_from([(0, 1), (2, 1)])
        self.directed_graph = g
        sep = np.full((3, 3, 2), -1, np.int32)
        sep[0, 2, 0] = sep[2, 0, 0] = 1
        self.separation_sets = sep
        self.pmax = np.array([[-1, 0.5, 0.012345], [-1, -1, 0.25], [-1, -1, -1]], np.float32)


class _Result:
    result = _Inner()


class _StubPC:
    calls = []

    def __init__(self, data, max_level, alpha, **kw):
        _StubPC.calls.append(dict(data=data, max_level=max_level, alpha=alpha, **kw))

    def set_distribution_specific_options(self, *a, **kw):
        _StubPC.options = (a, kw)
        return self

    def execute(self):
        return _Result()


@pytest.fixture
def stubbed(monkeypatch, tmp_path):
    _StubPC.calls = []
    monkeypatch.setattr(cli, "GaussianPC", _StubPC)
    monkeypatch.setattr(cli, "DiscretePC", _StubPC)
    monkeypatch.setattr(cli, "load_table", lambda path: np.zeros((10, 3)))
    written = {}
    monkeypatch.setattr(cli, "save_run", lambda *a: written.setdefault("args", a))
    return written


def test_missing_parameters_exit(stubbed):
    for argv in ([], ["-l", "3", "-o", "out", "--gaussian"], ["-d", "data", "-l", "3", "--gaussian"]):
        with pytest.raises(SystemExit):
            cli.gpucsl_cli(argv)


def test_exactly_one_distribution(stubbed):
    with pytest.raises(GpucslError):
        cli.gpucsl_cli(["-d", "data", "-o", "out", "-l", "3"])
    with pytest.raises(GpucslError):
        cli.gpucsl_cli(["-d", "data", "-o", "out", "-l", "3", "--gaussian", "--discrete"])


def test_discrete_rejects_correlation_matrix_and_several_gpus(stubbed):
    with pytest.raises(GpucslError):
        cli.gpucsl_cli(["-d", "data", "-o", "out", "-l", "3", "--discrete", "-c", "corr/path"])
    with pytest.raises(GpucslError):
        cli.gpucsl_cli(["-d", "data", "-o", "out", "--discrete", "-g", "0", "1"])


def test_alpha_and_max_level_reach_the_pc_class(stubbed):
    cli.gpucsl_cli(["-d", "data", "--gaussian", "-l", "5", "-o", "out", "-a", "0.6"])
    call = _StubPC.calls[-1]
    assert call["alpha"] == 0.6 and call["max_level"] == 5
    assert _StubPC.options[0][0] == [0]          # default device list: GPU 0
    assert stubbed["args"][0] == "out" and stubbed["args"][1] == "data"


@pytest.mark.parametrize("flags,level", [([], None), (["-v"], logging.INFO), (["-vv"], logging.DEBUG)])
def test_log_levels(stubbed, monkeypatch, flags, level):
    seen = {}
    monkeypatch.setattr(logging, "basicConfig", lambda level: seen.setdefault("level", level))
    cli.gpucsl_cli(["-d", "data", "--gaussian", "-l", "2", "-o", "out"] + flags)
    assert seen.get("level") == level


def test_path_helpers():
    assert csv_path(Path("foobar/foo.csv")) == Path("foobar/foo.csv")
    assert csv_path(Path("foobar/foo")) == Path("foobar/foo.csv")
    assert csv_path("foobar/foo.txt") == Path("foobar/foo.csv")
    assert dataset_stem("foobar/foo.csv") == "foo"
    assert load_correlation(None) is None


def test_missing_dataset_is_an_error(tmp_path):
    from gpucsl_amd.cli.cli_io import load_table

    with pytest.raises(GpucslError):
        load_table(tmp_path / "nothing_here")


def test_sepset_lines_row_major():
    sep = np.full((3, 3, 2), -1, np.int32)
    sep[2, 0, :2] = [1, 2]
    sep[0, 2, :2] = [1, 2]
    sep[1, 2, 0] = 0
    assert list(sepset_lines(sep)) == ["0 2 1 2", "1 2 0", "2 0 1 2"]
    assert list(sepset_lines(np.zeros((3, 3, 0), np.int32))) == []


def test_written_files(tmp_path):
    save_run(str(tmp_path / "out"), "some/dir/ds.csv", ["-d", "ds", "--gaussian"], _Inner())
    out = tmp_path / "out"
    assert (out / "ds_sepset.txt").read_text().splitlines() == ["0 2 1", "2 0 1"]
    rows = (out / "ds_pmax.csv").read_text().splitlines()
    assert rows[0] == "-1.0000,0.5000,0.0123"
    g = nx.read_gml(out / "ds.gml")
    assert sorted(g.edges) == [("0", "1"), ("2", "1")]
    cfg = (out / "ds_config.txt").read_text()
    assert cfg.startswith("The output for the dataset ds was generated") and "-d ds --gaussian" in cfg


def test_empty_sepsets_still_write_a_file(tmp_path):
    r = _Inner()
    r.separation_sets = np.zeros((3, 3, 0), np.int32)
    ResultFiles(tmp_path, "x").write(r, ["-d", "x"])
    assert (tmp_path / "x_sepset.txt").read_text() == ""
