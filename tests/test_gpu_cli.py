"""The command line end to end on the GPU: CSV in, CPDAG/sepsets/pmax files
out, checked against the FP64 oracle (GPU box only)."""
import networkx as nx
import numpy as np
import pytest

import oracle
from gpucsl_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _skeleton_from_gml(path, p):
    g = nx.read_gml(path)
    sk = np.zeros((p, p), np.uint16)
    for a, b in g.edges:
        sk[int(a), int(b)] = sk[int(b), int(a)] = 1
    return sk


def _sepset_lines(sep):
    p = sep.shape[0]
    out = []
    for i in range(p):
        for k in range(p):
            s = sep[i, k]
            if s.size and s[0] != -1:
                out.append(" ".join(map(str, [i, k] + [int(x) for x in s[s != -1]])))
    return out


def test_cli_gaussian(tmp_path):
    from gpucsl_amd.cli.cli import gpucsl_cli

    d = syn.gaussian_sem(25, 3000, 3.0, seed=21).data
    np.savetxt(tmp_path / "gauss.csv", d, delimiter=",")
    gpucsl_cli(["-d", str(tmp_path / "gauss"), "-o", str(tmp_path / "out"), "--gaussian", "-l", "3", "-a", "0.01"])
    X = np.loadtxt(tmp_path / "gauss.csv", delimiter=",")
    ref = oracle.gauss_skeleton(oracle.corrcoef(X), 3000, 0.01, 3)
    sk = _skeleton_from_gml(tmp_path / "out" / "gauss.gml", 25)
    assert (sk == ref["skeleton"]).all()
    assert (tmp_path / "out" / "gauss_sepset.txt").read_text().splitlines() == _sepset_lines(ref["sepsets"])
    pm = np.loadtxt(tmp_path / "out" / "gauss_pmax.csv", delimiter=",")
    iu = np.triu_indices(25, 1)
    assert np.abs(pm[iu] - ref["pmax"][iu]).max() <= 6e-5      # %.4f rounding + FP32 z
    assert (tmp_path / "out" / "gauss_config.txt").exists()


def test_cli_discrete(tmp_path):
    from gpucsl_amd.cli.cli import gpucsl_cli

    b = syn.discrete_bn(15, 3000, 2.0, seed=22)
    np.savetxt(tmp_path / "disc.csv", b.data, delimiter=",", fmt="%d")
    gpucsl_cli(["-d", str(tmp_path / "disc.csv"), "-o", str(tmp_path / "out"), "--discrete", "-l", "2"])
    from gpucsl_amd.discrete import ordinal_encode

    codes, dims = ordinal_encode(b.data)       # the encoding the discrete skeleton applies
    ref = oracle.disc_skeleton(np.ascontiguousarray(codes.T), dims, 0.05, 2)
    sk = _skeleton_from_gml(tmp_path / "out" / "disc.gml", 15)
    assert (sk == ref["skeleton"]).all()
    assert (tmp_path / "out" / "disc_sepset.txt").read_text().splitlines() == _sepset_lines(ref["sepsets"])
