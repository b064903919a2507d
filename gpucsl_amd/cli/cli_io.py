"""Files the command line reads and writes.

The file names and formats are the interface of gpucsl's command line
(gpucsl/cli/cli_io.py:39-101), so results can be diffed against it:

* input: a CSV table without a header row; a path without the ``.csv``
  suffix (or with another suffix) names ``<stem>.csv`` in the same folder;
* ``<stem>.gml``: the CPDAG, networkx GML;
* ``<stem>_sepset.txt``: one line ``i k s1 s2 ...`` per ordered pair (i, k)
  with a recorded separation set, pairs in row-major order (an empty file
  when the run had no levels above 0);
* ``<stem>_pmax.csv``: the pmax matrix, comma separated, ``%.4f``;
* ``<stem>_config.txt``: the command line the files came from.
"""
from __future__ import annotations

from pathlib import Path
from typing import Iterable, Iterator, Optional, Sequence

import networkx as nx
import numpy as np

from gpucsl_amd.cli.cli_util import error

CONFIG_HEADER = "The output for the dataset {stem} was generated using the following given parameters: \n\n"


def csv_path(path) -> Path:
    """The CSV file a dataset argument names: its suffix, if any, becomes .csv."""
    path = Path(path)
    return path if path.suffix == ".csv" else path.with_name(path.stem + ".csv")


def dataset_stem(path) -> str:
    """Output files are named after the dataset file without its suffix."""
    return Path(path).stem


def load_table(path) -> np.ndarray:
    """A header-less CSV table as an array (numeric or object columns)."""
    import pandas as pd

    target = csv_path(path).absolute()
    if not target.is_file():
        error(f"no such dataset file: {target}")
    return pd.read_csv(target, header=None).to_numpy()


def load_correlation(path: Optional[str]) -> Optional[np.ndarray]:
    """A cached correlation matrix, or None when the data's own is wanted."""
    return None if path is None else load_table(path)


def sepset_lines(separation_sets: np.ndarray) -> Iterator[str]:
    """``i k s1 s2 ...`` for every ordered pair whose first entry is set."""
    sep = np.asarray(separation_sets)
    if sep.ndim != 3 or sep.shape[2] == 0:
        return
    rows, cols = np.nonzero(sep[:, :, 0] != -1)
    for i, k in zip(rows.tolist(), cols.tolist()):
        members = sep[i, k]
        yield " ".join(str(v) for v in [i, k, *members[members != -1].tolist()])


class ResultFiles:
    """The four output files of one run, under one folder and stem."""

    def __init__(self, folder, stem: str):
        self.folder = Path(folder)
        self.stem = stem

    def path(self, suffix: str) -> Path:
        return self.folder / f"{self.stem}{suffix}"

    def write(self, result, argv: Sequence[str]) -> None:
        self.folder.mkdir(parents=True, exist_ok=True)
        nx.write_gml(result.directed_graph, self.path(".gml"))
        self._write_lines(self.path("_sepset.txt"), sepset_lines(result.separation_sets))
        np.savetxt(self.path("_pmax.csv"), np.asarray(result.pmax), delimiter=",", fmt="%.4f")
        with open(self.path("_config.txt"), "w", newline="") as f:
            f.write(CONFIG_HEADER.format(stem=self.stem))
            f.write(" ".join(argv) + "\n")

    @staticmethod
    def _write_lines(path: Path, lines: Iterable[str]) -> None:
        with open(path, "w", newline="") as f:
            for line in lines:
                f.write(line + "\n")


def save_run(out_dir, dataset_path, argv: Sequence[str], result) -> Path:
    """Writes the CPDAG, separation sets, pmax and configuration of a run."""
    files = ResultFiles(out_dir, dataset_stem(dataset_path))
    files.write(result, argv)
    print(f"successfully wrote results to: {files.folder}")
    return files.folder
