"""Input and output files of the command line, in the formats of
gpucsl/cli/cli_io.py:11-93: CSV datasets without header, `<stem>.gml`
(networkx), `<stem>_sepset.txt` (`i k s1 s2 ...` per separated ordered pair),
`<stem>_pmax.csv` (`%.4f`) and `<stem>_config.txt`."""
import os
from pathlib import Path

import networkx as nx
import numpy as np

from gpucsl_amd.cli.cli_util import error


def force_csv_suffix(path: Path) -> Path:
    if path.suffix != ".csv":
        path = path.parent / (path.stem + ".csv")
    return path


def get_file_name_stem(dataset_path: str) -> str:
    return Path(dataset_path).stem


def read_csv(path) -> np.ndarray:
    import pandas as pd

    try:
        path = force_csv_suffix(Path(path))
        return pd.read_csv(path.absolute(), header=None).to_numpy()
    except FileNotFoundError as err:
        error(str(err))


def read_correlation_matrix(path):
    """None means: compute the correlation matrix from the data."""
    if path is None:
        return None
    return read_csv(path)


def write_seperation_set(output_path: Path, file_name_stem: str, seperation_set) -> None:
    p = seperation_set.shape[0]
    with open(output_path / (file_name_stem + "_sepset.txt"), "w", newline="") as f:
        if seperation_set.size == 0:
            return  # an empty file still records that there are no separation sets
        for i in range(p):
            for k in range(p):
                s = seperation_set[i][k]
                if s[0] == -1:
                    continue
                items = s[s != -1]
                f.write(f"{i} {k} ")
                f.write(" ".join(map(str, items)))
                f.write("\n")


def write_directed_graph(output_path: Path, file_name_stem: str, directed_graph) -> None:
    nx.write_gml(directed_graph, output_path / (file_name_stem + ".gml"))


def write_pmax(output_path: Path, file_name_stem: str, pmax) -> None:
    np.savetxt(output_path / (file_name_stem + "_pmax.csv"), pmax, delimiter=",", fmt="%.4f")


def write_configuration(output_path: Path, file_name_stem: str, command_line_arguments) -> None:
    with open(output_path / (file_name_stem + "_config.txt"), "w", newline="") as f:
        f.write(f"The output for the dataset {file_name_stem} was generated using the following given "
                f"parameters: \n\n")
        f.write(" ".join(command_line_arguments))
        f.write("\n")


def write_results(output_path: Path, file_name_stem: str, result) -> None:
    if not os.path.isdir(output_path):
        os.makedirs(output_path)
    write_directed_graph(output_path, file_name_stem, result.directed_graph)
    write_seperation_set(output_path, file_name_stem, result.separation_sets)
    write_pmax(output_path, file_name_stem, result.pmax)
    print(f"successfully wrote results to: {output_path}")


def write_results_and_config(output_directory: str, dataset_path: str, command_line_arguments, result) -> None:
    output_path = Path(output_directory)
    stem = get_file_name_stem(dataset_path)
    write_results(output_path, stem, result)
    write_configuration(output_path, stem, command_line_arguments)
