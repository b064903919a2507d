"""``python -m gpucsl_amd``: one PC-stable run from a CSV file to result files.

The flags and the output files are gpucsl's command-line interface
(gpucsl/cli/cli.py:19-85, gpucsl/cli/command_line_parser.py:4-68); see
``cli_io`` for the file formats.  A run is parsed into a ``CliRun``, checked,
executed with ``GaussianPC`` or ``DiscretePC`` on one GPU, and written out.
"""
from __future__ import annotations

import logging
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

from gpucsl_amd.cli.cli_io import load_correlation, load_table, save_run
from gpucsl_amd.cli.cli_util import error
from gpucsl_amd.cli.command_line_parser import build_command_line_parser
from gpucsl_amd.pc.pc import DiscretePC, GaussianPC

_LOG_LEVELS = {0: None, 1: logging.INFO}


@dataclass(frozen=True)
class CliRun:
    kind: str                      # "gaussian" or "discrete"
    dataset: str
    output_directory: str
    max_level: Optional[int]
    alpha: float
    devices: Tuple[int, ...]
    sync_device: Optional[int]
    correlation_path: Optional[str]
    debug: bool
    verbosity: int

    @property
    def should_log(self) -> bool:
        return self.verbosity > 0


def parse_run(argv: Sequence[str]) -> CliRun:
    """Flags -> a checked run description (argparse exits on missing flags)."""
    ns = build_command_line_parser().parse_args(list(argv))
    chosen = [k for k, on in (("gaussian", ns.gaussian), ("discrete", ns.discrete)) if on]
    if len(chosen) != 1:
        error("choose the independence test: give exactly one of --gaussian and --discrete")
    run = CliRun(kind=chosen[0], dataset=ns.dataset, output_directory=ns.output_directory,
                 max_level=ns.max_level, alpha=ns.alpha, devices=tuple(ns.gpus) if ns.gpus else (0,),
                 sync_device=ns.sync_device, correlation_path=ns.correlation_matrix, debug=ns.debug,
                 verbosity=ns.verbose)
    if run.kind == "discrete":
        if run.correlation_path is not None:
            error("--correlation-matrix applies to the Gaussian test only; the discrete test reads the data")
        if len(run.devices) > 1:
            error("the discrete test runs on one GPU; pass a single index to --gpus")
    return run


def configure_logging(verbosity: int) -> None:
    level = _LOG_LEVELS.get(verbosity, logging.DEBUG)
    if level is not None:
        logging.basicConfig(level=level)


def execute(run: CliRun, data):
    """The PC result (CPDAG, separation sets, pmax) of one run."""
    if run.kind == "gaussian":
        pc = GaussianPC(data, run.max_level, run.alpha, is_debug=run.debug, should_log=run.should_log)
        pc.set_distribution_specific_options(list(run.devices), run.sync_device,
                                             load_correlation(run.correlation_path))
    else:
        pc = DiscretePC(data, run.max_level, run.alpha, is_debug=run.debug, should_log=run.should_log)
        pc.set_distribution_specific_options()
    return pc.execute().result


def gpucsl_cli(command_line_arguments: Sequence[str]) -> None:
    run = parse_run(command_line_arguments)
    configure_logging(run.verbosity)
    result = execute(run, load_table(run.dataset))
    save_run(run.output_directory, run.dataset, command_line_arguments, result)
