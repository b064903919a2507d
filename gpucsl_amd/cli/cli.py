"""The command line (gpucsl/cli/cli.py:19-85): read a CSV dataset, run
GaussianPC or DiscretePC on the GPU, write the CPDAG, separation sets, pmax
and the configuration."""
import logging

from gpucsl_amd.cli.cli_io import read_correlation_matrix, read_csv, write_results_and_config
from gpucsl_amd.cli.cli_util import error
from gpucsl_amd.cli.command_line_parser import build_command_line_parser
from gpucsl_amd.pc.pc import DiscretePC, GaussianPC


def set_log_level(level: int) -> None:
    if level == 1:
        logging.basicConfig(level=logging.INFO)
    if level > 1:
        logging.basicConfig(level=logging.DEBUG)


def gpucsl_cli(command_line_arguments):
    args = build_command_line_parser().parse_args(command_line_arguments)
    set_log_level(args.verbose)
    should_log = args.verbose > 0
    is_gaussian, is_discrete = args.gaussian, args.discrete
    gpus = args.gpus if args.gpus is not None else [0]

    if is_gaussian + is_discrete != 1:
        error("Please set exactly one of the options --gaussian/--discrete")
    if is_discrete and args.correlation_matrix is not None:
        error("Discrete independence test does not use a correlation matrix. Please check the arguments you are "
              "giving to gpucsl!")
    if is_discrete and len(gpus) > 1:
        error("Multi GPU independece test currently only supported for gaussian independece test")

    data = read_csv(args.dataset)
    pc_class = DiscretePC if is_discrete else GaussianPC
    pc = pc_class(data, args.max_level, args.alpha, is_debug=args.debug, should_log=should_log)
    if is_discrete:
        pc.set_distribution_specific_options()
    else:
        pc.set_distribution_specific_options(gpus, args.sync_device, read_correlation_matrix(args.correlation_matrix))
    result = pc.execute().result
    write_results_and_config(args.output_directory, args.dataset, command_line_arguments, result)
