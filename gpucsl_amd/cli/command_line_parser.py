"""Command-line flags of ``python -m gpucsl_amd``, the same as
gpucsl/cli/command_line_parser.py:4-68 (names, short forms, defaults, types)."""
import argparse


def build_command_line_parser():
    parser = argparse.ArgumentParser(description="PC-stable causal structure learning on an AMD Instinct GPU.")
    parser.add_argument("-a", "--alpha", help="Sets the alpha level", default=0.05, type=float)
    parser.add_argument("-b", "--debug", help="Accepted for compatibility; the kernels are prebuilt",
                        action="store_true")
    parser.add_argument("-c", "--correlation-matrix", help="Path to a cached correlation matrix")
    parser.add_argument("-d", "--dataset", help="The dataset to run", default="", required=True)
    parser.add_argument("--discrete", help="Uses a discrete conditional independence test", action="count",
                        default=0)
    parser.add_argument("--gaussian",
                        help="Uses a gaussian conditional independence test for data with a multivariate distribution",
                        action="count", default=0)
    parser.add_argument("-g", "--gpus", help="Indices of GPUs to use (one GPU in this build)", type=int, nargs="+")
    parser.add_argument("-l", "--max-level", help="Maximum size of seperation set used in pc algorithm", type=int)
    parser.add_argument("-o", "--output-directory", help="Output dir to store output files", required=True)
    parser.add_argument("-s", "--sync-device", help="Index of the GPU used to sync the others", type=int)
    parser.add_argument("-v", "--verbose",
                        help="Prints debug messages to stdout: -v logs info, -vv logs debug info",
                        action="count", default=0)
    return parser
