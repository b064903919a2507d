"""Flags of ``python -m gpucsl_amd``.

Names, short forms, defaults and value types are gpucsl's
(gpucsl/cli/command_line_parser.py:4-68), so existing invocations work
unchanged; the help texts describe this build.
"""
import argparse

# (flags, keyword arguments of add_argument)
_FLAGS = [
    (("-a", "--alpha"), dict(type=float, default=0.05, help="significance level of every CI test (default 0.05)")),
    (("-b", "--debug"), dict(action="store_true", help="run the bounds-checked debug build of the kernels")),
    (("-c", "--correlation-matrix"), dict(help="CSV file with a precomputed correlation matrix (Gaussian only)")),
    (("-d", "--dataset"), dict(required=True, default="", help="CSV file of observations, one column per variable")),
    (("--discrete",), dict(action="count", default=0, help="Pearson chi-square test on categorical data")),
    (("--gaussian",), dict(action="count", default=0, help="Fisher-z partial-correlation test on continuous data")),
    (("-g", "--gpus"), dict(type=int, nargs="+", help="GPU index to run on (one GPU)")),
    (("-l", "--max-level"), dict(type=int, help="largest separation-set size to test (default: p - 2)")),
    (("-o", "--output-directory"), dict(required=True, help="folder for the .gml, _sepset, _pmax and _config files")),
    (("-s", "--sync-device"), dict(type=int, help="accepted for compatibility; one GPU needs no sync device")),
    (("-v", "--verbose"), dict(action="count", default=0, help="-v: info log with level timings, -vv: debug log")),
]


def build_command_line_parser() -> argparse.ArgumentParser:
    parser = argparse.ArgumentParser(prog="python -m gpucsl_amd",
                                     description="PC-stable causal structure learning on an AMD Instinct GPU.")
    for flags, kwargs in _FLAGS:
        parser.add_argument(*flags, **kwargs)
    return parser
