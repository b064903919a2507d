"""``python -m gpucsl_amd ...`` (gpucsl/__main__.py): GpucslError exits with
status 1 without a stack trace."""
import sys

from gpucsl_amd.cli.cli import gpucsl_cli
from gpucsl_amd.cli.cli_util import GpucslError

if __name__ == "__main__":
    try:
        gpucsl_cli(sys.argv[1:])
    except GpucslError:
        sys.exit(1)
