"""Error and warning helpers (gpucsl/cli/cli_util.py:5-15)."""
import logging


class GpucslError(Exception):
    pass


def error(error_message: str):
    logging.error(error_message)
    raise GpucslError(error_message)


def warning(warning_message: str):
    logging.warning(warning_message)
