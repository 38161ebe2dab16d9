"""Errors and logging (role of python/astra/log.py and CLogger)."""
import logging
import sys

_logger = logging.getLogger("astra")


class AstraError(Exception):
    """Error raised by the ASTRA API (python/astra/log.py:154-168)."""

    def __init__(self, message="", append_log=False):
        super().__init__(message)


def debug(message):
    _logger.debug(message)


def info(message):
    _logger.info(message)


def warn(message):
    _logger.warning(message)


def error(message):
    _logger.error(message)


def setOutputScreen(fd=1, level=None):
    h = logging.StreamHandler(sys.stdout if fd == 1 else sys.stderr)
    _logger.addHandler(h)
    if level is not None:
        setLevel(level)


def setLevel(level):
    _logger.setLevel(level)
