"""Errors and logging (python/astra/log.py; the reference's CLogger).

Messages go through Python's `logging` under the logger "astra", to the
screen (stdout or stderr) and optionally to a file, each with its own
level and format.  Levels and screen ids are the reference's constants;
format strings use its substitutions: %f source file, %n line, %m message,
%d date, %t time, %l level name, %% a percent sign.
"""
import inspect
import logging
import sys
import time

STDOUT = 1
STDERR = 2

DEBUG = 0
INFO = 1
WARN = 2
ERROR = 3

_LEVELS = {DEBUG: logging.DEBUG, INFO: logging.INFO, WARN: logging.WARNING, ERROR: logging.ERROR}
_NAMES = {logging.DEBUG: "DEBUG", logging.INFO: "INFO", logging.WARNING: "WARN", logging.ERROR: "ERROR"}
_DEFAULT_FORMAT = "%d %t %f(%n): %l: %m\n"

_logger = logging.getLogger("astra")
_logger.setLevel(logging.DEBUG)
_logger.propagate = False


class AstraError(Exception):
    """Error raised by the ASTRA API (python/astra/log.py:154-168).  With
    append_log=True the native library's last error message is appended."""

    def __init__(self, message="", append_log=False):
        if append_log:
            last = _last_native_error()
            if last:
                message += ": " + last
        super().__init__(message)


def _last_native_error():
    try:
        from . import _lib
        return _lib.load().astra_amd_last_error().decode(errors="replace")
    except Exception:
        return ""


class _AstraFormatter(logging.Formatter):
    def __init__(self, fmt):
        super().__init__()
        self.astra_fmt = fmt

    def format(self, record):
        t = time.localtime(record.created)
        sub = {"f": record.pathname, "n": str(record.lineno), "m": record.getMessage(),
               "d": time.strftime("%Y-%m-%d", t), "t": time.strftime("%H:%M:%S", t),
               "l": _NAMES.get(record.levelno, record.levelname), "%": "%"}
        out, i, fmt = [], 0, self.astra_fmt
        while i < len(fmt):
            c = fmt[i]
            if c == "%" and i + 1 < len(fmt) and fmt[i + 1] in sub:
                out.append(sub[fmt[i + 1]])
                i += 2
            else:
                out.append(c)
                i += 1
        return "".join(out).rstrip("\n")


class _Sink:
    """One output (screen or file): a handler, its level, format and switch."""

    def __init__(self, handler, level):
        self.handler = handler
        self.level = level
        self.enabled = True
        handler.setFormatter(_AstraFormatter(_DEFAULT_FORMAT))
        self._apply()

    def _apply(self):
        self.handler.setLevel(_LEVELS.get(self.level, logging.DEBUG) if self.enabled else logging.CRITICAL + 1)


class _StdStream:
    """The current sys.stdout / sys.stderr (looked up at each write)."""

    def __init__(self, fd):
        self.fd = fd

    def write(self, s):
        (sys.stdout if self.fd == STDOUT else sys.stderr).write(s)

    def flush(self):
        (sys.stdout if self.fd == STDOUT else sys.stderr).flush()


_screen = _Sink(logging.StreamHandler(_StdStream(STDERR)), ERROR)
_logger.addHandler(_screen.handler)
_file = None


def _log(level, message):
    _logger.log(_LEVELS[level], message, stacklevel=3)


def debug(message):
    """Log a debug message."""
    _log(DEBUG, message)


def info(message):
    """Log an informational message."""
    _log(INFO, message)


def warn(message):
    """Log a warning."""
    _log(WARN, message)


def error(message):
    """Log an error message."""
    _log(ERROR, message)


def _set_enabled(screen=None, file=None):
    if screen is not None:
        _screen.enabled = screen
        _screen._apply()
    if file is not None and _file is not None:
        _file.enabled = file
        _file._apply()


def enable():
    """Enable logging to screen and file."""
    _set_enabled(True, True)


def enableScreen():
    _set_enabled(screen=True)


def enableFile():
    """Enable logging to the file (one has to be set with setOutputFile)."""
    _set_enabled(file=True)


def disable():
    """Disable all logging."""
    _set_enabled(False, False)


def disableScreen():
    _set_enabled(screen=False)


def disableFile():
    _set_enabled(file=False)


def setFormatScreen(fmt):
    """Format of screen messages (the substitutions of this module's doc)."""
    _screen.handler.setFormatter(_AstraFormatter(fmt))


def setFormatFile(fmt):
    """Format of file messages."""
    if _file is not None:
        _file.handler.setFormatter(_AstraFormatter(fmt))


def setOutputScreen(fd, level):
    """Screen output to STDOUT or STDERR at `level` (DEBUG ... ERROR)."""
    _screen.handler.setStream(_StdStream(STDOUT if fd == STDOUT else STDERR))
    _screen.level = level
    _screen._apply()


def setOutputFile(filename, level):
    """Log to file `filename` at `level` (replacing an earlier file)."""
    global _file
    if _file is not None:
        _logger.removeHandler(_file.handler)
        _file.handler.close()
    _file = _Sink(logging.FileHandler(filename), level)
    _logger.addHandler(_file.handler)


def setLevel(level):
    """Screen level (DEBUG ... ERROR)."""
    _screen.level = level
    _screen._apply()
