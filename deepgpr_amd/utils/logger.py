"""Package logger under the "DeepGPR" hierarchy; never configures the root logger."""
from __future__ import annotations

import logging
from typing import Union

from ..config.defaults import LOGGER_NAME

DEFAULT_LOG_FORMAT = "%(asctime)s %(levelname)-8s %(name)s: %(message)s"


def get_logger(name: str = LOGGER_NAME) -> logging.Logger:
    if name != LOGGER_NAME and not name.startswith(LOGGER_NAME + "."):
        name = f"{LOGGER_NAME}.{name.rsplit('.', 1)[-1]}"
    logger = logging.getLogger(name)
    root = logging.getLogger(LOGGER_NAME)
    if not root.handlers:
        root.addHandler(logging.NullHandler())
    return logger


def configure_logging(level: Union[int, str] = logging.INFO, fmt: str = DEFAULT_LOG_FORMAT, *,
                      stream=None) -> logging.Logger:
    """Attach a console handler to the package logger and set its level
    (src/DeepGPR/utils/logger.py:36-64: same arguments). A second call
    replaces the handler the first installed instead of adding another."""
    logger = logging.getLogger(LOGGER_NAME)
    for old in [h for h in logger.handlers if getattr(h, "_deepgpr_amd_console", False)]:
        logger.removeHandler(old)
    handler = logging.StreamHandler(stream)
    handler.setFormatter(logging.Formatter(fmt))
    handler._deepgpr_amd_console = True  # type: ignore[attr-defined]
    logger.addHandler(handler)
    logger.setLevel(level)
    return logger
