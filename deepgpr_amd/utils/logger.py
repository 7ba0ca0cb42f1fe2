"""Package logger under the "DeepGPR" hierarchy; never configures the root logger."""
from __future__ import annotations

import logging

from ..config.defaults import LOGGER_NAME


def get_logger(name: str = LOGGER_NAME) -> logging.Logger:
    if name != LOGGER_NAME and not name.startswith(LOGGER_NAME + "."):
        name = f"{LOGGER_NAME}.{name.rsplit('.', 1)[-1]}"
    logger = logging.getLogger(name)
    root = logging.getLogger(LOGGER_NAME)
    if not root.handlers:
        root.addHandler(logging.NullHandler())
    return logger
