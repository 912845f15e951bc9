"""Dotted-key configuration lookup (sources/config_read_helper.py:5-17)."""

from __future__ import annotations

from typing import Dict

_RED = "\033[31m"
_RESET = "\033[0m"


def try_read_param(config: Dict, tokens: str, default_val, reached_tokens: str = ""):
    """Walk ``config`` along the dot-separated ``tokens``.

    A missing key prints a red notice naming the token and the path reached so
    far and returns ``default_val``; lookups never raise (same contract as
    sources/config_read_helper.py:5-17).
    """
    node = config
    for token in tokens.split("."):
        try:
            node = node[token]
        except (KeyError, TypeError, IndexError):
            print(_RED + f"Did not found token \"{token}\" under \"{reached_tokens}\"!\n"
                  + f"Falling back to the default value of {default_val}.\n" + _RESET)
            return default_val
        reached_tokens += ("." if len(reached_tokens) > 0 else "") + token
    return node
