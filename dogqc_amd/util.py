"""Plan-file loading helpers.

``loadScript`` keeps the reference's plan/cfg file format (dogqc/util.py:4-13):
blank lines and lines whose first non-blank character is ``#`` are dropped and
the remaining text is returned for ``eval`` with ``alg``, ``scal``,
``Reduction`` and ``Type`` in scope; an empty file yields ``"{}"``.
"""
import os


def loadScript(path):
    with open(path, "r") as f:
        lines = f.read().splitlines()
    kept = [ln for ln in lines if ln.strip() != "" and not ln.lstrip().startswith("#")]
    text = os.linesep.join(kept)
    return text if text else "{}"


def listWrap(elem):
    return elem if isinstance(elem, list) else [elem]


def intToHuman(number):
    for div, suffix in ((10 ** 9, "B"), (10 ** 6, "M"), (10 ** 3, "K")):
        if number >= div:
            return "%.1f %s" % (number / div, suffix)
    return str(int(number))
