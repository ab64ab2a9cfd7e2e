"""Loads the committed fixtures under tests/golden/."""
import json
import os

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def tup(mode):
    """JSON lists -> the oracle's tuple representation (recursively)."""
    if isinstance(mode, list):
        return tuple(tup(m) for m in mode)
    return mode


def oracle_mode(m):
    """Fixture conv/padding modes to the oracle's representation."""
    if isinstance(m, str):
        return m
    tag = m[0]
    if tag == "const":
        return ("const", m[1])
    if tag == "custom" and len(m) == 3:
        return ("custom", list(m[1]), list(m[2]))
    if tag == "explicit" and len(m) == 3 and not isinstance(m[1][0][0], (str, list)):
        return ("explicit", [list(p) for p in m[1]], list(m[2]))
    if tag == "custom":
        return ("custom", [oracle_mode(b) for b in m[1]])
    if tag == "explicit":
        return ("explicit", [[oracle_mode(b[0]), oracle_mode(b[1])] for b in m[1]])
    raise ValueError(m)
